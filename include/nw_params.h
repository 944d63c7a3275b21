/*
 * nw_params.h -- parameter blocks shared by the host driver, the HIP kernels
 * and the CPU oracle.
 *
 * SimParams / SolventParams / BoundaryType are layout-identical to the
 * reference's structs (simulation/particles_kernel.cuh:18-73) so that a host
 * compiled against the reference header can hand its structs to this library
 * unchanged: 144 and 60 bytes, 4-byte aligned, same field order.  The CUDA
 * vector types of the reference are replaced by plain C structs of identical
 * size and alignment (three 4-byte members).
 *
 * NWExtParams is new: it selects the physics variants that the reference keeps
 * only in its archived code (WCA-LJ excluded volume, harmonic chain bonds,
 * active drive along the bond tangent; archived/vs_solution/fluid_substrate/
 * NWWormsKernel_LennardJones.h:11-81, NWWormsKernel_DriveForce.h:10-40,
 * NWWormsKernel_InternalForce.h:37-59) and the SRD variant in which beads take
 * part in the collision.  All-zero NWExtParams reproduces the reference.
 */
#ifndef NW_PARAMS_H
#define NW_PARAMS_H

#ifdef __cplusplus
extern "C" {
typedef bool nw_bool;
#else
typedef _Bool nw_bool;
#endif

typedef unsigned int uint;

typedef struct nw_float3 { float x, y, z; } nw_float3;
typedef struct nw_uint3 { unsigned int x, y, z; } nw_uint3;

/* simulation/particles_kernel.cuh:18-23 */
typedef enum BoundaryType {
    PERIODIC = 0,
    WALL = 1,
    WALL_NO_SLIP = 2
} BoundaryType;

/* simulation/particles_kernel.cuh:26-60 (field order and meaning unchanged) */
typedef struct SimParams {
    nw_float3 gravity;
    float particleRadius;
    float particleMass;

    nw_uint3 gridSize;
    unsigned int numCells;
    nw_float3 origin;
    nw_float3 boxSize;
    nw_float3 cellSize;

    BoundaryType boundaryX;
    BoundaryType boundaryY;
    BoundaryType boundaryZ;

    unsigned int numParticles;
    unsigned int numFilaments;
    unsigned int filamentSize;

    float bondSpringK;
    float bondSpringL;
    float bondBendingK;

    float activity;
    float reverseProbability;
    float kbT;
    float gamma;

    float spring;
    float damping;
    float shear;
    float attraction;
    float boundaryDamping;
} SimParams;

/* simulation/particles_kernel.cuh:62-73 */
typedef struct SolventParams {
    nw_bool enabled;
    nw_uint3 gridSize;
    unsigned int numCells;
    nw_float3 randOffset;   /* grid shift added to positions before hashing */
    nw_float3 cellSize;
    unsigned int numParticles; /* number of solvent particles */
    float alpha;            /* collision rotation angle (radians) */
    float kbT;              /* solvent temperature (isokinetic thermostat) */
    float particleMass;
} SolventParams;

/* ---- extensions (no counterpart in the reference's active code) ---- */

/* excluded-volume pair potential evaluated in collideFilaments */
enum {
    NW_PAIR_REF_DEM = 0, /* reference collideCell: Hookean bond if chain neighbours,
                            else DEM spring+dashpot+shear, plus extensile pair force */
    NW_PAIR_WCA = 1      /* WCA-LJ between beads of different worms (+ extensile if activity != 0) */
};

/* what filamentForces computes */
enum {
    NW_BONDED_REF = 0,   /* reference filamentKernel: cosine bending + tangents */
    NW_BONDED_CHAIN = 1  /* + harmonic chain bonds + active drive along the bond tangent */
};

/* neighbour-cell stencil of the pair scan */
enum {
    NW_STENCIL_REF27 = 0, /* 27 offsets wrapped per axis exactly as the reference does
                             (a grid with gridSize.z == 1 visits its only layer 3 times) */
    NW_STENCIL_2D = 1     /* the 9 cells of the bead's own z layer */
};

/* SRD thermostat applied before the collision (simulation/particleSystem.cpp:377-384) */
enum {
    NW_SRD_THERMO_ISOKINETIC = 0, /* reference: per-cell lambda = sqrt(kT / sum(m v^2 / 2)) */
    NW_SRD_THERMO_NONE = 1
};

typedef struct NWExtParams {
    unsigned int pairPotential;      /* NW_PAIR_* */
    unsigned int bondedMode;         /* NW_BONDED_* */
    unsigned int stencil;            /* NW_STENCIL_* */
    unsigned int srdCollideFilaments;/* 0: only solvent particles rotate (reference); 1: beads too */
    unsigned int srdThermostat;      /* NW_SRD_THERMO_* */
    float wcaEpsilon;                /* LJ epsilon */
    float wcaSigma;                  /* LJ sigma */
    float wcaCutoff;                 /* cutoff radius; WCA uses 2^(1/6) sigma */
    float driveForce;                /* f_a along unit(r[p+1]-r[p]) for p < n-1 */
    unsigned int collideKernel;      /* fused pair pass: NW_COLLIDE_AUTO / _GLOBAL / _LDS */
    unsigned int reserved[6];
} NWExtParams;

/* pair-force kernel of the fused step (identical results; different data paths) */
enum {
    NW_COLLIDE_AUTO = 0,   /* LDS tiles when the stencil has one z layer, else global */
    NW_COLLIDE_GLOBAL = 1, /* every candidate read through the cache hierarchy */
    NW_COLLIDE_LDS = 2     /* neighbour rows staged in LDS per 256-bead tile */
};

#ifdef __cplusplus
}
#endif
#endif

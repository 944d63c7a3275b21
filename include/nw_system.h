/*
 * nw_system.h -- C entry points of libnwhost.so, the host driver above the
 * C-ABI.  libnwhost.so holds ParticleSystem (a restatement of
 * simulation/particleSystem.{h,cpp}: same constructor defaults, setters,
 * per-step call order and .xyzv writer) and drives libnwamd.so only through
 * include/nw_abi.h.  These functions expose it to C and Python callers.
 */
#ifndef NW_SYSTEM_H
#define NW_SYSTEM_H

#include <stdint.h>
#include "nw_params.h"
#include "nw_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Zero fields take the reference's values (simulation/particleSystem.cpp:64-148). */
typedef struct NWSystemConfig {
    uint32_t numFilaments;
    uint32_t filamentSize;
    nw_uint3 gridSize;       /* MD cell grid */
    float cellSize;          /* 0: 2 * particleRadius (reference) */
    nw_float3 boxSize;       /* 0: gridSize * cellSize (reference) */
    uint32_t srd;            /* SRD solvent on */
    nw_uint3 srdGridSize;    /* 0: (gx/2, gy/2, max(gz/2, 1)) */
    float srdCellSize;       /* 0: 2 * cellSize */
    float srdDensity;        /* particles per SRD cell; 0: 10 */
    uint32_t srdGap;         /* steps between collisions; 0: 5 */
    float srdAlpha;          /* rotation angle (radians); 0: 2.26893 (130 deg) */
    uint32_t srdShiftCentred;/* grid shift in [-a/2, a/2) instead of [0, a) */
    uint32_t langevinFdt;    /* Langevin noise sqrt(2 gamma kT / dt) instead of sqrt(2 gamma kT) */
    uint64_t seed;           /* Philox key of every draw (device and host) */
    NWExtParams ext;         /* physics variants; zero = reference */
    uint32_t perCallSequence;/* 1: issue the reference's per-call C-ABI sequence every step;
                                0: the fused entry points (same results, fewer launches) */
    uint32_t solventSortEvery;/* k > 0: keep the solvent in SRD-cell order on the device, re-sorted
                                at reset and every k collisions (nwSortSolvent); arrays read or
                                written through the system stay in original particle order.
                                0: original order (the reference's, bit for bit) */
    uint32_t reserved[6];
} NWSystemConfig;

/* initial conditions */
enum {
    NW_INIT_REFERENCE = 0,   /* lattice of straight worms, half flipped, solvent +-5 (particleSystem.cpp:583-678) */
    NW_INIT_RANDOM = 1,      /* straight worms at uniform heads and angles, rejection-sampled (no overlap < sigma) */
    NW_INIT_LATTICE_RANDOM = 2 /* dense: rows of straight worms with random polarity */
};

typedef struct NWSystem NWSystem;

NWSystem *nwSystemCreate(const NWSystemConfig *cfg);
void nwSystemDestroy(NWSystem *sys);
/* SimParams / SolventParams of the system (read or overwrite; applied at the next update) */
SimParams *nwSystemParams(NWSystem *sys);
SolventParams *nwSystemSolventParams(NWSystem *sys);
/* initial state; kT sets Maxwellian velocities for the random modes */
void nwSystemReset(NWSystem *sys, int mode, float kT);
/* `steps` calls of ParticleSystem::update(dt); no host synchronisation */
void nwSystemUpdate(NWSystem *sys, float dt, uint32_t steps);
void nwSystemSync(NWSystem *sys);
/* counts */
uint32_t nwSystemNumParticles(NWSystem *sys);
uint32_t nwSystemNumSolvent(NWSystem *sys);
uint32_t nwSystemNumCells(NWSystem *sys);     /* cells allocated for cellStart/cellEnd */
uint64_t nwSystemStep(NWSystem *sys);
/* host <-> device copies of whole arrays (float4 records), in original
   particle order: 0 pos (N+Ns), 1 vel (N+Ns), 2 tangent (N), 3 force (N),
   4 forceOld (N) */
void nwSystemGetArray(NWSystem *sys, int which, float *host);
void nwSystemSetArray(NWSystem *sys, int which, const float *host);
/* device pointers of the same arrays (with solventSortEvery > 0 the solvent
   part is in the device's cell order) */
float *nwSystemDeviceArray(NWSystem *sys, int which);
/* RNG state (seed, generator offset) at this point of the run */
void nwSystemRngState(NWSystem *sys, uint64_t *seed, uint64_t *offset);
void nwSystemSetRngState(NWSystem *sys, uint64_t offset);
/* ------------------------------------------------------------------------
   Replica sets: many independent 2D periodic worm systems of uneven size
   (WCA + chain bonds + bending + drive + Langevin), stepped together by the
   batched fused step (nwBatchMdStep) or one replica per launch sequence.
   ------------------------------------------------------------------------ */
typedef struct NWReplicaSpec {
    uint32_t numFilaments, filamentSize;
    float boxSize;      /* side of the square periodic box */
    float driveForce;   /* f_a along the bond tangent */
    float bendingK;     /* kappa */
    float activity;     /* extensile pair strength (0: none) */
    uint64_t seed;      /* initial conditions and Philox noise */
    uint32_t reserved[4];
} NWReplicaSpec;

typedef struct NWReplicaPhysics {
    float bondK, bondL;           /* harmonic chain bonds */
    float wcaEpsilon, wcaSigma;   /* WCA between beads of different worms */
    float gamma, kT;              /* Langevin */
    float cellMin;                /* smallest MD cell side (>= 2^(1/6) sigma) */
    int langevinFdt;              /* noise sqrt(2 gamma kT / dt) instead of sqrt(2 gamma kT) */
    uint32_t reserved[8];
} NWReplicaPhysics;

enum {
    NW_REPLICAS_GRAPH = 0,          /* batched step, captured once as a hipGraph and replayed */
    NW_REPLICAS_LAUNCHES = 1,       /* batched step, launched kernel by kernel */
    NW_REPLICAS_ONE_PER_LAUNCH = 2  /* every replica through its own nwFusedMdStep per step */
};

typedef struct NWReplicaSet NWReplicaSet;
NWReplicaSet *nwReplicaSetCreate(const NWReplicaSpec *specs, uint32_t numReplicas, const NWReplicaPhysics *phys);
void nwReplicaSetDestroy(NWReplicaSet *set);
void nwReplicaSetReset(NWReplicaSet *set, int mode, float kT);  /* NW_INIT_RANDOM / NW_INIT_LATTICE_RANDOM */
void nwReplicaSetUpdate(NWReplicaSet *set, float dt, uint32_t steps, int form);
void nwReplicaSetSync(NWReplicaSet *set);
uint32_t nwReplicaSetNumReplicas(NWReplicaSet *set);
/* beads of one replica, or of all (replica < 0) */
uint64_t nwReplicaSetNumParticles(NWReplicaSet *set, int replica);
/* float4 records of one replica (which: 0 pos, 1 vel, 2 tangent, 3 force, 4 forceOld) */
void nwReplicaSetGetArray(NWReplicaSet *set, int which, int replica, float *host);
void nwReplicaSetSetArray(NWReplicaSet *set, int which, int replica, const float *host);
/* the replica's SimParams / NWExtParams and generator state */
void nwReplicaSetParams(NWReplicaSet *set, int replica, SimParams *p, NWExtParams *e);
void nwReplicaSetRngState(NWReplicaSet *set, int replica, uint64_t *seed, uint64_t *offset);
void nwReplicaSetObservables(NWReplicaSet *set, int replica, NWObservables *out);

/* order parameters of the current worm configuration (nwObservables) */
void nwSystemObservables(NWSystem *sys, NWObservables *out);
/* appends one .xyzv frame (particleSystem.cpp:481-532) */
void nwSystemWriteOutputs(NWSystem *sys, const char *fileName, int iteration, float dt);
/* runs the library's autotuner (nwAutotuneFused) on a copy of the system's
   state with its parameters, for steps of dt: the launch shapes of the
   fused steps it picks stay set for every later update.  cachePath / useCache
   as for nwAutotuneFused.  Returns what nwAutotuneFused returns (0 when the
   system does not take the fused steps). */
int nwSystemAutotune(NWSystem *sys, float dt, const char *cachePath, int useCache, NWTuning *out);

#ifdef __cplusplus
}
#endif
#endif

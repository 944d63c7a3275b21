/*
 * nw_oracle.h -- CPU restatement of the reference's worm time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path of
 * libnwamd.so and the CPU baseline timed beside it; the product never links,
 * loads or calls it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may use it.
 *
 * Every function restates one reference kernel (simulation/particles_kernel_impl.cuh)
 * or wrapper (simulation/particleSystem_cuda.cu) and cites it.  Floating-point
 * stages come in two precisions: _f64 (the oracle: float inputs, double
 * arithmetic and outputs) and _f32 (same algorithm in float; the CPU baseline).
 * Stages that produce integers (cell coordinates and hashes) use float
 * arithmetic in both, exactly as the reference, because their results must be
 * bit-identical.  Arrays are float4 records (4 floats per particle).
 *
 * How each function is pinned (closed forms, invariants, brute force, finite
 * differences, library routines, Philox known-answer vectors) is listed in
 * tests/test_oracle_*.py.
 */
#ifndef NW_ORACLE_H
#define NW_ORACLE_H

#include <stdint.h>
#include "nw_params.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- integer / bit-exact stages ---------------- */

/* calcHashD (particles_kernel_impl.cuh:312-342) with calcGridPos :231-238,
   calcGridHash :252-258, positionRelativeToBox :208-228, calcGridPosSRD
   :241-249, calcGridHashSRD :261-276.  Wrapping by mathematical modulo
   (== the reference's `& (G-1)` for power-of-two G).  Non-periodic SRD axes
   are clamped into the grid. */
void orc_calc_hash(const SimParams *p, const SolventParams *s, const float *pos, uint32_t n, int srd,
                   uint32_t *hash, uint32_t *index);

/* thrust::sort_by_key (particleSystem_cuda.cu:277-282): stable ascending
   (key -> value) sort, by bottom-up merge sort. */
void orc_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t n);

/* reorderDataAndFindCellStart (particleSystem_cuda.cu:171-205, kernel
   particles_kernel_impl.cuh:346-424).  vel/sortedVel, tan/sortedTan may be NULL. */
void orc_reorder(uint32_t *cellStart, uint32_t *cellEnd, float *sortedPos, float *sortedVel, float *sortedTan,
                 const uint32_t *hash, const uint32_t *index, const float *pos, const float *vel, const float *tan,
                 uint32_t n, uint32_t numCells);

/* reverseFilamentsKernel (particles_kernel_impl.cuh:820-857); uniform[w] is
   the worm's draw.  In place. */
void orc_reverse_filaments(const SimParams *p, const float *uniform, uint32_t W, float *force, float *forceOld,
                           float *vel, float *pos);

/* Philox4x32-10 (Salmon et al., SC'11) and the generator layout of the
   library: element e of a draw at `offset` is lane e%4 of block offset + e/4,
   block b = philox(counter=(lo32(b), hi32(b), 0, 0), key=(lo32(seed), hi32(seed))). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_random_uniform(uint64_t seed, uint64_t offset, uint32_t n, float *out);
/* normals: Box-Muller on lanes (0,1) and (2,3) of each block, evaluated in double */
void orc_random_gaussian_f64(uint64_t seed, uint64_t offset, uint32_t n, double mean, double stddev, double *out);

/* SRD grid shift of collision k (two uniforms in [0, 1) from the key seed ^
   0x5eed5eed5eed5eed, block k), times the SRD cell size; centred = 1 shifts
   into [-a/2, a/2).  Stands in for the reference's host rand() draw
   (simulation/particleSystem.cpp:345-346). */
void orc_srd_shift(uint64_t seed, uint64_t k, int centred, const float cell[3], float out[3]);

/* ---------------- floating-point stages (two precisions) ---------------- */
#define NW_ORACLE_DECLS(R, SFX)                                                                              \
    /* filament_integrator (particles_kernel_impl.cuh:71-142) */                                             \
    void orc_integrate_filaments_##SFX(const SimParams *p, const float *pos, const float *vel, const float *f,  \
                                        const float *fold, float dt, uint32_t n, R *posOut, R *velOut, R *fOut,  \
                                        R *foldOut);                                                         \
    /* solvent_integrator (particles_kernel_impl.cuh:145-205), particles [start, start+ns) */               \
    void orc_integrate_solvent_##SFX(const SimParams *p, const float *pos, const float *vel, float dt,        \
                                      uint32_t start, uint32_t ns, R *posOut, R *velOut);                    \
    /* collideFilamentsKernel + collideCell (particles_kernel_impl.cuh:427-556, 694-744); WCA mode from      \
       archived/vs_solution/fluid_substrate/NWWormsKernel_LennardJones.h:11-81.  force is the unsorted      \
       force array before the call; forceOut receives force + pair forces. */                                \
    void orc_collide_filaments_##SFX(const SimParams *p, const NWExtParams *e, const float *sPos,            \
                                      const float *sVel, const float *sTan, const uint32_t *sIdx,             \
                                      const uint32_t *cellStart, const uint32_t *cellEnd, uint32_t n,          \
                                      const float *force, R *forceOut);                                      \
    /* filamentKernel (particles_kernel_impl.cuh:751-818); chain bonds / drive from archived               \
       NWWormsKernel_InternalForce.h:37-59 and NWWormsKernel_DriveForce.h:10-40.  tangentOut is only        \
       written where the reference writes it. */                                                            \
    void orc_filament_forces_##SFX(const SimParams *p, const NWExtParams *e, const float *force,             \
                                    const float *pos, uint32_t W, R *forceOut, R *tangentOut);               \
    /* langevinKernel (particles_kernel_impl.cuh:859-874); noise = 4 values per bead */                      \
    void orc_langevin_##SFX(const SimParams *p, const float *force, const float *vel, const R *noise,          \
                             uint32_t n, R *forceOut);                                                        \
    /* cellCenterOfMomentumSRD (particles_kernel_impl.cuh:559-594) */                                        \
    void orc_srd_com_##SFX(const float *sVel, const uint32_t *cellStart, const uint32_t *cellEnd,            \
                            uint32_t numCells, R *comOut);                                                    \
    /* isokineticThermostatSRD (particles_kernel_impl.cuh:597-637); vel/sVel updated into the outputs */     \
    void orc_srd_isokinetic_##SFX(const SolventParams *s, const float *vel, const float *sVel,               \
                                   const uint32_t *cellStart, const uint32_t *cellEnd, const uint32_t *index,  \
                                   uint32_t numCells, uint32_t nTotal, R *velOut, R *sVelOut);               \
    /* collideSolventKernel + rotate2D (particles_kernel_impl.cuh:640-691); uniform[c] is cell c's draw */   \
    void orc_collide_solvent_##SFX(const SimParams *p, const SolventParams *s, const NWExtParams *e,         \
                                    const float *pos, const float *vel, const float *com,                    \
                                    const float *uniform, uint32_t numCells, uint32_t n, R *velOut);

NW_ORACLE_DECLS(double, f64)
NW_ORACLE_DECLS(float, f32)

/* Whole reference step (simulation/particleSystem.cpp:249-339 and :341-413)
   in FP32 with OpenMP, for the CPU baseline.  See nw_oracle.c for the buffer
   contract.  Returns the number of OpenMP threads used. */
typedef struct OrcSystem {
    SimParams p;
    SolventParams s;
    NWExtParams e;
    uint32_t numSolvent, srdGap, srdEnabled, langevinFdt, srdShiftCentred;
    float dt;
    uint64_t seed, offset;
    uint64_t step;
    float *pos, *vel, *force, *forceOld, *tangent;  /* (N + Ns) x 4 for pos/vel, N x 4 for the rest */
    float *sPos, *sVel, *sTan, *com, *rnd;          /* scratch */
    uint32_t *hash, *index, *cellStart, *cellEnd;
    uint32_t numCellsAlloc;
} OrcSystem;
int orc_system_step_f32(OrcSystem *sys, uint32_t steps);

#ifdef __cplusplus
}
#endif
#endif

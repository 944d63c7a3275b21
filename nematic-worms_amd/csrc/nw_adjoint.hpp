// nw_adjoint.hpp -- kernels of the adjoint of the fused MD step (nw_adjoint.hip)
#pragma once

#include "nw_bucket.hpp"
#include "nw_internal.hpp"

namespace nw {

// parameters of the gradient, in this order
enum : int { kAdjDrive, kAdjKappa, kAdjEpsilon, kAdjGamma, kAdjParams };

struct AdjStepArgs {
    const float4 *pos, *vel;        // r', v' of the step (bead order)
    const float4 *sPos;             // r' in cell order
    const uint32_t *ids, *off;      // the step's cell lists
    const float4 *lF;               // adjoint of the step's force
    float4 *lFout;                  // adjoint of the previous force (the other buffer)
    float4 *lr, *lv, *lG;           // updated in place
    double *partial;                // out: one row of kAdjParams per workgroup
    uint32_t W;
    int langevin;
    float lgvStd, gamma, dt;
    uint64_t seed, offLgv;          // the step's Langevin draws
};

uint32_t adj_step_blocks(uint32_t W, uint32_t n);
hipError_t launch_adj_step(const AdjStepArgs &a, const KP &kp, hipStream_t st);
// d S / d r of the bond-tangent nematic order: e the unit eigenvector of the
// largest eigenvalue of <1.5 b b - 0.5 I>, scale = 3 / bonds
hipError_t launch_bond_order_grad(const float4 *pos, uint32_t W, const double e[3], double scale, float4 *out,
                                  const KP &kp, hipStream_t st);
// adjoint of one SRD collision (cells, signs fixed) on the velocity adjoints
// lv of all NT particles, beads first: ids / off the collision's cell order
// and CSR offsets, hash the cell of every particle, com its (u, sign) per
// cell, mass.w the particle masses; cellSum: scratch of `cells` float4
hipError_t launch_srd_adj(const uint32_t *ids, const uint32_t *off, const uint32_t *hash, const float4 *com,
                          const float4 *mass, float4 *lv, float4 *cellSum, uint32_t N, uint32_t NT, uint32_t cells,
                          float cosA, float sinA, int rotBeads, hipStream_t st);
// scale * d(sum of v . t)/d(r, v) of the polar speed (dVel may be NULL)
hipError_t launch_polar_speed_grad(const float4 *pos, const float4 *vel, uint32_t W, double scale, float4 *dPos,
                                   float4 *dVel, const KP &kp, hipStream_t st);
// one adjoint step of every replica of a batch (a: concatenated arrays,
// off over the concatenated cells, ids global; the replicas' own W,
// Langevin and Philox from reps; a.partial: one row per front wavefront,
// added to)
hipError_t launch_adj_step_batch(const AdjStepArgs &a, const BatchRep *reps, uint32_t R, uint32_t waves,
                                 uint64_t step, hipStream_t st);
// out[r][k] = the sum of replica r's wavefront rows
hipError_t launch_adj_total_batch(const double *partial, const BatchRep *reps, uint32_t R, uint32_t waves,
                                  double *out, hipStream_t st);
// out[k] += the column sums of partial (rows x kAdjParams), in a fixed order
hipError_t launch_adj_total(const double *partial, uint64_t rows, double *out, hipStream_t st);

}  // namespace nw

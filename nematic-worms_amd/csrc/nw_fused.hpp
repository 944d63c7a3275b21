// nw_fused.hpp -- launchers of the fused step kernels (nw_fused.hip).
#pragma once

#include "nw_internal.hpp"

namespace nw {

// Library-owned scratch of the fused path.  `counts` (one bin per cell + an
// overflow bin) is zero between calls: every consumer clears the bins it read.
struct FusedScratch {
    uint32_t *counts = nullptr;   // cap_bins
    uint32_t *offsets = nullptr;  // cap_bins + 1 (CSR cell offsets)
    uint32_t *scan = nullptr;     // scan partials
    uint32_t *slot = nullptr;     // cap_n: histogram slot of each particle
    uint32_t *hashTmp = nullptr;  // cap_n
    uint64_t cap_n = 0, cap_bins = 0;
    bool dirty = true;
};

struct MdFrontArgs {
    float4 *pos, *vel, *force, *forceOld, *tangent;
    uint32_t *hash;
    float dt, lgvStd;
    int langevin;
    uint32_t W;
    uint64_t seed, offRev, offLgv;
};

hipError_t ensure_fused_scratch(FusedScratch &f, uint64_t n, uint64_t bins, hipStream_t st);
hipError_t launch_md_front(const MdFrontArgs &a, const KP &kp, FusedScratch &f, hipStream_t st);
hipError_t launch_scan_bins(FusedScratch &f, uint64_t bins, hipStream_t st);
hipError_t launch_scatter_ix(const uint32_t *hash, FusedScratch &f, uint32_t n, uint32_t *out, hipStream_t st);
hipError_t launch_md_cells(FusedScratch &f, uint32_t bins, uint32_t *ids, const float4 *pos, const float4 *vel,
                           const float4 *tan, float4 *sPos, float4 *sVel, float4 *sTan, hipStream_t st);
hipError_t launch_collide_csr(float4 *force, const float4 *sPos, const float4 *sVel, const float4 *sTan,
                              const uint32_t *sIdx, const FusedScratch &f, uint32_t n, const KP &kp,
                              hipStream_t st);
hipError_t launch_srd_front(const float4 *pos, uint32_t nt, uint32_t *hash, FusedScratch &f, const KP &kp,
                            hipStream_t st);
hipError_t launch_srd_cells(FusedScratch &f, uint32_t bins, uint32_t *ids, float4 *vel, float4 *com, int iso,
                            uint64_t seed, uint64_t offset, const KP &kp, hipStream_t st);
hipError_t launch_srd_rotate(float4 *pos, float4 *vel, const float4 *com, uint32_t N, uint32_t nt, float dt,
                             float cosA, float sinA, const KP &kp, hipStream_t st);

}  // namespace nw

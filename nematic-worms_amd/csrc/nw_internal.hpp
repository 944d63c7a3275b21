// nw_internal.hpp -- launchers shared between the C-ABI layer and the kernel
// translation units.  Every launcher enqueues on `st` and returns the launch
// status; none synchronises.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "nw_device.hpp"

namespace nw {

constexpr int kBlock = 256;

// Events that the next NW_LAUNCH on this thread attaches to its own dispatch
// packet (hipExtLaunchKernelGGL): the elapsed time between them is the
// kernel's execution alone.  A timing scope around a launcher that makes a
// single launch sets them; `used` tells it that the launch took them.
struct LaunchTiming {
    hipEvent_t start = nullptr, stop = nullptr;
    bool used = false;
};
LaunchTiming &launch_timing();

#define NW_LAUNCH(K, G, B, SH, ST, ...)                                                                  \
    do {                                                                                                 \
        ::nw::LaunchTiming &lt_ = ::nw::launch_timing();                                                 \
        if (lt_.start) {                                                                                 \
            hipExtLaunchKernelGGL(K, G, B, SH, ST, lt_.start, lt_.stop, 0u, __VA_ARGS__);                 \
            lt_.start = lt_.stop = nullptr;                                                              \
            lt_.used = true;                                                                             \
        } else {                                                                                         \
            hipLaunchKernelGGL(K, G, B, SH, ST, __VA_ARGS__);                                            \
        }                                                                                                \
    } while (0)

inline unsigned grid_for(uint64_t n, unsigned block = kBlock) {
    return unsigned((n + block - 1) / block);
}

// blocks i = x + 8k of a launch are spread round-robin over the 8 XCDs; map
// them so every XCD takes one contiguous run of work items (a bijection)
__device__ __forceinline__ uint32_t xcd_contiguous(uint32_t i, uint32_t G) {
    const uint32_t q = G >> 3, r = G & 7u;
    const uint32_t x = i & 7u, k = i >> 3;
    return x < r ? x * (q + 1u) + k : r * (q + 1u) + (x - r) * q + k;
}

// device anomaly flags (bit 0: sort key >= bound, bit 1: hash >= numCells in reorder)
uint32_t *device_flags();

// the parameters of the last setParameters / setSolventParameters /
// nwSetExtParameters, as the kernels receive them
KP current_kp();

// ---- nw_kernels.hip ----
hipError_t launch_integrate_filaments(float4 *pos, float4 *vel, float4 *f, float4 *fold, float dt,
                                      uint32_t n, const KP &kp, hipStream_t st);
hipError_t launch_integrate_solvent(float4 *pos, float4 *vel, float dt, uint32_t start, uint32_t ns,
                                    const KP &kp, hipStream_t st);
hipError_t launch_calc_hash(uint32_t *hash, uint32_t *index, const float4 *pos, uint32_t n, bool srd,
                            const KP &kp, hipStream_t st);
hipError_t launch_fill_u32(uint32_t *p, uint32_t value, uint64_t n, hipStream_t st);
hipError_t launch_reorder(uint32_t *cellStart, uint32_t *cellEnd, float4 *sPos, float4 *sVel, float4 *sTan,
                          const uint32_t *hash, const uint32_t *index, const float4 *pos, const float4 *vel,
                          const float4 *tan, uint32_t n, uint32_t numCells, hipStream_t st);
hipError_t launch_collide_filaments(float4 *force, const float4 *sPos, const float4 *sVel, const float4 *sTan,
                                    const uint32_t *sIdx, const uint32_t *cellStart, const uint32_t *cellEnd,
                                    uint32_t n, uint32_t numCells, const KP &kp, hipStream_t st);
hipError_t launch_filament_forces(float4 *force, float4 *tangent, const float4 *pos, uint32_t numFilaments,
                                  const KP &kp, hipStream_t st);
hipError_t launch_langevin(float4 *force, const float4 *vel, float stddev, uint32_t n, uint64_t seed,
                           uint64_t offset, const KP &kp, hipStream_t st);
hipError_t launch_randomize(float *out, uint32_t n, int gaussian, float mean, float stddev, uint64_t seed,
                            uint64_t offset, hipStream_t st);
hipError_t launch_srd_com(float4 *com, const float4 *sVel, const uint32_t *cellStart, const uint32_t *cellEnd,
                          uint32_t numCells, hipStream_t st);
hipError_t launch_srd_isokinetic(float4 *vel, float4 *sVel, const uint32_t *cellStart, const uint32_t *cellEnd,
                                 const uint32_t *index, uint32_t numCells, const KP &kp, hipStream_t st);
hipError_t launch_collide_solvent(const float4 *pos, float4 *vel, const float4 *com, uint32_t numCells,
                                  uint32_t n, float cosA, float sinA, uint64_t seed, uint64_t offset,
                                  const KP &kp, hipStream_t st);
hipError_t launch_reverse_filaments(float4 *force, float4 *forceOld, float4 *vel, float4 *pos,
                                    uint32_t numFilaments, uint64_t seed, uint64_t offset, const KP &kp,
                                    hipStream_t st);

// ---- nw_sort.hip ----
// Scratch sized for n keys / `bins` bins is grown by the caller (ensure_sort_scratch).
struct SortScratch {
    uint32_t *counts = nullptr;  // bins + 1
    uint32_t *offsets = nullptr; // bins + 2
    uint32_t *slot = nullptr;    // n
    uint2 *pairs = nullptr;      // n
    uint32_t *keys = nullptr;    // n (overflow keys / radix ping-pong)
    uint32_t *vals = nullptr;    // n (radix ping-pong)
    uint32_t *scan_tmp = nullptr;// scan partials
    uint64_t cap_n = 0, cap_bins = 0, cap_scan = 0;
};
hipError_t ensure_sort_scratch(SortScratch &s, uint64_t n, uint64_t bins);
hipError_t counting_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t n, uint32_t bound, SortScratch &s,
                               hipStream_t st);
hipError_t radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t n, uint32_t keyBits, SortScratch &s,
                            hipStream_t st);
// exclusive scan of in[0..m) into out[0..m]; out[m] = total.  tmp from scan_scratch_words(m).
uint64_t scan_scratch_words(uint64_t m);
hipError_t exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint64_t m, uint32_t *tmp, hipStream_t st);

}  // namespace nw

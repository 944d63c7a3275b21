// nw_observe.hip -- order parameters of the worm configuration on the device
// (gfx950): the sums behind the nematic order S and polar order P of the
// reference's analysis (analysis/analyze_correlations.py:13-38 with the
// tangents of analysis/afa/utils.py:35-49), the bond-tangent Q tensor and the
// mean speed along the tangent.
//
// Whole worms per wavefront (as in the step kernels): bead p of a worm reads
// its chain neighbours by cross-lane shuffles.  Every lane accumulates its
// terms in FP64; wave shuffles then an LDS pass reduce a workgroup to one
// row of partial sums, and a single workgroup adds the rows in a fixed order,
// so the result does not depend on scheduling.
#include "nw_internal.hpp"
#include "nw_observe.hpp"
#include "nw_physics.hpp"

namespace nw {

constexpr int kObsBlock = 256;

__device__ __forceinline__ double wave_sum(double v) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

__global__ __launch_bounds__(kObsBlock) void k_observe(const float4 *__restrict__ pos, const float4 *__restrict__ vel,
                                                       uint32_t W, double *__restrict__ partial, KP kp) {
    __shared__ double lsum[kObsBlock / 64][kObsSums];
    const SimParams &p = kp.p;
    const uint32_t n = p.filamentSize;
    const WaveWorm m = wave_worm(n, W);
    const float4 P = m.active ? pos[m.bead] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 V = m.active ? vel[m.bead] : make_float4(0.f, 0.f, 0.f, 0.f);
    const f3 r = mk(P);
    const f3 rn = shfl_down3(r, 1), rp = shfl_up3(r, 1);

    double acc[kObsSums];
#pragma unroll
    for (int k = 0; k < kObsSums; ++k) acc[k] = 0.0;
    if (m.active && n >= 2) {
        // bead tangent: unit(r[p+1] - r[p-1]), one-sided at the ends
        const f3 a = m.pp == 0 ? r : rp;
        const f3 b = m.pp + 1 == n ? r : rn;
        const f3 d = min_image(p, a, b);
        const double dx = d.x, dy = d.y, dz = d.z;
        const double inv = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
        const double tx = dx * inv, ty = dy * inv, tz = dz * inv;
        acc[kObsTxx] = tx * tx;
        acc[kObsTxy] = tx * ty;
        acc[kObsTxz] = tx * tz;
        acc[kObsTyy] = ty * ty;
        acc[kObsTyz] = ty * tz;
        acc[kObsTzz] = tz * tz;
        acc[kObsTx] = tx;
        acc[kObsTy] = ty;
        acc[kObsTz] = tz;
        acc[kObsBeads] = 1.0;
        acc[kObsVdotT] = double(V.x) * tx + double(V.y) * ty + double(V.z) * tz;
        acc[kObsV2] = double(V.x) * V.x + double(V.y) * V.y + double(V.z) * V.z;
        // bond tangent of the bond (p, p+1)
        if (m.pp + 1 < n) {
            const f3 e = min_image(p, r, rn);
            const double ex = e.x, ey = e.y, ez = e.z;
            const double ie = 1.0 / sqrt(ex * ex + ey * ey + ez * ez);
            const double bx = ex * ie, by = ey * ie, bz = ez * ie;
            acc[kObsBxx] = bx * bx;
            acc[kObsBxy] = bx * by;
            acc[kObsBxz] = bx * bz;
            acc[kObsByy] = by * by;
            acc[kObsByz] = by * bz;
            acc[kObsBzz] = bz * bz;
            acc[kObsBonds] = 1.0;
        }
    }
    const uint32_t wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kObsSums; ++k) {
        const double s = wave_sum(acc[k]);
        if ((threadIdx.x & 63u) == 0u) lsum[wid][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < kObsSums) {
        double s = 0.0;
        for (int w = 0; w < kObsBlock / 64; ++w) s += lsum[w][threadIdx.x];
        partial[uint64_t(blockIdx.x) * kObsSums + threadIdx.x] = s;
    }
}

// out[k] = sum over rows of partial[row][k], rows in order
// thread (g, k) adds rows g, g + kObsGroups, ... of sum k in order; thread k
// then adds the groups' results in group order (the same order every call)
constexpr uint32_t kObsGroups = 32;

__global__ __launch_bounds__(kObsGroups * kObsSums) void k_observe_total(const double *__restrict__ partial,
                                                                         uint32_t rows, double *__restrict__ out) {
    __shared__ double red[kObsGroups * kObsSums];
    const uint32_t g = threadIdx.x / kObsSums, k = threadIdx.x - g * kObsSums;
    double s = 0.0;
    for (uint32_t r = g; r < rows; r += kObsGroups) s += partial[uint64_t(r) * kObsSums + k];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < uint32_t(kObsSums)) {
        double t = 0.0;
        for (uint32_t q = 0; q < kObsGroups; ++q) t += red[q * kObsSums + threadIdx.x];
        out[threadIdx.x] = t;
    }
}

uint32_t observe_rows(uint32_t W, uint32_t n) {
    const uint32_t wpw = 64u / n;
    const uint64_t threads = ((uint64_t(W) + wpw - 1) / wpw) * 64u;
    return grid_for(threads, kObsBlock);
}

hipError_t launch_observe(const float4 *pos, const float4 *vel, uint32_t W, double *partial, double *out,
                          const KP &kp, hipStream_t st) {
    const uint32_t rows = observe_rows(W, kp.p.filamentSize);
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(k_observe, dim3(rows), dim3(kObsBlock), 0, st, pos, vel, W, partial, kp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_observe_total, dim3(1), dim3(kObsGroups * kObsSums), 0, st, partial, rows, out);
    return hipGetLastError();
}

}  // namespace nw

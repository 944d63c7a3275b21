// nw_device.hpp -- device-side building blocks of the worm time step (gfx950).
//
// Vector arithmetic keeps the operation order of the reference's helpers
// (common/helper_math.h: dot = x*x + y*y + z*z at :1248, length = sqrtf(dot)
// at :1291, float3 / float divides per component at :997) so that single
// force terms agree with the FP64 oracle to a few ulp.  The library is built
// with -ffp-contract=off; kernels that opt into FMA contraction say so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "nw_params.h"

namespace nw {

// --------------------------------------------------------------------------
// float3 arithmetic (plain struct, explicit operation order)
// --------------------------------------------------------------------------
struct f3 { float x, y, z; };

__device__ __forceinline__ f3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 mk(float4 a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ void operator+=(f3 &a, f3 b) { a.x += b.x; a.y += b.y; a.z += b.z; }
__device__ __forceinline__ void operator-=(f3 &a, f3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float length(f3 v) { return sqrtf(dot(v, v)); }
__device__ __forceinline__ float4 f4(f3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

// --------------------------------------------------------------------------
// Kernel-argument parameter bundle.  Travels by value (kernarg), so a launch
// never depends on a prior host->device upload.
// --------------------------------------------------------------------------
struct KP {
    SimParams p;
    SolventParams s;
    NWExtParams e;
    uint32_t fdivMagic;   // ceil(2^32 / filamentSize) (0: filamentSize 1)
    uint32_t gxMagic;     // the same for gridSize.x (cell index -> row)
    uint32_t layerMagic;  // and for gridSize.x * gridSize.y (-> layer)
    uint32_t pad;
};

// x / d from m = ceil(2^32 / d), d >= 2.  m d = 2^32 + e with 0 <= e < d,
// so x m / 2^32 = x / d + x e / (d 2^32) < x / d + 1 for x < 2^32: mulhi(x, m)
// is floor(x / d) or one more, and one compare takes the excess off (exact
// for x + d < 2^32; bead and cell ids stay below 2^31).  m = 0 stands for
// d = 1.  No branch: four VALU operations.
__device__ __forceinline__ uint32_t div_exact(uint32_t x, uint32_t m, uint32_t d) {
    const uint32_t q = __umulhi(x, m);
    return m == 0u ? x : q - (q * d > x ? 1u : 0u);
}

// ---------------------------------------------------------------------------
// Wave-wide scans over the 64 lanes of a wavefront, every lane active, in
// DPP lane moves (no LDS round trip): Hillis-Steele steps of 1, 2, 4 and 8
// lanes inside each row of 16 (row_shr), then the row totals carried forward
// by the broadcasts of lane 15 into rows 1 and 3 (row_bcast:15) and of lane
// 31 into rows 2 and 3 (row_bcast:31).  A lane with no source lane, or in a
// row the broadcast skips, reads 0: the identity of + and of max on uint32.
// ---------------------------------------------------------------------------
template <bool MAX>
__device__ __forceinline__ uint32_t dpp_step(uint32_t v, uint32_t w) {
    return MAX ? (v > w ? v : w) : v + w;
}
template <bool MAX>
__device__ __forceinline__ uint32_t wave_incl_scan_dpp(uint32_t v) {
    v = dpp_step<MAX>(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x111, 0xf, 0xf, false)));  // row_shr:1
    v = dpp_step<MAX>(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x112, 0xf, 0xf, false)));  // row_shr:2
    v = dpp_step<MAX>(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x114, 0xf, 0xf, false)));  // row_shr:4
    v = dpp_step<MAX>(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x118, 0xf, 0xf, false)));  // row_shr:8
    v = dpp_step<MAX>(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x142, 0xa, 0xf, false)));  // row_bcast:15
    v = dpp_step<MAX>(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x143, 0xc, 0xf, false)));  // row_bcast:31
    return v;
}
// inclusive prefix sum / running maximum of v over the lanes
__device__ __forceinline__ uint32_t wave_scan_add(uint32_t v) { return wave_incl_scan_dpp<false>(v); }
__device__ __forceinline__ uint32_t wave_scan_max(uint32_t v) { return wave_incl_scan_dpp<true>(v); }
// lane 63's value (of an inclusive scan: the wave's total), wave-uniform
__device__ __forceinline__ uint32_t wave_last_lane(uint32_t v) {
    return uint32_t(__builtin_amdgcn_readlane(int(v), 63));
}

// coordinates of cell c of the MD grid
__device__ __forceinline__ void cell_xyz(const KP &kp, uint32_t c, int &gx, int &gy, int &gz) {
    const uint32_t Gx = kp.p.gridSize.x, layer = kp.p.gridSize.x * kp.p.gridSize.y;
    const uint32_t z = div_exact(c, kp.layerMagic, layer), rem = c - z * layer;
    const uint32_t y = div_exact(rem, kp.gxMagic, Gx);
    gx = int(rem - y * Gx);
    gy = int(y);
    gz = int(z);
}

// i / filamentSize and i % filamentSize for bead ids.
__device__ __forceinline__ void worm_of(const KP &kp, uint32_t i, uint32_t &w, uint32_t &c) {
    const uint32_t n = kp.p.filamentSize;
    w = div_exact(i, kp.fdivMagic, n);
    c = i - w * n;
}

// --------------------------------------------------------------------------
// Minimum image (lengthPeriodic, particles_kernel_impl.cuh:279-301).
// Minimum-image difference b - a.  Same branch decisions as lengthPeriodic
// (wrap when |b - a| > L/2), but a wrapped component is evaluated against the
// box faces, (b - L/2) - (a + L/2), whose two subtractions are exact for
// positions near opposite faces; its rounding error then scales with the
// separation instead of with L (FP32 b - a - L loses ~ulp(L) = 6e-5 at L = 1000).
__device__ __forceinline__ float mi_axis(float b, float a, float L, bool periodic) {
    float d = b - a;
    if (periodic) {
        const float h = L / 2.0f;
        if (d > h) d = (b - h) - (a + h);
        else if (d < -h) d = (b + h) - (a - h);
    }
    return d;
}

__device__ __forceinline__ f3 min_image(const SimParams &p, f3 a, f3 b) {
    return mk(mi_axis(b.x, a.x, p.boxSize.x, p.boundaryX == PERIODIC),
              mi_axis(b.y, a.y, p.boxSize.y, p.boundaryY == PERIODIC),
              mi_axis(b.z, a.z, p.boxSize.z, p.boundaryZ == PERIODIC));
}

// The same value for the pair loops: pairs closer than half the box on every
// axis (all but a sliver near the periodic faces) take the plain difference,
// which is what mi_axis returns for them; the wrap tests run only when some
// lane holds a longer component.
__device__ __forceinline__ f3 min_image_pair(const SimParams &p, f3 a, f3 b) {
    const f3 d = mk(b.x - a.x, b.y - a.y, b.z - a.z);
    if (!(fabsf(d.x) > 0.5f * p.boxSize.x || fabsf(d.y) > 0.5f * p.boxSize.y || fabsf(d.z) > 0.5f * p.boxSize.z))
        return d;
    return min_image(p, a, b);
}

// unit(b - a) under the minimum image (particles_kernel_impl.cuh:303-309)
__device__ __forceinline__ f3 tangent_of(const SimParams &p, f3 a, f3 b) {
    const f3 rel = min_image_pair(p, a, b);
    const float d = length(rel);
    return rel / d;
}

// --------------------------------------------------------------------------
// Grid hashing (particles_kernel_impl.cuh:208-276).  The cell index uses the
// same FP32 sequence as the reference, floor((p - o) / c), with correctly
// rounded division.  Wrapping uses the mathematical modulo, which equals the
// reference's `& (G - 1)` for power-of-two G and stays correct otherwise.
// --------------------------------------------------------------------------
__device__ __forceinline__ int wrap_mod(int g, uint32_t G) {
    if ((G & (G - 1u)) == 0u) return g & int(G - 1u);
    const int Gi = int(G);
    // coordinates of positions in or next to the box: one add or subtract
    if (g >= -Gi && g < 2 * Gi) return g < 0 ? g + Gi : (g >= Gi ? g - Gi : g);
    const int m = g % Gi;
    return m < 0 ? m + Gi : m;
}

// floor(d / c) of the correctly rounded quotient, rc the hardware reciprocal
// of c.  t = d * rc is within 2^-22 |t| of d / c (reciprocal 1 ulp, one
// rounding each for the product and the quotient), so farther than
// 2^-19 |t| from an integer both have the same floor; closer (and for |t|
// below 2^-100, NaN or infinity, where the distance is NaN or 0) the
// quotient is formed exactly.  The distance to the nearest integer,
// |t - rint(t)|, is exact (t and rint(t) within a factor of two, or t
// small).
__device__ __forceinline__ int floor_div_rc(float d, float c, float rc) {
    const float t = d * rc;
    if (fabsf(t - rintf(t)) > fmaf(fabsf(t), 0x1p-19f, 0x1p-100f)) return int(floorf(t));
    return int(floorf(d / c));
}
__device__ __forceinline__ int floor_div(float d, float c) { return floor_div_rc(d, c, __builtin_amdgcn_rcpf(c)); }

__device__ __forceinline__ int cell_coord(float p, float o, float c) {
    return floor_div(p - o, c);
}

__device__ __forceinline__ uint32_t md_hash_wrapped(const SimParams &p, int gx, int gy, int gz) {
    const uint32_t x = uint32_t(wrap_mod(gx, p.gridSize.x));
    const uint32_t y = uint32_t(wrap_mod(gy, p.gridSize.y));
    const uint32_t z = uint32_t(wrap_mod(gz, p.gridSize.z));
    return (z * p.gridSize.y) * p.gridSize.x + y * p.gridSize.x + x;
}

// positionRelativeToBox (particles_kernel_impl.cuh:208-228), one axis: the
// coordinate relative to the origin o, moved by one box length L into
// [0, L) on a periodic axis
__device__ __forceinline__ float relative_to_box(float q, float o, float L, bool periodic) {
    float r = q - o;
    if (periodic) {
        if (r >= L) r -= L;
        else if (r < 0.0f) r += L;
    }
    return r;
}
__device__ __forceinline__ f3 position_relative_to_box(const SimParams &p, f3 q) {
    return mk(relative_to_box(q.x, p.origin.x, p.boxSize.x, p.boundaryX == PERIODIC),
              relative_to_box(q.y, p.origin.y, p.boxSize.y, p.boundaryY == PERIODIC),
              relative_to_box(q.z, p.origin.z, p.boxSize.z, p.boundaryZ == PERIODIC));
}

// SRD cell of a position (calcGridPosSRD :241-249 + calcGridHashSRD :261-276).
// Periodic axes wrap; a non-periodic axis is clamped into [0, G) (the
// reference leaves it unwrapped, which can index past the cell table).
// The grid is shifted by the collision's randOffset; SrdGrid holds that
// shift and the cell sizes' reciprocals, formed once for many positions.
struct SrdGrid {
    f3 shift;
    float rcx, rcy, rcz;
};
__device__ __forceinline__ SrdGrid srd_grid(const SolventParams &s, f3 shift) {
    return {shift, __builtin_amdgcn_rcpf(s.cellSize.x), __builtin_amdgcn_rcpf(s.cellSize.y),
            __builtin_amdgcn_rcpf(s.cellSize.z)};
}
__device__ __forceinline__ uint32_t srd_hash_g(const SimParams &p, const SolventParams &s, const SrdGrid &g, f3 q) {
    const f3 shifted = q + g.shift;
    const float rx = relative_to_box(shifted.x, p.origin.x, p.boxSize.x, p.boundaryX == PERIODIC);
    const float ry = relative_to_box(shifted.y, p.origin.y, p.boxSize.y, p.boundaryY == PERIODIC);
    int gx = floor_div_rc(rx, s.cellSize.x, g.rcx);
    int gy = floor_div_rc(ry, s.cellSize.y, g.rcy);
    if (p.boundaryX == PERIODIC) gx = wrap_mod(gx, s.gridSize.x);
    else gx = min(max(gx, 0), int(s.gridSize.x) - 1);
    if (p.boundaryY == PERIODIC) gy = wrap_mod(gy, s.gridSize.y);
    else gy = min(max(gy, 0), int(s.gridSize.y) - 1);
    // a grid of one layer puts every z in layer 0 (the wrap and the clamp
    // both give 0 for any coordinate), so its z is not computed
    int gz = 0;
    if (s.gridSize.z > 1u) {
        const float rz = relative_to_box(shifted.z, p.origin.z, p.boxSize.z, p.boundaryZ == PERIODIC);
        gz = floor_div_rc(rz, s.cellSize.z, g.rcz);
        if (p.boundaryZ == PERIODIC) gz = wrap_mod(gz, s.gridSize.z);
        else gz = min(max(gz, 0), int(s.gridSize.z) - 1);
    }
    return (uint32_t(gz) * s.gridSize.y) * s.gridSize.x + uint32_t(gy) * s.gridSize.x + uint32_t(gx);
}
__device__ __forceinline__ uint32_t srd_hash_at(const SimParams &p, const SolventParams &s, f3 shift, f3 q) {
    return srd_hash_g(p, s, srd_grid(s, shift), q);
}

__device__ __forceinline__ uint32_t srd_hash(const SimParams &p, const SolventParams &s, f3 q) {
    return srd_hash_at(p, s, mk(s.randOffset.x, s.randOffset.y, s.randOffset.z), q);
}

// --------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies with the
// published multipliers and Weyl key increments.
// --------------------------------------------------------------------------
struct u4 { uint32_t x, y, z, w; };

__host__ __device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b) {
    return uint32_t((uint64_t(a) * uint64_t(b)) >> 32);
}

__host__ __device__ __forceinline__ u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = mulhi32(M0, c.x), lo0 = M0 * c.x;
        const uint32_t hi1 = mulhi32(M1, c.z), lo1 = M1 * c.z;
        c = u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// Generator block b of stream `seed`: counter = (b_lo, b_hi, 0, 0).
__host__ __device__ __forceinline__ u4 philox_block(uint64_t seed, uint64_t b) {
    return philox4x32_10(u4{uint32_t(b), uint32_t(b >> 32), 0u, 0u}, uint32_t(seed), uint32_t(seed >> 32));
}

// 32-bit integer -> float uniform in (0, 1]: (x >> 8) * 2^-24 + 2^-24 (exact).
__device__ __forceinline__ float u01(uint32_t x) {
    return float(x >> 8) * 5.9604644775390625e-8f + 5.9604644775390625e-8f;
}

// Record i of a as one 16-byte load (read through a float4 whose w goes
// unused, the compiler may issue three or four 4-byte loads instead)
__device__ __forceinline__ float4 load4(const float4 *__restrict__ a, uint32_t i) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 x = reinterpret_cast<const v4 *>(a)[i];
    return make_float4(x.x, x.y, x.z, x.w);
}

// Record i of a, or zeros where !use, as one 16-byte load (a lane with !use
// reads record 0: a must hold at least one) and a select.  Written as
// `use ? a[i] : 0` the compiler splits the record into four 4-byte loads, each
// behind its own branch; here it keeps one load (which it may still guard by
// `use`: a caller that needs the loads unconditional pins them itself).
__device__ __forceinline__ float4 load4_or_zero(const float4 *__restrict__ a, uint32_t i, bool use) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 x = reinterpret_cast<const v4 *>(a)[use ? i : 0u];
    return use ? make_float4(x.x, x.y, x.z, x.w) : make_float4(0.f, 0.f, 0.f, 0.f);
}

// Box-Muller: two normals from two uniforms in (0, 1].
__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, float &z0, float &z1) {
    const float u1 = u01(a), u2 = u01(b);
    const float r = sqrtf(-2.0f * logf(u1));
    // angle 2 pi u2 as sin / cos of pi (2 u2): 2 u2 is exact, so no rounding
    // of the angle, and no range reduction
    float s, c;
    sincospif(2.0f * u2, &s, &c);
    z0 = r * c;
    z1 = r * s;
}

}  // namespace nw

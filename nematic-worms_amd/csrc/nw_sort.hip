// nw_sort.hip -- stable key/value sorts of cell hashes on gfx950.
//
// counting_sort_pairs: one digit of radix `bound` (a counting sort).  The
//   histogram's atomic slot order is arbitrary; a per-bin pass then restores
//   the stable order by sorting each bin's members by original position, so
//   the output equals std::stable_sort of (key, value) pairs exactly.  It also
//   leaves the exclusive bin offsets behind (cell start table for free).
// radix_sort_pairs: LSD radix sort, 8-bit digits, stable in-tile ranking via
//   wave ballots; general fallback for keys without a known bound.
// exclusive_scan_u32: tile reduce / recursive scan of tile sums / tile scan.
#include "nw_internal.hpp"

namespace nw {

// ---------------------------------------------------------------------------
// exclusive scan
// ---------------------------------------------------------------------------
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;  // 2048

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) { return wave_scan_add(v); }

// block-wide exclusive scan of one value per thread; returns the block total
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t &total) {
    __shared__ uint32_t wsum[kBlock / 64];
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    const uint32_t incl = wave_incl_scan(v);
    if (lane == 63u) wsum[wid] = incl;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        const uint32_t s = wsum[w];
        if (uint32_t(w) < wid) base += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return base + incl - v;
}

__global__ __launch_bounds__(kBlock) void k_tile_reduce(const uint32_t *__restrict__ in, uint64_t m,
                                                        uint32_t *__restrict__ sums) {
    const uint64_t base = uint64_t(blockIdx.x) * kScanTile + uint64_t(threadIdx.x) * kScanItems;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < m) s += in[base + k];
    uint32_t tot;
    (void)block_excl_scan(s, tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// scans one tile; carry = exclusive prefix of the tile (nullptr: 0).  Writes
// out[m] = carry + tile total from the tile that contains element m-1.
__global__ __launch_bounds__(kBlock) void k_tile_scan(const uint32_t *__restrict__ in, uint64_t m,
                                                      const uint32_t *__restrict__ carry, uint32_t *__restrict__ out) {
    const uint64_t base = uint64_t(blockIdx.x) * kScanTile + uint64_t(threadIdx.x) * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < m ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan(s, tot);
    const uint32_t c = carry ? carry[blockIdx.x] : 0u;
    run += c;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < m) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0 && uint64_t(blockIdx.x) == (m - 1) / kScanTile) out[m] = c + tot;
}

uint64_t scan_scratch_words(uint64_t m) {
    uint64_t words = 0;
    while (m > uint64_t(kScanTile)) {
        const uint64_t nb = (m + kScanTile - 1) / kScanTile;
        words += 2 * (nb + 1);
        m = nb;
    }
    return words + 2;
}

hipError_t exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint64_t m, uint32_t *tmp, hipStream_t st) {
    if (m == 0) return launch_fill_u32(out, 0u, 1, st);
    const uint64_t nb = (m + kScanTile - 1) / kScanTile;
    if (nb == 1) {
        hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(kBlock), 0, st, in, m, nullptr, out);
        return hipGetLastError();
    }
    uint32_t *sums = tmp;            // nb
    uint32_t *sumsScan = tmp + nb + 1; // nb + 1
    uint32_t *rest = sumsScan + nb + 1;
    hipLaunchKernelGGL(k_tile_reduce, dim3(unsigned(nb)), dim3(kBlock), 0, st, in, m, sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = exclusive_scan_u32(sums, sumsScan, nb, rest, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_tile_scan, dim3(unsigned(nb)), dim3(kBlock), 0, st, in, m, sumsScan, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// counting sort
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_count(const uint32_t *__restrict__ keys, uint32_t n, uint32_t bound,
                                                  uint32_t *__restrict__ counts, uint32_t *__restrict__ slot,
                                                  uint32_t *__restrict__ flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = keys[i];
    if (k >= bound) {
        atomicOr(flags, 1u);
        k = bound;
    }
    slot[i] = atomicAdd(&counts[k], 1u);
}

__global__ __launch_bounds__(kBlock) void k_scatter(const uint32_t *__restrict__ keys,
                                                    const uint32_t *__restrict__ vals, uint32_t n, uint32_t bound,
                                                    const uint32_t *__restrict__ offsets,
                                                    const uint32_t *__restrict__ slot, uint2 *__restrict__ pairs,
                                                    uint32_t *__restrict__ okeys) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = keys[i];
    const uint32_t b = k < bound ? k : bound;
    const uint32_t p = offsets[b] + slot[i];
    pairs[p] = make_uint2(i, vals[i]);
    if (b == bound) okeys[p] = k;
}

// One thread per bin: insertion sort of the bin's (position, value) pairs by
// position, then write values and keys back in place.  The overflow bin
// (keys >= bound) is ordered by (key, position).
__global__ __launch_bounds__(kBlock) void k_bin_fixup(uint2 *__restrict__ pairs, uint32_t *__restrict__ okeys,
                                                      const uint32_t *__restrict__ offsets, uint32_t bound,
                                                      uint32_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > bound) return;
    const uint32_t s = offsets[b], e = offsets[b + 1];
    if (s == e) return;
    if (b < bound) {
        for (uint32_t a = s + 1; a < e; ++a) {
            const uint2 x = pairs[a];
            uint32_t c = a;
            while (c > s) {
                const uint2 y = pairs[c - 1];
                if (y.x <= x.x) break;
                pairs[c] = y;
                --c;
            }
            pairs[c] = x;
        }
        for (uint32_t a = s; a < e; ++a) {
            vals[a] = pairs[a].y;
            keys[a] = b;
        }
    } else {
        for (uint32_t a = s + 1; a < e; ++a) {
            const uint2 x = pairs[a];
            const uint32_t xk = okeys[a];
            uint32_t c = a;
            while (c > s) {
                const uint2 y = pairs[c - 1];
                const uint32_t yk = okeys[c - 1];
                if (yk < xk || (yk == xk && y.x <= x.x)) break;
                pairs[c] = y;
                okeys[c] = yk;  // shifted alongside; only this thread owns the bin
                --c;
            }
            pairs[c] = x;
            okeys[c] = xk;
        }
        for (uint32_t a = s; a < e; ++a) {
            vals[a] = pairs[a].y;
            keys[a] = okeys[a];
        }
    }
}

hipError_t ensure_sort_scratch(SortScratch &s, uint64_t n, uint64_t bins) {
    hipError_t e = hipSuccess;
    if (n > s.cap_n) {
        uint64_t cap = n + n / 8 + 1024;
        if (s.slot) { (void)hipFree(s.slot); (void)hipFree(s.pairs); (void)hipFree(s.keys); (void)hipFree(s.vals); }
        s.slot = nullptr; s.pairs = nullptr; s.keys = nullptr; s.vals = nullptr;
        if ((e = hipMalloc(&s.slot, cap * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&s.pairs, cap * 8)) != hipSuccess) return e;
        if ((e = hipMalloc(&s.keys, cap * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&s.vals, cap * 4)) != hipSuccess) return e;
        s.cap_n = cap;
    }
    if (bins > s.cap_bins) {
        uint64_t cap = bins + bins / 8 + 1024;
        if (s.counts) { (void)hipFree(s.counts); (void)hipFree(s.offsets); }
        s.counts = nullptr; s.offsets = nullptr;
        if ((e = hipMalloc(&s.counts, (cap + 1) * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&s.offsets, (cap + 2) * 4)) != hipSuccess) return e;
        s.cap_bins = cap;
    }
    // radix sort scans 256 digits x tiles; counting sort scans bins + 1
    const uint64_t tiles = (n + 2047) / 2048 + 1;
    const uint64_t scanLen = bins + 1 > 256 * tiles ? bins + 1 : 256 * tiles;
    const uint64_t words = scan_scratch_words(scanLen) + 256 * tiles + 8;
    if (words > s.cap_scan) {
        uint64_t cap = words + words / 8 + 1024;
        if (s.scan_tmp) (void)hipFree(s.scan_tmp);
        s.scan_tmp = nullptr;
        if ((e = hipMalloc(&s.scan_tmp, cap * 4)) != hipSuccess) return e;
        s.cap_scan = cap;
    }
    return hipSuccess;
}

hipError_t counting_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t n, uint32_t bound, SortScratch &s,
                               hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipError_t e = ensure_sort_scratch(s, n, uint64_t(bound) + 1);
    if (e != hipSuccess) return e;
    const uint64_t bins = uint64_t(bound) + 1;  // last bin collects keys >= bound
    if ((e = launch_fill_u32(s.counts, 0u, bins, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_count, dim3(grid_for(n)), dim3(kBlock), 0, st, keys, n, bound, s.counts, s.slot,
                       device_flags());
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = exclusive_scan_u32(s.counts, s.offsets, bins, s.scan_tmp, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_scatter, dim3(grid_for(n)), dim3(kBlock), 0, st, keys, vals, n, bound, s.offsets, s.slot,
                       s.pairs, s.keys);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_bin_fixup, dim3(grid_for(bins)), dim3(kBlock), 0, st, s.pairs, s.keys, s.offsets, bound,
                       keys, vals);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// LSD radix sort, 8-bit digits.  Tile = 256 threads x 8 keys (thread-major
// runs of consecutive keys are processed in 8 rounds of 256 consecutive keys
// so that in-tile ranks follow input order).
// ---------------------------------------------------------------------------
constexpr int kRadixBits = 8;
constexpr int kRadixBins = 1 << kRadixBits;
constexpr int kRadixRounds = 8;
constexpr int kRadixTile = kBlock * kRadixRounds;

__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint32_t *__restrict__ keys, uint32_t n, int shift,
                                                       uint32_t *__restrict__ hist, uint32_t tiles) {
    __shared__ uint32_t h[kRadixBins];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = uint64_t(blockIdx.x) * kRadixTile;
#pragma unroll
    for (int r = 0; r < kRadixRounds; ++r) {
        const uint64_t i = base + uint64_t(r) * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (kRadixBins - 1)], 1u);
    }
    __syncthreads();
    hist[uint64_t(threadIdx.x) * tiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kBlock) void k_radix_scatter(const uint32_t *__restrict__ kin,
                                                          const uint32_t *__restrict__ vin, uint32_t n, int shift,
                                                          const uint32_t *__restrict__ offs, uint32_t tiles,
                                                          uint32_t *__restrict__ kout, uint32_t *__restrict__ vout) {
    __shared__ uint32_t base[kRadixBins];       // running output position per digit
    __shared__ uint32_t wcnt[kBlock / 64][kRadixBins];
    base[threadIdx.x] = offs[uint64_t(threadIdx.x) * tiles + blockIdx.x];
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < kRadixRounds; ++r) {
        for (int w = 0; w < kBlock / 64; ++w) wcnt[w][threadIdx.x] = 0;
        __syncthreads();
        const uint64_t i = uint64_t(blockIdx.x) * kRadixTile + uint64_t(r) * kBlock + threadIdx.x;
        const bool valid = i < n;
        const uint32_t k = valid ? kin[i] : 0u;
        const uint32_t v = valid ? vin[i] : 0u;
        const uint32_t d = (k >> shift) & (kRadixBins - 1);
        // lanes of this wave holding the same digit
        uint64_t same = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const uint64_t bb = __ballot((d >> b) & 1u);
            same &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t rankInWave = __popcll(same & lt);
        const uint32_t leader = same ? (__ffsll((unsigned long long)same) - 1) : 64;
        if (valid && lane == leader) wcnt[wid][d] = __popcll(same);
        __syncthreads();
        if (valid) {
            uint32_t before = 0;
            for (uint32_t w = 0; w < wid; ++w) before += wcnt[w][d];
            const uint32_t p = base[d] + before + rankInWave;
            kout[p] = k;
            vout[p] = v;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < kBlock / 64; ++w) add += wcnt[w][threadIdx.x];
        base[threadIdx.x] += add;
        __syncthreads();
    }
}

hipError_t radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t n, uint32_t keyBits, SortScratch &s,
                            hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (keyBits == 0) return hipSuccess;
    if (keyBits > 32) keyBits = 32;
    hipError_t e = ensure_sort_scratch(s, n, 1);
    if (e != hipSuccess) return e;
    const uint32_t tiles = uint32_t((uint64_t(n) + kRadixTile - 1) / kRadixTile);
    const uint64_t hlen = uint64_t(kRadixBins) * tiles;
    uint32_t *hist = s.scan_tmp;
    uint32_t *offs = s.counts;  // needs hlen + 1 words
    if (hlen + 1 > s.cap_bins) {
        if ((e = ensure_sort_scratch(s, n, hlen + 1)) != hipSuccess) return e;
        offs = s.counts;
        hist = s.scan_tmp;
    }
    uint32_t *scanTmp = s.scan_tmp + hlen + 8;
    uint32_t *kA = keys, *vA = vals, *kB = s.keys, *vB = s.vals;
    const int passes = int((keyBits + kRadixBits - 1) / kRadixBits);
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = pass * kRadixBits;
        hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kBlock), 0, st, kA, n, shift, hist, tiles);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = exclusive_scan_u32(hist, offs, hlen, scanTmp, st)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kBlock), 0, st, kA, vA, n, shift, offs, tiles, kB, vB);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        uint32_t *t;
        t = kA; kA = kB; kB = t;
        t = vA; vA = vB; vB = t;
    }
    if (kA != keys) {
        if ((e = hipMemcpyAsync(keys, kA, uint64_t(n) * 4, hipMemcpyDeviceToDevice, st)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(vals, vA, uint64_t(n) * 4, hipMemcpyDeviceToDevice, st)) != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace nw

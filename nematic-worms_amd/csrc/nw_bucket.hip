// nw_bucket.hip -- atomic-free two-level cell sort of the fused steps and the
// SRD collision over it (gfx950).
//
// Device-scope atomics serialise far from the CUs on MI355X (a few tens of
// G/s measured), so a histogram with one global counter per cell costs more
// than the data it sorts, and a global scan over a tile x bucket matrix costs
// launches.  Here every counter lives in LDS and no global scan is needed:
//   level 1   (k_tiles_staged, k_tiles) a tile of consecutive particles
//             counts its members per bucket in LDS, scans the counts and
//             writes the prefix as one row of P, then its (cell, id) pairs
//             grouped by bucket into its pair region (staged in LDS and
//             written out coalesced when the tile fits)
//   level 2   (k_bucket_cells) a workgroup per bucket, buckets mapped to
//             XCDs in contiguous runs so neighbours share L2 lines of P: the
//             column sum of P at the bucket is its first sorted slot, the
//             column differences are its segments in the tiles' regions; the
//             members are counted per cell in LDS, placed, each cell sorted
//             by id (the order of a stable sort by cell), written out with
//             the CSR offsets; MD gathers the sorted records, SRD sums every
//             cell in sorted order ([isokinetic] + centre of momentum)
//   rotate    (k_srd_rotate) particle order, coalesced: every particle
//             rotates about its cell's momentum, solvent streams
// Results are deterministic and equal to a stable sort by cell.
#include <algorithm>
#include <utility>
#include <vector>

#include "nw_bucket.hpp"
#include "nw_physics.hpp"

namespace nw {

constexpr int kTileBlock = 1024;

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
BucketPlan make_bucket_plan(uint64_t n, uint64_t cells, uint32_t tileItems, uint32_t members) {
    BucketPlan b{};
    b.n = uint32_t(n);
    b.cells = uint32_t(cells);
    // at most `members` members per bucket on average, at most 2048 cells per
    // bucket
    const double perCell = cells ? double(n) / double(cells) : 1.0;
    int S = 0;
    while (S < 11 && double(1u << (S + 1)) * perCell <= double(members)) ++S;
    b.shift = uint32_t(S);
    b.buckets = uint32_t((cells + (1u << S) - 1) >> S);
    if (tileItems == 0) {
        // P stays near 2^21 entries: tiles of n (buckets + 1) / 2^21, power of two
        const double want = double(n) * double(b.buckets + 1) / double(1u << 21);
        uint32_t T = 4096;
        while (T < (1u << 17) && double(T) < want) T <<= 1;
        // 8192 rather than 16384: the SRD workload's level 1 takes 35 us
        // against 50 us, for a level 2 of 88 us against 84 us
        if (T == 16384) T = 8192;
        tileItems = T;
    }
    b.tileItems = tileItems;
    b.tiles = uint32_t((n + tileItems - 1) / tileItems);
    return b;
}

Tuning &tuning() {
    static Tuning t;
    return t;
}

BucketPlan fused_plan(uint64_t n, uint64_t cells, bool srd) {
    const Tuning &t = tuning();
    const uint32_t items = srd ? t.srdTileItems : t.mdTileItems;
    const uint32_t members = srd ? t.srdMembers : t.mdMembers;
    const BucketPlan bp = make_bucket_plan(n, cells, items, members ? members : 1536u);
    if ((items || members) && *bucket_plan_problem(bp)) return make_bucket_plan(n, cells, 0);
    return bp;
}

const char *bucket_plan_problem(const BucketPlan &bp) {
    if (bp.buckets > kMaxBuckets) return "grid has too many cells for the fused sort";
    if (bp.tiles > kMaxTiles) return "too many particles for the fused sort";
    if (uint64_t(bp.tiles) * (bp.buckets + 1) > kMaxMatrix) return "particles x cells too large for the fused sort";
    if (bp.tileItems > (1u << kPairIdBits)) return "level-1 tile too large for the packed sort pairs";
    return "";
}

// level-2 LDS layout (bytes), shared by host (size) and device (carving)
struct CellsLayout {
    uint32_t wsum, cnt, cur, lid, lcell, segStart, segBase, tileOf, total;
};

__host__ __device__ inline uint32_t align16(uint32_t x) { return (x + 15u) & ~15u; }

// cap: members a bucket stages in LDS (more: the global-memory path)
__host__ __device__ inline CellsLayout cells_layout(const BucketPlan &bp, bool srd, uint32_t cap = kBucketCap) {
    CellsLayout L;
    const uint32_t nc = 1u << bp.shift;
    uint32_t o = 0;
    L.wsum = o; o += align16(4u * 2u * (kCellsBlockMax / 64));
    L.cnt = o;  o += align16(4u * (nc + 1));
    L.cur = o;  o += align16(4u * nc);
    L.lid = o;  o += align16(4u * cap);
    L.lcell = o; if (!srd) o += align16(2u * cap);
    // segment tables (the staged members themselves wait in registers)
    L.segStart = o; o += align16(4u * (bp.tiles + 1));
    L.segBase = o;  o += align16(4u * bp.tiles);
    L.tileOf = L.lid;  // phase A only: lid is first written after it
    L.total = o;
    return L;
}


// words of a staged level-1 tile's LDS before its pair stage: the bucket
// counts (B1), the scan's wave sums, the tile's bucket range (2), aligned to 4
__host__ __device__ constexpr uint32_t staged_stage_words(uint32_t B1, uint32_t TB) {
    return (B1 + TB / 64 + 2 + 3) & ~3u;
}

// ---------------------------------------------------------------------------
// level 1 as a kernel of its own: the cell of every particle (SRD: from the
// position, written to hash; otherwise read from hash), tile histogram, P
// row, pairs
// ---------------------------------------------------------------------------
template <bool FROM_POS>
__global__ __launch_bounds__(kTileBlock) void k_tiles(const float4 *__restrict__ pos, uint32_t *__restrict__ hash,
                                                      uint32_t *__restrict__ pairs, uint32_t *__restrict__ P,
                                                      uint32_t *__restrict__ range,
                                                      BucketPlan bp, KP kp) {
    // the whole row is written: the full bucket range
    if (range && threadIdx.x == 0) {
        range[2u * blockIdx.x] = 0u;
        range[2u * blockIdx.x + 1u] = bp.buckets - 1u;
    }
    extern __shared__ uint32_t lh[];
    uint32_t *wsum = lh + bp.buckets + 1;
    const uint32_t B1 = bp.buckets + 1;
    for (uint32_t b = threadIdx.x; b < B1; b += blockDim.x) lh[b] = 0u;
    __syncthreads();
    const uint32_t t = blockIdx.x;
    const uint32_t i0 = t * bp.tileItems;
    const uint32_t i1 = min(i0 + bp.tileItems, bp.n);
    for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        uint32_t h;
        if (FROM_POS) {
            h = srd_hash(kp.p, kp.s, mk(pos[i]));
            hash[i] = h;
        } else {
            h = min(hash[i], bp.cells - 1u);  // a key outside the grid sorts as its last cell
        }
        atomicAdd(&lh[h >> bp.shift], 1u);
    }
    __syncthreads();
    block_excl_scan(lh, bp.buckets, wsum);
    uint32_t *row = P + uint64_t(t) * B1;
    for (uint32_t b = threadIdx.x; b < B1; b += blockDim.x) row[b] = lh[b];
    __syncthreads();
    // second pass over the same particles (each thread re-reads its own cells)
    const uint32_t cellMask = (1u << bp.shift) - 1u;
    for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const uint32_t h = min(hash[i], bp.cells - 1u);
        const uint32_t slot = atomicAdd(&lh[h >> bp.shift], 1u);
        pairs[i0 + slot] = pack_pair(h & cellMask, i - i0);
    }
}

// The same level 1 with each thread's ITEMS cells and ranks kept in
// registers (one read of the cells) and the tile's pairs staged in LDS in
// bucket order, then written out with coalesced stores (tiles of
// 1024 x ITEMS particles).
// Tile t of one system: hash / pairs / P start at the system's own first
// entries and its cells start at cellBase in the hash values (zero for a
// single system).
template <bool FROM_POS, int ITEMS, int TB = kTileBlock>
__device__ __forceinline__ void tiles_staged_body(const float4 *__restrict__ pos, uint32_t *__restrict__ hash,
                                                  uint32_t *__restrict__ pairs, uint32_t *__restrict__ P,
                                                  uint32_t *__restrict__ range, const BucketPlan &bp, const KP &kp,
                                                  uint32_t t, uint32_t cellBase, uint32_t *lh) {
    const uint32_t B1 = bp.buckets + 1;
    uint32_t *wsum = lh + B1;
    uint32_t *brng = wsum + TB / 64;  // the tile's first and last bucket
    uint32_t *stage = lh + staged_stage_words(B1, TB);
    for (uint32_t b = threadIdx.x; b < B1; b += blockDim.x) lh[b] = 0u;
    if (threadIdx.x == 0) {
        brng[0] = 0xffffffffu;
        brng[1] = 0u;
    }
    __syncthreads();
    const uint32_t i0 = t * bp.tileItems;
    const uint32_t i1 = min(i0 + bp.tileItems, bp.n);
    uint32_t h[ITEMS], rk[ITEMS];
    const SrdGrid g = srd_grid(kp.s, mk(kp.s.randOffset.x, kp.s.randOffset.y, kp.s.randOffset.z));
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t i = i0 + uint32_t(k) * TB + threadIdx.x;
        h[k] = 0u;
        rk[k] = 0u;
        if (i < i1) {
            if (FROM_POS) {
                h[k] = srd_hash_g(kp.p, kp.s, g, mk(pos[i]));
                hash[i] = h[k];
            } else {
                // clamped: a key outside the grid sorts as its last cell
                h[k] = min(hash[i] - cellBase, bp.cells - 1u);
            }
            rk[k] = atomicAdd(&lh[h[k] >> bp.shift], 1u);
        }
    }
    if (range) {
        uint32_t lo = 0xffffffffu, hi = 0u;
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            if (i0 + uint32_t(k) * TB + threadIdx.x < i1) {
                lo = min(lo, h[k] >> bp.shift);
                hi = max(hi, h[k] >> bp.shift);
            }
        }
        // the wave's least and greatest (the least as the complement of the
        // greatest complement)
        lo = ~wave_last_lane(wave_scan_max(~lo));
        hi = wave_last_lane(wave_scan_max(hi));
        if ((threadIdx.x & 63u) == 0u) {
            atomicMin(&brng[0], lo);
            atomicMax(&brng[1], hi);
        }
    }
    __syncthreads();
    // With a range: the buckets outside [b0, b1] hold no member, so the
    // prefix is 0 below b0 and the tile's count above b1 -- only the entries
    // b0 .. b1 + 1 are scanned and written (the cells pass reads no other).
    const uint32_t b0 = range ? brng[0] : 0u, b1 = range ? brng[1] : bp.buckets - 1u;
    block_excl_scan(lh + b0, b1 - b0 + 1u, wsum);
    uint32_t *row = P + uint64_t(t) * B1;
    for (uint32_t b = b0 + threadIdx.x; b <= b1 + 1u; b += blockDim.x) row[b] = lh[b];
    if (range && threadIdx.x == 0) {
        range[2u * t] = b0;
        range[2u * t + 1u] = b1;
    }
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t i = i0 + uint32_t(k) * TB + threadIdx.x;
        if (i < i1) stage[lh[h[k] >> bp.shift] + rk[k]] = pack_pair(h[k] & ((1u << bp.shift) - 1u), i - i0);
    }
    __syncthreads();
    const uint32_t cnt = i1 - i0;
    for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) pairs[i0 + k] = stage[k];
}

template <bool FROM_POS, int ITEMS, int TB = kTileBlock>
__global__ __launch_bounds__(TB) void k_tiles_staged(const float4 *__restrict__ pos,
                                                             uint32_t *__restrict__ hash,
                                                             uint32_t *__restrict__ pairs, uint32_t *__restrict__ P,
                                                             uint32_t *__restrict__ range, BucketPlan bp, KP kp) {
    extern __shared__ __align__(16) uint32_t lh[];
    tiles_staged_body<FROM_POS, ITEMS, TB>(pos, hash, pairs, P, range, bp, kp, blockIdx.x, 0u, lh);
}

// the replica that owns item v of a launch over all replicas (offsets
// ascending, wave-uniform)
__device__ __forceinline__ uint32_t sort_rep_of(const SortRep *__restrict__ reps, uint32_t R, uint32_t v,
                                                bool byTile) {
    uint32_t lo = 0, hi = R;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((byTile ? reps[mid].tileOff : reps[mid].bucketOff) <= v) lo = mid;
        else hi = mid;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

__device__ __forceinline__ BucketPlan plan_of(const SortRep &r) {
    BucketPlan bp;
    bp.n = r.n;
    bp.cells = r.cells;
    bp.shift = r.shift;
    bp.buckets = r.buckets;
    bp.tileItems = kBatchTileItems;
    bp.tiles = r.tiles;
    return bp;
}

// level 1 of the batched sort: every replica's tiles in one launch, each
// against its own buckets and prefix matrix (256 threads, 16 beads each: the
// LDS stage, not the wave count, bounds the tiles a CU holds)
constexpr int kBatchTileBlock = 256;

__global__ __launch_bounds__(kBatchTileBlock) void k_tiles_batch(const uint32_t *__restrict__ hash,
                                                                 uint32_t *__restrict__ pairs,
                                                                 uint32_t *__restrict__ P,
                                                                 const SortRep *__restrict__ reps, uint32_t R) {
    extern __shared__ __align__(16) uint32_t lh[];
    const SortRep &r = reps[sort_rep_of(reps, R, blockIdx.x, true)];
    const BucketPlan bp = plan_of(r);
    tiles_staged_body<false, kBatchTileItems / kBatchTileBlock, kBatchTileBlock>(
        nullptr, const_cast<uint32_t *>(hash) + r.beadOffset, pairs + r.beadOffset, P + r.matOff, nullptr, bp, KP{},
        blockIdx.x - r.tileOff, r.cellOffset, lh);
}

// ---------------------------------------------------------------------------
// level 2
// ---------------------------------------------------------------------------
struct CellsArgs {
    const uint32_t *pairs;
    const uint32_t *P;
    // each tile's first and last bucket (NULL: the rows of P are whole)
    const uint32_t *tileRange;
    uint32_t *ids, *off, *cellOf;
    // added to the cells written to cellOf and to the slots written to off
    // (a replica's first cell and first bead; zero for a single system)
    uint32_t cellBase, slotBase;
    // the id of the system's first particle (a replica's first bead in the
    // concatenated arrays; zero for a single system)
    uint32_t idBase;
    // MD gathers (sPos NULL: none)
    const float4 *pos, *vel, *tan;
    float4 *sPos, *sVel, *sTan;
    // SRD cell sums
    const float4 *srdVel;
    float4 *com;
    float *lam;
    int iso;
    float kbT;
    uint64_t seed, offset;
    // batched replicas (NULL: one system): each cell's sign from its replica's stream
    const SrdRep *srdReps;
    uint32_t srdR, collision;
};

// the replica that owns SRD cell `cell` (cell offsets ascending)
__device__ __forceinline__ uint32_t srd_rep_of_cell(const SrdRep *__restrict__ reps, uint32_t R, uint32_t cell) {
    uint32_t lo = 0, hi = R;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (reps[mid].cellOffset <= cell) lo = mid;
        else hi = mid;
    }
    return lo;
}

// index of the largest t in [0, tiles) with segStart[t] <= k
__device__ __forceinline__ uint32_t seg_of(const uint32_t *segStart, uint32_t tiles, uint32_t k) {
    uint32_t lo = 0, hi = tiles;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (segStart[mid] <= k) lo = mid;
        else hi = mid;
    }
    return lo;
}


// MD level 2, staged bucket of m members: ids, cells and the gathered
// records of the sorted slots, in groups of G members per thread with the
// gathers of a group in flight together; past the last member a thread
// handles member m - 1 again (the same values to the same addresses)
template <bool POS, bool VEL, bool TAN, int BLOCK, int G>
__device__ __forceinline__ void md_gather_group(const CellsArgs &o, const uint32_t *lid, const uint16_t *lcell,
                                                uint32_t sb, uint32_t cell0, uint32_t m, uint32_t j0) {
    // records moved as plain 4-vectors (register values, never a copy of a
    // float4 object through memory)
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 *pos = reinterpret_cast<const v4 *>(o.pos), *vel = reinterpret_cast<const v4 *>(o.vel),
             *tan = reinterpret_cast<const v4 *>(o.tan);
    v4 *sPos = reinterpret_cast<v4 *>(o.sPos), *sVel = reinterpret_cast<v4 *>(o.sVel),
       *sTan = reinterpret_cast<v4 *>(o.sTan);
    const uint32_t tid = threadIdx.x;
    uint32_t *out = o.ids + sb;
    {
        uint32_t k[G], id[G];
        v4 P[G], V[G], Tn[G];
#pragma unroll
        for (int j = 0; j < G; ++j) {
            k[j] = min(tid + (j0 + uint32_t(j)) * BLOCK, m - 1u);
            id[j] = lid[k[j]];
        }
#pragma unroll
        for (int j = 0; j < G; ++j) {
            if (POS) P[j] = pos[id[j]];
            if (VEL) V[j] = vel[id[j]];
            if (TAN) Tn[j] = tan[id[j]];
        }
#pragma unroll
        for (int j = 0; j < G; ++j) {
            out[k[j]] = id[j];
            o.cellOf[sb + k[j]] = o.cellBase + cell0 + uint32_t(lcell[k[j]]);
            if (POS) sPos[sb + k[j]] = P[j];
            if (VEL) sVel[sb + k[j]] = V[j];
            if (TAN) sTan[sb + k[j]] = Tn[j];
        }
    }
}

template <bool POS, bool VEL, bool TAN, int BLOCK, int CAP>
__device__ __forceinline__ void md_gather_staged(const CellsArgs &o, const uint32_t *lid, const uint16_t *lcell,
                                                 uint32_t sb, uint32_t cell0, uint32_t m) {
    // groups of two members, a group's gathers in flight together (groups of
    // four measured about 0.4 us slower per launch at 800k beads)
    constexpr int PER = CAP / BLOCK, G = PER < 2 ? PER : 2;
#pragma unroll
    for (int g = 0; g < PER; g += G)
        if (m > uint32_t(g) * BLOCK) md_gather_group<POS, VEL, TAN, BLOCK, G>(o, lid, lcell, sb, cell0, m, g);
}

// Bucket b of one system (see CellsArgs for the bases of a replica), BLOCK
// threads; a bucket of at most CAP members is staged in LDS, each thread
// holding at most PER = CAP / BLOCK of them.
// Ascending sort of N values in registers (N a power of two): Batcher's
// odd-even merge network (63 compare-exchanges for 16 against the bitonic
// network's 80), every compare-exchange at a compile-time position
// (branch-free min / max).
template <int N>
__device__ __forceinline__ void sort_network(uint32_t (&v)[N]) {
#pragma unroll
    for (int p = 1; p < N; p <<= 1) {
#pragma unroll
        for (int k = p; k >= 1; k >>= 1) {
#pragma unroll
            for (int j = k % p; j + k < N; j += 2 * k) {
#pragma unroll
                for (int i = 0; i < k; ++i) {
                    const int lo = i + j, hi = i + j + k;
                    // pairs inside one block of 2p (the merge of two sorted halves)
                    if (hi < N && lo / (2 * p) == hi / (2 * p)) {
                        const uint32_t a = v[lo], b = v[hi];
                        v[lo] = min(a, b);
                        v[hi] = max(a, b);
                    }
                }
            }
        }
    }
}

// The ids arr[a .. z) of one cell in ascending order (the order of a stable
// sort by cell).  Up to NET ids are loaded together into registers and sorted
// by the network (padded with 0xffffffff, which no id reaches); a longer cell
// takes an insertion sort in place.
template <int NET>
__device__ __forceinline__ void sort_cell_ids(uint32_t *arr, uint32_t a, uint32_t z) {
    const uint32_t m = z - a;
    if (m <= 1u) return;
    if (NET > 1 && m <= uint32_t(NET)) {
        uint32_t v[NET > 1 ? NET : 2];
#pragma unroll
        for (int k = 0; k < NET; ++k) v[k] = uint32_t(k) < m ? arr[a + k] : 0xffffffffu;
        sort_network<NET>(reinterpret_cast<uint32_t (&)[NET]>(v));
#pragma unroll
        for (int k = 0; k < NET; ++k)
            if (uint32_t(k) < m) arr[a + k] = v[k];
        return;
    }
    for (uint32_t u = a + 1; u < z; ++u) {
        const uint32_t x = arr[u];
        uint32_t w = u;
        while (w > a && arr[w - 1] > x) {
            arr[w] = arr[w - 1];
            --w;
        }
        arr[w] = x;
    }
}

// The mean velocity of the ids[a .. z) of one SRD cell (a < z), summed in
// that order, and with the isokinetic thermostat on its factor lambda, which
// scales the velocities before the mean (isokineticThermostatSRD and
// cellCenterOfMomentumSRD, particles_kernel_impl.cuh:559-637).  The first
// kSumHead ids are read, then their velocities loaded together and kept in
// registers for both sums (a cell holds ~6 particles); the rest are read
// kSumTail at a time in each sum.  Slots past z load the last particle
// again, unused.
__device__ __forceinline__ void srd_cell_sums(const CellsArgs &o, const uint32_t *ids, uint32_t a, uint32_t z,
                                              f3 &mom, float &lambda) {
    constexpr uint32_t kSumHead = 8, kSumTail = 4;
    uint32_t I[kSumHead];
#pragma unroll
    for (uint32_t k = 0; k < kSumHead; ++k) I[k] = ids[min(a + k, z - 1u)];
    float4 H[kSumHead];
#pragma unroll
    for (uint32_t k = 0; k < kSumHead; ++k) H[k] = o.srdVel[I[k]];
    if (o.iso) {
        float E = 0.0f;
#pragma unroll
        for (uint32_t k = 0; k < kSumHead; ++k) {
            if (a + k < z) {
                const f3 v = mk(H[k]);
                E += 0.5f * H[k].w * dot(v, v);
            }
        }
        for (uint32_t u0 = a + kSumHead; u0 < z; u0 += kSumTail) {
            uint32_t J[kSumTail];
#pragma unroll
            for (uint32_t k = 0; k < kSumTail; ++k) J[k] = ids[min(u0 + k, z - 1u)];
            float4 V[kSumTail];
#pragma unroll
            for (uint32_t k = 0; k < kSumTail; ++k) V[k] = o.srdVel[J[k]];
#pragma unroll
            for (uint32_t k = 0; k < kSumTail; ++k) {
                if (u0 + k < z) {
                    const f3 v = mk(V[k]);
                    E += 0.5f * V[k].w * dot(v, v);
                }
            }
        }
        lambda = sqrtf(o.kbT / E);
    }
    float M = 0.0f;
#pragma unroll
    for (uint32_t k = 0; k < kSumHead; ++k) {
        if (a + k < z) {
            float4 W = H[k];
            if (o.iso) W = f4(lambda * mk(W), W.w);
            mom += W.w * mk(W);
            M += W.w;
        }
    }
    for (uint32_t u0 = a + kSumHead; u0 < z; u0 += kSumTail) {
        uint32_t J[kSumTail];
#pragma unroll
        for (uint32_t k = 0; k < kSumTail; ++k) J[k] = ids[min(u0 + k, z - 1u)];
        float4 V[kSumTail];
#pragma unroll
        for (uint32_t k = 0; k < kSumTail; ++k) V[k] = o.srdVel[J[k]];
#pragma unroll
        for (uint32_t k = 0; k < kSumTail; ++k) {
            if (u0 + k < z) {
                float4 W = V[k];
                if (o.iso) W = f4(lambda * mk(W), W.w);
                mom += W.w * mk(W);
                M += W.w;
            }
        }
    }
    if (M > 0.0f) mom = mom * (1.0f / M);
}

// The tile of each of a staged bucket's members (members are in tile order):
// tileOf holds, cleared elsewhere, the tile index at the first slot of every
// nonempty segment (written by the segment scan); a block-wide running
// maximum over the BLOCK * PER slots, PER consecutive slots per thread, fills
// the rest.  A segment can hold every member (a tile of particles in cell
// order), so no thread fills one segment alone.  Entries at the member count
// and above are left undefined.
template <int BLOCK, int PER>
__device__ __forceinline__ void fill_tile_of(uint16_t *tileOf, uint32_t *wsum) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
    uint32_t v[PER];
    uint32_t run = 0;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        run = max(run, uint32_t(tileOf[tid * PER + j]));
        v[j] = run;
    }
    const uint32_t incl = wave_scan_max(run);
    if (lane == 63u) wsum[wid] = incl;
    __syncthreads();
    uint32_t pre = __shfl_up(incl, 1, 64);
    if (lane == 0u) pre = 0u;
    for (uint32_t w = 0; w < wid; ++w) pre = max(pre, wsum[w]);
#pragma unroll
    for (int j = 0; j < PER; ++j) tileOf[tid * PER + j] = uint16_t(max(v[j], pre));
    __syncthreads();
}

template <bool SRD, int BLOCK, int CAP = int(kBucketCap)>
__device__ __forceinline__ void bucket_cells_body(const CellsArgs &o, const BucketPlan &bp, uint32_t b,
                                                  unsigned char *smem) {
    constexpr int PER = CAP / BLOCK;
    static_assert(PER * BLOCK == CAP, "the staged members are PER per thread");
    const CellsLayout L = cells_layout(bp, SRD, CAP);
    uint32_t *wsum = reinterpret_cast<uint32_t *>(smem + L.wsum);
    uint32_t *cnt = reinterpret_cast<uint32_t *>(smem + L.cnt);
    uint32_t *cur = reinterpret_cast<uint32_t *>(smem + L.cur);
    uint32_t *lid = reinterpret_cast<uint32_t *>(smem + L.lid);
    uint16_t *lcell = reinterpret_cast<uint16_t *>(smem + L.lcell);
    uint32_t *segStart = reinterpret_cast<uint32_t *>(smem + L.segStart);
    uint32_t *segBase = reinterpret_cast<uint32_t *>(smem + L.segBase);
    uint16_t *tileOf = reinterpret_cast<uint16_t *>(smem + L.tileOf);

    const uint32_t B1 = bp.buckets + 1;
    const uint32_t tid = threadIdx.x, bd = blockDim.x;
    const uint32_t cell0 = b << bp.shift;
    const uint32_t nc = min(1u << bp.shift, bp.cells - cell0);
    // cleared here, under the barriers of the segment pass: the cell counts
    // and the member tile marks (tileOf, CAP entries over lid)
    for (uint32_t c = tid; c < nc; c += bd) cnt[c] = 0u;
    for (uint32_t k = tid; k < uint32_t(CAP) / 2u; k += bd) reinterpret_cast<uint32_t *>(tileOf)[k] = 0u;

    // segments of this bucket in every tile; first sorted slot.  With tile
    // ranges, only the tiles whose members reach this bucket have their row
    // read: a tile wholly below it contributes all its members to the slot,
    // one wholly above none.
    uint32_t part = 0;
    for (uint32_t t = tid; t < bp.tiles; t += bd) {
        uint32_t lo = 0u, len = 0u;
        bool read = true;
        if (o.tileRange) {
            const uint32_t r0 = o.tileRange[2u * t], r1 = o.tileRange[2u * t + 1u];
            read = b >= r0 && b <= r1;
            if (b > r1) lo = min(bp.tileItems, bp.n - t * bp.tileItems);
        }
        if (read) {
            const uint32_t *row = o.P + uint64_t(t) * B1 + b;
            lo = row[0];
            len = row[1] - lo;
        }
        segBase[t] = lo;
        segStart[t] = len;
        part += lo;
    }
    part = wave_scan_add(part);
    if ((tid & 63u) == 63u) wsum[BLOCK / 64 + (tid >> 6)] = part;  // upper half of wsum
    __syncthreads();
    uint32_t sb = 0;
    for (uint32_t w = 0; w < BLOCK / 64; ++w) sb += wsum[BLOCK / 64 + w];
    block_excl_scan<false, true>(segStart, bp.tiles, wsum, nullptr, tileOf, uint32_t(CAP));
    const uint32_t m = segStart[bp.tiles];

    const bool staged = m <= uint32_t(CAP);
    const uint64_t T = bp.tileItems;
    // a staged bucket's members (cell in the bucket, id), held by their
    // threads from the count to the placement (member k = tid + j * BLOCK)
    uint2 q[PER];
    if (staged && m) {
        // the tile of every member
        fill_tile_of<BLOCK, PER>(tileOf, wsum);
        // every load of the thread's members in flight before the first use
        // (indices past the last member load the last member again, unused)
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const uint32_t k = min(tid + uint32_t(j) * BLOCK, m - 1u);
            const uint32_t t = tileOf[k];
            q[j] = unpack_pair(o.pairs[uint64_t(t) * T + segBase[t] + (k - segStart[t])], o.idBase + t * uint32_t(T));
        }
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const uint32_t k = tid + uint32_t(j) * BLOCK;
            if (k < m) atomicAdd(&cnt[q[j].x], 1u);
        }
    } else if (!staged) {
        for (uint32_t k = tid; k < m; k += bd) {
            const uint32_t t = seg_of(segStart, bp.tiles, k);
            atomicAdd(&cnt[o.pairs[uint64_t(t) * T + segBase[t] + (k - segStart[t])] >> kPairIdBits], 1u);
        }
    }
    __syncthreads();
    block_excl_scan<true>(cnt, nc, wsum, cur);

    uint32_t *out = o.ids + sb;
    if (staged) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const uint32_t k = tid + uint32_t(j) * BLOCK;
            if (k < m) {
                const uint32_t lc = q[j].x;
                const uint32_t slot = atomicAdd(&cur[lc], 1u);
                lid[slot] = q[j].y;
                if (!SRD) lcell[slot] = uint16_t(lc);
            }
        }
    } else {
        for (uint32_t k = tid; k < m; k += bd) {
            const uint32_t t = seg_of(segStart, bp.tiles, k);
            const uint2 q =
                unpack_pair(o.pairs[uint64_t(t) * T + segBase[t] + (k - segStart[t])], o.idBase + t * uint32_t(T));
            out[atomicAdd(&cur[q.x], 1u)] = q.y;
        }
    }
    __syncthreads();
    // SRD, one system: the cells' rotation signs, one Philox block per four
    // cells (uniform_at's draws) computed once -- by the first wave for a
    // bucket of up to 256 cells -- into cur, free after the placement
    if (SRD && !o.srdReps) {
        const uint32_t b0 = cell0 >> 2, nb = ((cell0 + nc - 1u) >> 2) - b0 + 1u;
        for (uint32_t t = tid; t < nb; t += bd) {
            const u4 r = philox_block(o.seed, o.offset + b0 + t);
            const uint32_t w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t c = ((b0 + t) << 2) + uint32_t(k);
                if (c >= cell0 && c < cell0 + nc) cur[c - cell0] = u01(w[k]) > 0.5f ? 1u : 0u;
            }
        }
    }
    // every cell in ascending id order; CSR offsets.  SRD cells hold ~6
    // particles (a network of 16 takes all but the rare crowded ones); MD
    // cells hold ~1 bead.  The two storages take separate calls so that the
    // staged ids are read as LDS, not through a generic pointer.
    for (uint32_t c = tid; c < nc; c += bd) {
        const uint32_t a = cnt[c], z = cnt[c + 1];
        if (staged) sort_cell_ids<SRD ? 16 : 1>(lid, a, z);
        else sort_cell_ids<SRD ? 16 : 1>(out, a, z);
        o.off[cell0 + c] = o.slotBase + sb + a;
    }
    if (tid == 0 && cell0 + nc == bp.cells) o.off[bp.cells] = o.slotBase + bp.n;
    __syncthreads();

    if (!SRD) {
        if (staged) {
            if (!o.sPos) md_gather_staged<false, false, false, BLOCK, CAP>(o, lid, lcell, sb, cell0, m);
            else if (o.sVel && o.sTan) md_gather_staged<true, true, true, BLOCK, CAP>(o, lid, lcell, sb, cell0, m);
            else if (o.sVel) md_gather_staged<true, true, false, BLOCK, CAP>(o, lid, lcell, sb, cell0, m);
            else if (o.sTan) md_gather_staged<true, false, true, BLOCK, CAP>(o, lid, lcell, sb, cell0, m);
            else md_gather_staged<true, false, false, BLOCK, CAP>(o, lid, lcell, sb, cell0, m);
            return;
        }
        for (uint32_t k = tid; k < m; k += bd) {
            const uint32_t id = out[k];
            o.cellOf[sb + k] = o.cellBase + cell0 + seg_of(cnt, nc, k);
            if (o.sPos) {
                o.sPos[sb + k] = o.pos[id];
                if (o.sVel) o.sVel[sb + k] = o.vel[id];
                if (o.sTan) o.sTan[sb + k] = o.tan[id];
            }
        }
        return;
    }

    // SRD: the sorted ids out.  The cell sums below read the velocities from
    // memory by id (with the solvent in cell order those reads are near
    // sequential, and without an LDS stage of them more buckets fit per CU).
    // No barrier after the store: the sums read the staged ids from LDS, and
    // an unstaged cell's ids from memory only by the thread that sorted them.
    if (staged && m)
        for (uint32_t k = tid; k < m; k += bd) out[k] = lid[k];
    // one thread per cell, sums in sorted order: the arithmetic of
    // isokineticThermostatSRD and cellCenterOfMomentumSRD
    // (particles_kernel_impl.cuh:559-637)
    for (uint32_t c = tid; c < nc; c += bd) {
        const uint32_t a = cnt[c], z = cnt[c + 1];
        f3 mom = mk(0.f, 0.f, 0.f);
        float lambda = 1.0f;
        if (a != z) {
            // one instance for both storages, through a generic pointer (two
            // instances specialised to LDS and global memory measured slower)
            srd_cell_sums(o, staged ? lid : out, a, z, mom, lambda);
        }
        const uint32_t cell = cell0 + c;
        bool up;
        if (o.srdReps) {
            const SrdRep &q = o.srdReps[srd_rep_of_cell(o.srdReps, o.srdR, cell)];
            const uint64_t offset = q.rngBase + uint64_t(o.collision) * q.cellBlocks;
            up = uniform_at(q.seed, offset, cell - q.cellOffset) > 0.5f;
        } else {
            up = cur[c] != 0u;
        }
        const float sgn = up ? 1.0f : -1.0f;
        o.com[cell] = f4(mom, sgn);
        if (o.iso) o.lam[cell] = lambda;
    }
}

template <bool SRD, int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(7, 8))) void k_bucket_cells(CellsArgs o, BucketPlan bp) {
    extern __shared__ __align__(16) unsigned char smem[];
    bucket_cells_body<SRD, BLOCK>(o, bp, xcd_contiguous(blockIdx.x, gridDim.x), smem);
}

// level 2 of the batched sort: every replica's buckets in one launch; the
// arrays are offset to the replica's first bead and cell, the bead ids in the
// pairs are global
template <int BLOCK, int CAP>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(7, 8))) void k_bucket_cells_batch(CellsArgs a, const SortRep *__restrict__ reps,
                                                              uint32_t R) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t v = xcd_contiguous(blockIdx.x, gridDim.x);
    const SortRep &r = reps[sort_rep_of(reps, R, v, false)];
    CellsArgs o = a;
    o.pairs = a.pairs + r.beadOffset;
    o.P = a.P + r.matOff;
    o.ids = a.ids + r.beadOffset;
    o.off = a.off + r.cellOffset;
    o.cellOf = a.cellOf + r.beadOffset;
    if (o.sPos) o.sPos = a.sPos + r.beadOffset;
    if (o.sVel) o.sVel = a.sVel + r.beadOffset;
    if (o.sTan) o.sTan = a.sTan + r.beadOffset;
    o.cellBase = r.cellOffset;
    o.slotBase = r.beadOffset;
    o.idBase = r.beadOffset;
    bucket_cells_body<false, BLOCK, CAP>(o, plan_of(r), v - r.bucketOff, smem);
}

// ---------------------------------------------------------------------------
// SRD rotation and streaming in particle order (collideSolventKernel
// :651-691, solvent_integrator :145-205); the isokinetic factor of the cell
// scales the velocity first, as isokineticThermostatSRD writes it back
// ---------------------------------------------------------------------------
constexpr int kRotatePer = 4;  // particles per thread, their loads in flight together

__global__ __launch_bounds__(kBlock) void k_srd_rotate(const uint32_t *__restrict__ hash,
                                                       const float4 *__restrict__ com, const float *__restrict__ lam,
                                                       float4 *__restrict__ pos, float4 *__restrict__ vel,
                                                       uint32_t i0, uint32_t NT, uint32_t N, float dt, float cosA,
                                                       float sinA, int iso, KP kp) {
    // particles i0 + blockIdx.x * kBlock * kRotatePer + j * kBlock + threadIdx.x;
    // past NT a thread loads particle NT - 1 again and stores nothing
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 *com4 = reinterpret_cast<const v4 *>(com);
    v4 *pos4 = reinterpret_cast<v4 *>(pos), *vel4 = reinterpret_cast<v4 *>(vel);
    const bool rotBeads = kp.e.srdCollideFilaments != 0;
    const uint32_t base = i0 + blockIdx.x * (kBlock * kRotatePer) + threadIdx.x;
    uint32_t c[kRotatePer];
    v4 V[kRotatePer], C[kRotatePer], Pp[kRotatePer];
    float l[kRotatePer];
#pragma unroll
    for (int j = 0; j < kRotatePer; ++j) c[j] = hash[min(base + uint32_t(j) * kBlock, NT - 1u)];
#pragma unroll
    for (int j = 0; j < kRotatePer; ++j) {
        const uint32_t i = min(base + uint32_t(j) * kBlock, NT - 1u);
        V[j] = vel4[i];
        Pp[j] = pos4[i];
        C[j] = com4[c[j]];
        l[j] = iso ? lam[c[j]] : 1.0f;
    }
#pragma unroll
    for (int j = 0; j < kRotatePer; ++j) {
        const uint32_t i = base + uint32_t(j) * kBlock;
        if (i >= NT) continue;
        const bool solvent = i >= N;
        if (!solvent && !rotBeads && !iso) continue;
        f3 v = mk(V[j].x, V[j].y, V[j].z);
        if (iso) v = l[j] * v;
        if (solvent || rotBeads) {
            const f3 u = mk(C[j].x, C[j].y, C[j].z);
            const float sn = C[j].w * sinA;
            const f3 d = v - u;
            const f3 rot = mk(d.x * cosA - d.y * sn, d.x * sn + d.y * cosA, d.z);
            v = u + rot;
        }
        if (solvent) {
            f3 r = mk(Pp[j].x, Pp[j].y, Pp[j].z);
            const f3 dr = dt * v;
            r += dr;
            apply_boundaries(kp.p, r, v, dr, 0.0f);
            pos4[i] = v4{r.x, r.y, r.z, Pp[j].w};
        }
        vel4[i] = v4{v.x, v.y, v.z, V[j].w};
    }
}

// ---------------------------------------------------------------------------
// batched SRD: particles of every replica in one launch, each with its
// replica's parameters (found by bisection: particles of a wavefront mostly
// share a replica)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t srd_rep_of_particle(const SrdRep *__restrict__ reps, uint32_t R, uint32_t i,
                                                        bool bead) {
    uint32_t lo = 0, hi = R;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((bead ? reps[mid].beadOffset : reps[mid].solvOffset) <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kBlock) void k_srd_hash_batch(const float4 *__restrict__ pos, uint32_t *__restrict__ hash,
                                                           const SrdRep *__restrict__ reps, uint32_t R, uint32_t Nb,
                                                           uint32_t NT, const float *__restrict__ shift) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NT) return;
    const uint32_t r = srd_rep_of_particle(reps, R, i, i < Nb);
    const SrdRep &q = reps[r];
    const f3 sh = mk(shift[3 * r], shift[3 * r + 1], shift[3 * r + 2]);
    hash[i] = q.cellOffset + srd_hash_at(q.kp.p, q.kp.s, sh, mk(load4(pos, i)));
}

__global__ __launch_bounds__(kBlock) void k_srd_rotate_batch(const uint32_t *__restrict__ hash,
                                                             const float4 *__restrict__ com, float4 *__restrict__ pos,
                                                             float4 *__restrict__ vel, const SrdRep *__restrict__ reps,
                                                             uint32_t R, uint32_t Nb, uint32_t NT, float dt,
                                                             float cosA, float sinA, int rotBeads) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NT) return;
    const bool solvent = i >= Nb;
    if (!solvent && !rotBeads) return;
    const float4 V = load4(vel, i), C = load4(com, hash[i]);
    // the arithmetic of k_srd_rotate
    f3 v = mk(V.x, V.y, V.z);
    const f3 u = mk(C.x, C.y, C.z);
    const float sn = C.w * sinA;
    const f3 d = v - u;
    const f3 rot = mk(d.x * cosA - d.y * sn, d.x * sn + d.y * cosA, d.z);
    v = u + rot;
    if (solvent) {
        const SimParams &p = reps[srd_rep_of_particle(reps, R, i, false)].kp.p;
        const float4 P = load4(pos, i);
        f3 r = mk(P.x, P.y, P.z);
        const f3 dr = dt * v;
        r += dr;
        apply_boundaries(p, r, v, dr, 0.0f);
        pos[i] = make_float4(r.x, r.y, r.z, P.w);
    }
    vel[i] = make_float4(v.x, v.y, v.z, V.w);
}

// ---------------------------------------------------------------------------
// scratch and launchers
// ---------------------------------------------------------------------------
template <typename T>
static hipError_t regrow(T *&p, uint64_t &cap, uint64_t need, uint64_t elem) {
    if (need <= cap) return hipSuccess;
    const uint64_t c = need + need / 8 + 1024;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const hipError_t e = hipMalloc(reinterpret_cast<void **>(&p), c * elem);
    if (e == hipSuccess) cap = c;
    return e;
}

hipError_t ensure_bucket_scratch(BucketScratch &s, const BucketPlan &bp) {
    hipError_t e;
    if (bp.n > s.cap_n) {
        const uint64_t c = uint64_t(bp.n) + bp.n / 8 + 1024;
        if (s.hash) (void)hipFree(s.hash);
        if (s.sortedCell) (void)hipFree(s.sortedCell);
        if (s.aux) (void)hipFree(s.aux);
        s.hash = s.sortedCell = s.aux = nullptr;
        s.cap_n = 0;
        if ((e = hipMalloc(&s.hash, c * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&s.sortedCell, c * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&s.aux, c * 4)) != hipSuccess) return e;
        s.cap_n = c;
    }
    if ((e = regrow(s.pairs, s.cap_pairs, uint64_t(bp.tiles) * bp.tileItems, 4)) != hipSuccess) return e;
    if ((e = regrow(s.P, s.cap_mat, uint64_t(bp.tiles) * (bp.buckets + 1), 4)) != hipSuccess) return e;
    if ((e = regrow(s.tileRange, s.cap_range, 2ull * bp.tiles, 4)) != hipSuccess) return e;
    if (uint64_t(bp.cells) + 2 > s.cap_cells) {
        const uint64_t c = uint64_t(bp.cells) + bp.cells / 8 + 1024;
        if (s.off) (void)hipFree(s.off);
        if (s.lam) (void)hipFree(s.lam);
        s.off = nullptr;
        s.lam = nullptr;
        s.cap_cells = 0;
        if ((e = hipMalloc(&s.off, (c + 2) * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&s.lam, c * 4)) != hipSuccess) return e;
        s.cap_cells = c;
    }
    return hipSuccess;
}

// LDS beyond 64 KiB is granted per kernel, once (outside any graph capture:
// the first launch of a size runs directly)
static hipError_t grant_lds(const void *fn, size_t lds) {
    static std::vector<std::pair<const void *, size_t>> granted;
    if (lds <= 65536) return hipSuccess;
    for (auto &g : granted)
        if (g.first == fn && g.second >= lds) return hipSuccess;
    const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e == hipSuccess) granted.emplace_back(fn, lds);
    return e;
}

template <bool FROM_POS, int ITEMS, int TB = kTileBlock>
static hipError_t launch_staged(const float4 *pos, uint32_t *hash, const BucketPlan &bp, BucketScratch &s,
                                const KP &kp, uint32_t *range, hipStream_t st) {
    static_assert(ITEMS * TB == 1024 || ITEMS * TB == 2048 || ITEMS * TB == 4096 || ITEMS * TB == 8192 ||
                      ITEMS * TB == 16384,
                  "tile size");
    const size_t lds = size_t(staged_stage_words(bp.buckets + 1, TB)) * sizeof(uint32_t) +
                       size_t(bp.tileItems) * sizeof(uint32_t);
    const void *fn = reinterpret_cast<const void *>(&k_tiles_staged<FROM_POS, ITEMS, TB>);
    hipError_t e = grant_lds(fn, lds);
    if (e != hipSuccess) return e;
    NW_LAUNCH((k_tiles_staged<FROM_POS, ITEMS, TB>), dim3(bp.tiles), dim3(TB), lds, st, pos, hash, s.pairs, s.P,
              range, bp, kp);
    return hipGetLastError();
}

// staged kernel when the tile is 4096 or 16384 particles and its pairs and
// histogram fit the 160 KiB of LDS; otherwise the two-pass kernel
// ranges: record every tile's bucket range and write only that part of its
// P row (the cells pass must then read them: o.tileRange)
static hipError_t launch_tiles(bool fromPos, const float4 *pos, uint32_t *hash, const BucketPlan &bp,
                               BucketScratch &s, const KP &kp, bool ranges, hipStream_t st) {
    if (bp.n == 0) return hipSuccess;
    uint32_t *range = ranges ? s.tileRange : nullptr;
    const size_t stagedLds = size_t(staged_stage_words(bp.buckets + 1, kTileBlock)) * 4 + size_t(bp.tileItems) * 4;
    if (stagedLds <= 160 * 1024) {
        // small tiles (more workgroups for a chip of 256 CUs) as 256 threads
        if (!fromPos && bp.tileItems == 1024) return launch_staged<false, 4, 256>(pos, hash, bp, s, kp, range, st);
        if (!fromPos && bp.tileItems == 2048) return launch_staged<false, 8, 256>(pos, hash, bp, s, kp, range, st);
        if (bp.tileItems == 4 * kTileBlock)
            return fromPos ? launch_staged<true, 4>(pos, hash, bp, s, kp, range, st)
                           : launch_staged<false, 4>(pos, hash, bp, s, kp, range, st);
        // the SRD tiles of 8192 as 512 threads x 16: three workgroups per CU
        // by LDS and registers against two of 1024 threads (SRD workload:
        // 708 tiles in one round, 35 -> 33 us)
        if (bp.tileItems == 8 * kTileBlock)
            return fromPos ? launch_staged<true, 16, 512>(pos, hash, bp, s, kp, range, st)
                           : launch_staged<false, 8>(pos, hash, bp, s, kp, range, st);
        if (bp.tileItems == 16 * kTileBlock)
            return fromPos ? launch_staged<true, 16>(pos, hash, bp, s, kp, range, st)
                           : launch_staged<false, 16>(pos, hash, bp, s, kp, range, st);
    }
    const size_t lds = (size_t(bp.buckets) + 1 + kTileBlock / 64) * sizeof(uint32_t);
    const void *fn = fromPos ? reinterpret_cast<const void *>(&k_tiles<true>)
                             : reinterpret_cast<const void *>(&k_tiles<false>);
    const hipError_t e = grant_lds(fn, lds);
    if (e != hipSuccess) return e;
    if (fromPos)
        NW_LAUNCH(k_tiles<true>, dim3(bp.tiles), dim3(kTileBlock), lds, st, pos, hash, s.pairs, s.P, range,
                  bp, kp);
    else
        NW_LAUNCH(k_tiles<false>, dim3(bp.tiles), dim3(kTileBlock), lds, st, pos, hash, s.pairs, s.P, range,
                  bp, kp);
    return hipGetLastError();
}

hipError_t launch_srd_tiles(const float4 *pos, uint32_t *hash, const BucketPlan &bp, BucketScratch &s,
                            const KP &kp, hipStream_t st) {
    return launch_tiles(true, pos, hash, bp, s, kp, true, st);
}

hipError_t launch_hash_tiles(const uint32_t *hash, const BucketPlan &bp, BucketScratch &s, hipStream_t st,
                             bool ranges) {
    return launch_tiles(false, nullptr, const_cast<uint32_t *>(hash), bp, s, KP{}, ranges, st);
}

// ---------------------------------------------------------------------------
// the batched sort
// ---------------------------------------------------------------------------
static size_t batch_tiles_lds(uint32_t maxBuckets) {
    return size_t(staged_stage_words(maxBuckets + 1, kBatchTileBlock)) * sizeof(uint32_t) +
           size_t(kBatchTileItems) * sizeof(uint32_t);
}

static CellsLayout batch_cells_layout(const BatchSortPlan &tot) {
    BucketPlan bp{};
    bp.shift = tot.maxShift;
    bp.tiles = tot.maxTiles;
    return cells_layout(bp, false, kBatchBucketCap);
}

const char *make_sort_reps(const uint32_t *beadOff, const uint32_t *cellOff, uint32_t R, SortRep *out,
                           BatchSortPlan &tot) {
    tot = BatchSortPlan{};
    for (uint32_t r = 0; r < R; ++r) {
        const uint32_t n = beadOff[r + 1] - beadOff[r], cells = cellOff[r + 1] - cellOff[r];
        if (cells == 0) return "a replica without cells";
        const BucketPlan bp = make_bucket_plan(n, cells, kBatchTileItems, kBatchMembers);
        const char *why = bucket_plan_problem(bp);
        if (*why) return why;
        SortRep &q = out[r];
        q.beadOffset = beadOff[r];
        q.n = n;
        q.cellOffset = cellOff[r];
        q.cells = cells;
        q.shift = bp.shift;
        q.buckets = bp.buckets;
        q.tiles = bp.tiles;
        q.tileOff = tot.tiles;
        q.bucketOff = tot.buckets;
        q.pad = 0;
        q.matOff = tot.matrix;
        tot.tiles += bp.tiles;
        tot.buckets += bp.buckets;
        tot.matrix += uint64_t(bp.tiles) * (bp.buckets + 1);
        tot.maxBuckets = std::max(tot.maxBuckets, bp.buckets);
        tot.maxShift = std::max(tot.maxShift, bp.shift);
        tot.maxTiles = std::max(tot.maxTiles, bp.tiles);
        tot.n = uint64_t(beadOff[r + 1]);
        tot.cells = uint64_t(cellOff[r + 1]);
    }
    if (batch_tiles_lds(tot.maxBuckets) > 160 * 1024) return "a replica with too many cells for the batched sort";
    if (batch_cells_layout(tot).total > 160 * 1024) return "a replica with too many particles for the batched sort";
    return "";
}

hipError_t ensure_batch_sort_scratch(BucketScratch &s, const BatchSortPlan &tot) {
    hipError_t e;
    if (tot.n > s.cap_n) {
        const uint64_t c = tot.n + tot.n / 8 + 1024;
        if (s.hash) (void)hipFree(s.hash);
        if (s.sortedCell) (void)hipFree(s.sortedCell);
        if (s.aux) (void)hipFree(s.aux);
        s.hash = s.sortedCell = s.aux = nullptr;
        s.cap_n = 0;
        if ((e = hipMalloc(&s.hash, c * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&s.sortedCell, c * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&s.aux, c * 4)) != hipSuccess) return e;
        s.cap_n = c;
    }
    if ((e = regrow(s.pairs, s.cap_pairs, tot.n, 4)) != hipSuccess) return e;
    if ((e = regrow(s.P, s.cap_mat, tot.matrix, 4)) != hipSuccess) return e;
    if (tot.cells + 2 > s.cap_cells) {
        const uint64_t c = tot.cells + tot.cells / 8 + 1024;
        if (s.off) (void)hipFree(s.off);
        if (s.lam) (void)hipFree(s.lam);
        s.off = nullptr;
        s.lam = nullptr;
        s.cap_cells = 0;
        if ((e = hipMalloc(&s.off, (c + 2) * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&s.lam, c * 4)) != hipSuccess) return e;
        s.cap_cells = c;
    }
    return hipSuccess;
}

hipError_t launch_hash_tiles_batch(const uint32_t *hash, const SortRep *reps, uint32_t R, const BatchSortPlan &tot,
                                   BucketScratch &s, hipStream_t st) {
    if (tot.tiles == 0) return hipSuccess;
    const size_t lds = batch_tiles_lds(tot.maxBuckets);
    const hipError_t e = grant_lds(reinterpret_cast<const void *>(&k_tiles_batch), lds);
    if (e != hipSuccess) return e;
    NW_LAUNCH(k_tiles_batch, dim3(tot.tiles), dim3(kBatchTileBlock), lds, st, hash, s.pairs, s.P, reps, R);
    return hipGetLastError();
}

hipError_t launch_md_cells_batch(const SortRep *reps, uint32_t R, const BatchSortPlan &tot, BucketScratch &s,
                                 uint32_t *ids, const float4 *pos, const float4 *vel, const float4 *tan, float4 *sPos,
                                 float4 *sVel, float4 *sTan, hipStream_t st) {
    if (tot.buckets == 0) return hipSuccess;
    CellsArgs o{};
    o.pairs = s.pairs;
    o.P = s.P;
    o.ids = ids;
    o.off = s.off;
    o.cellOf = s.sortedCell;
    o.pos = pos;
    o.vel = vel;
    o.tan = tan;
    o.sPos = sPos;
    o.sVel = sVel;
    o.sTan = sTan;
    const CellsLayout L = batch_cells_layout(tot);
    const void *fn = reinterpret_cast<const void *>(&k_bucket_cells_batch<kMdCellsBlock, int(kBatchBucketCap)>);
    const hipError_t e = grant_lds(fn, L.total);
    if (e != hipSuccess) return e;
    NW_LAUNCH((k_bucket_cells_batch<kMdCellsBlock, int(kBatchBucketCap)>), dim3(tot.buckets),
                       dim3(kMdCellsBlock), L.total, st, o, reps, R);
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_solvent_keys(const float4 *__restrict__ pos, uint32_t N, uint32_t Ns,
                                                         uint32_t *__restrict__ key, uint32_t *__restrict__ idx,
                                                         KP kp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Ns) return;
    key[i] = srd_hash_at(kp.p, kp.s, mk(0.f, 0.f, 0.f), mk(pos[N + i]));
    idx[i] = i;
}

__global__ __launch_bounds__(kBlock) void k_solvent_gather(const float4 *__restrict__ pos,
                                                           const float4 *__restrict__ vel,
                                                           const uint32_t *__restrict__ id,
                                                           const uint32_t *__restrict__ perm, uint32_t N, uint32_t Ns,
                                                           float4 *__restrict__ tp, float4 *__restrict__ tv,
                                                           uint32_t *__restrict__ tid) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= Ns) return;
    const uint32_t j = perm[k];
    tp[k] = pos[N + j];
    tv[k] = vel[N + j];
    tid[k] = id ? id[j] : j;
}

hipError_t launch_solvent_keys(const float4 *pos, uint32_t N, uint32_t Ns, uint32_t *key, uint32_t *idx,
                               const KP &kp, hipStream_t st) {
    if (Ns == 0) return hipSuccess;
    hipLaunchKernelGGL(k_solvent_keys, dim3(grid_for(Ns)), dim3(kBlock), 0, st, pos, N, Ns, key, idx, kp);
    return hipGetLastError();
}

hipError_t launch_solvent_gather(const float4 *pos, const float4 *vel, const uint32_t *id, const uint32_t *perm,
                                 uint32_t N, uint32_t Ns, float4 *tp, float4 *tv, uint32_t *tid, hipStream_t st) {
    if (Ns == 0) return hipSuccess;
    hipLaunchKernelGGL(k_solvent_gather, dim3(grid_for(Ns)), dim3(kBlock), 0, st, pos, vel, id, perm, N, Ns, tp, tv,
                       tid);
    return hipGetLastError();
}

// keys[k] = cell of slot k, vals[k] = the value of the particle in slot k
__global__ __launch_bounds__(kBlock) void k_sort_finish(uint32_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                        const uint32_t *__restrict__ pos,
                                                        const uint32_t *__restrict__ cell,
                                                        const uint32_t *__restrict__ valsIn, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    keys[k] = cell[k];
    vals[k] = valsIn[pos[k]];
}

hipError_t bucket_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t n, uint32_t bound, BucketScratch &s,
                             hipStream_t st) {
    if (n == 0) return hipSuccess;
    const BucketPlan bp = make_bucket_plan(n, bound, 0);
    if (*bucket_plan_problem(bp)) return hipErrorNotSupported;
    hipError_t e;
    if ((e = ensure_bucket_scratch(s, bp)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(s.aux, vals, size_t(n) * 4, hipMemcpyDeviceToDevice, st)) != hipSuccess) return e;
    if ((e = launch_hash_tiles(keys, bp, s, st)) != hipSuccess) return e;
    // sorted positions into s.hash (free here: the keys come from the caller)
    if ((e = launch_md_cells(bp, s, s.hash, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, st)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_sort_finish, dim3(grid_for(n)), dim3(kBlock), 0, st, keys, vals, s.hash, s.sortedCell, s.aux,
                       n);
    return hipGetLastError();
}

hipError_t launch_md_cells(const BucketPlan &bp, BucketScratch &s, uint32_t *ids, const float4 *pos,
                           const float4 *vel, const float4 *tan, float4 *sPos, float4 *sVel, float4 *sTan,
                           hipStream_t st) {
    if (bp.n == 0) return hipSuccess;
    CellsArgs o{};
    o.pairs = s.pairs;
    o.P = s.P;
    o.ids = ids;
    o.off = s.off;
    o.cellOf = s.sortedCell;
    o.pos = pos;
    o.vel = vel;
    o.tan = tan;
    o.sPos = sPos;
    o.sVel = sVel;
    o.sTan = sTan;
    const CellsLayout L = cells_layout(bp, false);
    NW_LAUNCH((k_bucket_cells<false, kMdCellsBlock>), dim3(bp.buckets), dim3(kMdCellsBlock), L.total, st, o,
                       bp);
    return hipGetLastError();
}

hipError_t launch_srd_cells(const BucketPlan &bp, BucketScratch &s, uint32_t *ids, const float4 *vel, float4 *com,
                            int iso, uint64_t seed, uint64_t offset, const KP &kp, hipStream_t st) {
    if (bp.n == 0) return hipSuccess;
    CellsArgs o{};
    o.pairs = s.pairs;
    o.P = s.P;
    o.tileRange = s.tileRange;
    o.ids = ids;
    o.off = s.off;
    o.cellOf = s.sortedCell;
    o.srdVel = vel;
    o.com = com;
    o.lam = s.lam;
    o.iso = iso;
    o.kbT = kp.s.kbT;
    o.seed = seed;
    o.offset = offset;
    const CellsLayout L = cells_layout(bp, true);
    NW_LAUNCH((k_bucket_cells<true, kSrdCellsBlock>), dim3(bp.buckets), dim3(kSrdCellsBlock), L.total, st,
                       o, bp);
    return hipGetLastError();
}

hipError_t launch_srd_cells_batch(const BucketPlan &bp, BucketScratch &s, uint32_t *ids, const float4 *vel,
                                  float4 *com, const SrdRep *reps, uint32_t R, uint32_t collision, hipStream_t st) {
    if (bp.n == 0) return hipSuccess;
    CellsArgs o{};
    o.pairs = s.pairs;
    o.P = s.P;
    o.tileRange = s.tileRange;
    o.ids = ids;
    o.off = s.off;
    o.cellOf = s.sortedCell;
    o.srdVel = vel;
    o.com = com;
    o.lam = s.lam;
    o.iso = 0;
    o.srdReps = reps;
    o.srdR = R;
    o.collision = collision;
    const CellsLayout L = cells_layout(bp, true);
    NW_LAUNCH((k_bucket_cells<true, kSrdCellsBlock>), dim3(bp.buckets), dim3(kSrdCellsBlock), L.total, st, o, bp);
    return hipGetLastError();
}

hipError_t launch_srd_hash_batch(const float4 *pos, uint32_t *hash, const SrdRep *reps, uint32_t R, uint32_t Nb,
                                 uint32_t NT, const float *shift, hipStream_t st) {
    if (NT == 0) return hipSuccess;
    hipLaunchKernelGGL(k_srd_hash_batch, dim3(grid_for(NT)), dim3(kBlock), 0, st, pos, hash, reps, R, Nb, NT, shift);
    return hipGetLastError();
}

hipError_t launch_srd_rotate_batch(const uint32_t *hash, const float4 *com, float4 *pos, float4 *vel,
                                   const SrdRep *reps, uint32_t R, uint32_t Nb, uint32_t NT, float dt, float cosA,
                                   float sinA, int rotBeads, hipStream_t st) {
    if (NT == 0) return hipSuccess;
    hipLaunchKernelGGL(k_srd_rotate_batch, dim3(grid_for(NT)), dim3(kBlock), 0, st, hash, com, pos, vel, reps, R, Nb,
                       NT, dt, cosA, sinA, rotBeads);
    return hipGetLastError();
}

hipError_t launch_srd_rotate(const uint32_t *hash, const float4 *com, const float *lam, float4 *pos, float4 *vel,
                             uint32_t NT, uint32_t N, float dt, float cosA, float sinA, int iso, const KP &kp,
                             hipStream_t st) {
    // beads that are neither rotated nor rescaled are not visited
    const uint32_t i0 = (kp.e.srdCollideFilaments != 0 || iso) ? 0u : min(N, NT);
    if (NT == i0) return hipSuccess;
    const uint32_t per = uint32_t(kBlock) * kRotatePer;
    NW_LAUNCH(k_srd_rotate, dim3((NT - i0 + per - 1) / per), dim3(kBlock), 0, st, hash, com, lam, pos, vel,
                       i0, NT, N, dt, cosA, sinA, iso, kp);
    return hipGetLastError();
}

}  // namespace nw

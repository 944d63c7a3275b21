// nw_bucket.hpp -- two-level cell sort, sorted SRD collision (nw_bucket.hip)
// and the fused MD kernels (nw_fused.hip).
#pragma once

#include "nw_internal.hpp"

namespace nw {

// Level 1 cuts the particles into tiles of tileItems consecutive ids; every
// tile counts its members per bucket (2^shift consecutive cells) and writes
// one row of the prefix matrix P[tile][bucket] (buckets + 1 entries: the
// exclusive prefix of its counts, the tile total last) plus its (cell, id)
// pairs grouped by bucket into its own region of the pair array.  Level 2
// runs one workgroup per bucket: the column sum of P[.][bucket] is the
// bucket's first output slot, the column differences locate its members.
struct BucketPlan {
    uint32_t n;          // particles
    uint32_t cells;      // cells of the grid
    uint32_t shift;      // log2(cells per bucket), <= 11
    uint32_t buckets;    // ceil(cells / 2^shift)
    uint32_t tileItems;  // particles per level-1 tile
    uint32_t tiles;      // ceil(n / tileItems)
};

constexpr uint32_t kBucketCap = 2048;     // bucket members staged in LDS (more: global-memory path)
constexpr uint32_t kMaxTiles = 4096;      // level-2 segment tables in LDS
constexpr uint32_t kMaxBuckets = 32767;   // level-1 LDS histogram (128 KiB of the 160 KiB)
constexpr uint64_t kMaxMatrix = 1ull << 24;  // entries of P

// A sort pair in 4 bytes: the particle's cell within its bucket (below
// 2^11, the largest bucket) in the high bits and its index within its
// level-1 tile (below 2^kPairIdBits) in the low bits; level 2 knows the
// bucket and the tile of every member and restores (cell in bucket, id).
constexpr uint32_t kPairIdBits = 21;
__host__ __device__ inline uint32_t pack_pair(uint32_t cellInBucket, uint32_t idInTile) {
    return (cellInBucket << kPairIdBits) | idInTile;
}
__host__ __device__ inline uint2 unpack_pair(uint32_t p, uint32_t tileFirstId) {
    return make_uint2(p >> kPairIdBits, tileFirstId + (p & ((1u << kPairIdBits) - 1u)));
}
constexpr int kMdCellsBlock = 256;        // level-2 workgroup, MD sort
constexpr int kSrdCellsBlock = 256;       // level-2 workgroup, SRD sort
constexpr int kCellsBlockMax = 1024;
constexpr int kPairBlock = 256;           // sorted beads per LDS pair-kernel workgroup

// tileItems == 0: chosen for the particle count and the cells
// members: the average bucket population aimed at (the LDS stage of a
// single system holds kBucketCap = 2048)
BucketPlan make_bucket_plan(uint64_t n, uint64_t cells, uint32_t tileItems, uint32_t members = 1536);
// empty string when the plan fits the limits above, else the reason
const char *bucket_plan_problem(const BucketPlan &bp);

// Launch shapes of the fused kernels (nwSetTuning); zero takes the built-in
// choice.  Every setting gives the same results bit for bit.
struct Tuning {
    uint32_t mdTileItems = 0;   // MD sort first level: beads per tile
    uint32_t mdMembers = 0;     // MD sort second level: bucket population aimed at
    uint32_t srdTileItems = 0;  // SRD sort first level: particles per tile
    uint32_t srdMembers = 0;    // SRD sort second level
    uint32_t pairBlock = 0;     // sorted beads per pair-kernel workgroup (128, 256, 512)
    uint32_t frontBlock = 0;    // threads per front workgroup (64 .. 1024)
    uint32_t pairStage = 0;     // beads staged per row run, on average: 2 pairBlock, else (0) pairBlock + 128
};
Tuning &tuning();
// the plan of the fused steps' cell sort (srd: the SRD collision's) with the
// tuning applied; a tuned plan that does not fit falls back to the built-in
BucketPlan fused_plan(uint64_t n, uint64_t cells, bool srd);

struct BucketScratch {
    uint32_t *hash = nullptr;        // cell per particle, particle order
    uint32_t *pairs = nullptr;       // packed (cell in bucket, index in tile), per tile grouped by bucket
    uint32_t *P = nullptr;           // tiles x (buckets + 1)
    uint32_t *tileRange = nullptr;   // per tile the first and last bucket of its members (NULL: full rows)
    uint32_t *sortedCell = nullptr;  // cell of each sorted slot
    uint32_t *off = nullptr;         // CSR cell offsets (cells + 1)
    float *lam = nullptr;            // isokinetic factor per SRD cell
    uint32_t *aux = nullptr;         // per particle: the values of a key-value sort
    uint64_t cap_n = 0, cap_pairs = 0, cap_mat = 0, cap_cells = 0, cap_range = 0;
};

hipError_t ensure_bucket_scratch(BucketScratch &s, const BucketPlan &bp);

struct MdFrontArgs {
    float4 *pos, *vel, *force, *forceOld, *tangent;
    const float4 *posIn, *velIn;  // the step's input state when not pos / vel (NULL: in place)
    uint32_t *hash;               // out: the cell of every bead
    float dt, lgvStd;
    int langevin;
    // nonzero: the new force goes to forceOld's buffer and force's buffer,
    // left as it is, becomes the old force (no reversal; the caller swaps)
    int swapForces;
    uint32_t W;
    uint64_t seed, offRev, offLgv;
};

// what one lane of the fused front produces for its bead
struct FrontLane {
    f3 r, v, acc, fold, tg;
    float pw, vw, fow;
    bool writeTan;
};

// integrate + reversal + Langevin + bonded + tangent + MD cell (a.tangent
// NULL: no tangents stored)
hipError_t launch_md_front(const MdFrontArgs &a, const KP &kp, hipStream_t st);
// the tangents the front stores, from the positions (worms of 1..64 beads)
hipError_t launch_filament_tangents(const float4 *pos, float4 *tangent, uint32_t W, const KP &kp, hipStream_t st);
// level 1 over precomputed cells
// ranges: also record every tile's bucket range (a cells pass reading
// o.tileRange must follow; the SRD cells launchers do)
hipError_t launch_hash_tiles(const uint32_t *hash, const BucketPlan &bp, BucketScratch &s, hipStream_t st,
                             bool ranges = false);
// stable sort of (keys, vals) by key, every key < bound (the per-call
// sortParticles after calcHash): the two-level bucket sort of the positions,
// then keys from the slots' cells and vals gathered by position; a key >=
// bound (never written by calcHash) sorts as bound - 1.
// hipErrorNotSupported when the bound does not fit the bucket sort.
hipError_t bucket_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t n, uint32_t bound, BucketScratch &s,
                             hipStream_t st);
// level 2 for MD: ids in stable cell order, CSR offsets, cell of every
// slot, gathered records (sVel / sTan may be NULL)
// the solvent's SRD cells on the unshifted grid (key) and indices (idx) of
// the Ns solvent particles after the N beads
hipError_t launch_solvent_keys(const float4 *pos, uint32_t N, uint32_t Ns, uint32_t *key, uint32_t *idx,
                               const KP &kp, hipStream_t st);
// tp / tv / tid[k] = the solvent particle perm[k]'s position, velocity and
// id (id NULL: perm[k] itself)
hipError_t launch_solvent_gather(const float4 *pos, const float4 *vel, const uint32_t *id, const uint32_t *perm,
                                 uint32_t N, uint32_t Ns, float4 *tp, float4 *tv, uint32_t *tid, hipStream_t st);
hipError_t launch_md_cells(const BucketPlan &bp, BucketScratch &s, uint32_t *ids, const float4 *pos,
                           const float4 *vel, const float4 *tan, float4 *sPos, float4 *sVel, float4 *sTan,
                           hipStream_t st);
// SRD: cell of every particle + level 1
hipError_t launch_srd_tiles(const float4 *pos, uint32_t *hash, const BucketPlan &bp, BucketScratch &s,
                            const KP &kp, hipStream_t st);
// SRD level 2: ids in stable cell order, per cell (centre-of-mass velocity,
// rotation sign) into com and the isokinetic factor into s.lam
hipError_t launch_srd_cells(const BucketPlan &bp, BucketScratch &s, uint32_t *ids, const float4 *vel, float4 *com,
                            int iso, uint64_t seed, uint64_t offset, const KP &kp, hipStream_t st);
// SRD rotation of every particle about its cell's momentum, solvent streaming
hipError_t launch_srd_rotate(const uint32_t *hash, const float4 *com, const float *lam, float4 *pos, float4 *vel,
                             uint32_t NT, uint32_t N, float dt, float cosA, float sinA, int iso, const KP &kp,
                             hipStream_t st);
// pair pass over CSR offsets; sortedCell may be NULL (cells recomputed from positions)
hipError_t launch_collide_off(float4 *force, const float4 *sPos, const float4 *sVel, const float4 *sTan,
                              const uint32_t *sIdx, const uint32_t *off, const uint32_t *sortedCell, uint32_t n,
                              const KP &kp, hipStream_t st);

// ---------------------------------------------------------------------------
// batched replicas (the fused MD step over many independent systems).  A
// device table of R records is followed by the R records' waveOffset values,
// then their blockOffset values, packed (replica_of_wave / replica_of_block
// read those); nwBatchSetReplicas builds it so.
// ---------------------------------------------------------------------------
struct BatchRep {
    KP kp;                 // the replica's parameters (box, grid, worms, force constants)
    uint32_t beadOffset;   // first bead in the concatenated arrays
    uint32_t cellOffset;   // first cell in the concatenated cell space
    uint32_t waveOffset;   // first wavefront of the front kernel
    uint32_t blockOffset;  // first workgroup of the pair kernel
    uint64_t seed;         // Philox key
    uint64_t rngBase;      // generator offset at step 0 of the batch
    uint64_t rngStride;    // blocks drawn per step (reversal + Langevin)
    uint32_t revBlocks;    // reversal blocks per step (0 without reversal)
    float lgvStd;          // sqrt(2 gamma kT) of the Langevin noise
    float lgvGamma;        // the Langevin gamma
    int32_t langevin;
    uint32_t numBeads;     // numFilaments x filamentSize
};

// the replica of a front wavefront / pair workgroup of a batched launch.
// The device table holds the R records, then their wavefront offsets, then
// their workgroup offsets, packed (ascending): the bisection reads 4 bytes per
// replica that stay in the scalar cache rather than one record's line per step.
__device__ __forceinline__ uint32_t replica_of_offset(const uint32_t *__restrict__ off, uint32_t R, uint32_t v) {
    uint32_t lo = 0, hi = R;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= v) lo = mid;
        else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ uint32_t replica_of_wave(const BatchRep *__restrict__ reps, uint32_t R, uint32_t wave) {
    return replica_of_offset(reinterpret_cast<const uint32_t *>(reps + R), R, wave);
}
__device__ __forceinline__ uint32_t replica_of_block(const BatchRep *__restrict__ reps, uint32_t R, uint32_t blk) {
    return replica_of_offset(reinterpret_cast<const uint32_t *>(reps + R) + R, R, blk);
}

// The batched sort: every replica sorted against its own plan (tiles of
// kBatchTileItems of its beads, buckets of its own cells, its own prefix
// matrix) with one launch per level for all replicas.  Replica r's pairs sit
// at its bead offset, its matrix at matOff; tiles and buckets of all
// replicas are numbered consecutively (tileOff, bucketOff).
constexpr uint32_t kBatchTileItems = 4096;
// the batched level 2 stages at most 1024 members per bucket and aims at 512
// (a smaller LDS stage: more buckets in flight per CU)
constexpr uint32_t kBatchBucketCap = 1024, kBatchMembers = 512;
struct SortRep {
    uint32_t beadOffset, n, cellOffset, cells;
    uint32_t shift, buckets, tiles, tileOff, bucketOff, pad;
    uint64_t matOff;
};
struct BatchSortPlan {
    uint32_t tiles = 0, buckets = 0, maxBuckets = 0, maxShift = 0, maxTiles = 0;
    uint64_t matrix = 0, n = 0, cells = 0;
};
// host: the replicas' plans from their bead and cell offsets (R + 1 entries
// each); returns "" or what does not fit
const char *make_sort_reps(const uint32_t *beadOff, const uint32_t *cellOff, uint32_t R, SortRep *out,
                           BatchSortPlan &tot);
hipError_t ensure_batch_sort_scratch(BucketScratch &s, const BatchSortPlan &tot);
hipError_t launch_hash_tiles_batch(const uint32_t *hash, const SortRep *reps, uint32_t R, const BatchSortPlan &tot,
                                   BucketScratch &s, hipStream_t st);
hipError_t launch_md_cells_batch(const SortRep *reps, uint32_t R, const BatchSortPlan &tot, BucketScratch &s,
                                 uint32_t *ids, const float4 *pos, const float4 *vel, const float4 *tan, float4 *sPos,
                                 float4 *sVel, float4 *sTan, hipStream_t st);

// SRD collisions of batched replicas: one cell sort over every replica's
// particles (all beads of the batch, then every replica's solvent) and the
// concatenated SRD cells; per replica its parameters, layout and Philox
// stream (collision c draws its rotation signs at rngBase + c cellBlocks)
struct SrdRep {
    KP kp;                             // SimParams (box, boundaries) and SolventParams (SRD grid)
    uint32_t beadOffset, numBeads;     // its beads: particles beadOffset ..
    uint32_t solvOffset, numSolvent;   // its solvent: particles solvOffset .. (after all beads)
    uint32_t cellOffset, cells;        // its SRD cells in the concatenated cell space
    uint64_t seed, rngBase, cellBlocks;
};
// cell of every particle (Nb beads, then the solvent; NT in all) with the
// collision's grid shift of replica r at shift[3 r ..]
hipError_t launch_srd_hash_batch(const float4 *pos, uint32_t *hash, const SrdRep *reps, uint32_t R, uint32_t Nb,
                                 uint32_t NT, const float *shift, hipStream_t st);
// level 2 of the batched SRD sort: launch_srd_cells with every cell's sign
// from its replica's stream at collision c (no isokinetic rescaling)
hipError_t launch_srd_cells_batch(const BucketPlan &bp, BucketScratch &s, uint32_t *ids, const float4 *vel,
                                  float4 *com, const SrdRep *reps, uint32_t R, uint32_t collision, hipStream_t st);
// rotation about the cell momentum (beads when rotBeads) and the solvent's
// streaming in its replica's box
hipError_t launch_srd_rotate_batch(const uint32_t *hash, const float4 *com, float4 *pos, float4 *vel,
                                   const SrdRep *reps, uint32_t R, uint32_t Nb, uint32_t NT, float dt, float cosA,
                                   float sinA, int rotBeads, hipStream_t st);

struct BatchArrays {
    float4 *pos, *vel, *force, *forceOld, *tangent;
    // nonzero: the new force goes to forceOld's buffer (pairs add there too)
    // and force's buffer is left as the old force (no replica reverses)
    int swapForces;
    uint32_t *hash;   // global cell of every bead
    uint32_t *ids;    // bead ids in cell order
    uint32_t *off;    // CSR offsets over the concatenated cells
    uint32_t *sCell;  // cell of every sorted slot
    float4 *sPos, *sVel, *sTan;
};

// exchanges the contents of two float4 arrays of n records
hipError_t launch_swap4(float4 *a, float4 *b, uint64_t n, hipStream_t st);
// sets the device step counter of a batch (replaying a stretch of steps)
hipError_t launch_set_step(uint64_t *ctr, uint64_t v, hipStream_t st);
hipError_t launch_md_front_batch(const BatchArrays &a, const BatchRep *reps, uint32_t R, uint32_t waves,
                                 const uint64_t *stepCtr, float dt, hipStream_t st);
// the last kernel of a batched step: advances *stepCtr
hipError_t launch_collide_batch(const BatchArrays &a, const BatchRep *reps, uint32_t R, uint32_t blocks, int mode,
                                bool tan, bool unit, uint64_t *stepCtr, hipStream_t st);

// ---------------------------------------------------------------------------
// block-level helpers shared by the level-1 kernels (whole workgroup calls)
// ---------------------------------------------------------------------------
// in-place exclusive scan of a[0..m) over the workgroup, a[m] = total;
// wsum holds blockDim/64 words; ends with a barrier.  COPY: copy[i] also
// receives the prefix of i.  MARK: marks[prefix] = i for every nonzero a[i]
// whose prefix is below markCap (the first slot of each nonempty segment).
template <bool COPY = false, bool MARK = false>
__device__ __forceinline__ void block_excl_scan(uint32_t *a, uint32_t m, uint32_t *wsum, uint32_t *copy = nullptr,
                                                uint16_t *marks = nullptr, uint32_t markCap = 0u) {
    const uint32_t T = blockDim.x;
    const uint32_t per = (m + T - 1) / T;
    const uint32_t b0 = threadIdx.x * per;
    uint32_t s = 0;
    for (uint32_t k = 0; k < per; ++k)
        if (b0 + k < m) s += a[b0 + k];
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    const uint32_t incl = wave_scan_add(s);
    if (lane == 63u) wsum[wid] = incl;
    __syncthreads();
    uint32_t run = incl - s;
    for (uint32_t w = 0; w < wid; ++w) run += wsum[w];
    for (uint32_t k = 0; k < per; ++k) {
        if (b0 + k < m) {
            const uint32_t v = a[b0 + k];
            a[b0 + k] = run;
            if (COPY) copy[b0 + k] = run;
            if (MARK && v != 0u && run < markCap) marks[run] = uint16_t(b0 + k);
            run += v;
        }
    }
    if (threadIdx.x == T - 1) {
        uint32_t tot = 0;
        for (uint32_t w = 0; w < T / 64; ++w) tot += wsum[w];
        a[m] = tot;
    }
    __syncthreads();
}

}  // namespace nw

// nw_fused.hip -- the fused MD step kernels (gfx950).
//
// The reference issues ~10 launches per MD step, several of them tiny; on
// MI355X each dispatch costs microseconds of fixed time, and the reorder
// gathers re-read what the sort just touched.  The fused step is
//   k_md_front      integrate + reversal + Langevin + bonded + tangent + cell
//                   hash (whole worms per wavefront, chain neighbours by
//                   cross-lane shuffles)
//   cell sort       nw_bucket.hip (stable, atomic-free, gathers the records)
//   k_collide_rows  pair forces, neighbour rows staged in LDS, one flattened
//                   candidate list per bead (k_collide_csr without LDS for
//                   stencils over several z layers)
// computing the per-call path's expressions in the same order, except that
// pair terms are accumulated one by one rather than per cell.
#include "nw_bucket.hpp"
#include "nw_physics.hpp"

namespace nw {

// cell coordinates of sorted bead i: from the sort's cell ids when given
// (already wrapped into the grid), else from the position
__device__ __forceinline__ void bead_cell(const KP &kp, const uint32_t *__restrict__ sCell, uint32_t i, f3 r, int &gx,
                                          int &gy, int &gz) {
    const SimParams &p = kp.p;
    if (sCell) {
        cell_xyz(kp, sCell[i], gx, gy, gz);
    } else {
        gx = wrap_mod(cell_coord(r.x, p.origin.x, p.cellSize.x), p.gridSize.x);
        gy = wrap_mod(cell_coord(r.y, p.origin.y, p.cellSize.y), p.gridSize.y);
        gz = wrap_mod(cell_coord(r.z, p.origin.z, p.cellSize.z), p.gridSize.z);
    }
}

// ===========================================================================
// MD front
// ===========================================================================
// One bead's share of the front: velocity Verlet, polarity reversal, Langevin
// and the bonded terms on the new positions (all lanes of the wave call it:
// chain neighbours and mirror beads arrive by shuffles).  Arrays are those
// of the bead's system (a replica's slice in the batched step).
__device__ __forceinline__ FrontLane md_front_lane(const KP &kp, const WaveWorm &m, const float4 *__restrict__ pos,
                                                   const float4 *__restrict__ vel, const float4 *__restrict__ force,
                                                   const float4 *__restrict__ forceOld, float dt, float lgvStd,
                                                   int langevin, uint64_t seed, uint64_t offRev, uint64_t offLgv,
                                                   bool wantTan) {
    const SimParams &p = kp.p;
    const uint32_t n = p.filamentSize;
    // Every lane loads its four records as 16-byte vectors (a lane past the
    // wave's worms reads bead 0, in bounds) and takes zeros by select.  A
    // load under the lane's condition was split into one 4-byte load per
    // component, each behind its own branch; the empty statements keep the
    // four loads unconditional, here.
    typedef float v4 __attribute__((ext_vector_type(4)));
    const uint32_t b = m.active ? m.bead : 0u;
    const v4 Pl = reinterpret_cast<const v4 *>(pos)[b], Vl = reinterpret_cast<const v4 *>(vel)[b];
    const v4 Fl = reinterpret_cast<const v4 *>(force)[b], FOl = reinterpret_cast<const v4 *>(forceOld)[b];
    // the random draws depend on the bead alone: formed while the loads are
    // in flight
    const bool rev = p.reverseProbability > 0.0f && m.active &&
                     !(uniform_at(seed, offRev, m.worm) > p.reverseProbability);
    f3 xi = mk(0.f, 0.f, 0.f);
    if (langevin) {
        const u4 rb = philox_block(seed, offLgv + m.bead);
        float z0, z1, z2, z3;
        box_muller(rb.x, rb.y, z0, z1);
        box_muller(rb.z, rb.w, z2, z3);
        (void)z3;
        xi = mk(z0 * lgvStd + 0.0f, z1 * lgvStd + 0.0f, z2 * lgvStd + 0.0f);
    }
    asm volatile("" ::"v"(Pl), "v"(Vl), "v"(Fl), "v"(FOl));
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 P = m.active ? make_float4(Pl.x, Pl.y, Pl.z, Pl.w) : zero4;
    const float4 V = m.active ? make_float4(Vl.x, Vl.y, Vl.z, Vl.w) : zero4;
    const float4 F = m.active ? make_float4(Fl.x, Fl.y, Fl.z, Fl.w) : zero4;
    const float4 FO = m.active ? make_float4(FOl.x, FOl.y, FOl.z, FOl.w) : zero4;

    // velocity Verlet (filament_integrator, particles_kernel_impl.cuh:71-142)
    FrontLane o;
    f3 r = mk(P), v = mk(V);
    const f3 ff = mk(F), fo = mk(FO);
    const f3 dv = 0.5f * (ff + fo) * dt;
    v += dv;
    const f3 dr = v * dt + 0.5f * ff * dt * dt;
    r += dr;
    apply_boundaries(p, r, v, dr, p.particleRadius);
    f3 fold = ff;
    float pw = P.w, vw = V.w, fow = FO.w;

    // polarity reversal (reverseFilamentsKernel :820-857): bead p <-> n-1-p
    // of the worm, taken from the mirror lane; the force is zero here
    if (p.reverseProbability > 0.0f) {
        const int src = rev ? int(m.wl * n + (n - 1u - m.pp)) : int(m.lane);
        r = mk(__shfl(r.x, src, 64), __shfl(r.y, src, 64), __shfl(r.z, src, 64));
        v = mk(__shfl(v.x, src, 64), __shfl(v.y, src, 64), __shfl(v.z, src, 64));
        fold = mk(__shfl(fold.x, src, 64), __shfl(fold.y, src, 64), __shfl(fold.z, src, 64));
        pw = __shfl(pw, src, 64);
        vw = __shfl(vw, src, 64);
        fow = __shfl(fow, src, 64);
    }

    // Langevin (langevinKernel :859-874), the force starts from zero
    f3 acc = mk(0.f, 0.f, 0.f);
    if (langevin) {
        f3 f = mk(0.f, 0.f, 0.f);
        f -= p.gamma * v;
        f += xi;
        acc = mk(0.0f + f.x, 0.0f + f.y, 0.0f + f.z);
    }

    // bonded forces and (wantTan) tangents on the new positions
    o.writeTan = bonded_lane(kp, r, m.pp, acc, o.tg, wantTan);
    o.r = r;
    o.v = v;
    o.acc = acc;
    o.fold = fold;
    o.pw = pw;
    o.vw = vw;
    o.fow = fow;
    return o;
}

// stores of an active lane; returns the bead's cell (calcHashD).  swap: the
// new force into forceOld's buffer; force's buffer already holds the old
// force (no reversal moved it)
__device__ __forceinline__ uint32_t md_front_store(const KP &kp, const WaveWorm &m, const FrontLane &o,
                                                   float4 *__restrict__ pos, float4 *__restrict__ vel,
                                                   float4 *__restrict__ force, float4 *__restrict__ forceOld,
                                                   float4 *__restrict__ tangent, bool swap = false) {
    const SimParams &p = kp.p;
    pos[m.bead] = f4(o.r, o.pw);
    vel[m.bead] = f4(o.v, o.vw);
    if (swap) {
        forceOld[m.bead] = f4(o.acc, 0.0f);
    } else {
        force[m.bead] = f4(o.acc, 0.0f);
        forceOld[m.bead] = f4(o.fold, o.fow);
    }
    if (o.writeTan && tangent) tangent[m.bead] = f4(o.tg, 0.0f);
    // one layer: the wrap puts every z in layer 0, so z is not computed
    return md_hash_wrapped(p, cell_coord(o.r.x, p.origin.x, p.cellSize.x), cell_coord(o.r.y, p.origin.y, p.cellSize.y),
                           p.gridSize.z > 1u ? cell_coord(o.r.z, p.origin.z, p.cellSize.z) : 0);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_md_front(MdFrontArgs a, KP kp) {
    const WaveWorm m = wave_worm(kp.p.filamentSize, a.W);
    const FrontLane o = md_front_lane(kp, m, a.posIn ? a.posIn : a.pos, a.velIn ? a.velIn : a.vel, a.force,
                                      a.forceOld, a.dt, a.lgvStd, a.langevin, a.seed, a.offRev, a.offLgv,
                                      a.tangent != nullptr);
    if (m.active)
        a.hash[m.bead] = md_front_store(kp, m, o, a.pos, a.vel, a.force, a.forceOld, a.tangent, a.swapForces != 0);
}

// The tangents of the worms' beads from their positions, as the front's
// bonded_lane writes them (the same arithmetic on the same values: bit for
// bit), for a caller whose fused steps skipped them.
__global__ __launch_bounds__(kBlock) void k_filament_tangents(const float4 *__restrict__ pos,
                                                              float4 *__restrict__ tangent, uint32_t W, KP kp) {
    const WaveWorm m = wave_worm(kp.p.filamentSize, W);
    const f3 r = m.active ? mk(pos[m.bead]) : mk(0.f, 0.f, 0.f);
    f3 acc = mk(0.f, 0.f, 0.f), tg = mk(0.f, 0.f, 0.f);
    const bool write = bonded_lane(kp, r, m.pp, acc, tg);
    if (m.active && write) tangent[m.bead] = f4(tg, 0.0f);
}

// ===========================================================================
// pair forces over CSR cell offsets
//
// Candidates are visited in stencil order (z rows, y rows, x cells, each
// cell's beads in sorted order) and every term goes straight into the bead's
// accumulator, multiplied by the stencil weight of its row (a cell counted
// k times by the reference's folded stencil contributes k x term; weight 1
// leaves the term unchanged).  The global and the LDS kernel below follow this
// order exactly and agree bit for bit.
// ===========================================================================
template <int MODE, bool TAN>
struct Owner {
    f3 r, v, t;
    uint32_t fil, ch;
};

// one candidate's terms (collideParticles / bond / WCA, then the extensile
// term), skipped for the bead itself and, in WCA mode, for its own worm
template <int MODE, bool TAN, bool UNIT = false>
__device__ __forceinline__ void add_pair(const KP &kp, const Owner<MODE, TAN> &o, float rc2, float w, f3 &acc,
                                         f3 r2, uint32_t fil2, uint32_t ch2, f3 v2, f3 t2) {
    const SimParams &p = kp.p;
    const bool same = o.fil == fil2;
    if (MODE == 0) {
        if (same && ch2 == o.ch) return;
        f3 t;
        if (same && ((o.ch + 1u == ch2) || (o.ch - 1u == ch2)))
            t = bond_hookean(p, o.r, r2, p.bondSpringK, p.bondSpringL);
        else
            t = dem_force(p, o.r, r2, o.v, v2);
        acc += UNIT ? t : w * t;
    } else {
        if (same) return;
        const f3 t = wca_force(p, kp.e, o.r, r2, rc2);
        acc += UNIT ? t : w * t;
    }
    if (TAN && !same) {
        const f3 t = extensile_force(p, o.r, r2, o.t, t2, p.activity);
        acc += UNIT ? t : w * t;
    }
}

// beadBase: the first bead id of the system (a replica's offset in the
// batched step; worm and chain indices are the system's own)
template <int MODE, bool TAN>
__device__ __forceinline__ void add_run_global(const KP &kp, const Owner<MODE, TAN> &o, float rc2, float w, f3 &acc,
                                               uint32_t a, uint32_t b, const float4 *__restrict__ sPos,
                                               const float4 *__restrict__ sVel, const float4 *__restrict__ sTan,
                                               const uint32_t *__restrict__ sIdx, uint32_t beadBase) {
    for (uint32_t j = a; j < b; ++j) {
        uint32_t fil2, ch2;
        worm_of(kp, sIdx[j] - beadBase, fil2, ch2);
        add_pair<MODE, TAN>(kp, o, rc2, w, acc, mk(sPos[j]), fil2, ch2, MODE == 0 ? mk(sVel[j]) : mk(0.f, 0.f, 0.f),
                            TAN ? mk(sTan[j]) : mk(0.f, 0.f, 0.f));
    }
}

// sorted-slot runs of the three x cells around cx in the row starting at
// cell `base`: one run, or two across the x wrap (Gx >= 3)
__device__ __forceinline__ void row_runs(const uint32_t *__restrict__ off, uint32_t base, int cx, uint32_t Gx,
                                         uint32_t &a0, uint32_t &a1, uint32_t &b0, uint32_t &b1) {
    const bool mid = cx >= 1 && cx + 1 < int(Gx);
    const bool left = cx == 0;
    const uint32_t ia0 = mid ? base + uint32_t(cx) - 1u : (left ? base + Gx - 1u : base + Gx - 2u);
    const uint32_t ia1 = mid ? base + uint32_t(cx) + 2u : base + Gx;
    a0 = off[ia0];
    a1 = off[ia1];
    b0 = 0u;
    b1 = 0u;
    if (!mid) {
        b0 = off[base];
        b1 = off[left ? base + 2u : base + 1u];
    }
}

// The candidates of a one-layer stencil (Gx >= 3) from global memory: the
// row pieces (three rows, one or two x runs each, in stencil order) first,
// then the flattened candidate list four at a time with the loads of the
// four issued together.  The same candidates, order and arithmetic as the
// row-by-row walk below.
constexpr int kGlobalBatch = 4;

template <int MODE, bool TAN>
__device__ __forceinline__ void add_layer_global(const KP &kp, const Owner<MODE, TAN> &o, float rc2, f3 &acc,
                                                 int gx, const AxisCells &ay, uint32_t zc, float wz,
                                                 const uint32_t *__restrict__ off, const float4 *__restrict__ sPos,
                                                 const float4 *__restrict__ sVel, const float4 *__restrict__ sTan,
                                                 const uint32_t *__restrict__ sIdx, uint32_t beadBase) {
    const SimParams &p = kp.p;
    const uint32_t Gx = p.gridSize.x;
    uint32_t ps[6], pl[6];  // piece start slot and length
    float pw[6];
#pragma unroll
    for (int yi = 0; yi < 3; ++yi) {
        ps[2 * yi] = pl[2 * yi] = ps[2 * yi + 1] = pl[2 * yi + 1] = 0u;
        pw[2 * yi] = pw[2 * yi + 1] = 0.0f;
        if (yi < ay.n) {
            uint32_t a0, a1, b0, b1;
            row_runs(off, (zc * p.gridSize.y + uint32_t(ay.c[yi])) * Gx, gx, Gx, a0, a1, b0, b1);
            ps[2 * yi] = a0;
            pl[2 * yi] = a1 - a0;
            ps[2 * yi + 1] = b0;
            pl[2 * yi + 1] = b1 - b0;
            pw[2 * yi] = pw[2 * yi + 1] = wz * ay.w[yi];
        }
    }
    uint32_t pre[7];  // first flattened index of each piece
    pre[0] = 0u;
#pragma unroll
    for (int k = 0; k < 6; ++k) pre[k + 1] = pre[k] + pl[k];
    const uint32_t K = pre[6];
    for (uint32_t t0 = 0; t0 < K; t0 += kGlobalBatch) {
        uint32_t I[kGlobalBatch];
        float4 P[kGlobalBatch], V[kGlobalBatch], T[kGlobalBatch];
        float W[kGlobalBatch];
#pragma unroll
        for (int u = 0; u < kGlobalBatch; ++u) {
            const uint32_t t = min(t0 + uint32_t(u), K - 1u);
            uint32_t j = 0u;
            float w = 0.0f;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                if (t >= pre[k] && t < pre[k + 1]) {
                    j = ps[k] + (t - pre[k]);
                    w = pw[k];
                }
            }
            I[u] = sIdx[j];
            P[u] = sPos[j];
            if (MODE == 0) V[u] = sVel[j];
            if (TAN) T[u] = sTan[j];
            W[u] = w;
        }
#pragma unroll
        for (int u = 0; u < kGlobalBatch; ++u) {
            if (t0 + uint32_t(u) < K) {
                uint32_t fil2, ch2;
                worm_of(kp, I[u] - beadBase, fil2, ch2);
                add_pair<MODE, TAN>(kp, o, rc2, W[u], acc, mk(P[u]), fil2, ch2,
                                    MODE == 0 ? mk(V[u]) : mk(0.f, 0.f, 0.f), TAN ? mk(T[u]) : mk(0.f, 0.f, 0.f));
            }
        }
    }
}

// every candidate of the owner from global memory, in stencil order (off:
// the CSR offsets of the system's cells, indexed by its own cell ids)
template <int MODE, bool TAN>
__device__ __forceinline__ void add_all_global(const KP &kp, const Owner<MODE, TAN> &o, float rc2, f3 &acc, int gx,
                                               int gy, int gz, const uint32_t *__restrict__ off,
                                               const float4 *__restrict__ sPos, const float4 *__restrict__ sVel,
                                               const float4 *__restrict__ sTan, const uint32_t *__restrict__ sIdx,
                                               uint32_t beadBase = 0) {
    const SimParams &p = kp.p;
    const uint32_t Gx = p.gridSize.x;
    const AxisCells ay = axis_cells_in(gy, p.gridSize.y, false);
    const AxisCells az = axis_cells_in(gz, p.gridSize.z, kp.e.stencil == NW_STENCIL_2D);
    if (az.n == 1 && Gx >= 3) {
        add_layer_global<MODE, TAN>(kp, o, rc2, acc, gx, ay, uint32_t(az.c[0]), az.w[0], off, sPos, sVel, sTan, sIdx,
                                    beadBase);
        return;
    }
#pragma unroll
    for (int zi = 0; zi < 3; ++zi) {
        if (zi >= az.n) break;
#pragma unroll
        for (int yi = 0; yi < 3; ++yi) {
            if (yi >= ay.n) break;
            const float wzy = az.w[zi] * ay.w[yi];
            const uint32_t base = (uint32_t(az.c[zi]) * p.gridSize.y + uint32_t(ay.c[yi])) * Gx;
            if (Gx >= 3) {
                uint32_t a0, a1, b0, b1;
                row_runs(off, base, gx, Gx, a0, a1, b0, b1);
                add_run_global<MODE, TAN>(kp, o, rc2, wzy, acc, a0, a1, sPos, sVel, sTan, sIdx, beadBase);
                add_run_global<MODE, TAN>(kp, o, rc2, wzy, acc, b0, b1, sPos, sVel, sTan, sIdx, beadBase);
            } else {
                const AxisCells ax = axis_cells_in(gx, Gx, false);
#pragma unroll
                for (int xi = 0; xi < 3; ++xi) {
                    if (xi >= ax.n) break;
                    const uint32_t h = base + uint32_t(ax.c[xi]);
                    add_run_global<MODE, TAN>(kp, o, rc2, wzy * ax.w[xi], acc, off[h], off[h + 1], sPos, sVel,
                                              sTan, sIdx, beadBase);
                }
            }
        }
    }
}

template <int MODE, bool TAN>
__device__ __forceinline__ Owner<MODE, TAN> load_owner(const KP &kp, uint32_t i, uint32_t me,
                                                       const float4 *__restrict__ sPos,
                                                       const float4 *__restrict__ sVel,
                                                       const float4 *__restrict__ sTan) {
    Owner<MODE, TAN> o;
    o.r = mk(sPos[i]);
    o.v = MODE == 0 ? mk(sVel[i]) : mk(0.f, 0.f, 0.f);
    o.t = TAN ? mk(sTan[i]) : mk(0.f, 0.f, 0.f);
    worm_of(kp, me, o.fil, o.ch);
    return o;
}

__device__ __forceinline__ void add_force(float4 *__restrict__ force, uint32_t me, f3 acc) {
    const float4 Fo = force[me];
    force[me] = make_float4(Fo.x + acc.x, Fo.y + acc.y, Fo.z + acc.z, Fo.w + 0.0f);
}

template <int MODE, bool TAN>
__global__ __launch_bounds__(kBlock) void k_collide_csr(float4 *__restrict__ force, const float4 *__restrict__ sPos,
                                                        const float4 *__restrict__ sVel,
                                                        const float4 *__restrict__ sTan,
                                                        const uint32_t *__restrict__ sIdx,
                                                        const uint32_t *__restrict__ off,
                                                        const uint32_t *__restrict__ sCell, uint32_t n, KP kp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t me = sIdx[i];
    const Owner<MODE, TAN> o = load_owner<MODE, TAN>(kp, i, me, sPos, sVel, sTan);
    int gx, gy, gz;
    bead_cell(kp, sCell, i, o.r, gx, gy, gz);
    f3 acc = mk(0.f, 0.f, 0.f);
    add_all_global<MODE, TAN>(kp, o, kp.e.wcaCutoff * kp.e.wcaCutoff, acc, gx, gy, gz, off, sPos, sVel, sTan, sIdx);
    add_force(force, me, acc);
}

// ===========================================================================
// pair forces with LDS row runs (stencils of one z layer, gridSize.x >= 3,
// cells of the sorted beads given)
//
// A workgroup owns BLOCK (128, 256 or 512; 256 unless tuned) consecutive
// sorted beads whose cells span the hash interval [h0, h1].  Every neighbour
// cell of those beads that does not wrap lies in one of three hash intervals
// [h0 - 1 + dy Gx, h1 + 1 + dy Gx], dy = -1, 0, 1, i.e. in three contiguous
// runs of the sorted arrays, staged back to back in one LDS pool with
// coalesced loads (position + worm id in one float4; velocity + chain index,
// tangent when the terms read them), beside the two cells an edge bead needs
// across the x wrap.  A bead whose stencil lies in the stage walks it as one
// flattened candidate list, so the lanes of a wavefront stay busy until the
// longest list ends rather than the longest run of every row; any other bead
// (a y wrap, an overfull stage) takes the global-memory walk.  Same order and
// arithmetic as k_collide_csr.
// ===========================================================================
// average staged beads per row run of a workgroup of BLOCK sorted beads (the
// pool holds three times as many; the batched kernel's stage; a single
// system's is tuned, pairStage)
template <int BLOCK>
constexpr int run_cap() {
    return BLOCK + 128;
}

// beads staged per wrap cell (a cell at the other end of a row)
constexpr int kWrapCap = 8;

// LDS of one pair-kernel workgroup: the three row runs back to back in a pool
// of 3 CAP records, then the two wrap cells (kWrapCap records each)
template <int MODE, bool TAN, int CAP>
struct RowsLds {
    static constexpr int kPool = 3 * CAP;
    static constexpr int kWrap0 = kPool;  // first slot of the wrap cells
    static constexpr int kSlots = kPool + 2 * kWrapCap;
    float4 pos[kSlots];
    float4 vel[MODE == 0 ? kSlots : 1];
    float4 tan[TAN ? kSlots : 1];
};

// One candidate of the LDS walk.  The WCA term alone is evaluated for every
// candidate and masked afterwards (no branch): a skipped one (the bead's own worm, or at
// or beyond the cutoff, or a non-finite position) adds +0, as the branching
// form does, which leaves the accumulator unchanged -- it starts at +0, and
// an FP32 sum is -0 only when both of its terms are.
template <int MODE, bool TAN, bool UNIT>
__device__ __forceinline__ void lds_pair(const KP &kp, const Owner<MODE, TAN> &o, float rc2, float wt, f3 &acc,
                                         const float4 &P, const float4 &V, const float4 &T) {
    if (MODE == 1 && !TAN) {
        const f3 rel = min_image_pair(kp.p, o.r, mk(P));
        const float rr = dot(rel, rel);
        const bool hit = __float_as_uint(P.w) != o.fil && rr < rc2;
        const f3 f = -wca_mag(kp.e, rr) * rel;
        const uint32_t m = hit ? 0xffffffffu : 0u;
        const f3 t = mk(__uint_as_float(__float_as_uint(f.x) & m), __uint_as_float(__float_as_uint(f.y) & m),
                        __uint_as_float(__float_as_uint(f.z) & m));
        acc += UNIT ? t : wt * t;
    } else {
        add_pair<MODE, TAN, UNIT>(kp, o, rc2, wt, acc, mk(P), __float_as_uint(P.w),
                                  MODE == 0 ? __float_as_uint(V.w) : 0u, MODE == 0 ? mk(V) : mk(0.f, 0.f, 0.f),
                                  TAN ? mk(T) : mk(0.f, 0.f, 0.f));
    }
}

// The walk over a bead's candidate list: NP pieces of consecutive LDS slots
// in stencil order, piece k = pieces[k STRIDE] of the given arrays (n slots
// from q), the pieces of stencil row r of weight wr.  Flattened index t lies in piece k for pre[k] <= t <
// pre[k + 1], at slot t + d[k].  Each record is read from LDS where it is
// evaluated: reading the next one ahead kept a second copy of the records
// live and moved it every step, slower than the wait the other waves cover.
template <int MODE, bool TAN, bool UNIT, int NP, int STRIDE, int CAP>
__device__ __forceinline__ void walk_lds(const KP &kp, const Owner<MODE, TAN> &o, float rc2, f3 &acc,
                                         const uint32_t (&q)[6], const uint32_t (&n)[6], float w0, float w1,
                                         float w2, const RowsLds<MODE, TAN, CAP> &L) {
    uint32_t pre[NP + 1], d[NP];
    pre[0] = 0u;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        pre[k + 1] = pre[k] + n[k * STRIDE];
        d[k] = q[k * STRIDE] - pre[k];  // mod 2^32
    }
    const uint32_t K = pre[NP];
    for (uint32_t t = 0; t < K; ++t) {
        // stencil row r holds pieces [r NP / 3, (r + 1) NP / 3)
        const float wt = UNIT ? 1.0f : (t < pre[NP / 3] ? w0 : (t < pre[2 * NP / 3] ? w1 : w2));
        uint32_t s = d[NP - 1];
#pragma unroll
        for (int k = NP - 2; k >= 0; --k) s = t < pre[k + 1] ? d[k] : s;
        s += t;
        const float4 Pc = L.pos[s];
        const float4 Vc = MODE == 0 ? L.vel[s] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 Tc = TAN ? L.tan[s] : make_float4(0.f, 0.f, 0.f, 0.f);
        lds_pair<MODE, TAN, UNIT>(kp, o, rc2, wt, acc, Pc, Vc, Tc);
    }
}

// The workgroup's sorted beads [i0, min(i0 + blockDim, iEnd)) of one system
// whose beads and cells start at beadBase / cellBase in the shared arrays
// (zero for a single system); offR = off + cellBase.
template <int MODE, bool TAN, bool UNIT, int BLOCK, int CAP = run_cap<BLOCK>()>
__device__ __forceinline__ void collide_rows_tile(const KP &kp, uint32_t i0, uint32_t iEnd, uint32_t beadBase,
                                                  uint32_t cellBase, float4 *__restrict__ force,
                                                  const float4 *__restrict__ sPos, const float4 *__restrict__ sVel,
                                                  const float4 *__restrict__ sTan, const uint32_t *__restrict__ sIdx,
                                                  const uint32_t *__restrict__ offR,
                                                  const uint32_t *__restrict__ sCell,
                                                  RowsLds<MODE, TAN, CAP> &L) {
    using Lds = RowsLds<MODE, TAN, CAP>;
    // records move as plain 4-vectors (register values)
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 *pos4 = reinterpret_cast<const v4 *>(sPos), *vel4 = reinterpret_cast<const v4 *>(sVel),
             *tan4 = reinterpret_cast<const v4 *>(sTan);
    const SimParams &p = kp.p;
    const uint32_t Gx = p.gridSize.x, Gy = p.gridSize.y;
    const uint32_t layer = Gx * Gy;
    const uint32_t C = layer * p.gridSize.z;
    const uint32_t i = i0 + threadIdx.x;
    const bool own = i < iEnd;
    const uint32_t ic = own ? i : iEnd - 1u;  // lanes past the end read the last bead, unused

    // the bead's own sorted records, in flight with the row bounds
    const uint32_t me = sIdx[ic];
    const uint32_t cOwn = sCell[ic] - cellBase;
    const v4 Pown = pos4[ic];
    const v4 Vown = MODE == 0 ? vel4[ic] : v4{0.f, 0.f, 0.f, 0.f};
    const v4 Town = TAN ? tan4[ic] : v4{0.f, 0.f, 0.f, 0.f};

    // The three row runs (the same for the whole workgroup: scalar loads),
    // back to back in the pool: run r, sorted slots [rs[r], rs[r] + rl[r]),
    // at LDS slots [rq[r], rq[r] + rl[r]); a run that would overflow the pool
    // is not staged (rl = 0).
    uint32_t rs[3], rl[3], rq[3];
    // A bead in the last cell of a row (x = Gx - 1) also needs cell x = 0 of
    // its three stencil rows, one in the first cell (x = 0) cell x = Gx - 1.
    // Two of those three cells lie in the row runs, and the third as well
    // except for the tile's first right-edge cell (cell 0 of the row above
    // it) and its last left-edge cell (cell Gx - 1 of the row below it; rows
    // wrap within their layer).  Those two cells are staged too: sorted slots
    // [ws[m], ws[m] + wl[m]) at LDS slots kWrap0 + m kWrapCap (wl = 0: not
    // staged).
    uint32_t ws[2] = {0u, 0u}, wl[2] = {0u, 0u};
    uint32_t used = 0u;
    {
        const uint32_t last = min(i0 + blockDim.x, iEnd) - 1u;
        const uint32_t h0 = sCell[i0] - cellBase, h1 = sCell[last] - cellBase;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            // the hash interval [h0 - 1 + (r - 1) Gx, h1 + 2 + (r - 1) Gx) within [0, C] (C < 2^31)
            const int lo = int(h0) - 1 + (r - 1) * int(Gx), hi = int(h1) + 2 + (r - 1) * int(Gx);
            const uint32_t clo = uint32_t(min(max(lo, 0), int(C))), chi = uint32_t(min(max(hi, 0), int(C)));
            const uint32_t s = offR[clo], e = offR[chi];
            const uint32_t n = e - s;
            rs[r] = s;
            rl[r] = used + n <= uint32_t(Lds::kPool) ? n : 0u;
            rq[r] = used;
            used += rl[r];
        }
        int x0, y0, z0, x1, y1, z1;
        cell_xyz(kp, h0, x0, y0, z0);
        cell_xyz(kp, h1, x1, y1, z1);
        // the first cell x = Gx - 1 at or after h0 (in h0's row)
        if (h0 + (Gx - 1u - uint32_t(x0)) <= h1) {
            const uint32_t y = uint32_t(y0) == 0u ? Gy - 1u : uint32_t(y0) - 1u;
            const uint32_t c = (uint32_t(z0) * Gy + y) * Gx;
            const uint32_t s = offR[c], e = offR[c + 1u];
            ws[0] = s;
            wl[0] = e - s <= uint32_t(kWrapCap) ? e - s : 0u;
        }
        // the last cell x = 0 at or before h1 (in h1's row)
        if (h1 - uint32_t(x1) >= h0) {
            const uint32_t y = uint32_t(y1) + 1u == Gy ? 0u : uint32_t(y1) + 1u;
            const uint32_t c = (uint32_t(z1) * Gy + y) * Gx + Gx - 1u;
            const uint32_t s = offR[c], e = offR[c + 1u];
            ws[1] = s;
            wl[1] = e - s <= uint32_t(kWrapCap) ? e - s : 0u;
        }
    }

    // in flight together: the bounds of the bead's three rows, the staged
    // records, then the old force (issued last: loads retire in issue order,
    // so the waits for the records before the barrier leave it in flight)
    int gx, gy, gzi;
    cell_xyz(kp, cOwn, gx, gy, gzi);
    const uint32_t gz = uint32_t(gzi);
    const AxisCells ay = axis_cells_in(gy, Gy, false);
    const float wz = kp.e.stencil == NW_STENCIL_2D ? 1.0f : axis_cells_in(gzi, p.gridSize.z, false).w[0];
    const bool inner = gx >= 1 && gx + 1 < int(Gx);
    // an inner bead's three stencil rows (cells gx - 1 .. gx + 1), read by
    // every lane from in-bounds cells (an edge bead's are unused: its six
    // pieces are loaded after the barrier, few beads, fewer registers held
    // across it)
    const bool edgeOk = !inner && ay.n == 3;
    const uint32_t gxc = inner ? uint32_t(gx) : 1u;
    uint32_t ra[3], rb[3];
#pragma unroll
    for (int yi = 0; yi < 3; ++yi) {
        // rows past ay.n hold the centre row (in bounds, unused): every row
        // is loaded, none copied from another's load (a copy would wait for
        // that load, and for every load issued before it)
        const uint32_t base = (gz * Gy + uint32_t(ay.c[yi])) * Gx;
        ra[yi] = offR[base + gxc - 1u];
        rb[yi] = offR[base + gxc + 2u];
    }
    // The staged records: thread x takes pool slots x + h BLOCK below
    // `used`, and in the last round threads kWT.. take the wrap cells' records
    // (past the pool).  Loads of unused records read bead 0 (in bounds).
    constexpr int RH = (Lds::kPool + 2 * kWrapCap + BLOCK - 1) / BLOCK;
    constexpr int kWT = BLOCK - 2 * kWrapCap;
    static_assert((RH - 1) * BLOCK + kWT >= Lds::kPool, "wrap records past the pool");
    v4 SP[RH], SV[RH], ST[RH];
    uint32_t SI[RH], SD[RH];
#pragma unroll
    for (int h = 0; h < RH; ++h) {
        const uint32_t k = threadIdx.x + uint32_t(h) * BLOCK;
        uint32_t j = k < rq[1] ? rs[0] + k : (k < rq[2] ? rs[1] + (k - rq[1]) : rs[2] + (k - rq[2]));
        bool st = k < used;
        uint32_t dst = k;
        if (h == RH - 1) {
            const uint32_t wk = threadIdx.x - uint32_t(kWT), c = wk % uint32_t(kWrapCap);
            const bool wt = threadIdx.x >= uint32_t(kWT), m1 = wk >= uint32_t(kWrapCap);
            const bool wv = wt && c < (m1 ? wl[1] : wl[0]);
            j = wv ? (m1 ? ws[1] : ws[0]) + c : j;
            st = wt ? wv : st;
            dst = wt ? uint32_t(Lds::kWrap0) + wk : dst;
        }
        j = st ? j : 0u;
        SD[h] = st ? dst : ~0u;
        SP[h] = pos4[j];
        SI[h] = sIdx[j];
        if (MODE == 0) SV[h] = vel4[j];
        if (TAN) ST[h] = tan4[j];
    }
    const v4 Fo = reinterpret_cast<const v4 *>(force)[me];
    // The records are stored below under a branch per record; these empty
    // statements use them here, so their loads stay above, issued together
    // (the compiler would otherwise move a record's loads into its branch,
    // to issue there after every other load and wait for them all)
#pragma unroll
    for (int h = 0; h < RH; ++h) {
        asm volatile("" ::"v"(SP[h]), "v"(SI[h]));
        if constexpr (MODE == 0) asm volatile("" ::"v"(SV[h]));
        if constexpr (TAN) asm volatile("" ::"v"(ST[h]));
    }
#pragma unroll
    for (int h = 0; h < RH; ++h) {
        const uint32_t dst = SD[h];
        if (dst != ~0u) {
            uint32_t fil, ch;
            worm_of(kp, SI[h] - beadBase, fil, ch);
            L.pos[dst] = make_float4(SP[h].x, SP[h].y, SP[h].z, __uint_as_float(fil));
            if (MODE == 0) L.vel[dst] = make_float4(SV[h].x, SV[h].y, SV[h].z, __uint_as_float(ch));
            if (TAN) L.tan[dst] = make_float4(ST[h].x, ST[h].y, ST[h].z, ST[h].w);
        }
    }
    __syncthreads();
    if (!own) return;

    Owner<MODE, TAN> o;
    o.r = mk(Pown.x, Pown.y, Pown.z);
    o.v = mk(Vown.x, Vown.y, Vown.z);
    o.t = mk(Town.x, Town.y, Town.z);
    worm_of(kp, me - beadBase, o.fil, o.ch);
    const float rc2 = kp.e.wcaCutoff * kp.e.wcaCutoff;

    // The bead's candidate pieces in stencil order, two per stencil row: an
    // inner bead's row is one piece (cells gx - 1 .. gx + 1), an edge bead's
    // two (x = 0: cell Gx - 1, then cells 0, 1; x = Gx - 1: cells Gx - 2,
    // Gx - 1, then cell 0).  A bead with a piece outside the stage takes the
    // global walk.
    uint32_t pq[6] = {0u, 0u, 0u, 0u, 0u, 0u}, pn[6] = {0u, 0u, 0u, 0u, 0u, 0u};
    float rw[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int yi = 0; yi < 3; ++yi)
        if (yi < ay.n) rw[yi] = wz * ay.w[yi];
    bool lds = false;
    if (inner) {
        // stencil row yi (no y wrap) lies in the hash interval of run yi
        bool ok = true;
#pragma unroll
        for (int yi = 0; yi < 3; ++yi) {
            if (yi < ay.n) {
                const uint32_t a = ra[yi], b = rb[yi];
                const uint32_t u = a - rs[yi], v = b - rs[yi];  // in the run: u <= v <= rl (mod 2^32)
                ok = ok && (a == b || (u <= rl[yi] && v <= rl[yi]));
                pq[2 * yi] = rq[yi] + u;
                pn[2 * yi] = b - a;
            }
        }
        lds = ok;
    } else if (edgeOk) {
        const bool left = gx == 0;
        bool ok = true;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const uint32_t base = (gz * Gy + uint32_t(ay.c[k / 2])) * Gx;
            const uint32_t a = (k & 1) ? offR[base] : offR[left ? base + Gx - 1u : base + Gx - 2u];
            const uint32_t b = (k & 1) ? offR[left ? base + 2u : base + 1u] : offR[base + Gx];
            bool hit = a == b;
            uint32_t q = 0u;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const uint32_t u = a - rs[r], v = b - rs[r];
                const bool in = u <= rl[r] && v <= rl[r];
                q = !hit && in ? rq[r] + u : q;
                hit = hit || in;
            }
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                const bool in = a == ws[m] && b - a == wl[m];
                q = !hit && in ? uint32_t(Lds::kWrap0) + uint32_t(m) * kWrapCap : q;
                hit = hit || in;
            }
            pq[k] = q;
            pn[k] = b - a;
            ok = ok && hit;
        }
        lds = ok;
    }

    f3 acc = mk(0.f, 0.f, 0.f);
    // a wavefront with an edge bead on the LDS walk walks six pieces for all
    // of its beads, any other three
    if (__any(lds && !inner)) {
        if (lds) walk_lds<MODE, TAN, UNIT, 6, 1>(kp, o, rc2, acc, pq, pn, rw[0], rw[1], rw[2], L);
    } else {
        if (lds) walk_lds<MODE, TAN, UNIT, 3, 2>(kp, o, rc2, acc, pq, pn, rw[0], rw[1], rw[2], L);
    }
    if (!lds) add_all_global<MODE, TAN>(kp, o, rc2, acc, gx, gy, int(gz), offR, sPos, sVel, sTan, sIdx, beadBase);
    force[me] = make_float4(Fo.x + acc.x, Fo.y + acc.y, Fo.z + acc.z, Fo.w + 0.0f);
}

// The stage of a single system's pair workgroup: BIG pools six times its
// beads, else 3 (BLOCK + 128) (a pool of three runs overflows in a handful of
// tiles of a relaxed 800k system); a stage whose LDS would pass the 64 KiB a
// workgroup's static LDS may take falls back to the smaller one.
template <int MODE, bool TAN, int BLOCK, bool BIG>
constexpr int pair_stage() {
    constexpr int rec = 16 * (1 + (MODE == 0 ? 1 : 0) + (TAN ? 1 : 0));
    constexpr int big = 2 * BLOCK;
    return BIG && (3 * big + 2 * kWrapCap) * rec <= 65536 ? big : run_cap<BLOCK>();
}

// waves per SIMD the pair kernels are compiled for: the WCA-only kernel
// fits eight (64 VGPRs; its walk waits on LDS and needs the waves to hide
// it), the larger records leave the compiler its own choice
template <int MODE, bool TAN>
constexpr int pair_waves() {
    return MODE == 1 && !TAN ? 8 : 1;
}

template <int MODE, bool TAN, bool UNIT, int BLOCK, int CAP>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(pair_waves<MODE, TAN>()))) void k_collide_rows(float4 *__restrict__ force, const float4 *__restrict__ sPos,
                                                         const float4 *__restrict__ sVel,
                                                         const float4 *__restrict__ sTan,
                                                         const uint32_t *__restrict__ sIdx,
                                                         const uint32_t *__restrict__ off,
                                                         const uint32_t *__restrict__ sCell, uint32_t n, KP kp) {
    __shared__ RowsLds<MODE, TAN, CAP> L;
    // neighbouring tiles share row runs: a contiguous range of tiles per XCD
    // keeps their re-reads in that XCD's L2
    const uint32_t tile = xcd_contiguous(blockIdx.x, gridDim.x);
    collide_rows_tile<MODE, TAN, UNIT, BLOCK, CAP>(kp, tile * blockDim.x, n, 0u, 0u, force, sPos, sVel, sTan, sIdx, off, sCell,
                                       L);
}

// ===========================================================================
// batched replicas: every replica keeps its own parameters, grid, worms and
// Philox stream; beads, cells, front wavefronts and pair workgroups of the
// replicas are concatenated, so one launch of each kernel steps them all.
// A wavefront / workgroup belongs to one replica (found by bisection over
// the descriptor table, wave-uniform, so the replica's parameters arrive by
// scalar loads).
// ===========================================================================

__global__ __launch_bounds__(kBlock) void k_md_front_batch(BatchArrays a, const BatchRep *__restrict__ reps,
                                                           uint32_t R, uint32_t waves,
                                                           const uint64_t *__restrict__ stepCtr, float dt) {
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (wave >= waves) return;
    const uint32_t r = __builtin_amdgcn_readfirstlane(replica_of_wave(reps, R, wave));
    const BatchRep &rep = reps[r];
    const KP &kp = rep.kp;
    const WaveWorm m = wave_worm_at(wave - rep.waveOffset, kp.p.filamentSize, kp.p.numFilaments);
    // the replica's generator position at this step
    const uint64_t offRev = rep.rngBase + stepCtr[0] * rep.rngStride;
    const uint64_t offLgv = offRev + rep.revBlocks;
    float4 *pos = a.pos + rep.beadOffset, *vel = a.vel + rep.beadOffset;
    float4 *force = a.force + rep.beadOffset, *forceOld = a.forceOld + rep.beadOffset;
    const FrontLane o = md_front_lane(kp, m, pos, vel, force, forceOld, dt, rep.lgvStd, rep.langevin, rep.seed,
                                      offRev, offLgv, a.tangent != nullptr);
    if (m.active) {
        const uint32_t h = md_front_store(kp, m, o, pos, vel, force, forceOld,
                                          a.tangent ? a.tangent + rep.beadOffset : nullptr, a.swapForces != 0);
        a.hash[rep.beadOffset + m.bead] = rep.cellOffset + h;
    }
}

template <int MODE, bool TAN, bool UNIT>
__global__ __launch_bounds__(kPairBlock) __attribute__((amdgpu_waves_per_eu(pair_waves<MODE, TAN>()))) void k_collide_rows_batch(BatchArrays a, const BatchRep *__restrict__ reps,
                                                               uint32_t R, uint64_t *__restrict__ stepCtr) {
    __shared__ RowsLds<MODE, TAN, run_cap<kPairBlock>()> L;
    // the step's last kernel advances the step counter the next front reads
    if (blockIdx.x == 0 && threadIdx.x == 0) stepCtr[0] = stepCtr[0] + 1u;
    // tiles in contiguous runs per XCD (neighbouring tiles share row runs)
    const uint32_t tile = xcd_contiguous(blockIdx.x, gridDim.x);
    const uint32_t r = __builtin_amdgcn_readfirstlane(replica_of_block(reps, R, tile));
    const BatchRep &rep = reps[r];
    const uint32_t i0 = rep.beadOffset + (tile - rep.blockOffset) * blockDim.x;
    const uint32_t iEnd = rep.beadOffset + rep.kp.p.numFilaments * rep.kp.p.filamentSize;
    collide_rows_tile<MODE, TAN, UNIT, kPairBlock>(rep.kp, i0, iEnd, rep.beadOffset, rep.cellOffset,
                                       a.swapForces ? a.forceOld : a.force, a.sPos, a.sVel,
                                       a.sTan, a.ids, a.off + rep.cellOffset, a.sCell, L);
}

// ===========================================================================
// launchers
// ===========================================================================
hipError_t launch_md_front(const MdFrontArgs &a, const KP &kp, hipStream_t st) {
    const uint32_t n = kp.p.filamentSize;
    const uint32_t wpw = 64u / n;
    const uint64_t threads = ((uint64_t(a.W) + wpw - 1) / wpw) * 64u;
    // the workgroup size changes only how the wavefronts are grouped
    switch (tuning().frontBlock) {
        case 64: NW_LAUNCH(k_md_front<64>, dim3(grid_for(threads, 64)), dim3(64), 0, st, a, kp); break;
        case 128: NW_LAUNCH(k_md_front<128>, dim3(grid_for(threads, 128)), dim3(128), 0, st, a, kp); break;
        case 512: NW_LAUNCH(k_md_front<512>, dim3(grid_for(threads, 512)), dim3(512), 0, st, a, kp); break;
        case 1024: NW_LAUNCH(k_md_front<1024>, dim3(grid_for(threads, 1024)), dim3(1024), 0, st, a, kp); break;
        default: NW_LAUNCH(k_md_front<kBlock>, dim3(grid_for(threads)), dim3(kBlock), 0, st, a, kp); break;
    }
    return hipGetLastError();
}

hipError_t launch_filament_tangents(const float4 *pos, float4 *tangent, uint32_t W, const KP &kp, hipStream_t st) {
    const uint32_t n = kp.p.filamentSize;
    if (W == 0 || n < 1 || n > 64) return hipSuccess;
    const uint32_t wpw = 64u / n;
    const uint64_t threads = ((uint64_t(W) + wpw - 1) / wpw) * 64u;
    NW_LAUNCH(k_filament_tangents, dim3(grid_for(threads)), dim3(kBlock), 0, st, pos, tangent, W, kp);
    return hipGetLastError();
}

hipError_t launch_collide_off(float4 *force, const float4 *sPos, const float4 *sVel, const float4 *sTan,
                              const uint32_t *sIdx, const uint32_t *off, const uint32_t *sCell, uint32_t n,
                              const KP &kp, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const bool oneLayer = kp.e.stencil == NW_STENCIL_2D || kp.p.gridSize.z == 1;
    const bool ldsOk = oneLayer && kp.p.gridSize.x >= 3 && sCell != nullptr;
    const bool lds = ldsOk && kp.e.collideKernel != NW_COLLIDE_GLOBAL;
    const bool wca = kp.e.pairPotential == NW_PAIR_WCA;
    const bool tan = !wca || kp.p.activity != 0.0f;
    // sorted beads per LDS workgroup: 128, 256 or 512 (512 only for the WCA
    // pair term, whose stage fits the static LDS of a workgroup)
    const uint32_t tb = tuning().pairBlock;
    const int pb = tb == 128u ? 128 : (tb == 512u && wca ? 512 : kPairBlock);
    // the larger stage when pairStage asks for it
    const bool big = tuning().pairStage > uint32_t(pb + 128);
    const dim3 g(grid_for(n, lds ? uint32_t(pb) : uint32_t(kBlock))), b(lds ? pb : kBlock);
#define NW_COLLIDE(K, ...) NW_LAUNCH((K<__VA_ARGS__>), g, b, 0, st, force, sPos, sVel, sTan, sIdx, off, sCell, n, kp)
#define NW_ROWS_B(M, T, U, B)                                                                            \
    do {                                                                                                 \
        if (big) NW_COLLIDE(k_collide_rows, M, T, U, B, (pair_stage<M, T, B, true>()));                  \
        else NW_COLLIDE(k_collide_rows, M, T, U, B, (pair_stage<M, T, B, false>()));                     \
    } while (0)
#define NW_COLLIDE_ROWS(M, T, U)                                                                         \
    do {                                                                                                 \
        if (pb == 128) NW_ROWS_B(M, T, U, 128);                                                          \
        else if (pb == 512) NW_ROWS_B(M, T, U, (M == 1 ? 512 : kPairBlock));                             \
        else NW_ROWS_B(M, T, U, kPairBlock);                                                             \
    } while (0)
    if (lds) {
        // every row weight is 1 for the one-layer stencil over >= 3 rows
        const bool unit = kp.e.stencil == NW_STENCIL_2D && kp.p.gridSize.y >= 3;
        if (unit) {
            if (wca && !tan) NW_COLLIDE_ROWS(1, false, true);
            else if (wca) NW_COLLIDE_ROWS(1, true, true);
            else NW_COLLIDE_ROWS(0, true, true);
        } else {
            if (wca && !tan) NW_COLLIDE_ROWS(1, false, false);
            else if (wca) NW_COLLIDE_ROWS(1, true, false);
            else NW_COLLIDE_ROWS(0, true, false);
        }
#undef NW_COLLIDE_ROWS
#undef NW_ROWS_B
    } else {
        if (wca && !tan) NW_COLLIDE(k_collide_csr, 1, false);
        else if (wca) NW_COLLIDE(k_collide_csr, 1, true);
        else NW_COLLIDE(k_collide_csr, 0, true);
    }
#undef NW_COLLIDE
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_swap4(float4 *__restrict__ a, float4 *__restrict__ b, uint64_t n) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 x = a[i], y = b[i];
    a[i] = y;
    b[i] = x;
}

hipError_t launch_swap4(float4 *a, float4 *b, uint64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_swap4, dim3(grid_for(n)), dim3(kBlock), 0, st, a, b, n);
    return hipGetLastError();
}

// the batch's device step counter (one lane stores)
__global__ void k_set_step(uint64_t *__restrict__ ctr, uint64_t v) {
    if (threadIdx.x == 0) ctr[0] = v;
}

hipError_t launch_set_step(uint64_t *ctr, uint64_t v, hipStream_t st) {
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, st, ctr, v);
    return hipGetLastError();
}

hipError_t launch_md_front_batch(const BatchArrays &a, const BatchRep *reps, uint32_t R, uint32_t waves,
                                 const uint64_t *stepCtr, float dt, hipStream_t st) {
    if (waves == 0) return hipSuccess;
    NW_LAUNCH(k_md_front_batch, dim3(grid_for(uint64_t(waves) * 64u)), dim3(kBlock), 0, st, a, reps, R,
                       waves, stepCtr, dt);
    return hipGetLastError();
}

hipError_t launch_collide_batch(const BatchArrays &a, const BatchRep *reps, uint32_t R, uint32_t blocks, int mode,
                                bool tan, bool unit, uint64_t *stepCtr, hipStream_t st) {
    if (blocks == 0) return hipSuccess;
    const dim3 g(blocks), b(kPairBlock);
#define NW_BATCH_COLLIDE(...) NW_LAUNCH((k_collide_rows_batch<__VA_ARGS__>), g, b, 0, st, a, reps, R, stepCtr)
    if (mode == 1 && !tan) {
        if (unit) NW_BATCH_COLLIDE(1, false, true);
        else NW_BATCH_COLLIDE(1, false, false);
    } else if (mode == 1) {
        if (unit) NW_BATCH_COLLIDE(1, true, true);
        else NW_BATCH_COLLIDE(1, true, false);
    } else {
        if (unit) NW_BATCH_COLLIDE(0, true, true);
        else NW_BATCH_COLLIDE(0, true, false);
    }
#undef NW_BATCH_COLLIDE
    return hipGetLastError();
}

}  // namespace nw

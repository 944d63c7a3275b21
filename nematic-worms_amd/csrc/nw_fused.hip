// nw_fused.hip -- the fused MD step kernels (gfx950).
//
// The reference issues ~10 launches per MD step, several of them tiny; on
// MI355X each dispatch costs microseconds of fixed time, and the reorder
// gathers re-read what the sort just touched.  The fused step is
//   k_md_front      integrate + reversal + Langevin + bonded + tangent + cell
//                   hash (whole worms per wavefront, chain neighbours by
//                   cross-lane shuffles)
//   cell sort       nw_bucket.hip (stable, atomic-free, gathers the records)
//   k_collide_lds   pair forces, neighbour rows staged in LDS (k_collide_csr
//                   without LDS for stencils over several z layers)
// computing the per-call path's expressions in the same order.
#include "nw_bucket.hpp"
#include "nw_physics.hpp"

namespace nw {

// cell coordinates of sorted bead i: from the sort's cell ids when given
// (already wrapped into the grid), else from the position
__device__ __forceinline__ void bead_cell(const KP &kp, const uint32_t *__restrict__ sCell, uint32_t i, f3 r, int &gx,
                                          int &gy, int &gz) {
    const SimParams &p = kp.p;
    if (sCell) {
        const uint32_t c = sCell[i];
        const uint32_t Gx = p.gridSize.x, layer = p.gridSize.x * p.gridSize.y;
        const uint32_t z = c / layer, rem = c - z * layer;
        const uint32_t y = rem / Gx;
        gx = int(rem - y * Gx);
        gy = int(y);
        gz = int(z);
    } else {
        gx = wrap_mod(cell_coord(r.x, p.origin.x, p.cellSize.x), p.gridSize.x);
        gy = wrap_mod(cell_coord(r.y, p.origin.y, p.cellSize.y), p.gridSize.y);
        gz = wrap_mod(cell_coord(r.z, p.origin.z, p.cellSize.z), p.gridSize.z);
    }
}

// ===========================================================================
// MD front
// ===========================================================================
constexpr int kFrontBlock = 1024;  // one level-1 tile of the cell sort

__global__ __launch_bounds__(kFrontBlock) void k_md_front(MdFrontArgs a, KP kp) {
    extern __shared__ uint32_t lh[];  // bucket histogram (buckets + 1), wave sums
    const SimParams &p = kp.p;
    const uint32_t n = p.filamentSize;
    const BucketPlan &bp = a.bp;
    const uint32_t B1 = bp.buckets + 1;
    uint32_t *wsum = lh + B1;
    for (uint32_t b = threadIdx.x; b < B1; b += blockDim.x) lh[b] = 0u;

    const WaveWorm m = wave_worm(n, a.W);
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 P = m.active ? a.pos[m.bead] : zero4;
    const float4 V = m.active ? a.vel[m.bead] : zero4;
    const float4 F = m.active ? a.force[m.bead] : zero4;
    const float4 FO = m.active ? a.forceOld[m.bead] : zero4;

    // velocity Verlet (filament_integrator, particles_kernel_impl.cuh:71-142)
    f3 r = mk(P), v = mk(V);
    const f3 ff = mk(F), fo = mk(FO);
    const f3 dv = 0.5f * (ff + fo) * a.dt;
    v += dv;
    const f3 dr = v * a.dt + 0.5f * ff * a.dt * a.dt;
    r += dr;
    apply_boundaries(p, r, v, dr, p.particleRadius);
    f3 fold = ff;
    float pw = P.w, vw = V.w, fow = FO.w;

    // polarity reversal (reverseFilamentsKernel :820-857): bead p <-> n-1-p
    // of the worm, taken from the mirror lane; the force is zero here
    if (p.reverseProbability > 0.0f) {
        const bool rev = m.active && !(uniform_at(a.seed, a.offRev, m.worm) > p.reverseProbability);
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
    if (a.langevin) {
        const u4 b = philox_block(a.seed, a.offLgv + m.bead);
        float z0, z1, z2, z3;
        box_muller(b.x, b.y, z0, z1);
        box_muller(b.z, b.w, z2, z3);
        (void)z3;
        const f3 xi = mk(z0 * a.lgvStd + 0.0f, z1 * a.lgvStd + 0.0f, z2 * a.lgvStd + 0.0f);
        f3 f = mk(0.f, 0.f, 0.f);
        f -= p.gamma * v;
        f += xi;
        acc = mk(0.0f + f.x, 0.0f + f.y, 0.0f + f.z);
    }

    // bonded forces and tangents on the new positions
    f3 tg;
    const bool writeTan = bonded_lane(kp, r, m.pp, acc, tg);

    uint32_t h = 0, rank = 0;
    __syncthreads();  // histogram cleared
    if (m.active) {
        a.pos[m.bead] = f4(r, pw);
        a.vel[m.bead] = f4(v, vw);
        a.force[m.bead] = f4(acc, 0.0f);
        a.forceOld[m.bead] = f4(fold, fow);
        if (writeTan) a.tangent[m.bead] = f4(tg, 0.0f);
        // cell of the new position (calcHashD)
        h = md_hash_wrapped(p, cell_coord(r.x, p.origin.x, p.cellSize.x), cell_coord(r.y, p.origin.y, p.cellSize.y),
                            cell_coord(r.z, p.origin.z, p.cellSize.z));
        if (a.hash) a.hash[m.bead] = h;
        rank = atomicAdd(&lh[h >> bp.shift], 1u);
    }
    __syncthreads();
    // level 1 of the cell sort: this workgroup's beads are tile blockIdx.x
    block_excl_scan(lh, bp.buckets, wsum);
    uint32_t *row = a.P + uint64_t(blockIdx.x) * B1;
    for (uint32_t b = threadIdx.x; b < B1; b += blockDim.x) row[b] = lh[b];
    if (m.active) a.pairs[uint64_t(blockIdx.x) * bp.tileItems + lh[h >> bp.shift] + rank] = make_uint2(h, m.bead);
}

// ===========================================================================
// pair forces over CSR cell offsets
// ===========================================================================
template <int MODE>
struct PairCtx {
    f3 r, v, t;
    uint32_t i, fil, ch;
    bool useTan;
    float rc2;
};

template <int MODE>
__device__ __forceinline__ f3 scan_run(const KP &kp, const PairCtx<MODE> &c, uint32_t lo, uint32_t hi,
                                       const float4 *__restrict__ sPos, const float4 *__restrict__ sVel,
                                       const float4 *__restrict__ sTan, const uint32_t *__restrict__ sIdx) {
    const SimParams &p = kp.p;
    f3 fc = mk(0.f, 0.f, 0.f);
    for (uint32_t j = lo; j < hi; ++j) {
        if (j == c.i) continue;
        const f3 r2 = mk(sPos[j]);
        uint32_t fil2, ch2;
        worm_of(kp, sIdx[j], fil2, ch2);
        const bool same = c.fil == fil2;
        if (MODE == 0) {
            const f3 v2 = mk(sVel[j]);
            if (same && ((c.ch + 1u == ch2) || (c.ch - 1u == ch2)))
                fc += bond_hookean(p, c.r, r2, p.bondSpringK, p.bondSpringL);
            else
                fc += dem_force(p, c.r, r2, c.v, v2);
        } else if (!same) {
            fc += wca_force(p, kp.e, c.r, r2, c.rc2);
        }
        if (c.useTan && !same) fc += extensile_force(p, c.r, r2, c.t, mk(sTan[j]), p.activity);
    }
    return fc;
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_collide_csr(float4 *__restrict__ force, const float4 *__restrict__ sPos,
                                                        const float4 *__restrict__ sVel,
                                                        const float4 *__restrict__ sTan,
                                                        const uint32_t *__restrict__ sIdx,
                                                        const uint32_t *__restrict__ off,
                                                        const uint32_t *__restrict__ sCell, uint32_t n, KP kp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SimParams &p = kp.p;
    PairCtx<MODE> c;
    c.i = i;
    c.r = mk(sPos[i]);
    c.useTan = MODE == 0 || p.activity != 0.0f;
    c.v = MODE == 0 ? mk(sVel[i]) : mk(0.f, 0.f, 0.f);
    c.t = c.useTan ? mk(sTan[i]) : mk(0.f, 0.f, 0.f);
    c.rc2 = kp.e.wcaCutoff * kp.e.wcaCutoff;
    const uint32_t me = sIdx[i];
    worm_of(kp, me, c.fil, c.ch);

    int gx, gy, gz;
    bead_cell(kp, sCell, i, c.r, gx, gy, gz);
    const uint32_t Gx = p.gridSize.x;
    const AxisCells ay = axis_cells(gy, p.gridSize.y, false);
    const AxisCells az = axis_cells(gz, p.gridSize.z, kp.e.stencil == NW_STENCIL_2D);
    const int cx = gx;

    f3 acc = mk(0.f, 0.f, 0.f);
    for (int zi = 0; zi < az.n; ++zi) {
        for (int yi = 0; yi < ay.n; ++yi) {
            const float wzy = az.w[zi] * ay.w[yi];
            const uint32_t base = (uint32_t(az.c[zi]) * p.gridSize.y + uint32_t(ay.c[yi])) * Gx;
            f3 fr = mk(0.f, 0.f, 0.f);
            if (Gx >= 3) {
                // cells cx-1, cx, cx+1 in this order: one run, or two across the x wrap
                uint32_t a0, a1, b0 = 0, b1 = 0;
                if (cx >= 1 && cx + 1 < int(Gx)) {
                    a0 = off[base + cx - 1];
                    a1 = off[base + cx + 2];
                } else if (cx == 0) {
                    a0 = off[base + Gx - 1];
                    a1 = off[base + Gx];
                    b0 = off[base];
                    b1 = off[base + 2];
                } else {
                    a0 = off[base + Gx - 2];
                    a1 = off[base + Gx];
                    b0 = off[base];
                    b1 = off[base + 1];
                }
                fr = scan_run<MODE>(kp, c, a0, a1, sPos, sVel, sTan, sIdx);
                fr += scan_run<MODE>(kp, c, b0, b1, sPos, sVel, sTan, sIdx);
            } else {
                const AxisCells ax = axis_cells(gx, Gx, false);
                for (int xi = 0; xi < ax.n; ++xi) {
                    const uint32_t h = base + uint32_t(ax.c[xi]);
                    const f3 fc = scan_run<MODE>(kp, c, off[h], off[h + 1], sPos, sVel, sTan, sIdx);
                    fr += ax.w[xi] * fc;
                }
            }
            if (wzy == 1.0f) acc += fr;
            else acc += wzy * fr;
        }
    }
    float4 Fo = force[me];
    force[me] = make_float4(Fo.x + acc.x, Fo.y + acc.y, Fo.z + acc.z, Fo.w + 0.0f);
}

// ===========================================================================
// pair forces with LDS tiles (stencils of one z layer, gridSize.x, .y >= 3)
//
// A workgroup owns 256 consecutive sorted beads whose cells span the hash
// interval [h0, h1].  Every neighbour cell of those beads that does not wrap
// lies in one of three hash intervals [h0 - 1 + dy Gx, h1 + 1 + dy Gx],
// dy = -1, 0, 1, i.e. in three contiguous runs of the sorted arrays.  The
// workgroup copies the runs into LDS with coalesced loads (positions and worm
// ids, plus velocities / tangents when the pair terms read them), then every
// candidate run that lies inside a staged run is read from LDS; the rest
// (x or y wrap, a run longer than the tile) is read from global memory.  The
// candidate order and arithmetic are those of k_collide_csr: bitwise equal.
// ===========================================================================
constexpr int kTileCap = 384;  // staged beads per row run

template <int MODE, bool TAN>
__global__ __launch_bounds__(kBlock) void k_collide_lds(float4 *__restrict__ force, const float4 *__restrict__ sPos,
                                                        const float4 *__restrict__ sVel,
                                                        const float4 *__restrict__ sTan,
                                                        const uint32_t *__restrict__ sIdx,
                                                        const uint32_t *__restrict__ off,
                                                        const uint32_t *__restrict__ sCell, uint32_t n, KP kp) {
    __shared__ float4 lpos[3 * kTileCap];
    __shared__ uint32_t lfil[3 * kTileCap];
    __shared__ uint32_t lch[MODE == 0 ? 3 * kTileCap : 1];
    __shared__ float4 lvel[MODE == 0 ? 3 * kTileCap : 1];
    __shared__ float4 ltan[TAN ? 3 * kTileCap : 1];
    __shared__ uint32_t rs[3], rl[3];

    const SimParams &p = kp.p;
    const uint32_t Gx = p.gridSize.x;
    const uint32_t C = p.gridSize.x * p.gridSize.y * p.gridSize.z;
    const uint32_t i0 = blockIdx.x * blockDim.x;
    const uint32_t i = i0 + threadIdx.x;

    // hash interval of the tile (sorted: first and last bead)
    if (threadIdx.x < 3) {
        const uint32_t last = min(i0 + blockDim.x, n) - 1u;
        const f3 a = mk(sPos[i0]), b = mk(sPos[last]);
        const uint32_t h0 = md_hash_wrapped(p, cell_coord(a.x, p.origin.x, p.cellSize.x),
                                            cell_coord(a.y, p.origin.y, p.cellSize.y),
                                            cell_coord(a.z, p.origin.z, p.cellSize.z));
        const uint32_t h1 = md_hash_wrapped(p, cell_coord(b.x, p.origin.x, p.cellSize.x),
                                            cell_coord(b.y, p.origin.y, p.cellSize.y),
                                            cell_coord(b.z, p.origin.z, p.cellSize.z));
        const int dy = int(threadIdx.x) - 1;
        const int64_t lo = int64_t(h0) - 1 + int64_t(dy) * Gx;
        const int64_t hi = int64_t(h1) + 1 + int64_t(dy) * Gx;  // inclusive
        const uint32_t clo = uint32_t(lo < 0 ? 0 : (lo >= int64_t(C) ? C : lo));
        const uint32_t chi = uint32_t(hi + 1 < 0 ? 0 : (hi + 1 > int64_t(C) ? C : hi + 1));
        const uint32_t s = off[clo], e = off[chi];
        rs[threadIdx.x] = s;
        rl[threadIdx.x] = (e > s && e - s <= uint32_t(kTileCap)) ? e - s : 0u;
    }
    __syncthreads();
    for (int r = 0; r < 3; ++r) {
        const uint32_t s = rs[r], l = rl[r];
        for (uint32_t k = threadIdx.x; k < l; k += blockDim.x) {
            const uint32_t j = s + k;
            lpos[r * kTileCap + k] = sPos[j];
            uint32_t fil, ch;
            worm_of(kp, sIdx[j], fil, ch);
            lfil[r * kTileCap + k] = fil;
            if (MODE == 0) {
                lch[r * kTileCap + k] = ch;
                lvel[r * kTileCap + k] = sVel[j];
            }
            if (TAN) ltan[r * kTileCap + k] = sTan[j];
        }
    }
    __syncthreads();
    if (i >= n) return;

    const f3 r0 = mk(sPos[i]);
    const f3 v0 = MODE == 0 ? mk(sVel[i]) : mk(0.f, 0.f, 0.f);
    const f3 t0 = TAN ? mk(sTan[i]) : mk(0.f, 0.f, 0.f);
    const float rc2 = kp.e.wcaCutoff * kp.e.wcaCutoff;
    const uint32_t me = sIdx[i];
    uint32_t fil, ch;
    worm_of(kp, me, fil, ch);

    // one candidate, from LDS or global; terms added to fc in scan_run's order
    auto pair = [&](f3 &fc, const f3 r2, uint32_t fil2, uint32_t ch2, const f3 v2, const f3 t2) {
        const bool same = fil == fil2;
        if (MODE == 0) {
            if (same && ((ch + 1u == ch2) || (ch - 1u == ch2)))
                fc += bond_hookean(p, r0, r2, p.bondSpringK, p.bondSpringL);
            else
                fc += dem_force(p, r0, r2, v0, v2);
        } else if (!same) {
            fc += wca_force(p, kp.e, r0, r2, rc2);
        }
        if (TAN && !same) fc += extensile_force(p, r0, r2, t0, t2, p.activity);
    };
    auto run = [&](uint32_t a, uint32_t b) {
        f3 fc = mk(0.f, 0.f, 0.f);
        if (a >= b) return fc;
        int src = -1;
        for (int r = 0; r < 3; ++r)
            if (rl[r] && a >= rs[r] && b <= rs[r] + rl[r]) src = r;
        if (src >= 0) {
            const uint32_t base = uint32_t(src) * kTileCap - rs[src];
            for (uint32_t j = a; j < b; ++j) {
                if (j == i) continue;
                const uint32_t q = base + j;
                pair(fc, mk(lpos[q]), lfil[q], MODE == 0 ? lch[q] : 0u,
                     MODE == 0 ? mk(lvel[q]) : mk(0.f, 0.f, 0.f), TAN ? mk(ltan[q]) : mk(0.f, 0.f, 0.f));
            }
        } else {
            for (uint32_t j = a; j < b; ++j) {
                if (j == i) continue;
                uint32_t fil2, ch2;
                worm_of(kp, sIdx[j], fil2, ch2);
                pair(fc, mk(sPos[j]), fil2, ch2, MODE == 0 ? mk(sVel[j]) : mk(0.f, 0.f, 0.f),
                     TAN ? mk(sTan[j]) : mk(0.f, 0.f, 0.f));
            }
        }
        return fc;
    };

    int gx, gy, gz;
    bead_cell(kp, sCell, i, r0, gx, gy, gz);
    const AxisCells ay = axis_cells(gy, p.gridSize.y, false);
    const AxisCells az = axis_cells(gz, p.gridSize.z, kp.e.stencil == NW_STENCIL_2D);
    const int cx = gx;
    f3 acc = mk(0.f, 0.f, 0.f);
    for (int yi = 0; yi < ay.n; ++yi) {
        const float wzy = az.w[0] * ay.w[yi];
        const uint32_t base = (uint32_t(az.c[0]) * p.gridSize.y + uint32_t(ay.c[yi])) * Gx;
        uint32_t a0, a1, b0 = 0, b1 = 0;
        if (cx >= 1 && cx + 1 < int(Gx)) {
            a0 = off[base + cx - 1];
            a1 = off[base + cx + 2];
        } else if (cx == 0) {
            a0 = off[base + Gx - 1];
            a1 = off[base + Gx];
            b0 = off[base];
            b1 = off[base + 2];
        } else {
            a0 = off[base + Gx - 2];
            a1 = off[base + Gx];
            b0 = off[base];
            b1 = off[base + 1];
        }
        f3 fr = run(a0, a1);
        fr += run(b0, b1);
        if (wzy == 1.0f) acc += fr;
        else acc += wzy * fr;
    }
    float4 Fo = force[me];
    force[me] = make_float4(Fo.x + acc.x, Fo.y + acc.y, Fo.z + acc.z, Fo.w + 0.0f);
}

// ===========================================================================
// launchers
// ===========================================================================
hipError_t launch_md_front(const MdFrontArgs &a, const KP &kp, hipStream_t st) {
    const uint32_t n = kp.p.filamentSize;
    const uint32_t wpw = 64u / n;
    const uint64_t threads = ((uint64_t(a.W) + wpw - 1) / wpw) * 64u;
    const size_t lds = (size_t(a.bp.buckets) + 1 + kFrontBlock / 64) * sizeof(uint32_t);
    hipLaunchKernelGGL(k_md_front, dim3(grid_for(threads, kFrontBlock)), dim3(kFrontBlock), lds, st, a, kp);
    return hipGetLastError();
}

hipError_t launch_collide_off(float4 *force, const float4 *sPos, const float4 *sVel, const float4 *sTan,
                              const uint32_t *sIdx, const uint32_t *off, const uint32_t *sCell, uint32_t n,
                              const KP &kp, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const bool oneLayer = kp.e.stencil == NW_STENCIL_2D || kp.p.gridSize.z == 1;
    const bool ldsOk = oneLayer && kp.p.gridSize.x >= 3;
    const bool lds = ldsOk && kp.e.collideKernel != NW_COLLIDE_GLOBAL;
    const bool wca = kp.e.pairPotential == NW_PAIR_WCA;
    const bool tan = !wca || kp.p.activity != 0.0f;
    const dim3 g(grid_for(n)), b(kBlock);
    if (lds) {
        if (wca && !tan)
            hipLaunchKernelGGL((k_collide_lds<1, false>), g, b, 0, st, force, sPos, sVel, sTan, sIdx, off, sCell, n,
                               kp);
        else if (wca)
            hipLaunchKernelGGL((k_collide_lds<1, true>), g, b, 0, st, force, sPos, sVel, sTan, sIdx, off, sCell, n, kp);
        else
            hipLaunchKernelGGL((k_collide_lds<0, true>), g, b, 0, st, force, sPos, sVel, sTan, sIdx, off, sCell, n, kp);
    } else if (wca) {
        hipLaunchKernelGGL(k_collide_csr<1>, g, b, 0, st, force, sPos, sVel, sTan, sIdx, off, sCell, n, kp);
    } else {
        hipLaunchKernelGGL(k_collide_csr<0>, g, b, 0, st, force, sPos, sVel, sTan, sIdx, off, sCell, n, kp);
    }
    return hipGetLastError();
}

}  // namespace nw

// nw_fused.hip -- the worm step as few, fat kernels (gfx950).
//
// The reference issues ~10 launches per MD step and ~9 per SRD collision,
// several of them tiny; on MI355X each dispatch costs a few microseconds of
// fixed time, and the reorder gathers re-read what the sort just touched.
// This file computes the same step (bit for bit where the per-call path is
// bit-exact) in:
//   MD :  k_md_front   integrate + reversal + Langevin + bonded + tangent +
//                      cell hash + histogram (whole worms per wavefront)
//         scan         cell offsets (CSR) from the histogram
//         k_scatter_ix bead ids into their cells (arbitrary order in a cell)
//         k_md_cells   per cell: stable order restored, positions gathered,
//                      histogram bin cleared for the next step
//         k_collide_csr pair forces; a row of three cells is one run of beads
//   SRD:  k_srd_front  shifted SRD hash + histogram of beads and solvent
//         scan, k_scatter_ix
//         k_srd_cells  per cell: stable order, isokinetic rescale, centre of
//                      momentum in sorted order, the cell's rotation sign
//         k_srd_rotate rotation about the cell momentum + solvent streaming
#include "nw_fused.hpp"
#include "nw_physics.hpp"

namespace nw {

// ===========================================================================
// MD front
// ===========================================================================
__global__ __launch_bounds__(kBlock) void k_md_front(float4 *__restrict__ pos, float4 *__restrict__ vel,
                                                     float4 *__restrict__ force, float4 *__restrict__ forceOld,
                                                     float4 *__restrict__ tangent, uint32_t *__restrict__ hash,
                                                     uint32_t *__restrict__ slot, uint32_t *__restrict__ counts,
                                                     float dt, float lgvStd, int langevin, uint32_t W, uint64_t seed,
                                                     uint64_t offRev, uint64_t offLgv, KP kp) {
    const SimParams &p = kp.p;
    const uint32_t n = p.filamentSize;
    const WaveWorm m = wave_worm(n, W);
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 P = m.active ? pos[m.bead] : zero4;
    const float4 V = m.active ? vel[m.bead] : zero4;
    const float4 F = m.active ? force[m.bead] : zero4;
    const float4 FO = m.active ? forceOld[m.bead] : zero4;

    // velocity Verlet (filament_integrator, particles_kernel_impl.cuh:71-142)
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
        const bool rev = m.active && !(uniform_at(seed, offRev, m.worm) > p.reverseProbability);
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
        const u4 b = philox_block(seed, offLgv + m.bead);
        float z0, z1, z2, z3;
        box_muller(b.x, b.y, z0, z1);
        box_muller(b.z, b.w, z2, z3);
        (void)z3;
        const f3 xi = mk(z0 * lgvStd + 0.0f, z1 * lgvStd + 0.0f, z2 * lgvStd + 0.0f);
        f3 f = mk(0.f, 0.f, 0.f);
        f -= p.gamma * v;
        f += xi;
        acc = mk(0.0f + f.x, 0.0f + f.y, 0.0f + f.z);
    }

    // bonded forces and tangents on the new positions
    f3 tg;
    const bool writeTan = bonded_lane(kp, r, m.pp, acc, tg);
    if (!m.active) return;

    pos[m.bead] = f4(r, pw);
    vel[m.bead] = f4(v, vw);
    force[m.bead] = f4(acc, 0.0f);
    forceOld[m.bead] = f4(fold, fow);
    if (writeTan) tangent[m.bead] = f4(tg, 0.0f);

    // cell of the new position and its histogram slot
    const uint32_t h = md_hash_wrapped(p, cell_coord(r.x, p.origin.x, p.cellSize.x),
                                       cell_coord(r.y, p.origin.y, p.cellSize.y),
                                       cell_coord(r.z, p.origin.z, p.cellSize.z));
    hash[m.bead] = h;
    slot[m.bead] = atomicAdd(&counts[h], 1u);
}

// ===========================================================================
// shared: scatter ids into their bins
// ===========================================================================
__global__ __launch_bounds__(kBlock) void k_scatter_ix(const uint32_t *__restrict__ hash,
                                                       const uint32_t *__restrict__ slot,
                                                       const uint32_t *__restrict__ offsets, uint32_t n,
                                                       uint32_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[offsets[hash[i]] + slot[i]] = i;
}

// ascending insertion sort of ids[s, e) (a cell's members: a handful)
__device__ __forceinline__ void sort_segment(uint32_t *__restrict__ ids, uint32_t s, uint32_t e) {
    for (uint32_t a = s + 1; a < e; ++a) {
        const uint32_t x = ids[a];
        uint32_t c = a;
        while (c > s) {
            const uint32_t y = ids[c - 1];
            if (y <= x) break;
            ids[c] = y;
            --c;
        }
        ids[c] = x;
    }
}

// ===========================================================================
// MD cells: stable order + gather + clear histogram
// ===========================================================================
__global__ __launch_bounds__(kBlock) void k_md_cells(const uint32_t *__restrict__ offsets,
                                                     uint32_t *__restrict__ counts, uint32_t bins,
                                                     uint32_t *__restrict__ ids, const float4 *__restrict__ pos,
                                                     const float4 *__restrict__ vel, const float4 *__restrict__ tan,
                                                     float4 *__restrict__ sPos, float4 *__restrict__ sVel,
                                                     float4 *__restrict__ sTan) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= bins) return;
    counts[c] = 0u;
    const uint32_t s = offsets[c], e = offsets[c + 1];
    sort_segment(ids, s, e);
    for (uint32_t j = s; j < e; ++j) {
        const uint32_t b = ids[j];
        sPos[j] = pos[b];
        if (sVel) sVel[j] = vel[b];
        if (sTan) sTan[j] = tan[b];
    }
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
                                                        const uint32_t *__restrict__ off, uint32_t n, KP kp) {
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

    const int gx = cell_coord(c.r.x, p.origin.x, p.cellSize.x);
    const int gy = cell_coord(c.r.y, p.origin.y, p.cellSize.y);
    const int gz = cell_coord(c.r.z, p.origin.z, p.cellSize.z);
    const uint32_t Gx = p.gridSize.x;
    const AxisCells ay = axis_cells(gy, p.gridSize.y, false);
    const AxisCells az = axis_cells(gz, p.gridSize.z, kp.e.stencil == NW_STENCIL_2D);
    const int cx = wrap_mod(gx, Gx);

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
// SRD
// ===========================================================================
__global__ __launch_bounds__(kBlock) void k_srd_front(const float4 *__restrict__ pos, uint32_t n,
                                                      uint32_t *__restrict__ hash, uint32_t *__restrict__ slot,
                                                      uint32_t *__restrict__ counts, KP kp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t h = srd_hash(kp.p, kp.s, mk(pos[i]));
    hash[i] = h;
    slot[i] = atomicAdd(&counts[h], 1u);
}

// per SRD cell: members in stable order, optional isokinetic rescale
// (isokineticThermostatSRD :604-637), centre of momentum in sorted order
// (cellCenterOfMomentumSRD :559-594), rotation sign in com.w
__global__ __launch_bounds__(kBlock) void k_srd_cells(const uint32_t *__restrict__ offsets,
                                                      uint32_t *__restrict__ counts, uint32_t bins,
                                                      uint32_t *__restrict__ ids, float4 *__restrict__ vel,
                                                      float4 *__restrict__ com, int isokinetic, uint64_t seed,
                                                      uint64_t offset, KP kp) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= bins) return;
    counts[c] = 0u;
    if (c + 1u == bins) return;  // the overflow bin: always empty here
    const uint32_t s = offsets[c], e = offsets[c + 1];
    sort_segment(ids, s, e);
    f3 m = mk(0.f, 0.f, 0.f);
    if (s != e) {
        float lambda = 1.0f;
        if (isokinetic) {
            float E = 0.0f;
            for (uint32_t j = s; j < e; ++j) {
                const float4 V = vel[ids[j]];
                const f3 v = mk(V);
                E += 0.5f * V.w * dot(v, v);
            }
            lambda = sqrtf(kp.s.kbT / E);
        }
        float M = 0.0f;
        for (uint32_t j = s; j < e; ++j) {
            const uint32_t b = ids[j];
            float4 V = vel[b];
            if (isokinetic) {
                V = f4(lambda * mk(V), V.w);
                vel[b] = V;
            }
            m += V.w * mk(V);
            M += V.w;
        }
        if (M > 0.0f) m = m * (1.0f / M);
    }
    const float sgn = uniform_at(seed, offset, c) > 0.5f ? 1.0f : -1.0f;
    com[c] = f4(m, sgn);
}

// collideSolventKernel (:651-691) for every particle that rotates, then
// solvent_integrator (:145-205) for the solvent
__global__ __launch_bounds__(kBlock) void k_srd_rotate(float4 *__restrict__ pos, float4 *__restrict__ vel,
                                                       const float4 *__restrict__ com, uint32_t N, uint32_t nt,
                                                       float dt, float cosA, float sinA, KP kp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    const bool solvent = i >= N;
    const bool rotate = solvent || kp.e.srdCollideFilaments;
    if (!rotate) return;
    const float4 P = pos[i];
    const float4 V = vel[i];
    f3 r = mk(P), v = mk(V);
    const float4 C = com[srd_hash(kp.p, kp.s, r)];
    const f3 u = mk(C);
    const float s = C.w * sinA;
    const f3 d = v - u;
    const f3 rot = mk(d.x * cosA - d.y * s, d.x * s + d.y * cosA, d.z);
    v = u + rot;
    if (solvent) {
        const f3 dr = dt * v;
        r += dr;
        apply_boundaries(kp.p, r, v, dr, 0.0f);
        pos[i] = f4(r, P.w);
    }
    vel[i] = f4(v, V.w);
}

// ===========================================================================
// launchers
// ===========================================================================
hipError_t ensure_fused_scratch(FusedScratch &f, uint64_t n, uint64_t bins, hipStream_t st) {
    hipError_t e;
    if (n > f.cap_n) {
        const uint64_t cap = n + n / 8 + 1024;
        if (f.slot) { (void)hipFree(f.slot); (void)hipFree(f.hashTmp); }
        f.slot = f.hashTmp = nullptr;
        if ((e = hipMalloc(&f.slot, cap * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&f.hashTmp, cap * 4)) != hipSuccess) return e;
        f.cap_n = cap;
    }
    if (bins > f.cap_bins) {
        const uint64_t cap = bins + bins / 8 + 1024;
        if (f.counts) { (void)hipFree(f.counts); (void)hipFree(f.offsets); (void)hipFree(f.scan); }
        f.counts = f.offsets = f.scan = nullptr;
        if ((e = hipMalloc(&f.counts, cap * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&f.offsets, (cap + 1) * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&f.scan, scan_scratch_words(cap) * 4)) != hipSuccess) return e;
        f.cap_bins = cap;
        f.dirty = true;
    }
    if (f.dirty) {
        // the bins are kept at zero by the consumers; clear once after (re)allocation
        if ((e = launch_fill_u32(f.counts, 0u, f.cap_bins, st)) != hipSuccess) return e;
        f.dirty = false;
    }
    return hipSuccess;
}

hipError_t launch_md_front(const MdFrontArgs &a, const KP &kp, FusedScratch &f, hipStream_t st) {
    const uint32_t n = kp.p.filamentSize;
    const uint32_t wpw = 64u / n;
    const uint64_t threads = ((uint64_t(a.W) + wpw - 1) / wpw) * 64u;
    hipLaunchKernelGGL(k_md_front, dim3(grid_for(threads)), dim3(kBlock), 0, st, a.pos, a.vel, a.force, a.forceOld,
                       a.tangent, a.hash, f.slot, f.counts, a.dt, a.lgvStd, a.langevin, a.W, a.seed, a.offRev,
                       a.offLgv, kp);
    return hipGetLastError();
}

hipError_t launch_scan_bins(FusedScratch &f, uint64_t bins, hipStream_t st) {
    return exclusive_scan_u32(f.counts, f.offsets, bins, f.scan, st);
}

hipError_t launch_scatter_ix(const uint32_t *hash, FusedScratch &f, uint32_t n, uint32_t *out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_ix, dim3(grid_for(n)), dim3(kBlock), 0, st, hash, f.slot, f.offsets, n, out);
    return hipGetLastError();
}

hipError_t launch_md_cells(FusedScratch &f, uint32_t bins, uint32_t *ids, const float4 *pos, const float4 *vel,
                           const float4 *tan, float4 *sPos, float4 *sVel, float4 *sTan, hipStream_t st) {
    hipLaunchKernelGGL(k_md_cells, dim3(grid_for(bins)), dim3(kBlock), 0, st, f.offsets, f.counts, bins, ids, pos,
                       vel, tan, sPos, sVel, sTan);
    return hipGetLastError();
}

hipError_t launch_collide_csr(float4 *force, const float4 *sPos, const float4 *sVel, const float4 *sTan,
                              const uint32_t *sIdx, const FusedScratch &f, uint32_t n, const KP &kp,
                              hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (kp.e.pairPotential == NW_PAIR_WCA)
        hipLaunchKernelGGL(k_collide_csr<1>, dim3(grid_for(n)), dim3(kBlock), 0, st, force, sPos, sVel, sTan, sIdx,
                           f.offsets, n, kp);
    else
        hipLaunchKernelGGL(k_collide_csr<0>, dim3(grid_for(n)), dim3(kBlock), 0, st, force, sPos, sVel, sTan, sIdx,
                           f.offsets, n, kp);
    return hipGetLastError();
}

hipError_t launch_srd_front(const float4 *pos, uint32_t nt, uint32_t *hash, FusedScratch &f, const KP &kp,
                            hipStream_t st) {
    if (nt == 0) return hipSuccess;
    hipLaunchKernelGGL(k_srd_front, dim3(grid_for(nt)), dim3(kBlock), 0, st, pos, nt, hash, f.slot, f.counts, kp);
    return hipGetLastError();
}

hipError_t launch_srd_cells(FusedScratch &f, uint32_t bins, uint32_t *ids, float4 *vel, float4 *com, int iso,
                            uint64_t seed, uint64_t offset, const KP &kp, hipStream_t st) {
    hipLaunchKernelGGL(k_srd_cells, dim3(grid_for(bins)), dim3(kBlock), 0, st, f.offsets, f.counts, bins, ids, vel,
                       com, iso, seed, offset, kp);
    return hipGetLastError();
}

hipError_t launch_srd_rotate(float4 *pos, float4 *vel, const float4 *com, uint32_t N, uint32_t nt, float dt,
                             float cosA, float sinA, const KP &kp, hipStream_t st) {
    if (nt == 0) return hipSuccess;
    hipLaunchKernelGGL(k_srd_rotate, dim3(grid_for(nt)), dim3(kBlock), 0, st, pos, vel, com, N, nt, dt, cosA, sinA,
                       kp);
    return hipGetLastError();
}

}  // namespace nw

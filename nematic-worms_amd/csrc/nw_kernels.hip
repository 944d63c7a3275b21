// nw_kernels.hip -- the per-step kernels of the worm time step on gfx950.
//
// Layout: float4 records per bead/solvent particle in HBM (pos.w = radius,
// vel.w = mass), worm-contiguous (bead = worm * n + p).  Every kernel writes
// only slots it owns, so results are deterministic and atomic-free.
#include "nw_internal.hpp"
#include "nw_physics.hpp"

namespace nw {

__device__ uint32_t g_flags;  // anomaly bits, read by nwGetDeviceFlags

uint32_t *device_flags() {
    void *p = nullptr;
    (void)hipGetSymbolAddress(&p, HIP_SYMBOL(g_flags));
    return static_cast<uint32_t *>(p);
}

// ===========================================================================
// Integrators (particles_kernel_impl.cuh:34-205)
// ===========================================================================
// filament_integrator (particles_kernel_impl.cuh:71-142): 128 B/bead stream.
__global__ __launch_bounds__(kBlock) void k_integrate_filaments(float4 *__restrict__ pos, float4 *__restrict__ vel,
                                                                float4 *__restrict__ f, float4 *__restrict__ fold,
                                                                float dt, uint32_t n, KP kp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 P = pos[i], V = vel[i], F = f[i], FO = fold[i];
    f3 r = mk(P), v = mk(V);
    const f3 ff = mk(F), fo = mk(FO);
    const f3 dv = 0.5f * (ff + fo) * dt;
    v += dv;
    const f3 dr = v * dt + 0.5f * ff * dt * dt;
    r += dr;
    apply_boundaries(kp.p, r, v, dr, kp.p.particleRadius);
    pos[i] = f4(r, P.w);
    vel[i] = f4(v, V.w);
    f[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    fold[i] = f4(ff, FO.w);
}

// solvent_integrator (particles_kernel_impl.cuh:145-205): ballistic streaming.
__global__ __launch_bounds__(kBlock) void k_integrate_solvent(float4 *__restrict__ pos, float4 *__restrict__ vel,
                                                              float dt, uint32_t start, uint32_t ns, KP kp) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ns) return;
    const uint32_t i = start + t;
    const float4 P = pos[i], V = vel[i];
    f3 r = mk(P), v = mk(V);
    const f3 dr = dt * v;
    r += dr;
    apply_boundaries(kp.p, r, v, dr, 0.0f);
    pos[i] = f4(r, P.w);
    vel[i] = f4(v, V.w);
}

// ===========================================================================
// Hash (particles_kernel_impl.cuh:312-342)
// ===========================================================================
__global__ __launch_bounds__(kBlock) void k_calc_hash(uint32_t *__restrict__ hash, uint32_t *__restrict__ index,
                                                      const float4 *__restrict__ pos, uint32_t n, int srd, KP kp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 P = pos[i];
    const f3 q = mk(P);
    uint32_t h;
    if (srd) {
        h = srd_hash(kp.p, kp.s, q);
    } else {
        const int gx = cell_coord(q.x, kp.p.origin.x, kp.p.cellSize.x);
        const int gy = cell_coord(q.y, kp.p.origin.y, kp.p.cellSize.y);
        const int gz = cell_coord(q.z, kp.p.origin.z, kp.p.cellSize.z);
        h = md_hash_wrapped(kp.p, gx, gy, gz);
    }
    hash[i] = h;
    index[i] = i;
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t *__restrict__ p, uint32_t v, uint64_t n) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ===========================================================================
// Reorder + cell bounds (particles_kernel_impl.cuh:346-424).  The previous
// hash is read straight from memory (a cache hit for the neighbouring lane).
// ===========================================================================
__global__ __launch_bounds__(kBlock) void k_reorder(uint32_t *__restrict__ cellStart, uint32_t *__restrict__ cellEnd,
                                                    float4 *__restrict__ sPos, float4 *__restrict__ sVel,
                                                    float4 *__restrict__ sTan, const uint32_t *__restrict__ hash,
                                                    const uint32_t *__restrict__ index,
                                                    const float4 *__restrict__ pos, const float4 *__restrict__ vel,
                                                    const float4 *__restrict__ tan, uint32_t n, uint32_t numCells) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t h = hash[i];
    const uint32_t prev = i > 0 ? hash[i - 1] : 0u;
    if (i == 0 || h != prev) {
        if (h < numCells) cellStart[h] = i;
        else atomicOr(&g_flags, 2u);
        if (i > 0 && prev < numCells) cellEnd[prev] = i;
    }
    if (i == n - 1) {
        if (h < numCells) cellEnd[h] = i + 1;
    }
    const uint32_t s = index[i];
    sPos[i] = pos[s];
    if (vel != nullptr && sVel != nullptr) sVel[i] = vel[s];
    if (tan != nullptr && sTan != nullptr) sTan[i] = tan[s];
}

// ===========================================================================
// Pair forces (collideFilamentsKernel :694-744, collideCell :492-556,
// collideParticles :427-462, bondHookean :466-473, extensileForce :477-487).
//
// The 27 reference offsets are folded per axis into distinct wrapped cells
// with multiplicities (a grid one cell deep in z visits that layer 3 times
// upstream: multiplicity 3).  A cell row y..z is scanned as one contiguous
// run of the sorted arrays when its three x-cells are consecutive hashes.
// ===========================================================================
template <int MODE>  // 0: reference DEM/bond/extensile, 1: WCA (+extensile)
__global__ __launch_bounds__(kBlock) void k_collide(float4 *__restrict__ force, const float4 *__restrict__ sPos,
                                                    const float4 *__restrict__ sVel,
                                                    const float4 *__restrict__ sTan,
                                                    const uint32_t *__restrict__ sIdx,
                                                    const uint32_t *__restrict__ cellStart,
                                                    const uint32_t *__restrict__ cellEnd, uint32_t n, KP kp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SimParams &p = kp.p;
    const f3 r = mk(sPos[i]);
    const bool useVel = MODE == 0;
    const bool useTan = MODE == 0 || p.activity != 0.0f;
    const f3 v = useVel ? mk(sVel[i]) : mk(0.f, 0.f, 0.f);
    const f3 t = useTan ? mk(sTan[i]) : mk(0.f, 0.f, 0.f);
    const uint32_t myIdx = sIdx[i];
    uint32_t fil, ch;
    worm_of(kp, myIdx, fil, ch);

    const int gx = cell_coord(r.x, p.origin.x, p.cellSize.x);
    const int gy = cell_coord(r.y, p.origin.y, p.cellSize.y);
    const int gz = cell_coord(r.z, p.origin.z, p.cellSize.z);
    const bool twoD = kp.e.stencil == NW_STENCIL_2D;
    const AxisCells ax = axis_cells(gx, p.gridSize.x, false);
    const AxisCells ay = axis_cells(gy, p.gridSize.y, false);
    const AxisCells az = axis_cells(gz, p.gridSize.z, twoD);
    const float rc2 = kp.e.wcaCutoff * kp.e.wcaCutoff;

    // the neighbour cells by selects over scalars (a runtime index into the
    // AxisCells arrays would put them in scratch memory)
    const int cx0 = ax.c[0], cx1 = ax.c[1], cx2 = ax.c[2], cy0 = ay.c[0], cy1 = ay.c[1], cy2 = ay.c[2];
    const int cz0 = az.c[0], cz1 = az.c[1], cz2 = az.c[2];
    const float wx0 = ax.w[0], wx1 = ax.w[1], wx2 = ax.w[2], wy0 = ay.w[0], wy1 = ay.w[1], wy2 = ay.w[2];
    const float wz0 = az.w[0], wz1 = az.w[1], wz2 = az.w[2];
    f3 force_acc = mk(0.f, 0.f, 0.f);
    for (int zi = 0; zi < az.n; ++zi) {
        const int cz = zi == 0 ? cz0 : (zi == 1 ? cz1 : cz2);
        const float wz = zi == 0 ? wz0 : (zi == 1 ? wz1 : wz2);
        for (int yi = 0; yi < ay.n; ++yi) {
            const int cy = yi == 0 ? cy0 : (yi == 1 ? cy1 : cy2);
            const float wy = yi == 0 ? wy0 : (yi == 1 ? wy1 : wy2);
            const uint32_t rowBase = (uint32_t(cz) * p.gridSize.y + uint32_t(cy)) * p.gridSize.x;
            for (int xi = 0; xi < ax.n; ++xi) {
                const float wgt = wz * wy * (xi == 0 ? wx0 : (xi == 1 ? wx1 : wx2));
                const uint32_t h = rowBase + uint32_t(xi == 0 ? cx0 : (xi == 1 ? cx1 : cx2));
                const uint32_t start = cellStart[h];
                if (start == 0xffffffffu) continue;
                const uint32_t end = cellEnd[h];
                f3 fc = mk(0.f, 0.f, 0.f);
                for (uint32_t j = start; j < end; ++j) {
                    if (j == i) continue;
                    const f3 r2 = mk(sPos[j]);
                    const uint32_t idx2 = sIdx[j];
                    uint32_t fil2, ch2;
                    worm_of(kp, idx2, fil2, ch2);
                    const bool same = fil == fil2;
                    if (MODE == 0) {
                        const f3 v2 = mk(sVel[j]);
                        if (same && ((ch + 1u == ch2) || (ch - 1u == ch2))) {
                            fc += bond_hookean(p, r, r2, p.bondSpringK, p.bondSpringL);
                        } else {
                            fc += dem_force(p, r, r2, v, v2);
                        }
                    } else {
                        if (!same) fc += wca_force(p, kp.e, r, r2, rc2);
                    }
                    if (useTan && !same) {
                        const f3 t2 = mk(sTan[j]);
                        fc += extensile_force(p, r, r2, t, t2, p.activity);
                    }
                }
                if (wgt == 1.0f) force_acc += fc;
                else force_acc += wgt * fc;
            }
        }
    }
    float4 F = force[myIdx];
    F.x += force_acc.x;
    F.y += force_acc.y;
    F.z += force_acc.z;
    F.w += 0.0f;
    force[myIdx] = F;
}

// ===========================================================================
// Bonded forces (filamentKernel :751-818).  One bead per lane; whole worms
// are packed into each 64-lane wavefront (floor(64/n) worms, the remaining
// lanes idle) and chain neighbours come from cross-lane shuffles instead of
// re-reading memory.  Lane p computes the bending triple (p, p+1, p+2) and
// the bond (p, p+1); the shares of beads p+1, p+2 are shuffled back.
// ===========================================================================
__global__ __launch_bounds__(kBlock) void k_filament_wave(float4 *__restrict__ force, float4 *__restrict__ tangent,
                                                          const float4 *__restrict__ pos, uint32_t W, KP kp) {
    const uint32_t n = kp.p.filamentSize;
    const WaveWorm m = wave_worm(n, W);
    const float4 P = m.active ? pos[m.bead] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 F = m.active ? force[m.bead] : make_float4(0.f, 0.f, 0.f, 0.f);
    f3 acc = mk(F.x, F.y, F.z), tg;
    const bool writeTan = bonded_lane(kp, mk(P), m.pp, acc, tg);
    if (!m.active) return;
    if (n < 3 && kp.e.bondedMode != NW_BONDED_CHAIN) return;
    force[m.bead] = make_float4(acc.x, acc.y, acc.z, F.w + 0.0f);
    if (writeTan) tangent[m.bead] = f4(tg, 0.0f);
}

// Worms longer than a wavefront: one thread per worm walks the chain.
__global__ __launch_bounds__(kBlock) void k_filament_serial(float4 *__restrict__ force, float4 *__restrict__ tangent,
                                                            const float4 *__restrict__ pos, uint32_t W, KP kp) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    const SimParams &p = kp.p;
    const uint32_t n = p.filamentSize;
    const uint32_t head = w * n;
    if (n >= 3) {
        for (uint32_t q = 0; q + 2u < n; ++q) {
            const uint32_t i = head + q;
            const f3 ri = mk(pos[i]), rj = mk(pos[i + 1]), rk = mk(pos[i + 2]);
            const f3 r_ij = min_image(p, ri, rj), r_jk = min_image(p, rj, rk);
            const float d_ij = length(r_ij);
            const float d_jk = length(r_jk);
            const float d_ij_jk = dot(r_ij, r_jk);
            const float A = p.bondBendingK / (d_ij * d_jk);
            const f3 B = r_ij * d_ij_jk / dot(r_ij, r_ij);
            const f3 C = r_jk * d_ij_jk / dot(r_jk, r_jk);
            const f3 fi = -A * (r_jk - B);
            const f3 fk = -A * (C - r_ij);
            const f3 fj = mk(-fi.x - fk.x, -fi.y - fk.y, -fi.z - fk.z);
            float4 a = force[i], b = force[i + 1], c = force[i + 2];
            force[i] = make_float4(a.x + fi.x, a.y + fi.y, a.z + fi.z, a.w + 0.0f);
            force[i + 1] = make_float4(b.x + fj.x, b.y + fj.y, b.z + fj.z, b.w + 0.0f);
            force[i + 2] = make_float4(c.x + fk.x, c.y + fk.y, c.z + fk.z, c.w + 0.0f);
            tangent[i + 1] = f4(tangent_of(p, ri, rk), 0.0f);
            if (q == 0) tangent[i] = f4(tangent_of(p, ri, rj), 0.0f);
            else if (q == n - 3) tangent[i + 2] = f4(tangent_of(p, rj, rk), 0.0f);
        }
    }
    if (kp.e.bondedMode == NW_BONDED_CHAIN && n >= 2) {
        for (uint32_t q = 0; q + 1u < n; ++q) {
            const uint32_t i = head + q;
            const f3 d = min_image(p, mk(pos[i]), mk(pos[i + 1]));
            const float len = length(d);
            const f3 fb = (p.bondSpringK * (len - p.bondSpringL) / len) * d;
            const f3 dv = (kp.e.driveForce / len) * d;
            float4 a = force[i], b = force[i + 1];
            force[i] = make_float4(a.x + fb.x + dv.x, a.y + fb.y + dv.y, a.z + fb.z + dv.z, a.w);
            force[i + 1] = make_float4(b.x - fb.x, b.y - fb.y, b.z - fb.z, b.w);
        }
    }
}

// ===========================================================================
// Langevin (langevinKernel :859-874) with the normals drawn in registers:
// bead i uses generator block offset + i, i.e. exactly the four values that
// randomizeGaussian(buf, 4N, 0, stddev) places in buf[4i..4i+3].
// ===========================================================================
__global__ __launch_bounds__(kBlock) void k_langevin(float4 *__restrict__ force, const float4 *__restrict__ vel,
                                                     float stddev, uint32_t n, uint64_t seed, uint64_t offset,
                                                     KP kp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u4 b = philox_block(seed, offset + i);
    float z0, z1, z2, z3;
    box_muller(b.x, b.y, z0, z1);
    box_muller(b.z, b.w, z2, z3);
    (void)z3;
    const f3 xi = mk(z0 * stddev + 0.0f, z1 * stddev + 0.0f, z2 * stddev + 0.0f);
    f3 f = mk(0.f, 0.f, 0.f);
    f -= kp.p.gamma * mk(vel[i]);
    f += xi;
    float4 F = force[i];
    force[i] = make_float4(F.x + f.x, F.y + f.y, F.z + f.z, F.w + 0.0f);
}

// randomizeUniform / randomizeGaussian: element e comes from block
// offset + e/4, lane e%4 (normals: Box-Muller on lanes (0,1) and (2,3)).
__global__ __launch_bounds__(kBlock) void k_randomize(float *__restrict__ out, uint32_t n, int gaussian, float mean,
                                                      float stddev, uint64_t seed, uint64_t offset) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t e0 = uint64_t(t) * 4u;
    if (e0 >= n) return;
    const u4 b = philox_block(seed, offset + t);
    float v[4];
    if (gaussian) {
        box_muller(b.x, b.y, v[0], v[1]);
        box_muller(b.z, b.w, v[2], v[3]);
        for (int k = 0; k < 4; ++k) v[k] = v[k] * stddev + mean;
    } else {
        v[0] = u01(b.x); v[1] = u01(b.y); v[2] = u01(b.z); v[3] = u01(b.w);
    }
    for (int k = 0; k < 4; ++k)
        if (e0 + k < n) out[e0 + k] = v[k];
}

// ===========================================================================
// SRD (particles_kernel_impl.cuh:559-691).  One thread per SRD cell walks its
// sorted segment in order, so sums follow the reference's summation order.
// ===========================================================================
__global__ __launch_bounds__(kBlock) void k_srd_com(float4 *__restrict__ com, const float4 *__restrict__ sVel,
                                                    const uint32_t *__restrict__ cellStart,
                                                    const uint32_t *__restrict__ cellEnd, uint32_t numCells) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= numCells) return;
    const uint32_t start = cellStart[c];
    f3 m = mk(0.f, 0.f, 0.f);
    if (start != 0xffffffffu) {
        const uint32_t end = cellEnd[c];
        float M = 0.0f;
        for (uint32_t i = start; i < end; ++i) {
            const float4 V = sVel[i];
            m += V.w * mk(V);
            M += V.w;
        }
        if (M > 0.0f) m = m * (1.0f / M);
    }
    com[c] = f4(m, 0.0f);
}

__global__ __launch_bounds__(kBlock) void k_srd_isokinetic(float4 *__restrict__ vel, float4 *__restrict__ sVel,
                                                           const uint32_t *__restrict__ cellStart,
                                                           const uint32_t *__restrict__ cellEnd,
                                                           const uint32_t *__restrict__ index, uint32_t numCells,
                                                           KP kp) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= numCells) return;
    const uint32_t start = cellStart[c];
    if (start == 0xffffffffu) return;
    const uint32_t end = cellEnd[c];
    float E = 0.0f;
    for (uint32_t i = start; i < end; ++i) {
        const float4 V = sVel[i];
        const f3 v = mk(V);
        E += 0.5f * V.w * dot(v, v);
    }
    const float lambda = sqrtf(kp.s.kbT / E);
    for (uint32_t i = start; i < end; ++i) {
        const uint32_t idx = index[i];
        const float4 V = vel[idx];
        const float4 nv = f4(lambda * mk(V), sVel[i].w);
        vel[idx] = nv;
        sVel[i] = nv;
    }
}

__global__ __launch_bounds__(kBlock) void k_collide_solvent(const float4 *__restrict__ pos, float4 *__restrict__ vel,
                                                            const float4 *__restrict__ com, uint32_t numCells,
                                                            uint32_t n, float cosA, float sinA, uint64_t seed,
                                                            uint64_t offset, KP kp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!kp.e.srdCollideFilaments && i < kp.p.numParticles) return;
    const f3 r = mk(pos[i]);
    const float4 V = vel[i];
    const f3 v = mk(V);
    const uint32_t cell = srd_hash(kp.p, kp.s, r);
    if (cell >= numCells) return;
    const f3 u = mk(com[cell]);
    const float uni = uniform_at(seed, offset, cell);
    const float sgn = uni > 0.5f ? 1.0f : -1.0f;
    const float s = sgn * sinA;
    const f3 d = v - u;
    const f3 rot = mk(d.x * cosA - d.y * s, d.x * s + d.y * cosA, d.z);
    const f3 nv = u + rot;
    vel[i] = f4(nv, V.w);
}

// ===========================================================================
// Polarity reversal (reverseFilamentsKernel :820-857)
// ===========================================================================
__global__ __launch_bounds__(kBlock) void k_reverse(float4 *__restrict__ force, float4 *__restrict__ forceOld,
                                                    float4 *__restrict__ vel, float4 *__restrict__ pos, uint32_t W,
                                                    uint64_t seed, uint64_t offset, KP kp) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    if (uniform_at(seed, offset, w) > kp.p.reverseProbability) return;
    const uint32_t n = kp.p.filamentSize;
    const uint32_t head = w * n;
    for (uint32_t q = 0; q < n / 2; ++q) {
        const uint32_t i = head + q, j = head + n - 1 - q;
        float4 t;
        t = pos[i]; pos[i] = pos[j]; pos[j] = t;
        t = vel[i]; vel[i] = vel[j]; vel[j] = t;
        t = force[i]; force[i] = force[j]; force[j] = t;
        t = forceOld[i]; forceOld[i] = forceOld[j]; forceOld[j] = t;
    }
}

// ===========================================================================
// launchers
// ===========================================================================
hipError_t launch_integrate_filaments(float4 *pos, float4 *vel, float4 *f, float4 *fold, float dt, uint32_t n,
                                      const KP &kp, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_integrate_filaments, dim3(grid_for(n)), dim3(kBlock), 0, st, pos, vel, f, fold, dt, n, kp);
    return hipGetLastError();
}

hipError_t launch_integrate_solvent(float4 *pos, float4 *vel, float dt, uint32_t start, uint32_t ns,
                                    const KP &kp, hipStream_t st) {
    if (ns == 0) return hipSuccess;
    hipLaunchKernelGGL(k_integrate_solvent, dim3(grid_for(ns)), dim3(kBlock), 0, st, pos, vel, dt, start, ns, kp);
    return hipGetLastError();
}

hipError_t launch_calc_hash(uint32_t *hash, uint32_t *index, const float4 *pos, uint32_t n, bool srd,
                            const KP &kp, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_calc_hash, dim3(grid_for(n)), dim3(kBlock), 0, st, hash, index, pos, n, int(srd), kp);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t value, uint64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(kBlock), 0, st, p, value, n);
    return hipGetLastError();
}

hipError_t launch_reorder(uint32_t *cellStart, uint32_t *cellEnd, float4 *sPos, float4 *sVel, float4 *sTan,
                          const uint32_t *hash, const uint32_t *index, const float4 *pos, const float4 *vel,
                          const float4 *tan, uint32_t n, uint32_t numCells, hipStream_t st) {
    hipError_t e = launch_fill_u32(cellStart, 0xffffffffu, numCells, st);
    if (e != hipSuccess || n == 0) return e;
    hipLaunchKernelGGL(k_reorder, dim3(grid_for(n)), dim3(kBlock), 0, st, cellStart, cellEnd, sPos, sVel, sTan, hash,
                       index, pos, vel, tan, n, numCells);
    return hipGetLastError();
}

hipError_t launch_collide_filaments(float4 *force, const float4 *sPos, const float4 *sVel, const float4 *sTan,
                                    const uint32_t *sIdx, const uint32_t *cellStart, const uint32_t *cellEnd,
                                    uint32_t n, uint32_t numCells, const KP &kp, hipStream_t st) {
    (void)numCells;
    if (n == 0) return hipSuccess;
    if (kp.e.pairPotential == NW_PAIR_WCA)
        hipLaunchKernelGGL(k_collide<1>, dim3(grid_for(n)), dim3(kBlock), 0, st, force, sPos, sVel, sTan, sIdx,
                           cellStart, cellEnd, n, kp);
    else
        hipLaunchKernelGGL(k_collide<0>, dim3(grid_for(n)), dim3(kBlock), 0, st, force, sPos, sVel, sTan, sIdx,
                           cellStart, cellEnd, n, kp);
    return hipGetLastError();
}

hipError_t launch_filament_forces(float4 *force, float4 *tangent, const float4 *pos, uint32_t W, const KP &kp,
                                  hipStream_t st) {
    const uint32_t n = kp.p.filamentSize;
    if (W == 0 || n == 0) return hipSuccess;
    if (n <= 64) {
        const uint32_t wpw = 64u / n;
        const uint64_t waves = (uint64_t(W) + wpw - 1) / wpw;
        const uint64_t threads = waves * 64u;
        hipLaunchKernelGGL(k_filament_wave, dim3(grid_for(threads)), dim3(kBlock), 0, st, force, tangent, pos, W, kp);
    } else {
        hipLaunchKernelGGL(k_filament_serial, dim3(grid_for(W)), dim3(kBlock), 0, st, force, tangent, pos, W, kp);
    }
    return hipGetLastError();
}

hipError_t launch_langevin(float4 *force, const float4 *vel, float stddev, uint32_t n, uint64_t seed,
                           uint64_t offset, const KP &kp, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_langevin, dim3(grid_for(n)), dim3(kBlock), 0, st, force, vel, stddev, n, seed, offset, kp);
    return hipGetLastError();
}

hipError_t launch_randomize(float *out, uint32_t n, int gaussian, float mean, float stddev, uint64_t seed,
                            uint64_t offset, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint64_t blocks = (uint64_t(n) + 3) / 4;
    hipLaunchKernelGGL(k_randomize, dim3(grid_for(blocks)), dim3(kBlock), 0, st, out, n, gaussian, mean, stddev,
                       seed, offset);
    return hipGetLastError();
}

hipError_t launch_srd_com(float4 *com, const float4 *sVel, const uint32_t *cellStart, const uint32_t *cellEnd,
                          uint32_t numCells, hipStream_t st) {
    if (numCells == 0) return hipSuccess;
    hipLaunchKernelGGL(k_srd_com, dim3(grid_for(numCells)), dim3(kBlock), 0, st, com, sVel, cellStart, cellEnd,
                       numCells);
    return hipGetLastError();
}

hipError_t launch_srd_isokinetic(float4 *vel, float4 *sVel, const uint32_t *cellStart, const uint32_t *cellEnd,
                                 const uint32_t *index, uint32_t numCells, const KP &kp, hipStream_t st) {
    if (numCells == 0) return hipSuccess;
    hipLaunchKernelGGL(k_srd_isokinetic, dim3(grid_for(numCells)), dim3(kBlock), 0, st, vel, sVel, cellStart,
                       cellEnd, index, numCells, kp);
    return hipGetLastError();
}

hipError_t launch_collide_solvent(const float4 *pos, float4 *vel, const float4 *com, uint32_t numCells, uint32_t n,
                                  float cosA, float sinA, uint64_t seed, uint64_t offset, const KP &kp,
                                  hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_collide_solvent, dim3(grid_for(n)), dim3(kBlock), 0, st, pos, vel, com, numCells, n, cosA,
                       sinA, seed, offset, kp);
    return hipGetLastError();
}

hipError_t launch_reverse_filaments(float4 *force, float4 *forceOld, float4 *vel, float4 *pos, uint32_t W,
                                    uint64_t seed, uint64_t offset, const KP &kp, hipStream_t st) {
    if (W == 0) return hipSuccess;
    hipLaunchKernelGGL(k_reverse, dim3(grid_for(W)), dim3(kBlock), 0, st, force, forceOld, vel, pos, W, seed, offset,
                       kp);
    return hipGetLastError();
}

}  // namespace nw

// nw_physics.hpp -- device physics shared by the per-call kernels
// (nw_kernels.hip) and the fused step kernels (nw_fused.hip): boundaries,
// pair terms, stencil folding, cross-lane shuffles, per-element uniforms.
#pragma once

#include "nw_device.hpp"

namespace nw {

// ---------------------------------------------------------------------------
// boundaries (particles_kernel_impl.cuh:34-68)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void periodic_bc(float &r, float L0, float L) {
    if (r < L0) r += L;
    else if (r > L + L0) r -= L;
}

__device__ __forceinline__ void damped_wall(float &r, float &v, float coef, float L0, float L, float R) {
    if (r < L0 + R) {
        r = L0 + R;
        v *= coef;
    } else if (r > (L + L0) - R) {
        r = (L + L0) - R;
        v *= coef;
    }
}

// noslipWallBoundary: undo the whole step displacement, pin the axis at the
// wall and keep only the reflected normal velocity component.
__device__ __forceinline__ void noslip_wall(f3 &pos, int axis, f3 &vel, f3 dr, f3 N, float L0, float L, float R) {
    float r = axis == 0 ? pos.x : pos.y;
    bool hit = false;
    float target = 0.0f;
    if (r < L0 + R) {
        hit = true;
        target = L0 + R;
    } else if (r > (L + L0) - R) {
        hit = true;
        target = (L + L0) - R;
    }
    if (hit) {
        pos -= dr;
        if (axis == 0) pos.x = target; else pos.y = target;
        vel = mk(-vel.x * N.x, -vel.y * N.y, -vel.z * N.z);
    }
}

__device__ __forceinline__ void apply_boundaries(const SimParams &p, f3 &pos, f3 &vel, f3 dr, float R) {
    if (p.boundaryX == PERIODIC) periodic_bc(pos.x, p.origin.x, p.boxSize.x);
    else if (p.boundaryX == WALL) damped_wall(pos.x, vel.x, p.boundaryDamping, p.origin.x, p.boxSize.x, R);
    else if (p.boundaryX == WALL_NO_SLIP)
        noslip_wall(pos, 0, vel, dr, mk(1.f, 0.f, 0.f), p.origin.x, p.boxSize.x, R);

    if (p.boundaryY == PERIODIC) periodic_bc(pos.y, p.origin.y, p.boxSize.y);
    else if (p.boundaryY == WALL) damped_wall(pos.y, vel.y, p.boundaryDamping, p.origin.y, p.boxSize.y, R);
    else if (p.boundaryY == WALL_NO_SLIP)
        noslip_wall(pos, 1, vel, dr, mk(0.f, 1.f, 0.f), p.origin.y, p.boxSize.y, R);

    // the reference has no no-slip branch for z
    if (p.boundaryZ == PERIODIC) periodic_bc(pos.z, p.origin.z, p.boxSize.z);
    else if (p.boundaryZ == WALL) damped_wall(pos.z, vel.z, p.boundaryDamping, p.origin.z, p.boxSize.z, R);
}

// ---------------------------------------------------------------------------
// pair terms and the neighbour stencil (particles_kernel_impl.cuh:427-556)
// ---------------------------------------------------------------------------
struct AxisCells {
    int c[3];
    float w[3];
    int n;
};

__device__ __forceinline__ AxisCells axis_cells(int g, uint32_t G, bool onlyCentre) {
    AxisCells a;
    if (onlyCentre) {
        a.c[0] = wrap_mod(g, G);
        a.w[0] = 1.0f;
        a.n = 1;
        return a;
    }
    const int m = wrap_mod(g - 1, G), c = wrap_mod(g, G), p = wrap_mod(g + 1, G);
    if (G >= 3) {
        a.c[0] = m; a.c[1] = c; a.c[2] = p;
        a.w[0] = a.w[1] = a.w[2] = 1.0f;
        a.n = 3;
    } else if (G == 2) {
        // offsets -1 and +1 land on the same cell
        a.c[0] = m; a.w[0] = 2.0f;
        a.c[1] = c; a.w[1] = 1.0f;
        a.n = 2;
    } else {
        a.c[0] = c; a.w[0] = 3.0f;
        a.n = 1;
    }
    return a;
}

// axis_cells of a coordinate already in [0, G): the neighbours wrap by
// selects (the same cells as wrap_mod gives)
__device__ __forceinline__ AxisCells axis_cells_in(int g, uint32_t G, bool onlyCentre) {
    AxisCells a = {{g, g, g}, {0.f, 0.f, 0.f}, 0};  // unused entries: the centre cell, weight 0
    if (onlyCentre) {
        a.c[0] = g;
        a.w[0] = 1.0f;
        a.n = 1;
        return a;
    }
    const int Gi = int(G);
    const int m = g == 0 ? Gi - 1 : g - 1, c = g, p = g + 1 == Gi ? 0 : g + 1;
    if (G >= 3) {
        a.c[0] = m; a.c[1] = c; a.c[2] = p;
        a.w[0] = a.w[1] = a.w[2] = 1.0f;
        a.n = 3;
    } else if (G == 2) {
        a.c[0] = m; a.w[0] = 2.0f;
        a.c[1] = c; a.w[1] = 1.0f;
        a.n = 2;
    } else {
        a.c[0] = c; a.w[0] = 3.0f;
        a.n = 1;
    }
    return a;
}

// DEM soft sphere (collideParticles)
__device__ __forceinline__ f3 dem_force(const SimParams &p, f3 rA, f3 rB, f3 vA, f3 vB) {
    const f3 rel = min_image_pair(p, rA, rB);
    const float dist = length(rel);
    const float cd = p.particleRadius + p.particleRadius;
    f3 force = mk(0.f, 0.f, 0.f);
    if (dist < cd) {
        const f3 nrm = rel / dist;
        const f3 relVel = vB - vA;
        const f3 tanVel = relVel - (dot(relVel, nrm) * nrm);
        force = -p.spring * (cd - dist) * nrm;
        force += p.damping * relVel;
        force += p.shear * tanVel;
    }
    return force;
}

__device__ __forceinline__ f3 bond_hookean(const SimParams &p, f3 rA, f3 rB, float k, float L) {
    const f3 rel = min_image_pair(p, rA, rB);
    const float dist = length(rel);
    const f3 nrm = rel / dist;
    return -k * (L - dist) * nrm;
}

__device__ __forceinline__ f3 extensile_force(const SimParams &p, f3 rA, f3 rB, f3 tA, f3 tB, float A) {
    const f3 rel = min_image_pair(p, rA, rB);
    const float dist = length(rel);
    if (dot(tA, tB) < -0.5f) return -A * (tA - tB) / (2.0f * dist);
    return mk(0.f, 0.f, 0.f);
}

// 24 eps (2 s^12/r^14 - s^6/r^8) at rr = r^2, with the hardware reciprocal
// (1 ulp): the force is a product of powers of 1/r^2
__device__ __forceinline__ float wca_mag(const NWExtParams &e, float rr) {
    const float inv2 = __builtin_amdgcn_rcpf(rr);
    const float s2 = e.wcaSigma * e.wcaSigma * inv2;
    const float s6 = s2 * s2 * s2;
    return 24.0f * e.wcaEpsilon * inv2 * s6 * (2.0f * s6 - 1.0f);
}

// WCA-LJ force on A from B: F_A = -24 eps (2 s^12/r^14 - s^6/r^8) (rB - rA), r < rc
__device__ __forceinline__ f3 wca_force(const SimParams &p, const NWExtParams &e, f3 rA, f3 rB, float rc2) {
    const f3 rel = min_image_pair(p, rA, rB);
    const float rr = dot(rel, rel);
    if (rr < rc2) return -wca_mag(e, rr) * rel;
    return mk(0.f, 0.f, 0.f);
}

// ---------------------------------------------------------------------------
// cross-lane exchange of chain neighbours
// ---------------------------------------------------------------------------
__device__ __forceinline__ f3 shfl_down3(f3 a, int d) {
    return mk(__shfl_down(a.x, d, 64), __shfl_down(a.y, d, 64), __shfl_down(a.z, d, 64));
}
__device__ __forceinline__ f3 shfl_up3(f3 a, int d) {
    return mk(__shfl_up(a.x, d, 64), __shfl_up(a.y, d, 64), __shfl_up(a.z, d, 64));
}

// ---------------------------------------------------------------------------
// Whole worms packed per 64-lane wavefront: floor(64/n) worms per wave, bead
// p of worm w on lane (w mod wpw) * n + p, the remaining lanes idle.
// ---------------------------------------------------------------------------
struct WaveWorm {
    uint32_t lane, wl, pp, worm, bead;
    bool active;
};

// lanes of wavefront `wave` of a system of W worms of n beads
__device__ __forceinline__ WaveWorm wave_worm_at(uint32_t wave, uint32_t n, uint32_t W) {
    WaveWorm m;
    m.lane = threadIdx.x & 63u;
    const uint32_t wpw = 64u / n;
    m.wl = m.lane / n;
    m.pp = m.lane - m.wl * n;
    m.worm = wave * wpw + m.wl;
    m.active = m.wl < wpw && m.worm < W;
    m.bead = m.worm * n + m.pp;
    return m;
}

__device__ __forceinline__ WaveWorm wave_worm(uint32_t n, uint32_t W) {
    WaveWorm m;
    m.lane = threadIdx.x & 63u;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t wpw = 64u / n;
    m.wl = m.lane / n;
    m.pp = m.lane - m.wl * n;
    m.worm = wave * wpw + m.wl;
    m.active = m.wl < wpw && m.worm < W;
    m.bead = m.worm * n + m.pp;
    return m;
}

// Bonded forces on bead pp (filamentKernel, particles_kernel_impl.cuh:751-818,
// and in chain mode the archived harmonic bonds + drive) added to acc in the
// reference's order: triple pp-2 (its k share), triple pp-1 (j share), triple
// pp (i share), then -bond(pp-1), +bond(pp), +drive.  Lane pp evaluates the
// triple (pp, pp+1, pp+2) and the bond (pp, pp+1) once; the shares of the
// other beads arrive by shuffles.  Every lane of the wave must call this.
// Returns whether the tangent tg is written (the reference writes none for
// n < 3 and no tail tangent for n == 3).
// wantTan = false: the tangent is neither computed nor reported (a caller
// that stores none)
__device__ __forceinline__ bool bonded_lane(const KP &kp, f3 r, uint32_t pp, f3 &acc, f3 &tg, bool wantTan = true) {
    const SimParams &p = kp.p;
    const uint32_t n = p.filamentSize;
    const f3 rn1 = shfl_down3(r, 1);
    const f3 rn2 = shfl_down3(r, 2);
    const f3 rp1 = shfl_up3(r, 1);

    f3 fi = mk(0.f, 0.f, 0.f), fj = fi, fk = fi;
    const bool hasTriple = n >= 3 && pp + 2u < n;
    if (hasTriple) {
        // min_image_pair: min_image's value, the wrap tests only when some
        // lane's chain spans half the box
        const f3 r_ij = min_image_pair(p, r, rn1);
        const f3 r_jk = min_image_pair(p, rn1, rn2);
        const float d_ij = length(r_ij);
        const float d_jk = length(r_jk);
        const float d_ij_jk = dot(r_ij, r_jk);
        const float A = p.bondBendingK / (d_ij * d_jk);
        const f3 B = r_ij * d_ij_jk / dot(r_ij, r_ij);
        const f3 C = r_jk * d_ij_jk / dot(r_jk, r_jk);
        fi = -A * (r_jk - B);
        fk = -A * (C - r_ij);
        fj = mk(-fi.x - fk.x, -fi.y - fk.y, -fi.z - fk.z);
    }
    const f3 fk_from2 = shfl_up3(fk, 2);
    const f3 fj_from1 = shfl_up3(fj, 1);

    const bool chain = kp.e.bondedMode == NW_BONDED_CHAIN;
    f3 fb = mk(0.f, 0.f, 0.f), drive = fb;
    if (chain && pp + 1u < n) {
        const f3 d = min_image_pair(p, r, rn1);
        const float len = length(d);
        fb = (p.bondSpringK * (len - p.bondSpringL) / len) * d;
        drive = (kp.e.driveForce / len) * d;
    }
    const f3 fb_prev = shfl_up3(fb, 1);

    if (pp >= 2u) acc += fk_from2;
    if (pp >= 1u) acc += fj_from1;
    if (hasTriple) acc += fi;
    if (chain) {
        if (pp >= 1u) acc -= fb_prev;
        if (pp + 1u < n) {
            acc += fb;
            acc += drive;
        }
    }
    if (n < 3 || !wantTan) return false;
    if (pp == 0u) tg = tangent_of(p, r, rn1);
    else if (pp + 1u < n) tg = tangent_of(p, rp1, rn1);
    else {
        tg = tangent_of(p, rp1, r);
        return n > 3u;
    }
    return true;
}

// element e of a uniform draw at `offset` (lane e%4 of block offset + e/4)
__device__ __forceinline__ float uniform_at(uint64_t seed, uint64_t offset, uint32_t e) {
    const u4 b = philox_block(seed, offset + (e >> 2));
    const uint32_t k = e & 3u;
    const uint32_t x = k == 0 ? b.x : k == 1 ? b.y : k == 2 ? b.z : b.w;
    return u01(x);
}

}  // namespace nw

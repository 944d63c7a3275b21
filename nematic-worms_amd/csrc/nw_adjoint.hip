// nw_adjoint.hip -- the adjoint of the fused MD step (gfx950): reverse-mode
// derivatives of a loss of the final configuration with respect to the drive
// f_a, the bending stiffness kappa, the WCA epsilon and the Langevin gamma.
//
// One step maps (r, v, F, G) to
//   v' = v + (F + G) dt / 2,  r' = r + v' dt + F dt^2 / 2,  G' = F,
//   F' = -gamma v' + sigma(gamma) z + B(r'; kappa, f_a) + P(r'; epsilon)
// (z the replayed Philox normals, sigma = sqrt(2 gamma kT), B the bonded
// terms, P the WCA pairs; periodic wrapping has unit derivative).  With the
// adjoints (lr, lv, lF, lG) of the step's outputs:
//   R = lr + (dB/dr' + dP/dr')^T lF          adjoint of r'
//   V = lv - gamma lF + dt R                 adjoint of v'
//   lr <- R, lv <- V, lF <- lG + dt^2 R / 2 + dt V / 2, lG <- dt V / 2
//   grad += (dF'/dtheta)^T lF.
// Bending, bonds and WCA are conservative: their Jacobian is symmetric, so
// J^T lF = J lF, the derivative of the force along lF, evaluated with dual
// numbers through the forward expressions.  The drive is not; its transpose
// product is written out.  The bonded kernel runs whole worms per wavefront
// (chain neighbours by shuffles), the pair kernel walks the forward step's
// cell lists of r'.  Parameter partials are FP64 per workgroup, added in a
// fixed order at the end.
#include "nw_adjoint.hpp"
#include "nw_physics.hpp"

namespace nw {

// ---------------------------------------------------------------------------
// dual numbers: value and derivative along the adjoint direction
// ---------------------------------------------------------------------------
struct d1 {
    float v, t;
};
struct d3 {
    f3 v, t;
};

__device__ __forceinline__ d1 operator*(d1 a, d1 b) { return {a.v * b.v, a.t * b.v + a.v * b.t}; }
__device__ __forceinline__ d1 operator+(d1 a, d1 b) { return {a.v + b.v, a.t + b.t}; }
__device__ __forceinline__ d1 operator-(d1 a, d1 b) { return {a.v - b.v, a.t - b.t}; }
__device__ __forceinline__ d1 operator/(d1 a, d1 b) {
    const float q = a.v / b.v;
    return {q, (a.t - q * b.t) / b.v};
}
__device__ __forceinline__ d1 dsqrt(d1 a) {
    const float s = sqrtf(a.v);
    return {s, a.t / (2.0f * s)};
}
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return {a.v + b.v, a.t + b.t}; }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return {a.v - b.v, a.t - b.t}; }
__device__ __forceinline__ d3 operator-(d3 a) { return {mk(-a.v.x, -a.v.y, -a.v.z), mk(-a.t.x, -a.t.y, -a.t.z)}; }
__device__ __forceinline__ d3 operator*(d3 a, d1 s) { return {a.v * s.v, a.t * s.v + a.v * s.t}; }
__device__ __forceinline__ d3 operator*(d1 s, d3 a) { return a * s; }
__device__ __forceinline__ d3 operator/(d3 a, d1 s) {
    const f3 q = a.v / s.v;
    return {q, (a.t - q * s.t) / s.v};
}
__device__ __forceinline__ d1 ddot(d3 a, d3 b) { return {dot(a.v, b.v), dot(a.t, b.v) + dot(a.v, b.t)}; }
__device__ __forceinline__ d1 dconst(float c) { return {c, 0.0f}; }

// difference b - a under the minimum image (the wrap is piecewise constant)
__device__ __forceinline__ d3 dmin_image(const SimParams &p, d3 a, d3 b) {
    return {min_image(p, a.v, b.v), b.t - a.t};
}

__device__ __forceinline__ d3 dshfl_down(d3 a, int d) { return {shfl_down3(a.v, d), shfl_down3(a.t, d)}; }
__device__ __forceinline__ d3 dshfl_up(d3 a, int d) { return {shfl_up3(a.v, d), shfl_up3(a.t, d)}; }

// candidates of the adjoint's pair walk loaded together
constexpr int kAdjBatch = 4;

__device__ __forceinline__ double wave_sum_d(double v) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// the workgroup's FP64 partials (kAdjParams of them) into out[blockIdx]
__device__ __forceinline__ void block_partials(double (&acc)[kAdjParams], double *__restrict__ out) {
    __shared__ double ls[kBlock / 64][kAdjParams];
#pragma unroll
    for (int k = 0; k < kAdjParams; ++k) {
        const double s = wave_sum_d(acc[k]);
        if ((threadIdx.x & 63u) == 0u) ls[threadIdx.x >> 6][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < kAdjParams) {
        double s = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) s += ls[w][threadIdx.x];
        out[uint64_t(blockIdx.x) * kAdjParams + threadIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------
// One adjoint step, whole worms per wavefront (bead order):
//   bonded:  Rb = (dB/dr')^T lF by dual numbers through the forward bending
//            and bond expressions (chain neighbours by shuffles), the drive's
//            transpose product written out; kappa, f_a, gamma partials
//   pairs:   Rp = sum_j w H_ij (lF_i - lF_j), H = f I + 2 f' d d^T for
//            F_i = -f(|d|^2) d, d = r_j - r_i, over the forward step's cell
//            lists of r' (the bead's cell recomputed as the forward hashed
//            it; w the stencil weight of the row); epsilon partial
//   update:  R = lr + Rb + Rp, V = lv - gamma lF + dt R, then the new
//            (lr, lv, lF, lG) of the bead (lF into the other buffer: the
//            neighbours still read the old one)
// ---------------------------------------------------------------------------
// The lanes m of one system: a's bead arrays (pos, vel, lF, lFout, lr, lv,
// lG) and off start at the system's first bead and cell; ids and sPos are the
// step's sorted slots, whose ids exceed the system's bead index by idBase.
__device__ __forceinline__ void adj_step_lanes(const AdjStepArgs &a, const KP &kp, const WaveWorm &m,
                                               uint32_t idBase, double (&acc)[kAdjParams]) {
    const SimParams &p = kp.p;
    const uint32_t n = p.filamentSize;
    const f3 rv = mk(load4_or_zero(a.pos, m.bead, m.active));
    const f3 lam = mk(load4_or_zero(a.lF, m.bead, m.active));
    const d3 r = {rv, lam};
    const uint32_t pp = m.pp;

    // bending (filamentKernel): lane pp owns the triple (pp, pp+1, pp+2)
    const d3 rn1 = dshfl_down(r, 1), rn2 = dshfl_down(r, 2);
    d3 fi = {mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 0.f)}, fj = fi, fk = fi;
    const bool hasTriple = n >= 3 && pp + 2u < n;
    if (hasTriple) {
        const d3 r_ij = dmin_image(p, r, rn1);
        const d3 r_jk = dmin_image(p, rn1, rn2);
        const d1 d_ij = dsqrt(ddot(r_ij, r_ij));
        const d1 d_jk = dsqrt(ddot(r_jk, r_jk));
        const d1 d_ij_jk = ddot(r_ij, r_jk);
        const d1 A = dconst(p.bondBendingK) / (d_ij * d_jk);
        const d3 B = r_ij * d_ij_jk / ddot(r_ij, r_ij);
        const d3 C = r_jk * d_ij_jk / ddot(r_jk, r_jk);
        fi = -(A * (r_jk - B));
        fk = -(A * (C - r_ij));
        fj = -(fi + fk);
    }
    const d3 fk2 = dshfl_up(fk, 2), fj1 = dshfl_up(fj, 1);
    d3 bend = {mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 0.f)};
    if (pp >= 2u) bend = bend + fk2;
    if (pp >= 1u) bend = bend + fj1;
    if (hasTriple) bend = bend + fi;

    // chain bonds (conservative) and drive (its transpose product by hand)
    d3 fb = {mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 0.f)};
    f3 driveT = mk(0.f, 0.f, 0.f);  // share of bead pp+1: f_a Hd lF_pp
    f3 that = mk(0.f, 0.f, 0.f);    // unit bond (pp, pp+1)
    if (pp + 1u < n) {
        const d3 d = dmin_image(p, r, rn1);
        const d1 len = dsqrt(ddot(d, d));
        fb = (dconst(p.bondSpringK) * (len - dconst(p.bondSpringL)) / len) * d;
        that = d.v / len.v;
        // Hd lF = (lF - t (t . lF)) / len
        const f3 hl = (lam - that * dot(that, lam)) / len.v;
        driveT = kp.e.driveForce * hl;
    }
    const d3 fbPrev = dshfl_up(fb, 1);
    const f3 driveFromPrev = shfl_up3(driveT, 1);
    d3 bond = {mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 0.f)};
    if (pp >= 1u) bond = bond - fbPrev;
    if (pp + 1u < n) bond = bond + fb;

    if (m.active) {
        f3 R = bend.t + bond.t;
        if (pp + 1u < n) R -= driveT;
        if (pp >= 1u) R += driveFromPrev;
        // parameter partials lF . dF/dtheta (F the force this step produced)
        if (pp + 1u < n) acc[kAdjDrive] = double(dot(lam, that));
        if (p.bondBendingK != 0.0f) acc[kAdjKappa] = double(dot(lam, bend.v)) / double(p.bondBendingK);
        if (a.langevin) {
            const f3 v = mk(load4(a.vel, m.bead));
            const u4 b = philox_block(a.seed, a.offLgv + m.bead);
            float z0, z1, z2, z3;
            box_muller(b.x, b.y, z0, z1);
            box_muller(b.z, b.w, z2, z3);
            (void)z3;
            const f3 xi = mk(z0 * a.lgvStd + 0.0f, z1 * a.lgvStd + 0.0f, z2 * a.lgvStd + 0.0f);
            // F_lgv = -gamma v + sigma(gamma) z, d sigma / d gamma = sigma / (2 gamma)
            acc[kAdjGamma] = -double(dot(lam, v)) + double(dot(lam, xi)) / (2.0 * double(a.gamma));
        }

        // WCA pairs over the cell lists of r'
        const uint32_t me = m.bead;
        const uint32_t Gx = p.gridSize.x, Gy = p.gridSize.y;
        const int gx = wrap_mod(cell_coord(rv.x, p.origin.x, p.cellSize.x), Gx);
        const int gy = wrap_mod(cell_coord(rv.y, p.origin.y, p.cellSize.y), Gy);
        const int gz = wrap_mod(cell_coord(rv.z, p.origin.z, p.cellSize.z), p.gridSize.z);
        const AxisCells ay = axis_cells_in(gy, Gy, false);
        const float wz = kp.e.stencil == NW_STENCIL_2D ? 1.0f : axis_cells_in(gz, p.gridSize.z, false).w[0];
        const float rc2 = kp.e.wcaCutoff * kp.e.wcaCutoff;
        const float eps = kp.e.wcaEpsilon, s2c = kp.e.wcaSigma * kp.e.wcaSigma;
        const uint32_t fil = m.worm;
        f3 Rp = mk(0.f, 0.f, 0.f);
        double epsPart = 0.0;
        // the six candidate runs (two per stencil row, the row's three cells
        // with the periodic wrap), flattened; candidates in groups of
        // kAdjBatch with all their loads in flight together (every slot of a
        // run holds a bead of this system, so a group past the last candidate
        // reloads the last one and discards it)
        uint32_t ps[6], pl[6];
        float pw[6];
#pragma unroll
        for (int yi = 0; yi < 3; ++yi) {
            ps[2 * yi] = pl[2 * yi] = ps[2 * yi + 1] = pl[2 * yi + 1] = 0u;
            pw[2 * yi] = pw[2 * yi + 1] = 0.0f;
            if (yi < ay.n) {
                const uint32_t base = (uint32_t(gz) * Gy + uint32_t(ay.c[yi])) * Gx;
                const bool mid = gx >= 1 && gx + 1 < int(Gx), left = gx == 0;
                const uint32_t ia0 = mid ? base + gx - 1 : (left ? base + Gx - 1 : base + Gx - 2);
                const uint32_t ia1 = mid ? base + gx + 2 : base + Gx;
                const uint32_t ib1 = mid ? base : (left ? base + 2 : base + 1);
                const uint32_t a0 = a.off[ia0], a1 = a.off[ia1], b0 = a.off[base], b1 = a.off[ib1];
                ps[2 * yi] = a0;
                pl[2 * yi] = a1 - a0;
                ps[2 * yi + 1] = b0;
                pl[2 * yi + 1] = b1 - b0;
                pw[2 * yi] = pw[2 * yi + 1] = wz * ay.w[yi];
            }
        }
        uint32_t pre[7];
        pre[0] = 0u;
#pragma unroll
        for (int k = 0; k < 6; ++k) pre[k + 1] = pre[k] + pl[k];
        const uint32_t K = pre[6];
        for (uint32_t t0 = 0; t0 < K; t0 += kAdjBatch) {
            uint32_t I[kAdjBatch];
            float4 Q[kAdjBatch], LF[kAdjBatch];
            float W[kAdjBatch];
#pragma unroll
            for (int u = 0; u < kAdjBatch; ++u) {
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
                I[u] = a.ids[j] - idBase;
                Q[u] = a.sPos[j];
                W[u] = w;
            }
#pragma unroll
            for (int u = 0; u < kAdjBatch; ++u) LF[u] = a.lF[I[u]];
#pragma unroll
            for (int u = 0; u < kAdjBatch; ++u) {
                if (t0 + uint32_t(u) >= K) continue;
                uint32_t fil2, ch2;
                worm_of(kp, I[u], fil2, ch2);
                if (fil2 == fil) continue;
                const f3 d = min_image_pair(p, rv, mk(Q[u]));
                const float rr = dot(d, d);
                if (!(rr < rc2)) continue;
                const float uu = __builtin_amdgcn_rcpf(rr);
                const float s6 = s2c * uu * (s2c * uu) * (s2c * uu);
                const float f = 24.0f * eps * uu * s6 * (2.0f * s6 - 1.0f);
                const float fp = -24.0f * eps * uu * uu * s6 * (14.0f * s6 - 4.0f);  // df / d(r^2)
                const f3 dl = lam - mk(LF[u]);                                       // lF_i - lF_j
                Rp += W[u] * (f * dl + (2.0f * fp * dot(d, dl)) * d);                // w H dl
                epsPart += double(W[u]) * double(dot(lam, d)) * double(f);
            }
        }
        acc[kAdjEpsilon] = -epsPart / double(eps);  // F is linear in epsilon
        (void)me;

        // adjoint of the step for this bead
        const f3 Rt = mk(load4(a.lr, m.bead)) + R + Rp;
        // the drag -gamma v' is part of F' only with the Langevin thermostat on
        const f3 V = mk(load4(a.lv, m.bead)) - (a.langevin ? a.gamma : 0.0f) * lam + a.dt * Rt;
        const f3 lG = mk(load4(a.lG, m.bead));
        a.lr[m.bead] = f4(Rt, 0.0f);
        a.lv[m.bead] = f4(V, 0.0f);
        a.lFout[m.bead] = f4(lG + (0.5f * a.dt * a.dt) * Rt + (0.5f * a.dt) * V, 0.0f);
        a.lG[m.bead] = f4((0.5f * a.dt) * V, 0.0f);
    }
}

__global__ __launch_bounds__(kBlock) void k_adj_step(AdjStepArgs a, KP kp) {
    const WaveWorm m = wave_worm(kp.p.filamentSize, a.W);
    double acc[kAdjParams] = {0.0, 0.0, 0.0, 0.0};
    adj_step_lanes(a, kp, m, 0u, acc);
    block_partials(acc, a.partial);  // every thread, once (a workgroup barrier inside)
}

// One adjoint step of every replica of a batch (a: the concatenated arrays;
// the wavefronts of the batched front kernel).  Each wavefront adds its FP64
// partials to its own row of a.partial (one row of kAdjParams per
// wavefront, accumulated over the steps in step order).
__global__ __launch_bounds__(kBlock) void k_adj_step_batch(AdjStepArgs a0, const BatchRep *__restrict__ reps,
                                                           uint32_t R, uint32_t waves, uint64_t step) {
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (wave >= waves) return;
    const uint32_t r = __builtin_amdgcn_readfirstlane(replica_of_wave(reps, R, wave));
    const BatchRep &rep = reps[r];
    const KP &kp = rep.kp;
    const WaveWorm m = wave_worm_at(wave - rep.waveOffset, kp.p.filamentSize, kp.p.numFilaments);
    AdjStepArgs a = a0;
    const uint32_t b0 = rep.beadOffset;
    a.pos += b0;
    a.vel += b0;
    a.lF += b0;
    a.lFout += b0;
    a.lr += b0;
    a.lv += b0;
    a.lG += b0;
    a.off += rep.cellOffset;
    a.W = kp.p.numFilaments;
    a.langevin = rep.langevin;
    a.lgvStd = rep.lgvStd;
    a.gamma = kp.p.gamma;
    a.seed = rep.seed;
    a.offLgv = rep.rngBase + step * rep.rngStride + rep.revBlocks;
    double acc[kAdjParams] = {0.0, 0.0, 0.0, 0.0};
    adj_step_lanes(a, kp, m, b0, acc);
#pragma unroll
    for (int k = 0; k < kAdjParams; ++k) {
        const double t = wave_sum_d(acc[k]);
        if ((threadIdx.x & 63u) == 0u) a0.partial[uint64_t(wave) * kAdjParams + k] += t;
    }
}

// out[r][k] = sum of the rows of replica r's wavefronts, in a fixed order
constexpr uint32_t kRepTotalBlock = 256;
__global__ __launch_bounds__(kRepTotalBlock) void k_adj_total_batch(const double *__restrict__ partial,
                                                                    const BatchRep *__restrict__ reps, uint32_t R,
                                                                    uint32_t waves, double *__restrict__ out) {
    __shared__ double red[kAdjParams][kRepTotalBlock];
    const uint32_t r = blockIdx.x, t = threadIdx.x;
    const uint32_t lo = reps[r].waveOffset, hi = r + 1 < R ? reps[r + 1].waveOffset : waves;
    double acc[kAdjParams] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t w = lo + t; w < hi; w += kRepTotalBlock)
#pragma unroll
        for (int k = 0; k < kAdjParams; ++k) acc[k] += partial[uint64_t(w) * kAdjParams + k];
#pragma unroll
    for (int k = 0; k < kAdjParams; ++k) red[k][t] = acc[k];
    __syncthreads();
    for (uint32_t h = kRepTotalBlock / 2; h >= 1u; h >>= 1) {
        if (t < h)
#pragma unroll
            for (int k = 0; k < kAdjParams; ++k) red[k][t] += red[k][t + h];
        __syncthreads();
    }
    if (t < uint32_t(kAdjParams)) out[uint64_t(r) * kAdjParams + t] = red[t][0];
}

// ---------------------------------------------------------------------------
// d S / d r of the bond-tangent order S = largest eigenvalue of
// <1.5 b b - 0.5 I>: with e its unit eigenvector, dS/db = 3 (e . b) e / Nb,
// through b = d / |d|, d = r[p+1] - r[p]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_bond_order_grad(const float4 *__restrict__ pos, uint32_t W,
                                                            float ex, float ey, float ez, float scale,
                                                            float4 *__restrict__ out, KP kp) {
    const SimParams &p = kp.p;
    const uint32_t n = p.filamentSize;
    const WaveWorm m = wave_worm(n, W);
    const f3 r = mk(load4_or_zero(pos, m.bead, m.active));
    const f3 rn = shfl_down3(r, 1);
    f3 g = mk(0.f, 0.f, 0.f);
    if (m.active && m.pp + 1u < n) {
        const f3 d = min_image(p, r, rn);
        const float len = sqrtf(dot(d, d));
        const f3 b = d / len;
        const f3 e = mk(ex, ey, ez);
        const f3 gb = (scale * dot(e, b)) * e;
        g = (gb - b * dot(b, gb)) / len;  // dS/dd
    }
    const f3 gPrev = shfl_up3(g, 1);
    if (!m.active) return;
    f3 o = mk(0.f, 0.f, 0.f);
    if (m.pp + 1u < n) o -= g;
    if (m.pp >= 1u) o += gPrev;
    out[m.bead] = f4(o, 0.0f);
}

// ---------------------------------------------------------------------------
// The adjoint of an SRD collision with its cells and rotation signs held
// fixed.  The collision is linear in the velocities: for particle i of cell c
// (mass m_i, cell mass M_c, rotation R_c by s_c alpha in the xy plane)
//   u_c = sum_j m_j v_j / M_c,  v_i' = u_c + R_c (v_i - u_c)  (rotated)
//                               v_i' = v_i                    (not rotated:
//                               beads when they only join the momentum sum)
// so with L_c the sum of the adjoints of the rotated members
//   lv_i <- (R_c^T or I) lv_i' + (m_i / M_c) (I - R_c^T) L_c.
// ---------------------------------------------------------------------------
// per cell, members in the forward step's cell order: (L_c, M_c)
__global__ __launch_bounds__(kBlock) void k_srd_adj_cells(const uint32_t *__restrict__ ids,
                                                          const uint32_t *__restrict__ off,
                                                          const float4 *__restrict__ lv,
                                                          const float4 *__restrict__ mass,
                                                          float4 *__restrict__ cellSum, uint32_t N, uint32_t cells,
                                                          int rotBeads) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const uint32_t lo = off[c], hi = off[c + 1];
    if (lo == hi) return;  // no particle reads an empty cell
    f3 L = mk(0.f, 0.f, 0.f);
    float M = 0.0f;
    for (uint32_t k = lo; k < hi; ++k) {
        const uint32_t i = ids[k];
        M += mass[i].w;
        if (rotBeads || i >= N) L += mk(lv[i]);
    }
    cellSum[c] = f4(L, M);
}

// R^T x for the rotation by the angle with cosine cs and sine sn (xy plane)
__device__ __forceinline__ f3 rot_t(f3 x, float cs, float sn) {
    return mk(x.x * cs + x.y * sn, -x.x * sn + x.y * cs, x.z);
}

// per particle, in place
__global__ __launch_bounds__(kBlock) void k_srd_adj_apply(const uint32_t *__restrict__ hash,
                                                          const float4 *__restrict__ com,
                                                          const float4 *__restrict__ cellSum,
                                                          const float4 *__restrict__ mass, float4 *__restrict__ lv,
                                                          uint32_t N, uint32_t NT, float cosA, float sinA,
                                                          int rotBeads) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= NT) return;
    const uint32_t c = hash[j];
    const float4 S = cellSum[c];
    const float sn = com[c].w * sinA;  // the cell's rotation sign times sin(alpha)
    const f3 l = mk(lv[j]);
    f3 out = (rotBeads || j >= N) ? rot_t(l, cosA, sn) : l;
    const f3 Lc = mk(S.x, S.y, S.z);
    out += (mass[j].w / S.w) * (Lc - rot_t(Lc, cosA, sn));
    lv[j] = f4(out, 0.0f);
}

// ---------------------------------------------------------------------------
// d P / d (r, v) of the polar speed P = <v . t> over the beads of worms of
// >= 2 beads (t = unit(r[p+1] - r[p-1]), one-sided at the worm ends, minimum
// image): dP/dv = t / Nb, dP/dd = (v - t (t . v)) / (|d| Nb) through
// d = b - a (a = r[p-1] or r[p], b = r[p+1] or r[p])
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_polar_speed_grad(const float4 *__restrict__ pos,
                                                             const float4 *__restrict__ vel, uint32_t W,
                                                             float scale, float4 *__restrict__ dPos,
                                                             float4 *__restrict__ dVel, KP kp) {
    const SimParams &p = kp.p;
    const uint32_t n = p.filamentSize;
    const WaveWorm m = wave_worm(n, W);
    const f3 r = mk(load4_or_zero(pos, m.bead, m.active));
    const f3 v = mk(load4_or_zero(vel, m.bead, m.active));
    const f3 rn = shfl_down3(r, 1), rp = shfl_up3(r, 1);
    f3 g = mk(0.f, 0.f, 0.f), t = g;
    if (m.active && n >= 2u) {
        const f3 a = m.pp == 0u ? r : rp;
        const f3 b = m.pp + 1u == n ? r : rn;
        const f3 d = min_image(p, a, b);
        const float len = sqrtf(dot(d, d));
        t = d / len;
        g = (scale / len) * (v - t * dot(t, v));
    }
    const f3 gPrev = shfl_up3(g, 1), gNext = shfl_down3(g, 1);
    if (!m.active) return;
    f3 o = mk(0.f, 0.f, 0.f);
    if (n >= 2u) {
        if (m.pp >= 1u) o += gPrev;       // b of bead pp - 1
        if (m.pp + 1u < n) o -= gNext;    // a of bead pp + 1
        if (m.pp == 0u) o -= g;           // a of the first bead
        if (m.pp + 1u == n) o += g;       // b of the last bead
    }
    dPos[m.bead] = f4(o, 0.0f);
    if (dVel) dVel[m.bead] = f4(scale * t, 0.0f);
}

// out[k] += sum over rows of partial[row][k]: thread t adds the entries
// t, t + kTotalBlock, ... of the row-major table (all of parameter
// t mod kAdjParams) in order, then a fixed tree over the threads of each
// parameter -- the same order on every call
constexpr uint32_t kTotalBlock = 1024;
static_assert(kTotalBlock % kAdjParams == 0 && (kAdjParams & (kAdjParams - 1)) == 0,
              "each thread keeps one parameter; the tree halves down to kAdjParams");

__global__ __launch_bounds__(kTotalBlock) void k_adj_total(const double *__restrict__ partial, uint64_t rows,
                                                           double *__restrict__ out) {
    __shared__ double red[kTotalBlock];
    const uint32_t t = threadIdx.x;
    const uint64_t n = rows * kAdjParams;
    double s = 0.0;
#pragma unroll 8
    for (uint64_t e = t; e < n; e += kTotalBlock) s += partial[e];
    red[t] = s;
    __syncthreads();
    for (uint32_t h = kTotalBlock / 2; h >= uint32_t(kAdjParams); h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    if (t < uint32_t(kAdjParams)) out[t] = out[t] + red[t];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
uint32_t adj_step_blocks(uint32_t W, uint32_t n) {
    const uint32_t wpw = 64u / n;
    return grid_for(((uint64_t(W) + wpw - 1) / wpw) * 64u);
}

hipError_t launch_adj_step(const AdjStepArgs &a, const KP &kp, hipStream_t st) {
    const uint32_t blocks = adj_step_blocks(a.W, kp.p.filamentSize);
    if (!blocks) return hipSuccess;
    NW_LAUNCH(k_adj_step, dim3(blocks), dim3(kBlock), 0, st, a, kp);
    return hipGetLastError();
}

hipError_t launch_bond_order_grad(const float4 *pos, uint32_t W, const double e[3], double scale, float4 *out,
                                  const KP &kp, hipStream_t st) {
    const uint32_t blocks = adj_step_blocks(W, kp.p.filamentSize);
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_bond_order_grad, dim3(blocks), dim3(kBlock), 0, st, pos, W, float(e[0]), float(e[1]),
                       float(e[2]), float(scale), out, kp);
    return hipGetLastError();
}

hipError_t launch_srd_adj(const uint32_t *ids, const uint32_t *off, const uint32_t *hash, const float4 *com,
                          const float4 *mass, float4 *lv, float4 *cellSum, uint32_t N, uint32_t NT, uint32_t cells,
                          float cosA, float sinA, int rotBeads, hipStream_t st) {
    if (NT == 0) return hipSuccess;
    hipLaunchKernelGGL(k_srd_adj_cells, dim3(grid_for(cells)), dim3(kBlock), 0, st, ids, off, lv, mass, cellSum, N,
                       cells, rotBeads);
    hipLaunchKernelGGL(k_srd_adj_apply, dim3(grid_for(NT)), dim3(kBlock), 0, st, hash, com, cellSum, mass, lv, N, NT,
                       cosA, sinA, rotBeads);
    return hipGetLastError();
}

hipError_t launch_polar_speed_grad(const float4 *pos, const float4 *vel, uint32_t W, double scale, float4 *dPos,
                                   float4 *dVel, const KP &kp, hipStream_t st) {
    const uint32_t blocks = adj_step_blocks(W, kp.p.filamentSize);
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_polar_speed_grad, dim3(blocks), dim3(kBlock), 0, st, pos, vel, W, float(scale), dPos, dVel,
                       kp);
    return hipGetLastError();
}

hipError_t launch_adj_step_batch(const AdjStepArgs &a, const BatchRep *reps, uint32_t R, uint32_t waves,
                                 uint64_t step, hipStream_t st) {
    if (waves == 0) return hipSuccess;
    NW_LAUNCH(k_adj_step_batch, dim3(grid_for(uint64_t(waves) * 64u)), dim3(kBlock), 0, st, a, reps, R,
                       waves, step);
    return hipGetLastError();
}

hipError_t launch_adj_total_batch(const double *partial, const BatchRep *reps, uint32_t R, uint32_t waves,
                                  double *out, hipStream_t st) {
    if (R == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_total_batch, dim3(R), dim3(kRepTotalBlock), 0, st, partial, reps, R, waves, out);
    return hipGetLastError();
}

hipError_t launch_adj_total(const double *partial, uint64_t rows, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_total, dim3(1), dim3(kTotalBlock), 0, st, partial, rows, out);
    return hipGetLastError();
}

}  // namespace nw

/*
 * nw_oracle_impl.h -- floating-point stages of the oracle, instantiated twice
 * by nw_oracle.c with R = double (suffix f64) and R = float (suffix f32).
 * TEST INFRASTRUCTURE ONLY (see nw_oracle.h).
 *
 * Operation order follows the reference's helper_math.h (dot = x*x+y*y+z*z,
 * length = sqrt(dot), vector/scalar divides per component) so that the f32
 * instantiation evaluates the same expressions as the reference kernels.
 */
#ifndef R
#error "define R and SFX before including nw_oracle_impl.h"
#endif

#define CAT_(a, b) a##_##b
#define CAT(a, b) CAT_(a, b)
#define F(name) CAT(name, SFX)

typedef struct { R x, y, z; } F(v3);
#define V3 F(v3)

static inline V3 F(mk)(R x, R y, R z) { V3 v = {x, y, z}; return v; }
static inline V3 F(ld)(const float *a, uint32_t i) { return F(mk)((R)a[4 * (size_t)i], (R)a[4 * (size_t)i + 1], (R)a[4 * (size_t)i + 2]); }
static inline V3 F(add)(V3 a, V3 b) { return F(mk)(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 F(sub)(V3 a, V3 b) { return F(mk)(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 F(scl)(R s, V3 a) { return F(mk)(s * a.x, s * a.y, s * a.z); }
static inline V3 F(divs)(V3 a, R s) { return F(mk)(a.x / s, a.y / s, a.z / s); }
static inline R F(dot)(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline R F(sqrt_)(R x) { return (R)sqrt((double)x); }
static inline R F(len)(V3 a) { return F(sqrt_)(F(dot)(a, a)); }

/* lengthPeriodic (particles_kernel_impl.cuh:279-301): minimum image b - a.
   A component is wrapped when |b - a| > L/2 as upstream; the wrapped value is
   formed as (b - L/2) - (a + L/2) (in double this equals b - a - L exactly). */
static inline R F(mi_axis)(R b, R a, R L, int periodic) {
    R d = b - a;
    if (periodic) {
        const R h = L / (R)2;
        if (d > h) d = (b - h) - (a + h);
        else if (d < -h) d = (b + h) - (a - h);
    }
    return d;
}
static inline V3 F(min_image)(const SimParams *p, V3 a, V3 b) {
    return F(mk)(F(mi_axis)(b.x, a.x, (R)p->boxSize.x, p->boundaryX == PERIODIC),
                 F(mi_axis)(b.y, a.y, (R)p->boxSize.y, p->boundaryY == PERIODIC),
                 F(mi_axis)(b.z, a.z, (R)p->boxSize.z, p->boundaryZ == PERIODIC));
}

/* getTangent (particles_kernel_impl.cuh:303-309) */
static inline V3 F(tangent)(const SimParams *p, V3 a, V3 b) {
    const V3 rel = F(min_image)(p, a, b);
    const R d = F(len)(rel);
    return F(divs)(rel, d);
}

static inline void F(st4)(R *out, uint32_t i, V3 v, R w) {
    out[4 * (size_t)i] = v.x; out[4 * (size_t)i + 1] = v.y; out[4 * (size_t)i + 2] = v.z; out[4 * (size_t)i + 3] = w;
}

/* ---- boundaries (particles_kernel_impl.cuh:34-68) ---- */
static inline void F(periodic_bc)(R *r, R L0, R L) {
    if (*r < L0) *r += L;
    else if (*r > L + L0) *r -= L;
}
static inline void F(damped_wall)(R *r, R *v, R coef, R L0, R L, R Rad) {
    if (*r < L0 + Rad) { *r = L0 + Rad; *v *= coef; }
    else if (*r > (L + L0) - Rad) { *r = (L + L0) - Rad; *v *= coef; }
}
/* noslipWallBoundary: pos -= dr, the wall axis pinned, vel = -vel * N per component */
static inline void F(noslip_wall)(V3 *pos, int axis, V3 *vel, V3 dr, V3 N, R L0, R L, R Rad) {
    const R r = axis == 0 ? pos->x : pos->y;
    int hit = 0; R target = 0;
    if (r < L0 + Rad) { hit = 1; target = L0 + Rad; }
    else if (r > (L + L0) - Rad) { hit = 1; target = (L + L0) - Rad; }
    if (hit) {
        *pos = F(sub)(*pos, dr);
        if (axis == 0) pos->x = target; else pos->y = target;
        *vel = F(mk)(-vel->x * N.x, -vel->y * N.y, -vel->z * N.z);
    }
}
static inline void F(boundaries)(const SimParams *p, V3 *pos, V3 *vel, V3 dr, R Rad) {
    if (p->boundaryX == PERIODIC) F(periodic_bc)(&pos->x, (R)p->origin.x, (R)p->boxSize.x);
    else if (p->boundaryX == WALL) F(damped_wall)(&pos->x, &vel->x, (R)p->boundaryDamping, (R)p->origin.x, (R)p->boxSize.x, Rad);
    else if (p->boundaryX == WALL_NO_SLIP) F(noslip_wall)(pos, 0, vel, dr, F(mk)(1, 0, 0), (R)p->origin.x, (R)p->boxSize.x, Rad);
    if (p->boundaryY == PERIODIC) F(periodic_bc)(&pos->y, (R)p->origin.y, (R)p->boxSize.y);
    else if (p->boundaryY == WALL) F(damped_wall)(&pos->y, &vel->y, (R)p->boundaryDamping, (R)p->origin.y, (R)p->boxSize.y, Rad);
    else if (p->boundaryY == WALL_NO_SLIP) F(noslip_wall)(pos, 1, vel, dr, F(mk)(0, 1, 0), (R)p->origin.y, (R)p->boxSize.y, Rad);
    if (p->boundaryZ == PERIODIC) F(periodic_bc)(&pos->z, (R)p->origin.z, (R)p->boxSize.z);
    else if (p->boundaryZ == WALL) F(damped_wall)(&pos->z, &vel->z, (R)p->boundaryDamping, (R)p->origin.z, (R)p->boxSize.z, Rad);
}

void F(orc_integrate_filaments)(const SimParams *p, const float *pos, const float *vel, const float *f,
                                const float *fold, float dt_, uint32_t n, R *posOut, R *velOut, R *fOut,
                                R *foldOut) {
    const R dt = (R)dt_;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        V3 r = F(ld)(pos, i), v = F(ld)(vel, i);
        const V3 ff = F(ld)(f, i), fo = F(ld)(fold, i);
        const R pw = (R)pos[4 * (size_t)i + 3], vw = (R)vel[4 * (size_t)i + 3], fow = (R)fold[4 * (size_t)i + 3];
        /* v += 0.5 (f + f_old) dt ; r += v dt + 0.5 f dt^2 */
        const V3 dv = F(scl)(dt, F(scl)((R)0.5, F(add)(ff, fo)));
        v = F(add)(v, dv);
        const V3 dr = F(add)(F(scl)(dt, v), F(scl)(dt, F(scl)(dt, F(scl)((R)0.5, ff))));
        r = F(add)(r, dr);
        F(boundaries)(p, &r, &v, dr, (R)p->particleRadius);
        F(st4)(posOut, i, r, pw);
        F(st4)(velOut, i, v, vw);
        F(st4)(fOut, i, F(mk)(0, 0, 0), 0);
        F(st4)(foldOut, i, ff, fow);
    }
}

void F(orc_integrate_solvent)(const SimParams *p, const float *pos, const float *vel, float dt_, uint32_t start,
                              uint32_t ns, R *posOut, R *velOut) {
    const R dt = (R)dt_;
#pragma omp parallel for schedule(static)
    for (int64_t tt = 0; tt < (int64_t)ns; ++tt) {
        const uint32_t i = start + (uint32_t)tt;
        V3 r = F(ld)(pos, i), v = F(ld)(vel, i);
        const V3 dr = F(scl)(dt, v);
        r = F(add)(r, dr);
        F(boundaries)(p, &r, &v, dr, (R)0);
        F(st4)(posOut, i, r, (R)pos[4 * (size_t)i + 3]);
        F(st4)(velOut, i, v, (R)vel[4 * (size_t)i + 3]);
    }
}

/* ---- pair terms (particles_kernel_impl.cuh:427-487; archived LJ) ---- */
static inline V3 F(dem)(const SimParams *p, V3 rA, V3 rB, V3 vA, V3 vB) {
    const V3 rel = F(min_image)(p, rA, rB);
    const R dist = F(len)(rel);
    const R cd = (R)p->particleRadius + (R)p->particleRadius;
    V3 force = F(mk)(0, 0, 0);
    if (dist < cd) {
        const V3 nrm = F(divs)(rel, dist);
        const V3 relVel = F(sub)(vB, vA);
        const V3 tanVel = F(sub)(relVel, F(scl)(F(dot)(relVel, nrm), nrm));
        force = F(scl)(-(R)p->spring * (cd - dist), nrm);
        force = F(add)(force, F(scl)((R)p->damping, relVel));
        force = F(add)(force, F(scl)((R)p->shear, tanVel));
    }
    return force;
}
static inline V3 F(bond)(const SimParams *p, V3 rA, V3 rB, R k, R L) {
    const V3 rel = F(min_image)(p, rA, rB);
    const R dist = F(len)(rel);
    const V3 nrm = F(divs)(rel, dist);
    return F(scl)(-k * (L - dist), nrm);
}
static inline V3 F(extensile)(const SimParams *p, V3 rA, V3 rB, V3 tA, V3 tB, R A) {
    const V3 rel = F(min_image)(p, rA, rB);
    const R dist = F(len)(rel);
    if (F(dot)(tA, tB) < (R)-0.5) return F(divs)(F(scl)(-A, F(sub)(tA, tB)), (R)2 * dist);
    return F(mk)(0, 0, 0);
}
/* WCA-LJ: F_A = -24 eps (2 s^12 / r^14 - s^6 / r^8) (rB - rA) for r^2 < rc^2 */
static inline V3 F(wca)(const SimParams *p, const NWExtParams *e, V3 rA, V3 rB) {
    const V3 rel = F(min_image)(p, rA, rB);
    const R rr = F(dot)(rel, rel);
    const R rc = (R)e->wcaCutoff;
    if (rr < rc * rc) {
        const R inv2 = (R)1 / rr;
        const R s2 = (R)e->wcaSigma * (R)e->wcaSigma * inv2;
        const R s6 = s2 * s2 * s2;
        const R fmag = (R)24 * (R)e->wcaEpsilon * inv2 * s6 * ((R)2 * s6 - (R)1);
        return F(scl)(-fmag, rel);
    }
    return F(mk)(0, 0, 0);
}

static inline int F(wrapm)(int g, uint32_t G) {
    int m = g % (int)G;
    return m < 0 ? m + (int)G : m;
}

void F(orc_collide_filaments)(const SimParams *p, const NWExtParams *e, const float *sPos, const float *sVel,
                              const float *sTan, const uint32_t *sIdx, const uint32_t *cellStart,
                              const uint32_t *cellEnd, uint32_t n, const float *force, R *forceOut) {
    const uint32_t fs = p->filamentSize;
    const int wcaMode = e->pairPotential == NW_PAIR_WCA;
    const int useTan = !wcaMode || p->activity != 0.0f;
    const int zlo = e->stencil == NW_STENCIL_2D ? 0 : -1, zhi = e->stencil == NW_STENCIL_2D ? 0 : 1;
    /* outputs start as the incoming force array */
    for (uint32_t i = 0; i < n; ++i)
        for (int c = 0; c < 4; ++c) forceOut[4 * (size_t)i + c] = (R)force[4 * (size_t)i + c];
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const V3 r = F(ld)(sPos, i);
        const V3 v = wcaMode ? F(mk)(0, 0, 0) : F(ld)(sVel, i);
        const V3 t = useTan ? F(ld)(sTan, i) : F(mk)(0, 0, 0);
        const uint32_t me = sIdx[i];
        const uint32_t fil = me / fs, ch = me % fs;
        /* calcGridPos in float, as the reference */
        const float px = sPos[4 * (size_t)i], py = sPos[4 * (size_t)i + 1], pz = sPos[4 * (size_t)i + 2];
        const int gx = (int)floorf((px - p->origin.x) / p->cellSize.x);
        const int gy = (int)floorf((py - p->origin.y) / p->cellSize.y);
        const int gz = (int)floorf((pz - p->origin.z) / p->cellSize.z);
        V3 acc = F(mk)(0, 0, 0);
        for (int z = zlo; z <= zhi; ++z)
            for (int y = -1; y <= 1; ++y)
                for (int x = -1; x <= 1; ++x) {
                    const uint32_t cx = (uint32_t)F(wrapm)(gx + x, p->gridSize.x);
                    const uint32_t cy = (uint32_t)F(wrapm)(gy + y, p->gridSize.y);
                    const uint32_t cz = (uint32_t)F(wrapm)(gz + z, p->gridSize.z);
                    const uint32_t h = (cz * p->gridSize.y) * p->gridSize.x + cy * p->gridSize.x + cx;
                    const uint32_t start = cellStart[h];
                    V3 fc = F(mk)(0, 0, 0);
                    if (start != 0xffffffffu) {
                        const uint32_t end = cellEnd[h];
                        for (uint32_t j = start; j < end; ++j) {
                            if (j == i) continue;
                            const V3 r2 = F(ld)(sPos, j);
                            const uint32_t o = sIdx[j];
                            const uint32_t fil2 = o / fs, ch2 = o % fs;
                            const int same = fil == fil2;
                            if (!wcaMode) {
                                const V3 v2 = F(ld)(sVel, j);
                                if (same && ((ch + 1u == ch2) || (ch - 1u == ch2)))
                                    fc = F(add)(fc, F(bond)(p, r, r2, (R)p->bondSpringK, (R)p->bondSpringL));
                                else
                                    fc = F(add)(fc, F(dem)(p, r, r2, v, v2));
                            } else if (!same) {
                                fc = F(add)(fc, F(wca)(p, e, r, r2));
                            }
                            if (useTan && !same)
                                fc = F(add)(fc, F(extensile)(p, r, r2, t, F(ld)(sTan, j), (R)p->activity));
                        }
                    }
                    acc = F(add)(acc, fc);
                }
        R *o = forceOut + 4 * (size_t)me;
        o[0] += acc.x; o[1] += acc.y; o[2] += acc.z;
    }
}

void F(orc_filament_forces)(const SimParams *p, const NWExtParams *e, const float *force, const float *pos,
                            uint32_t W, R *forceOut, R *tangentOut) {
    const uint32_t n = p->filamentSize;
    const R kb = (R)p->bondBendingK;
    const size_t N = (size_t)W * n;
    for (size_t i = 0; i < 4 * N; ++i) forceOut[i] = (R)force[i];
#pragma omp parallel for schedule(static)
    for (int64_t ww = 0; ww < (int64_t)W; ++ww) {
        const uint32_t head = (uint32_t)ww * n;
        if (n >= 3) {
            for (uint32_t q = 0; q + 2u < n; ++q) {
                const uint32_t i = head + q;
                const V3 ri = F(ld)(pos, i), rj = F(ld)(pos, i + 1), rk = F(ld)(pos, i + 2);
                const V3 r_ij = F(min_image)(p, ri, rj), r_jk = F(min_image)(p, rj, rk);
                const R d_ij = F(len)(r_ij);
                const R d_jk = F(len)(r_jk);
                const R d_ij_jk = F(dot)(r_ij, r_jk);
                const R A = kb / (d_ij * d_jk);
                const V3 B = F(divs)(F(scl)(d_ij_jk, r_ij), F(dot)(r_ij, r_ij));
                const V3 C = F(divs)(F(scl)(d_ij_jk, r_jk), F(dot)(r_jk, r_jk));
                const V3 fi = F(scl)(-A, F(sub)(r_jk, B));
                const V3 fk = F(scl)(-A, F(sub)(C, r_ij));
                const V3 fj = F(mk)(-fi.x - fk.x, -fi.y - fk.y, -fi.z - fk.z);
                R *a = forceOut + 4 * (size_t)i;
                a[0] += fi.x; a[1] += fi.y; a[2] += fi.z;
                a[4] += fj.x; a[5] += fj.y; a[6] += fj.z;
                a[8] += fk.x; a[9] += fk.y; a[10] += fk.z;
                F(st4)(tangentOut, i + 1, F(tangent)(p, ri, rk), 0);
                if (q == 0) F(st4)(tangentOut, i, F(tangent)(p, ri, rj), 0);
                else if (q == n - 3) F(st4)(tangentOut, i + 2, F(tangent)(p, rj, rk), 0);
            }
        }
        if (e && e->bondedMode == NW_BONDED_CHAIN && n >= 2) {
            for (uint32_t q = 0; q + 1u < n; ++q) {
                const uint32_t i = head + q;
                const V3 d = F(min_image)(p, F(ld)(pos, i), F(ld)(pos, i + 1));
                const R len = F(len)(d);
                /* archived InterForceKernel: f_i += K (r - L) / r * dr, f_{i+1} -= the same */
                const V3 fb = F(scl)((R)p->bondSpringK * (len - (R)p->bondSpringL) / len, d);
                /* archived DriveForceKernel: f_i += f_a * unit(r_{i+1} - r_i), i < n-1 */
                const V3 dv = F(scl)((R)e->driveForce / len, d);
                R *a = forceOut + 4 * (size_t)i;
                a[0] += fb.x + dv.x; a[1] += fb.y + dv.y; a[2] += fb.z + dv.z;
                a[4] -= fb.x; a[5] -= fb.y; a[6] -= fb.z;
            }
        }
    }
}

void F(orc_langevin)(const SimParams *p, const float *force, const float *vel, const R *noise, uint32_t n,
                     R *forceOut) {
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const size_t i = (size_t)ii;
        const V3 v = F(ld)(vel, (uint32_t)i);
        V3 f = F(mk)(0, 0, 0);
        f = F(sub)(f, F(scl)((R)p->gamma, v));
        f = F(add)(f, F(mk)(noise[4 * i], noise[4 * i + 1], noise[4 * i + 2]));
        forceOut[4 * i] = (R)force[4 * i] + f.x;
        forceOut[4 * i + 1] = (R)force[4 * i + 1] + f.y;
        forceOut[4 * i + 2] = (R)force[4 * i + 2] + f.z;
        forceOut[4 * i + 3] = (R)force[4 * i + 3];
    }
}

void F(orc_srd_com)(const float *sVel, const uint32_t *cellStart, const uint32_t *cellEnd, uint32_t numCells,
                    R *comOut) {
#pragma omp parallel for schedule(static)
    for (int64_t cc = 0; cc < (int64_t)numCells; ++cc) {
        const uint32_t c = (uint32_t)cc;
        V3 m = F(mk)(0, 0, 0);
        if (cellStart[c] != 0xffffffffu) {
            R M = 0;
            for (uint32_t i = cellStart[c]; i < cellEnd[c]; ++i) {
                const R w = (R)sVel[4 * (size_t)i + 3];
                m = F(add)(m, F(scl)(w, F(ld)(sVel, i)));
                M += w;
            }
            if (M > 0) m = F(scl)((R)1 / M, m);
        }
        F(st4)(comOut, c, m, 0);
    }
}

void F(orc_srd_isokinetic)(const SolventParams *s, const float *vel, const float *sVel, const uint32_t *cellStart,
                           const uint32_t *cellEnd, const uint32_t *index, uint32_t numCells, uint32_t nTotal,
                           R *velOut, R *sVelOut) {
    for (size_t i = 0; i < 4 * (size_t)nTotal; ++i) { velOut[i] = (R)vel[i]; sVelOut[i] = (R)sVel[i]; }
#pragma omp parallel for schedule(static)
    for (int64_t cc = 0; cc < (int64_t)numCells; ++cc) {
        const uint32_t c = (uint32_t)cc;
        if (cellStart[c] == 0xffffffffu) continue;
        R E = 0;
        for (uint32_t i = cellStart[c]; i < cellEnd[c]; ++i) {
            const V3 v = F(ld)(sVel, i);
            E += (R)0.5 * (R)sVel[4 * (size_t)i + 3] * F(dot)(v, v);
        }
        const R lambda = F(sqrt_)((R)s->kbT / E);
        for (uint32_t i = cellStart[c]; i < cellEnd[c]; ++i) {
            const uint32_t idx = index[i];
            const V3 v = F(ld)(vel, idx);
            const V3 nv = F(scl)(lambda, v);
            F(st4)(velOut, idx, nv, (R)sVel[4 * (size_t)i + 3]);
            F(st4)(sVelOut, i, nv, (R)sVel[4 * (size_t)i + 3]);
        }
    }
}

void F(orc_collide_solvent)(const SimParams *p, const SolventParams *s, const NWExtParams *e, const float *pos,
                            const float *vel, const float *com, const float *uniform, uint32_t numCells, uint32_t n,
                            R *velOut) {
    const R ca = (R)cos((double)s->alpha), sa = (R)sin((double)s->alpha);
    const int all = e && e->srdCollideFilaments;
    for (size_t i = 0; i < 4 * (size_t)n; ++i) velOut[i] = (R)vel[i];
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        if (!all && i < p->numParticles) continue;
        uint32_t cell;
        orc_calc_hash(p, s, pos + 4 * (size_t)i, 1, 1, &cell, &(uint32_t){0});
        if (cell >= numCells) continue;
        const V3 u = F(ld)(com, cell);
        const R sgn = uniform[cell] > 0.5f ? (R)1 : (R)-1;
        const R sn = sgn * sa;
        const V3 d = F(sub)(F(ld)(vel, i), u);
        const V3 rot = F(mk)(d.x * ca - d.y * sn, d.x * sn + d.y * ca, d.z);
        F(st4)(velOut, i, F(add)(u, rot), (R)vel[4 * (size_t)i + 3]);
    }
}

#undef V3
#undef F
#undef CAT
#undef CAT_

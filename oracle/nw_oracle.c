/*
 * nw_oracle.c -- CPU restatement of the reference's worm time step.
 * TEST INFRASTRUCTURE ONLY: the checker of libnwamd.so and the CPU baseline
 * (see nw_oracle.h).  Built by oracle/Makefile with -ffp-contract=off.
 */
#include "nw_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ hash */
static int wrap_mod(int g, uint32_t G) {
    int m = g % (int)G;
    return m < 0 ? m + (int)G : m;
}

static int clampi(int g, uint32_t G) {
    if (g < 0) return 0;
    if (g >= (int)G) return (int)G - 1;
    return g;
}

/* positionRelativeToBox (particles_kernel_impl.cuh:208-228) */
static void rel_to_box(const SimParams *p, float q[3]) {
    q[0] = q[0] - p->origin.x;
    q[1] = q[1] - p->origin.y;
    q[2] = q[2] - p->origin.z;
    const float L[3] = {p->boxSize.x, p->boxSize.y, p->boxSize.z};
    const BoundaryType b[3] = {p->boundaryX, p->boundaryY, p->boundaryZ};
    for (int a = 0; a < 3; ++a) {
        if (b[a] != PERIODIC) continue;
        if (q[a] >= L[a]) q[a] -= L[a];
        else if (q[a] < 0.0f) q[a] += L[a];
    }
}

void orc_calc_hash(const SimParams *p, const SolventParams *s, const float *pos, uint32_t n, int srd,
                   uint32_t *hash, uint32_t *index) {
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const size_t i = (size_t)ii;
        const float *q = pos + 4 * i;
        uint32_t h;
        if (srd) {
            /* calcGridPosSRD :241-249, calcGridHashSRD :261-276 */
            float r[3] = {q[0] + s->randOffset.x, q[1] + s->randOffset.y, q[2] + s->randOffset.z};
            rel_to_box(p, r);
            int g[3] = {(int)floorf(r[0] / s->cellSize.x), (int)floorf(r[1] / s->cellSize.y),
                        (int)floorf(r[2] / s->cellSize.z)};
            const uint32_t G[3] = {s->gridSize.x, s->gridSize.y, s->gridSize.z};
            const BoundaryType b[3] = {p->boundaryX, p->boundaryY, p->boundaryZ};
            for (int a = 0; a < 3; ++a) g[a] = b[a] == PERIODIC ? wrap_mod(g[a], G[a]) : clampi(g[a], G[a]);
            h = ((uint32_t)g[2] * G[1]) * G[0] + (uint32_t)g[1] * G[0] + (uint32_t)g[0];
        } else {
            /* calcGridPos :231-238, calcGridHash :252-258 */
            const int gx = (int)floorf((q[0] - p->origin.x) / p->cellSize.x);
            const int gy = (int)floorf((q[1] - p->origin.y) / p->cellSize.y);
            const int gz = (int)floorf((q[2] - p->origin.z) / p->cellSize.z);
            const uint32_t x = (uint32_t)wrap_mod(gx, p->gridSize.x);
            const uint32_t y = (uint32_t)wrap_mod(gy, p->gridSize.y);
            const uint32_t z = (uint32_t)wrap_mod(gz, p->gridSize.z);
            h = (z * p->gridSize.y) * p->gridSize.x + y * p->gridSize.x + x;
        }
        hash[i] = h;
        index[i] = (uint32_t)i;
    }
}

/* ------------------------------------------------------------------ sort */
void orc_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t n) {
    if (n < 2) return;
    uint32_t *tk = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *tv = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *ak = keys, *av = vals, *bk = tk, *bv = tv;
    for (uint32_t width = 1; width < n; width *= 2) {
        for (uint32_t lo = 0; lo < n; lo += 2 * width) {
            const uint32_t mid = lo + width < n ? lo + width : n;
            const uint32_t hi = lo + 2 * width < n ? lo + 2 * width : n;
            uint32_t a = lo, b = mid, o = lo;
            while (a < mid && b < hi) {
                /* take from the left run on ties: stability */
                if (ak[b] < ak[a]) { bk[o] = ak[b]; bv[o] = av[b]; ++b; }
                else { bk[o] = ak[a]; bv[o] = av[a]; ++a; }
                ++o;
            }
            while (a < mid) { bk[o] = ak[a]; bv[o] = av[a]; ++a; ++o; }
            while (b < hi) { bk[o] = ak[b]; bv[o] = av[b]; ++b; ++o; }
        }
        uint32_t *t;
        t = ak; ak = bk; bk = t;
        t = av; av = bv; bv = t;
    }
    if (ak != keys) {
        memcpy(keys, ak, sizeof(uint32_t) * n);
        memcpy(vals, av, sizeof(uint32_t) * n);
    }
    free(tk);
    free(tv);
}

/* --------------------------------------------------------------- reorder */
void orc_reorder(uint32_t *cellStart, uint32_t *cellEnd, float *sortedPos, float *sortedVel, float *sortedTan,
                 const uint32_t *hash, const uint32_t *index, const float *pos, const float *vel, const float *tan,
                 uint32_t n, uint32_t numCells) {
    for (uint32_t c = 0; c < numCells; ++c) cellStart[c] = 0xffffffffu;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t h = hash[i];
        if (i == 0 || h != hash[i - 1]) {
            if (h < numCells) cellStart[h] = i;
            if (i > 0 && hash[i - 1] < numCells) cellEnd[hash[i - 1]] = i;
        }
        if (i == n - 1 && h < numCells) cellEnd[h] = i + 1;
    }
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const size_t i = (size_t)ii, s = index[i];
        memcpy(sortedPos + 4 * i, pos + 4 * s, 16);
        if (vel && sortedVel) memcpy(sortedVel + 4 * i, vel + 4 * s, 16);
        if (tan && sortedTan) memcpy(sortedTan + 4 * i, tan + 4 * s, 16);
    }
}

/* --------------------------------------------------------------- reverse */
void orc_reverse_filaments(const SimParams *p, const float *uniform, uint32_t W, float *force, float *forceOld,
                           float *vel, float *pos) {
    const uint32_t n = p->filamentSize;
    for (uint32_t w = 0; w < W; ++w) {
        if (uniform[w] > p->reverseProbability) continue;
        for (uint32_t q = 0; q < n / 2; ++q) {
            const size_t i = (size_t)w * n + q, j = (size_t)w * n + n - 1 - q;
            float t[4];
            float *arr[4] = {pos, vel, force, forceOld};
            for (int a = 0; a < 4; ++a) {
                memcpy(t, arr[a] + 4 * i, 16);
                memcpy(arr[a] + 4 * i, arr[a] + 4 * j, 16);
                memcpy(arr[a] + 4 * j, t, 16);
            }
        }
    }
}

/* ---------------------------------------------------------------- Philox */
static uint32_t mulhilo32(uint32_t a, uint32_t b, uint32_t *hi) {
    const uint64_t prod = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(prod >> 32);
    return (uint32_t)prod;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, hi1;
        const uint32_t lo0 = mulhilo32(0xD2511F53u, c0, &hi0);
        const uint32_t lo1 = mulhilo32(0xCD9E8D57u, c2, &hi1);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void block_of(uint64_t seed, uint64_t b, uint32_t out[4]) {
    const uint32_t ctr[4] = {(uint32_t)b, (uint32_t)(b >> 32), 0u, 0u};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr, key, out);
}

/* (x >> 8) * 2^-24 + 2^-24: uniform on (0, 1], exact in float */
static float u01(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-8f + 5.9604644775390625e-8f; }

void orc_random_uniform(uint64_t seed, uint64_t offset, uint32_t n, float *out) {
    for (uint32_t e = 0; e < n; ++e) {
        uint32_t b[4];
        block_of(seed, offset + e / 4, b);
        out[e] = u01(b[e % 4]);
    }
}

void orc_random_gaussian_f64(uint64_t seed, uint64_t offset, uint32_t n, double mean, double stddev,
                             double *out) {
    const double twopi = 6.283185307179586476925286766559;
    for (uint32_t e0 = 0; e0 < n; e0 += 4) {
        uint32_t b[4];
        block_of(seed, offset + e0 / 4, b);
        double z[4];
        for (int k = 0; k < 2; ++k) {
            const double u1 = (double)u01(b[2 * k]), u2 = (double)u01(b[2 * k + 1]);
            const double r = sqrt(-2.0 * log(u1));
            z[2 * k] = r * cos(twopi * u2);
            z[2 * k + 1] = r * sin(twopi * u2);
        }
        for (int k = 0; k < 4 && e0 + k < n; ++k) out[e0 + k] = z[k] * stddev + mean;
    }
}

/* ------------------------------------------------- FP stages (two precisions) */
#define R double
#define SFX f64
#include "nw_oracle_impl.h"
#undef R
#undef SFX
#define R float
#define SFX f32
#include "nw_oracle_impl.h"
#undef R
#undef SFX

/* ----------------------------------------------------------- whole step */
/* SRD grid shift for collision number k: two Philox uniforms from a key
   derived from the seed; the reference draws it with the host rand()
   (simulation/particleSystem.cpp:345-346). Shift in [0, a) (centred=0) or
   [-a/2, a/2) (centred=1). */
void orc_srd_shift(uint64_t seed, uint64_t k, int centred, const float cell[3], float out[3]) {
    uint32_t b[4];
    block_of(seed ^ 0x5eed5eed5eed5eedull, k, b);
    for (int a = 0; a < 2; ++a) {
        const float u = (float)(b[a] >> 8) * 5.9604644775390625e-8f; /* [0, 1) */
        out[a] = (centred ? u - 0.5f : u) * cell[a];
    }
    out[2] = 0.0f;
}

/*
 * One reference update (simulation/particleSystem.cpp:249-339) per step, FP32.
 * sys->pos/vel hold N + Ns particles; force/forceOld/tangent N beads.
 * Scratch: sPos/sVel (N + Ns), sTan (N), com (Cs), rnd (max(4N, Cs, W)),
 * hash/index (N + Ns), cellStart/cellEnd (numCellsAlloc >= max(C, Cs)).
 */
int orc_system_step_f32(OrcSystem *sys, uint32_t steps) {
    SimParams *p = &sys->p;
    const uint32_t N = p->numParticles, W = p->numFilaments, Ns = sys->numSolvent;
    const uint32_t NT = N + (sys->srdEnabled ? Ns : 0);
    const int needVT = sys->e.pairPotential == NW_PAIR_REF_DEM || p->activity != 0.0f;
    for (uint32_t st = 0; st < steps; ++st) {
        orc_integrate_filaments_f32(p, sys->pos, sys->vel, sys->force, sys->forceOld, sys->dt, N, sys->pos,
                                    sys->vel, sys->force, sys->forceOld);
        if (p->reverseProbability > 0.0f) {
            orc_random_uniform(sys->seed, sys->offset, W, sys->rnd);
            sys->offset += (W + 3) / 4;
            orc_reverse_filaments(p, sys->rnd, W, sys->force, sys->forceOld, sys->vel, sys->pos);
        }
        if (!sys->srdEnabled) {
            const float kT = sys->langevinFdt ? p->kbT / sys->dt : p->kbT;
            const float sd = sqrtf(2.0f * p->gamma * kT);
            /* normals in double, rounded to float as the FP32 path */
            double *z = (double *)malloc(sizeof(double) * 4 * (size_t)N);
            orc_random_gaussian_f64(sys->seed, sys->offset, 4 * N, 0.0, (double)sd, z);
            for (size_t i = 0; i < 4 * (size_t)N; ++i) sys->rnd[i] = (float)z[i];
            free(z);
            sys->offset += N;
            orc_langevin_f32(p, sys->force, sys->vel, sys->rnd, N, sys->force);
        }
        orc_filament_forces_f32(p, &sys->e, sys->force, sys->pos, W, sys->force, sys->tangent);
        orc_calc_hash(p, &sys->s, sys->pos, N, 0, sys->hash, sys->index);
        orc_sort_pairs(sys->hash, sys->index, N);
        orc_reorder(sys->cellStart, sys->cellEnd, sys->sPos, needVT ? sys->sVel : NULL, needVT ? sys->sTan : NULL,
                    sys->hash, sys->index, sys->pos, sys->vel, sys->tangent, N, sys->numCellsAlloc);
        orc_collide_filaments_f32(p, &sys->e, sys->sPos, sys->sVel, sys->sTan, sys->index, sys->cellStart,
                                  sys->cellEnd, N, sys->force, sys->force);
        if (sys->srdEnabled && (sys->step % sys->srdGap) == 0) {
            const float cell[3] = {sys->s.cellSize.x, sys->s.cellSize.y, sys->s.cellSize.z};
            float sh[3];
            orc_srd_shift(sys->seed, sys->step / sys->srdGap, (int)sys->srdShiftCentred, cell, sh);
            sys->s.randOffset.x = sh[0];
            sys->s.randOffset.y = sh[1];
            sys->s.randOffset.z = sh[2];
            const uint32_t Cs = sys->s.numCells;
            orc_calc_hash(p, &sys->s, sys->pos, NT, 1, sys->hash, sys->index);
            orc_sort_pairs(sys->hash, sys->index, NT);
            orc_reorder(sys->cellStart, sys->cellEnd, sys->sPos, sys->sVel, NULL, sys->hash, sys->index, sys->pos,
                        sys->vel, NULL, NT, sys->numCellsAlloc);
            if (sys->e.srdThermostat == NW_SRD_THERMO_ISOKINETIC)
                orc_srd_isokinetic_f32(&sys->s, sys->vel, sys->sVel, sys->cellStart, sys->cellEnd, sys->index, Cs,
                                       NT, sys->vel, sys->sVel);
            orc_srd_com_f32(sys->sVel, sys->cellStart, sys->cellEnd, Cs, sys->com);
            orc_random_uniform(sys->seed, sys->offset, Cs, sys->rnd);
            sys->offset += (Cs + 3) / 4;
            orc_collide_solvent_f32(p, &sys->s, &sys->e, sys->pos, sys->vel, sys->com, sys->rnd, Cs, NT, sys->vel);
            orc_integrate_solvent_f32(p, sys->pos, sys->vel, sys->dt * (float)sys->srdGap, N, Ns, sys->pos,
                                      sys->vel);
        }
        sys->step++;
    }
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

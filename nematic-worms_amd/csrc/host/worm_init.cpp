// worm_init.cpp -- see worm_init.h.
#include "worm_init.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

namespace nwhost {

float HostStream::next() {
    uint32_t b[4];
    nwPhiloxBlock(key, draws / 4, b);
    const uint32_t x = b[draws % 4];
    draws++;
    return float(x >> 8) * 5.9604644775390625e-8f;
}

void place_worms(const SimParams &p, float sigma, int mode, float kT, HostStream &rng, float *hPos, float *hVel,
                 float *hTangent) {
    const uint n = p.filamentSize, W = p.numFilaments;
    const bool twoD = p.gridSize.z == 1;
    const float Lx = p.boxSize.x, Ly = p.boxSize.y, Lz = p.boxSize.z;
    const float zmid = p.origin.z + 0.5f * Lz;
    const float bond = p.bondSpringL;
    auto frand = [&]() { return rng.next(); };
    auto wrap = [](float x, float o, float L) {
        float r = std::fmod(x - o, L);
        if (r < 0) r += L;
        return o + r;
    };
    const float twopi = 6.28318530717958647692f;

    if (mode == NW_INIT_RANDOM) {
        // uniform heads and headings, rejected when any bead comes closer than
        // sigma to a bead of an already placed worm (host cell grid of size >= sigma)
        const int gx = std::max(1, int(Lx / sigma)), gy = std::max(1, int(Ly / sigma));
        const float cx = Lx / gx, cy = Ly / gy;
        std::vector<std::vector<uint>> grid(size_t(gx) * gy);
        auto cellOf = [&](float x, float y, int &ix, int &iy) {
            ix = std::min(gx - 1, std::max(0, int((x - p.origin.x) / cx)));
            iy = std::min(gy - 1, std::max(0, int((y - p.origin.y) / cy)));
        };
        std::vector<float> cand(3 * n);
        for (uint w = 0; w < W; ++w) {
            bool placed = false;
            for (int attempt = 0; attempt < 20000 && !placed; ++attempt) {
                const float hx = p.origin.x + frand() * Lx, hy = p.origin.y + frand() * Ly;
                const float th = frand() * twopi;
                bool ok = true;
                for (uint q = 0; q < n && ok; ++q) {
                    const float x = wrap(hx + q * bond * std::cos(th), p.origin.x, Lx);
                    const float y = wrap(hy + q * bond * std::sin(th), p.origin.y, Ly);
                    cand[3 * q] = x;
                    cand[3 * q + 1] = y;
                    cand[3 * q + 2] = twoD ? zmid : p.origin.z + frand() * Lz;
                    int ix, iy;
                    cellOf(x, y, ix, iy);
                    for (int dy = -1; dy <= 1 && ok; ++dy)
                        for (int dx = -1; dx <= 1 && ok; ++dx) {
                            const int jx = (ix + dx + gx) % gx, jy = (iy + dy + gy) % gy;
                            for (uint j : grid[size_t(jy) * gx + jx]) {
                                float ddx = hPos[4 * size_t(j)] - x, ddy = hPos[4 * size_t(j) + 1] - y;
                                ddx -= Lx * std::round(ddx / Lx);
                                ddy -= Ly * std::round(ddy / Ly);
                                if (ddx * ddx + ddy * ddy < sigma * sigma) {
                                    ok = false;
                                    break;
                                }
                            }
                        }
                }
                if (!ok) continue;
                for (uint q = 0; q < n; ++q) {
                    const size_t i = size_t(w) * n + q;
                    hPos[4 * i] = cand[3 * q];
                    hPos[4 * i + 1] = cand[3 * q + 1];
                    hPos[4 * i + 2] = cand[3 * q + 2];
                    hPos[4 * i + 3] = p.particleRadius;
                    int ix, iy;
                    cellOf(cand[3 * q], cand[3 * q + 1], ix, iy);
                    grid[size_t(iy) * gx + ix].push_back(uint(i));
                    hTangent[4 * i] = std::cos(th);
                    hTangent[4 * i + 1] = std::sin(th);
                    hTangent[4 * i + 2] = 0.f;
                    hTangent[4 * i + 3] = 0.f;
                }
                placed = true;
            }
            if (!placed) throw std::runtime_error("resetRandom: could not place worm without overlap");
        }
    } else {
        // rows of straight worms along x with random polarity
        const float len = (n - 1) * bond + sigma;
        const uint perRow = std::max(1u, uint(Lx / len));
        const uint rows = (W + perRow - 1) / perRow;
        const float dy = Ly / rows;
        for (uint w = 0; w < W; ++w) {
            const uint row = w / perRow, col = w % perRow;
            const bool flip = frand() < 0.5f;
            const float x0 = p.origin.x + col * (Lx / perRow) + 0.5f * sigma;
            const float y0 = p.origin.y + (row + 0.5f) * dy;
            for (uint q = 0; q < n; ++q) {
                const uint qq = flip ? n - 1 - q : q;
                const size_t i = size_t(w) * n + q;
                hPos[4 * i] = wrap(x0 + qq * bond, p.origin.x, Lx);
                hPos[4 * i + 1] = y0;
                hPos[4 * i + 2] = twoD ? zmid : p.origin.z + 0.5f * Lz;
                hPos[4 * i + 3] = p.particleRadius;
                hTangent[4 * i] = flip ? -1.f : 1.f;
                hTangent[4 * i + 1] = 0.f;
                hTangent[4 * i + 2] = 0.f;
                hTangent[4 * i + 3] = 0.f;
            }
        }
    }
    // Maxwellian velocities at kT (unit mass), Box-Muller on host draws
    auto normal = [&]() {
        const float u1 = 1.0f - frand(), u2 = frand();
        return std::sqrt(-2.0f * std::log(u1)) * std::cos(twopi * u2);
    };
    const float sv = std::sqrt(kT / p.particleMass);
    for (uint i = 0; i < W * n; ++i) {
        float *V = &hVel[4 * size_t(i)];
        V[0] = sv * normal();
        V[1] = sv * normal();
        V[2] = twoD ? 0.f : sv * normal();
        V[3] = p.particleMass;
    }
}

}  // namespace nwhost

// worm_init.h -- host initial conditions of worm systems (shared by the
// single-system driver and the replica set).
#pragma once

#include <cstdint>

#include "nw_abi.h"
#include "nw_system.h"

namespace nwhost {

// host uniform stream: draw d is lane d % 4 of Philox block d / 4 under
// `key`, a 24-bit uniform in [0, 1)
struct HostStream {
    uint64_t key = 0;
    uint64_t draws = 0;
    float next();
};

// W x n worms of p: mode NW_INIT_RANDOM places uniform heads and headings,
// rejecting any worm with a bead closer than sigma to an earlier worm's bead;
// NW_INIT_LATTICE_RANDOM lays straight worms in rows along x with random
// polarity.  Velocities are Maxwellian at kT (bead mass p.particleMass).
// pos / vel / tangent: float4 per bead.  Throws std::runtime_error when a
// worm cannot be placed.
void place_worms(const SimParams &p, float sigma, int mode, float kT, HostStream &rng, float *pos, float *vel,
                 float *tangent);

}  // namespace nwhost

// system_capi.cpp -- C entry points of libnwhost.so (include/nw_system.h).
#include <cstring>
#include <cstdio>
#include <new>
#include <stdexcept>

#include "particle_system.h"
#include "replica_set.h"

struct NWSystem {
    ParticleSystem *ps;
};

struct NWReplicaSet {
    ReplicaSet *rs;
};

extern "C" {

NWSystem *nwSystemCreate(const NWSystemConfig *cfg) {
    if (!cfg || cfg->numFilaments == 0 || cfg->filamentSize == 0 || cfg->gridSize.x == 0 || cfg->gridSize.y == 0 ||
        cfg->gridSize.z == 0) {
        fprintf(stderr, "nwSystemCreate: invalid configuration\n");
        return nullptr;
    }
    NWSystem *s = new (std::nothrow) NWSystem;
    if (!s) return nullptr;
    s->ps = new ParticleSystem(*cfg);
    return s;
}

void nwSystemDestroy(NWSystem *sys) {
    if (!sys) return;
    delete sys->ps;
    delete sys;
}

SimParams *nwSystemParams(NWSystem *sys) { return &sys->ps->params(); }
SolventParams *nwSystemSolventParams(NWSystem *sys) { return &sys->ps->solvent(); }

void nwSystemReset(NWSystem *sys, int mode, float kT) {
    try {
        if (mode == NW_INIT_REFERENCE) sys->ps->reset();
        else sys->ps->resetRandom(mode, kT);
    } catch (const std::exception &e) {
        fprintf(stderr, "nwSystemReset: %s\n", e.what());
    }
}

void nwSystemUpdate(NWSystem *sys, float dt, uint32_t steps) {
    for (uint32_t i = 0; i < steps; ++i) sys->ps->update(dt);
}

void nwSystemSync(NWSystem *) { threadSync(); }

uint32_t nwSystemNumParticles(NWSystem *sys) { return uint32_t(sys->ps->getNumParticles()); }
uint32_t nwSystemNumSolvent(NWSystem *sys) { return sys->ps->getNumSolvent(); }
uint32_t nwSystemNumCells(NWSystem *sys) { return sys->ps->getNumCellsAlloc(); }
uint64_t nwSystemStep(NWSystem *sys) { return sys->ps->getStep(); }

// through the driver's accessors: arrays in original particle order
void nwSystemGetArray(NWSystem *sys, int which, float *host) {
    const auto a = static_cast<ParticleSystem::ParticleArray>(which);
    const size_t n = (a == ParticleSystem::POSITION || a == ParticleSystem::VELOCITY)
                         ? size_t(sys->ps->getNumParticles()) + sys->ps->getNumSolvent()
                         : size_t(sys->ps->getNumParticles());
    if (n) std::memcpy(host, sys->ps->getArray(a), n * 16);
}

void nwSystemSetArray(NWSystem *sys, int which, const float *host) {
    const auto a = static_cast<ParticleSystem::ParticleArray>(which);
    const size_t n = (a == ParticleSystem::POSITION || a == ParticleSystem::VELOCITY)
                         ? size_t(sys->ps->getNumParticles()) + sys->ps->getNumSolvent()
                         : size_t(sys->ps->getNumParticles());
    if (n) sys->ps->setArray(a, host, 0, int(n));
}

float *nwSystemDeviceArray(NWSystem *sys, int which) {
    return sys->ps->deviceArray(static_cast<ParticleSystem::ParticleArray>(which));
}

void nwSystemRngState(NWSystem *sys, uint64_t *seed, uint64_t *offset) {
    if (seed) *seed = sys->ps->seed();
    if (offset) *offset = sys->ps->rngOffset();
}

void nwSystemSetRngState(NWSystem *sys, uint64_t offset) { sys->ps->setRngOffset(offset); }

void nwSystemObservables(NWSystem *sys, NWObservables *out) { sys->ps->observables(out); }

void nwSystemWriteOutputs(NWSystem *sys, const char *fileName, int iteration, float dt) {
    sys->ps->writeOutputs(fileName, iteration, dt);
}

int nwSystemAutotune(NWSystem *sys, float dt, const char *cachePath, int useCache, NWTuning *out) {
    return sys->ps->autotune(dt, cachePath, useCache, out);
}

// ---- replica sets ----------------------------------------------------------
NWReplicaSet *nwReplicaSetCreate(const NWReplicaSpec *specs, uint32_t numReplicas, const NWReplicaPhysics *phys) {
    if ((numReplicas && !specs) || !phys) {
        fprintf(stderr, "nwReplicaSetCreate: NULL argument\n");
        return nullptr;
    }
    try {
        NWReplicaSet *s = new NWReplicaSet;
        s->rs = new ReplicaSet(specs, numReplicas, *phys);
        return s;
    } catch (const std::exception &e) {
        fprintf(stderr, "nwReplicaSetCreate: %s\n", e.what());
        return nullptr;
    }
}

void nwReplicaSetDestroy(NWReplicaSet *set) {
    if (!set) return;
    delete set->rs;
    delete set;
}

void nwReplicaSetReset(NWReplicaSet *set, int mode, float kT) {
    try {
        set->rs->reset(mode, kT);
    } catch (const std::exception &e) {
        fprintf(stderr, "nwReplicaSetReset: %s\n", e.what());
    }
}

void nwReplicaSetUpdate(NWReplicaSet *set, float dt, uint32_t steps, int form) { set->rs->update(dt, steps, form); }
void nwReplicaSetSync(NWReplicaSet *set) { set->rs->sync(); }
uint32_t nwReplicaSetNumReplicas(NWReplicaSet *set) { return set->rs->numReplicas(); }
uint64_t nwReplicaSetNumParticles(NWReplicaSet *set, int replica) { return set->rs->numParticles(replica); }

void nwReplicaSetGetArray(NWReplicaSet *set, int which, int replica, float *host) {
    set->rs->getArray(which, replica, host);
}

void nwReplicaSetSetArray(NWReplicaSet *set, int which, int replica, const float *host) {
    set->rs->setArray(which, replica, host);
}

void nwReplicaSetParams(NWReplicaSet *set, int replica, SimParams *p, NWExtParams *e) {
    set->rs->params(replica, p, e);
}

void nwReplicaSetRngState(NWReplicaSet *set, int replica, uint64_t *seed, uint64_t *offset) {
    set->rs->rngState(replica, seed, offset);
}

void nwReplicaSetObservables(NWReplicaSet *set, int replica, NWObservables *out) {
    set->rs->observables(replica, out);
}

}  // extern "C"

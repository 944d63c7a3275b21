// replica_set.cpp -- see replica_set.h.
#include "replica_set.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "worm_init.h"

namespace {
constexpr uint64_t kHostKey = 0xC0FFEE1234567ull;  // host uniform stream (initial conditions)
}

ReplicaSet::ReplicaSet(const NWReplicaSpec *specs, uint32_t numReplicas, const NWReplicaPhysics &phys)
    : m_phys(phys) {
    m_reps.resize(numReplicas);
    m_off.assign(numReplicas + 1, 0);
    const float cellMin = phys.cellMin > 0.0f ? phys.cellMin : std::pow(2.0f, 1.0f / 6.0f) * phys.wcaSigma;
    for (uint32_t r = 0; r < numReplicas; ++r) {
        const NWReplicaSpec &q = specs[r];
        if (q.filamentSize < 1 || q.filamentSize > 64) throw std::invalid_argument("replica worms of 1..64 beads");
        NWReplicaParams &R = m_reps[r];
        R = NWReplicaParams{};
        SimParams &p = R.params;
        const float L = q.boxSize;
        const uint32_t G = std::max(3u, uint32_t(std::floor(L / cellMin)));
        const float c = L / float(G);
        p.gridSize = nw_uint3{G, G, 1};
        p.numCells = G * G;
        p.cellSize = nw_float3{c, c, c};
        p.boxSize = nw_float3{L, L, c};
        p.origin = nw_float3{-L / 2.0f, -L / 2.0f, -c / 2.0f};
        p.boundaryX = p.boundaryY = p.boundaryZ = PERIODIC;
        p.numFilaments = q.numFilaments;
        p.filamentSize = q.filamentSize;
        p.numParticles = q.numFilaments * q.filamentSize;
        p.particleRadius = 0.5f;
        p.particleMass = 1.0f;
        p.bondSpringK = phys.bondK;
        p.bondSpringL = phys.bondL;
        p.bondBendingK = q.bendingK;
        p.activity = q.activity;
        p.reverseProbability = 0.0f;
        p.kbT = phys.kT;
        p.gamma = phys.gamma;
        p.spring = 0.5f;
        p.damping = 0.02f;
        p.shear = 0.1f;
        p.boundaryDamping = -0.5f;
        NWExtParams &e = R.ext;
        e.pairPotential = NW_PAIR_WCA;
        e.bondedMode = NW_BONDED_CHAIN;
        e.stencil = NW_STENCIL_2D;
        e.wcaEpsilon = phys.wcaEpsilon;
        e.wcaSigma = phys.wcaSigma;
        e.wcaCutoff = std::pow(2.0f, 1.0f / 6.0f) * phys.wcaSigma;
        e.driveForce = q.driveForce;
        R.seed = q.seed;
        R.offset = 0;
        R.langevin = 1;
        R.langevinGamma = phys.gamma;
        R.langevinKbT = phys.kT;
        m_off[r + 1] = m_off[r] + uint64_t(q.numFilaments) * q.filamentSize;
    }
    m_N = m_off[numReplicas];
    const size_t bytes = size_t(m_N) * 16;
    float **arrs[] = {&m_dPos, &m_dVel, &m_dForce, &m_dForceOld, &m_dTangent, &m_dSortedPos, &m_dSortedTangent};
    for (float **a : arrs) allocateArray(reinterpret_cast<void **>(a), std::max<size_t>(bytes, 16));
    allocateArray(reinterpret_cast<void **>(&m_dIndex), std::max<size_t>(size_t(m_N) * 4, 4));
    std::vector<float> zeros(size_t(m_N) * 4, 0.0f);
    if (m_N) {
        copyArrayToDevice(m_dForce, zeros.data(), 0, int(bytes));
        copyArrayToDevice(m_dForceOld, zeros.data(), 0, int(bytes));
        copyArrayToDevice(m_dTangent, zeros.data(), 0, int(bytes));
    }
}

ReplicaSet::~ReplicaSet() {
    float *arrs[] = {m_dPos, m_dVel, m_dForce, m_dForceOld, m_dTangent, m_dSortedPos, m_dSortedTangent};
    for (float *a : arrs)
        if (a) freeArray(a);
    if (m_dIndex) freeArray(m_dIndex);
    if (m_batchCurrent) nwBatchSetReplicas(nullptr, 0);
}

float *ReplicaSet::_array(int which) const {
    switch (which) {
        case 0: return m_dPos;
        case 1: return m_dVel;
        case 2: return m_dTangent;
        case 3: return m_dForce;
        case 4: return m_dForceOld;
        default: throw std::invalid_argument("array: 0 pos, 1 vel, 2 tangent, 3 force, 4 forceOld");
    }
}

uint64_t ReplicaSet::numParticles(int replica) const {
    if (replica < 0) return m_N;
    return m_off.at(size_t(replica) + 1) - m_off.at(size_t(replica));
}

void ReplicaSet::reset(int mode, float kT) {
    for (uint32_t r = 0; r < numReplicas(); ++r) {
        const SimParams &p = m_reps[r].params;
        const size_t n = size_t(numParticles(int(r)));
        std::vector<float> pos(n * 4), vel(n * 4), tan(n * 4, 0.0f), zero(n * 4, 0.0f);
        nwhost::HostStream rng;
        rng.key = m_reps[r].seed ^ kHostKey;
        nwhost::place_worms(p, m_reps[r].ext.wcaSigma, mode, kT, rng, pos.data(), vel.data(), tan.data());
        setArray(0, int(r), pos.data());
        setArray(1, int(r), vel.data());
        setArray(2, int(r), tan.data());
        setArray(3, int(r), zero.data());
        setArray(4, int(r), zero.data());
    }
}

void ReplicaSet::_uploadBatch(float dt) {
    std::vector<NWReplicaParams> reps = m_reps;
    for (NWReplicaParams &R : reps) R.langevinKbT = m_phys.langevinFdt ? m_phys.kT / dt : m_phys.kT;
    nwBatchSetReplicas(reps.data(), numReplicas());
    m_batchCurrent = true;
    m_batchDt = dt;
    m_batchSteps = 0;
}

void ReplicaSet::update(float dt, uint32_t steps, int form) {
    if (steps == 0 || m_N == 0) return;
    if (form == NW_REPLICAS_GRAPH || form == NW_REPLICAS_LAUNCHES) {
        if (!m_batchCurrent || m_batchDt != dt) _uploadBatch(dt);
        nwBatchMdStep(m_dPos, m_dVel, m_dForce, m_dForceOld, m_dTangent, m_dIndex, m_dSortedPos, nullptr,
                      m_dSortedTangent, dt, steps, form == NW_REPLICAS_GRAPH ? 1 : 0);
        m_batchSteps += steps;
        for (uint32_t r = 0; r < numReplicas(); ++r) nwBatchRngState(r, &m_reps[r].offset);
        // the library's batch now starts from these offsets: keep it (its
        // step counter and the offsets advance together)
        return;
    }
    // one replica per launch sequence: the reference's way of running many systems
    m_batchCurrent = false;
    const float kTn = m_phys.langevinFdt ? m_phys.kT / dt : m_phys.kT;
    for (uint32_t k = 0; k < steps; ++k) {
        for (uint32_t r = 0; r < numReplicas(); ++r) {
            NWReplicaParams &R = m_reps[r];
            const size_t o = size_t(m_off[r]) * 4;
            setParameters(&R.params);
            nwSetExtParameters(&R.ext);
            nwRngSetSeed(R.seed, R.offset);
            const bool tan = R.params.activity != 0.0f;
            nwFusedMdStep(m_dPos + o, m_dVel + o, m_dForce + o, m_dForceOld + o, m_dTangent + o, nullptr,
                          m_dIndex + m_off[r], m_dSortedPos + o, nullptr, tan ? m_dSortedTangent + o : nullptr, dt,
                          R.langevinGamma, kTn, R.langevin, R.params.numFilaments);
            nwRngGetState(nullptr, &R.offset);
        }
    }
}

void ReplicaSet::sync() { threadSync(); }

void ReplicaSet::getArray(int which, int replica, float *host) {
    const size_t o = size_t(m_off.at(size_t(replica))) * 16, n = size_t(numParticles(replica)) * 16;
    if (n) copyArrayFromDevice(host, reinterpret_cast<char *>(_array(which)) + o, nullptr, int(n));
}

void ReplicaSet::setArray(int which, int replica, const float *host) {
    const size_t o = size_t(m_off.at(size_t(replica))) * 16, n = size_t(numParticles(replica)) * 16;
    if (n) copyArrayToDevice(_array(which), host, int(o), int(n));
}

void ReplicaSet::params(int replica, SimParams *p, NWExtParams *e) const {
    if (p) *p = m_reps.at(size_t(replica)).params;
    if (e) *e = m_reps.at(size_t(replica)).ext;
}

void ReplicaSet::rngState(int replica, uint64_t *seed, uint64_t *offset) const {
    if (seed) *seed = m_reps.at(size_t(replica)).seed;
    if (offset) *offset = m_reps.at(size_t(replica)).offset;
}

void ReplicaSet::observables(int replica, NWObservables *out) {
    const NWReplicaParams &R = m_reps.at(size_t(replica));
    setParameters(const_cast<SimParams *>(&R.params));
    const size_t o = size_t(m_off[size_t(replica)]) * 4;
    nwObservables(m_dPos + o, m_dVel + o, R.params.numFilaments, out);
}

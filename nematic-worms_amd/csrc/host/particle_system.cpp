// particle_system.cpp -- see particle_system.h.  Every device operation goes
// through the C-ABI of include/nw_abi.h (libnwamd.so).
#include "particle_system.h"
#include "worm_init.h"

#include <algorithm>
#include <utility>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>

namespace {
constexpr uint64_t kHostKey = 0xC0FFEE1234567ull;      // host uniform stream (initial conditions)
constexpr uint64_t kShiftKey = 0x5eed5eed5eed5eedull;  // SRD grid shift stream

float u01_open(uint32_t x) { return float(x >> 8) * 5.9604644775390625e-8f; }  // [0, 1)

nw_float3 f3(float x, float y, float z) { return nw_float3{x, y, z}; }
}  // namespace

// --------------------------------------------------------------------------
// construction (particleSystem.cpp:39-111, 124-148, 172-223)
// --------------------------------------------------------------------------
void ParticleSystem::_defaults(uint numFilaments, uint filamentSize, nw_uint3 gridSize, bool srd) {
    m_numFilaments = numFilaments;
    m_filamentSize = filamentSize;
    m_numParticles = numFilaments * filamentSize;
    m_gridSize = gridSize;
    m_solventEnabled = srd;

    m_params.numFilaments = m_numFilaments;
    m_params.filamentSize = m_filamentSize;
    m_params.numParticles = m_numParticles;
    m_params.particleRadius = 0.5f;
    m_params.particleMass = 1.0f;

    m_numGridCells = gridSize.x * gridSize.y * gridSize.z;
    m_params.gridSize = gridSize;
    m_params.numCells = m_numGridCells;
    const float cellSize = m_params.particleRadius * 2.0f;
    m_params.cellSize = f3(cellSize, cellSize, cellSize);
    m_params.boxSize = f3(gridSize.x * cellSize, gridSize.y * cellSize, gridSize.z * cellSize);
    m_params.origin = f3(-m_params.boxSize.x / 2.0f, -m_params.boxSize.y / 2.0f, -m_params.boxSize.z / 2.0f);
    m_params.boundaryX = m_params.boundaryY = m_params.boundaryZ = PERIODIC;

    m_params.bondSpringK = 57.f;
    m_params.bondSpringL = m_params.particleRadius * 0.8f;
    m_params.bondBendingK = 0.001f;
    m_params.spring = 0.5f;
    m_params.damping = 0.02f;
    m_params.shear = 0.1f;
    m_params.attraction = 0.0f;
    m_params.boundaryDamping = -0.5f;
    m_params.activity = 0.1f;
    m_params.reverseProbability = 0.0f;
    m_params.kbT = 1.0f;
    m_params.gamma = 0.01f;
    m_params.gravity = f3(0.f, 0.f, 0.f);
}

void ParticleSystem::_initSolventParams(const NWSystemConfig *cfg) {
    if (!m_solventEnabled) return;
    m_solvent.enabled = true;
    m_solvent.alpha = (cfg && cfg->srdAlpha != 0.0f) ? cfg->srdAlpha : 2.26893f;  // 130 deg
    nw_uint3 g{m_params.gridSize.x / 2, m_params.gridSize.y / 2, std::max(m_params.gridSize.z / 2, 1u)};
    if (cfg && cfg->srdGridSize.x) g = cfg->srdGridSize;
    m_solvent.gridSize = g;
    if (cfg && cfg->srdCellSize != 0.0f) {
        // cubic cells of the given size; an axis one cell deep spans the box
        m_solvent.cellSize = f3(cfg->srdCellSize, cfg->srdCellSize,
                                g.z == 1 ? m_params.boxSize.z : cfg->srdCellSize);
    } else {
        m_solvent.cellSize = f3(m_params.cellSize.x * 2, m_params.cellSize.y * 2, m_params.cellSize.z * 2);
    }
    m_solvent.numCells = g.x * g.y * g.z;
    m_solvent.randOffset = f3(0.f, 0.f, 0.f);
    m_solvent.particleMass = 1.0f;
    m_solvent.kbT = m_params.kbT;
    const float density = (cfg && cfg->srdDensity != 0.0f) ? cfg->srdDensity : 10.f;
    m_solventUpdateGap = (cfg && cfg->srdGap) ? cfg->srdGap : 5;
    m_numSolvent = uint(density * m_solvent.numCells);
    m_numSolventCells = m_solvent.numCells;
    m_solvent.numParticles = m_numSolvent;
}

ParticleSystem::ParticleSystem(uint numFilaments, uint filamentSize, nw_uint3 gridSize, bool srdSolvent) {
    _defaults(numFilaments, filamentSize, gridSize, srdSolvent);
    _initSolventParams(nullptr);
    _initialize();
}

ParticleSystem::ParticleSystem(const NWSystemConfig &cfg) {
    _defaults(cfg.numFilaments, cfg.filamentSize, cfg.gridSize, cfg.srd != 0);
    if (cfg.cellSize != 0.0f) {
        const float c = cfg.cellSize;
        m_params.cellSize = f3(c, c, c);
        m_params.boxSize = f3(m_gridSize.x * c, m_gridSize.y * c, m_gridSize.z * c);
    }
    if (cfg.boxSize.x != 0.0f) {
        // box given: cells tile it exactly
        m_params.boxSize = cfg.boxSize;
        m_params.cellSize = f3(cfg.boxSize.x / m_gridSize.x, cfg.boxSize.y / m_gridSize.y,
                               cfg.boxSize.z / m_gridSize.z);
    }
    m_params.origin = f3(-m_params.boxSize.x / 2.0f, -m_params.boxSize.y / 2.0f, -m_params.boxSize.z / 2.0f);
    m_ext = cfg.ext;
    m_seed = cfg.seed;
    m_shiftCentred = cfg.srdShiftCentred != 0;
    m_langevinFdt = cfg.langevinFdt != 0;
    m_perCall = cfg.perCallSequence != 0;
    m_solventSortEvery = cfg.srd ? cfg.solventSortEvery : 0u;
    _initSolventParams(&cfg);
    _initialize();
}

ParticleSystem::~ParticleSystem() { _finalize(); }

void ParticleSystem::_initialize() {
    const size_t N = m_numParticles, NT = size_t(m_numParticles) + m_numSolvent;
    m_hPos.assign(NT * 4, 0.f);
    m_hVel.assign(NT * 4, 0.f);
    m_hForce.assign(N * 4, 0.f);
    m_hTangent.assign(N * 4, 0.f);
    m_numCellsAlloc = std::max(m_numGridCells, m_solventEnabled ? m_numSolventCells : 0u);
    m_hCellStart.assign(m_numCellsAlloc, 0);
    m_hCellEnd.assign(m_numCellsAlloc, 0);

    const size_t memSize = sizeof(float) * 4 * N, memSizeN = sizeof(float) * 4 * NT;
    // random: 4N normals, W uniforms or Cs uniforms (the reference sizes it 4N only)
    const size_t rndFloats = std::max<size_t>({4 * N, m_numFilaments, m_solventEnabled ? m_numSolventCells : 0, 4});
    allocateArray((void **)&m_dPos, memSizeN);
    allocateArray((void **)&m_dVel, memSizeN);
    allocateArray((void **)&m_dTangent, memSize);
    allocateArray((void **)&m_dForce, memSize);
    allocateArray((void **)&m_dForceOld, memSize);
    allocateArray((void **)&m_dRandom, rndFloats * sizeof(float));
    allocateArray((void **)&m_dSortedPos, memSizeN);
    allocateArray((void **)&m_dSortedVel, memSizeN);
    allocateArray((void **)&m_dSortedTangent, memSize);
    allocateArray((void **)&m_dGridParticleHash, NT * sizeof(uint));
    allocateArray((void **)&m_dGridParticleIndex, NT * sizeof(uint));
    allocateArray((void **)&m_dCellStart, size_t(m_numCellsAlloc) * sizeof(uint));
    allocateArray((void **)&m_dCellEnd, size_t(m_numCellsAlloc) * sizeof(uint));
    if (m_solventEnabled) allocateArray((void **)&m_dSolventCellCOM, size_t(m_numSolventCells) * 16);
    if (m_solventEnabled && m_numSolvent) allocateArray((void **)&m_dSolventId, size_t(m_numSolvent) * sizeof(uint));
    _solventOrderReset();

    // the reference never clears force / forceOld before the first step
    _upload(m_dForce, m_hForce, N);
    _upload(m_dForceOld, m_hForce, N);
    _upload(m_dTangent, m_hTangent, N);
    std::vector<float> zeros(size_t(m_numCellsAlloc), 0.f);
    copyArrayToDevice(m_dCellEnd, zeros.data(), 0, int(zeros.size() * sizeof(float)));
    m_bInitialized = true;
}

void ParticleSystem::_finalize() {
    if (!m_bInitialized) return;
    float *fa[] = {m_dPos, m_dVel, m_dTangent, m_dForce, m_dForceOld, m_dRandom, m_dSortedPos, m_dSortedVel,
                   m_dSortedTangent, m_dSolventCellCOM};
    for (float *p : fa)
        if (p) freeArray(p);
    uint *ua[] = {m_dGridParticleHash, m_dGridParticleIndex, m_dCellStart, m_dCellEnd, m_dSolventId};
    for (uint *p : ua)
        if (p) freeArray(p);
    m_bInitialized = false;
}

void ParticleSystem::_upload(float *dev, const std::vector<float> &host, size_t count4) {
    if (count4) copyArrayToDevice(dev, host.data(), 0, int(count4 * 4 * sizeof(float)));
}

// The device solvent in original order again (after new positions and
// velocities arrive in that order), then sorted when re-sorting is on.
void ParticleSystem::_solventOrderReset() {
    m_solventPermuted = false;
    if (!m_dSolventId) return;
    m_hSolventId.resize(m_numSolvent);
    for (uint k = 0; k < m_numSolvent; ++k) m_hSolventId[k] = k;
    copyArrayToDevice(m_dSolventId, m_hSolventId.data(), 0, int(m_numSolvent * sizeof(uint)));
}

void ParticleSystem::_sortSolvent() {
    if (!m_dSolventId) return;
    setParameters(&m_params);
    setSolventParameters(&m_solvent);
    nwSortSolvent(m_dPos, m_dVel, m_dSolventId, m_numParticles, m_numSolvent);
    m_solventPermuted = true;
}

void ParticleSystem::_maybeSortSolvent() {
    if (m_solventSortEvery && m_solventUpdateCount % m_solventSortEvery == 0) _sortSolvent();
}

// --------------------------------------------------------------------------
// the step (particleSystem.cpp:249-339)
// --------------------------------------------------------------------------
void ParticleSystem::observables(NWObservables *out) {
    setParameters(&m_params);
    nwObservables(m_dPos, m_dVel, m_numFilaments, out);
}

// the launch shapes of the fused steps for this system's sizes, timed on a
// copy of its state with the arguments _updateFused passes
int ParticleSystem::autotune(float deltaTime, const char *cachePath, int useCache, NWTuning *out) {
    if (!fusedActive()) return 0;
    setParameters(&m_params);
    nwSetExtParameters(&m_ext);
    if (m_solventEnabled) setSolventParameters(&m_solvent);
    const float kT = m_langevinFdt ? m_params.kbT / deltaTime : m_params.kbT;
    return nwAutotuneFused(m_dPos, m_dVel, m_dForce, m_dForceOld, m_dTangent, m_numFilaments,
                           m_solventEnabled ? m_numSolvent : 0u, deltaTime, m_params.gamma, kT,
                           m_solventEnabled ? 0 : 1, deltaTime * float(m_solventUpdateGap),
                           m_ext.srdThermostat == NW_SRD_THERMO_ISOKINETIC ? 1 : 0, 0, cachePath, useCache, out);
}

bool ParticleSystem::fusedActive() const {
    if (m_perCall || m_filamentSize < 1 || m_filamentSize > 64) return false;
    const uint mdCells = m_params.gridSize.x * m_params.gridSize.y * m_params.gridSize.z;
    if (!nwFusedPlanFits(m_numParticles, mdCells, m_filamentSize)) return false;
    return !m_solventEnabled || nwFusedPlanFits(m_numParticles + m_numSolvent, m_numSolventCells, 0);
}

void ParticleSystem::update(float deltaTime) {
    if (fusedActive()) {
        _updateFused(deltaTime);
        return;
    }
    _materializeTangents();
    nwRngSetSeed(m_seed, m_rngOffset);
    setParameters(&m_params);
    nwSetExtParameters(&m_ext);

    integrateFilaments(m_dPos, m_dVel, m_dForce, m_dForceOld, deltaTime, m_numParticles);

    if (m_params.reverseProbability > 0.0f)
        reverseFilaments(m_dForce, m_dForceOld, m_dVel, m_dPos, m_dRandom, m_numFilaments);

    if (!m_solventEnabled) {
        const float kT = m_langevinFdt ? m_params.kbT / deltaTime : m_params.kbT;
        langevinThermostat(m_dForce, m_dVel, m_dRandom, m_params.gamma, kT, m_numParticles);
    }

    filamentForces(m_dForce, m_dTangent, m_dPos, float(m_numFilaments));

    calcHash(m_dGridParticleHash, m_dGridParticleIndex, m_dPos, int(m_numParticles));
    sortParticles(m_dGridParticleHash, m_dGridParticleIndex, m_numParticles);

    // the WCA pair pass reads positions only; skip gathering what it does not read
    const bool needVT = m_ext.pairPotential == NW_PAIR_REF_DEM || m_params.activity != 0.0f;
    reorderDataAndFindCellStart(m_dCellStart, m_dCellEnd, m_dSortedPos, needVT ? m_dSortedVel : nullptr,
                                needVT ? m_dSortedTangent : nullptr, m_dGridParticleHash, m_dGridParticleIndex,
                                m_dPos, needVT ? m_dVel : nullptr, needVT ? m_dTangent : nullptr, m_numParticles,
                                m_numGridCells);

    collideFilaments(m_dForce, m_dSortedPos, m_dSortedVel, m_dSortedTangent, m_dGridParticleIndex, m_dCellStart,
                     m_dCellEnd, m_numParticles, m_numGridCells);

    if (m_solventEnabled && (m_updateCount % m_solventUpdateGap) == 0) _solventUpdate(deltaTime * m_solventUpdateGap);

    m_updateCount++;
    uint64_t seed = 0;
    nwRngGetState(&seed, &m_rngOffset);
}

// The same step through the fused entry points: one call for the MD part,
// one per SRD collision.
void ParticleSystem::_updateFused(float deltaTime) {
    nwRngSetSeed(m_seed, m_rngOffset);
    setParameters(&m_params);
    nwSetExtParameters(&m_ext);
    const float kT = m_langevinFdt ? m_params.kbT / deltaTime : m_params.kbT;
    const bool needVel = m_ext.pairPotential == NW_PAIR_REF_DEM;
    const bool needTan = needVel || m_params.activity != 0.0f;
    // a step whose pair term reads no tangents stores none; they are rebuilt
    // from the positions when the array is read (nwFilamentTangents)
    float *tan = needTan ? m_dTangent : nullptr;
    m_tangentStale = !needTan;
    if (m_params.reverseProbability == 0.0f) {
        // double-buffered forces: the new force lands in the old force's
        // buffer, then the two swap roles (both start cleared, w = 0)
        nwFusedMdStepSwapForces(m_dPos, m_dVel, m_dForce, m_dForceOld, tan, m_dGridParticleHash,
                                m_dGridParticleIndex, m_dSortedPos, needVel ? m_dSortedVel : nullptr,
                                needTan ? m_dSortedTangent : nullptr, deltaTime, m_params.gamma, kT,
                                m_solventEnabled ? 0 : 1, m_numFilaments);
        std::swap(m_dForce, m_dForceOld);
    } else {
        nwFusedMdStep(m_dPos, m_dVel, m_dForce, m_dForceOld, tan, m_dGridParticleHash, m_dGridParticleIndex,
                      m_dSortedPos, needVel ? m_dSortedVel : nullptr, needTan ? m_dSortedTangent : nullptr,
                      deltaTime, m_params.gamma, kT, m_solventEnabled ? 0 : 1, m_numFilaments);
    }
    if (m_solventEnabled && (m_updateCount % m_solventUpdateGap) == 0) {
        _srdShift();
        nwFusedSrdStep(m_dPos, m_dVel, m_dSolventCellCOM, m_dGridParticleHash, m_dGridParticleIndex,
                       deltaTime * m_solventUpdateGap, m_numParticles, m_numSolvent,
                       m_ext.srdThermostat == NW_SRD_THERMO_ISOKINETIC ? 1 : 0);
        m_solventUpdateCount++;
        _maybeSortSolvent();
    }
    m_updateCount++;
    uint64_t seed = 0;
    nwRngGetState(&seed, &m_rngOffset);
}

// SRD grid shift of this collision (particleSystem.cpp:345-349)
void ParticleSystem::_srdShift() {
    uint32_t b[4];
    nwPhiloxBlock(m_seed ^ kShiftKey, m_solventUpdateCount, b);
    const float ux = u01_open(b[0]), uy = u01_open(b[1]);
    m_solvent.randOffset.x = (m_shiftCentred ? ux - 0.5f : ux) * m_solvent.cellSize.x;
    m_solvent.randOffset.y = (m_shiftCentred ? uy - 0.5f : uy) * m_solvent.cellSize.y;
    m_solvent.randOffset.z = 0.0f;
    setSolventParameters(&m_solvent);
}

// particleSystem.cpp:341-413
void ParticleSystem::_solventUpdate(float dt) {
    _srdShift();

    const uint NT = m_numParticles + m_numSolvent;
    calcHash(m_dGridParticleHash, m_dGridParticleIndex, m_dPos, int(NT), true);
    sortParticles(m_dGridParticleHash, m_dGridParticleIndex, NT);
    reorderDataAndFindCellStart(m_dCellStart, m_dCellEnd, m_dSortedPos, m_dSortedVel, nullptr, m_dGridParticleHash,
                                m_dGridParticleIndex, m_dPos, m_dVel, nullptr, NT, m_numCellsAlloc);
    if (m_ext.srdThermostat == NW_SRD_THERMO_ISOKINETIC)
        solventIsokineticThermostat(m_dVel, m_dSortedVel, m_dCellStart, m_dCellEnd, m_dGridParticleIndex,
                                    m_numSolventCells);
    solventCellCentorOfMomentum(m_dSolventCellCOM, m_dSortedVel, m_dCellStart, m_dCellEnd, m_dGridParticleIndex,
                                m_numSolventCells);
    collideSolvent(m_dPos, m_dVel, m_dSolventCellCOM, m_dRandom, m_numSolventCells, NT);
    integrateSolvent(m_dPos, m_dVel, dt, m_numParticles, m_numSolvent);
    m_solventUpdateCount++;
    _maybeSortSolvent();
}

// --------------------------------------------------------------------------
// initial conditions
// --------------------------------------------------------------------------
float ParticleSystem::_frand() {
    uint32_t b[4];
    nwPhiloxBlock(m_seed ^ kHostKey, m_hostDraws / 4, b);
    const uint32_t x = b[m_hostDraws % 4];
    m_hostDraws++;
    return u01_open(x);
}

// The reference's start (particleSystem.cpp:583-678), restated: every worm
// gets a slot of n bond lengths plus one bead diameter along x; slots fill a
// row of the box (as many as fit), rows sit one diameter apart in y and
// layers one diameter apart in z, the first row and layer one diameter in.
// A worm lies straight in its slot, head at the slot's left end; then every
// worm, in order, draws its heading, and a reversed worm (draw >= 1/2) takes
// its lattice points from tail to head -- except an odd worm's middle bead,
// which keeps its point and its +x tangent.  Velocities are uniform in
// [-0.1, 0.1) per component.  Host draws come in the reference's order:
// three per bead, one per worm, then the solvent's six per particle.
namespace {
struct WormLattice {
    float slot, radius, bond;
    uint perRow, perLayer;
    // lattice point k (0 .. n-1 along the slot) of worm w, relative to the
    // origin, in the reference's operation order
    void point(uint w, uint k, float out[3]) const {
        const uint col = w % perRow, row = (w / perRow) % (perLayer / perRow), layer = w / perLayer;
        const float headX = float(col) * slot + radius;
        out[0] = headX + float(k) * bond;
        out[1] = 2.f * radius * float(row + 1);
        out[2] = 2.f * radius * float(layer + 1);
    }
};
}  // namespace

void ParticleSystem::reset() {
    const SimParams &p = m_params;
    const uint n = p.filamentSize;
    WormLattice lat;
    lat.radius = p.particleRadius;
    lat.bond = p.bondSpringL;
    lat.slot = float(n) * lat.bond + 2.0f * lat.radius;
    lat.perRow = uint(std::max(1, int(std::floor(p.boxSize.x / lat.slot))));
    lat.perLayer = lat.perRow * uint(std::max(1, int(std::floor(p.boxSize.y / (2 * lat.radius)))));

    auto symmetric = [this](float amp) { return amp * (_frand() * 2.0f - 1.0f); };
    for (uint i = 0; i < m_numParticles; ++i) {
        float *V = &m_hVel[4 * size_t(i)];
        V[0] = symmetric(0.1f);
        V[1] = symmetric(0.1f);
        V[2] = symmetric(0.1f);
        V[3] = p.particleMass;
    }
    for (uint w = 0; w < p.numFilaments; ++w) {
        const bool reversed = !(_frand() < 0.5f);
        for (uint q = 0; q < n; ++q) {
            // a reversed worm's middle bead (odd n) keeps its place and tangent
            const bool moved = reversed && 2 * q + 1 != n;
            const uint k = moved ? n - 1 - q : q;
            float r[3];
            lat.point(w, k, r);
            const size_t i = size_t(w) * n + q;
            float *P = &m_hPos[4 * i], *T = &m_hTangent[4 * i];
            P[0] = r[0] + p.origin.x;
            P[1] = r[1] + p.origin.y;
            P[2] = r[2] + p.origin.z;
            P[3] = p.particleRadius;
            const float sgn = moved ? -1.0f : 1.0f;
            T[0] = sgn;
            T[1] = T[2] = T[3] = sgn * 0.0f;
        }
    }
    if (m_solventEnabled) {
        const float v0 = 5.0f;  // the reference's solvent speed scale
        for (uint s = 0; s < m_numSolvent; s++) {
            float *P = &m_hPos[4 * (size_t(m_numParticles) + s)], *V = &m_hVel[4 * (size_t(m_numParticles) + s)];
            P[0] = p.boxSize.x * _frand() + p.origin.x;
            P[1] = p.boxSize.y * _frand() + p.origin.y;
            P[2] = p.boxSize.z * _frand() + p.origin.z;
            P[3] = 0.0f;
            V[0] = symmetric(v0);
            V[1] = symmetric(v0);
            V[2] = symmetric(v0);
            V[3] = m_solvent.particleMass;
        }
    }
    const size_t NT = size_t(m_numParticles) + m_numSolvent;
    _upload(m_dPos, m_hPos, NT);
    _upload(m_dVel, m_hVel, NT);
    _upload(m_dTangent, m_hTangent, m_numParticles);
    m_tangentStale = false;
    _solventOrderReset();
    if (m_solventSortEvery) _sortSolvent();
}

void ParticleSystem::resetRandom(int mode, float kT) {
    const SimParams &p = m_params;
    const bool twoD = p.gridSize.z == 1;
    const float Lx = p.boxSize.x, Ly = p.boxSize.y, Lz = p.boxSize.z;
    const float zmid = p.origin.z + 0.5f * Lz;
    const float sigma = m_ext.pairPotential == NW_PAIR_WCA ? m_ext.wcaSigma : 2.0f * p.particleRadius;
    const float twopi = 6.28318530717958647692f;
    nwhost::HostStream rng;
    rng.key = m_seed ^ kHostKey;
    rng.draws = m_hostDraws;
    nwhost::place_worms(p, sigma, mode, kT, rng, m_hPos.data(), m_hVel.data(), m_hTangent.data());
    m_hostDraws = rng.draws;
    auto normal = [&]() {
        const float u1 = 1.0f - _frand(), u2 = _frand();
        return std::sqrt(-2.0f * std::log(u1)) * std::cos(twopi * u2);
    };
    if (m_solventEnabled) {
        const float ss = std::sqrt(kT / m_solvent.particleMass);
        for (uint s = 0; s < m_numSolvent; ++s) {
            const size_t i = size_t(m_numParticles) + s;
            m_hPos[4 * i] = p.origin.x + _frand() * Lx;
            m_hPos[4 * i + 1] = p.origin.y + _frand() * Ly;
            m_hPos[4 * i + 2] = twoD ? zmid : p.origin.z + _frand() * Lz;
            m_hPos[4 * i + 3] = 0.f;
            m_hVel[4 * i] = ss * normal();
            m_hVel[4 * i + 1] = ss * normal();
            m_hVel[4 * i + 2] = twoD ? 0.f : ss * normal();
            m_hVel[4 * i + 3] = m_solvent.particleMass;
        }
    }
    const size_t NT = size_t(m_numParticles) + m_numSolvent;
    std::fill(m_hForce.begin(), m_hForce.end(), 0.f);
    _upload(m_dPos, m_hPos, NT);
    _upload(m_dVel, m_hVel, NT);
    _upload(m_dTangent, m_hTangent, m_numParticles);
    m_tangentStale = false;
    _upload(m_dForce, m_hForce, m_numParticles);
    _upload(m_dForceOld, m_hForce, m_numParticles);
    _solventOrderReset();
    if (m_solventSortEvery) _sortSolvent();
}

// --------------------------------------------------------------------------
// arrays, dumps, output
// --------------------------------------------------------------------------
void ParticleSystem::_materializeTangents() {
    if (!m_tangentStale) return;
    setParameters(&m_params);
    nwFilamentTangents(m_dTangent, m_dPos, m_numFilaments);
    m_tangentStale = false;
}

float *ParticleSystem::deviceArray(ParticleArray a) {
    if (a == TANGENT) _materializeTangents();
    switch (a) {
        case POSITION: return m_dPos;
        case VELOCITY: return m_dVel;
        case TANGENT: return m_dTangent;
        case FORCE: return m_dForce;
        case FORCE_OLD: return m_dForceOld;
    }
    return nullptr;
}

float *ParticleSystem::getArray(ParticleArray a) {
    const bool full = a == POSITION || a == VELOCITY;
    const size_t n = full ? size_t(m_numParticles) + m_numSolvent : m_numParticles;
    std::vector<float> &h = a == POSITION ? m_hPos : a == VELOCITY ? m_hVel : a == TANGENT ? m_hTangent : m_hForce;
    if (h.size() < 4 * n) h.resize(4 * n);
    if (n) copyArrayFromDevice(h.data(), deviceArray(a), nullptr, int(n * 16));
    if (full && m_solventPermuted) {
        // slot k of the device solvent holds original particle m_hSolventId[k]
        copyArrayFromDevice(m_hSolventId.data(), m_dSolventId, nullptr, int(m_numSolvent * sizeof(uint)));
        const size_t N = m_numParticles;
        std::vector<float> tmp(h.begin() + 4 * N, h.begin() + 4 * (N + m_numSolvent));
        for (size_t k = 0; k < m_numSolvent; ++k)
            std::copy(&tmp[4 * k], &tmp[4 * k] + 4, &h[4 * (N + m_hSolventId[k])]);
    }
    return h.data();
}

void ParticleSystem::setArray(ParticleArray a, const float *data, int start, int count) {
    if (count <= 0) return;
    if (a == TANGENT) _materializeTangents();  // a partial write keeps the rest current
    const bool full = a == POSITION || a == VELOCITY;
    if (full && m_solventPermuted && size_t(start) + size_t(count) > m_numParticles) {
        // the whole array in original order, the new values in, then back in slot order
        const float *g = getArray(a);
        std::vector<float> cur(g, g + 4 * (size_t(m_numParticles) + m_numSolvent));
        std::copy(data, data + 4 * size_t(count), cur.begin() + 4 * size_t(start));
        const size_t N = m_numParticles;
        std::vector<float> dev(cur.begin(), cur.end());
        for (size_t k = 0; k < m_numSolvent; ++k)
            std::copy(&cur[4 * (N + m_hSolventId[k])], &cur[4 * (N + m_hSolventId[k])] + 4, &dev[4 * (N + k)]);
        copyArrayToDevice(deviceArray(a), dev.data(), 0, int(dev.size() * sizeof(float)));
        return;
    }
    copyArrayToDevice(deviceArray(a), data, start * 4 * int(sizeof(float)), count * 4 * int(sizeof(float)));
}

void ParticleSystem::dumpGrid() {
    copyArrayFromDevice(m_hCellStart.data(), m_dCellStart, nullptr, int(sizeof(uint) * m_numGridCells));
    copyArrayFromDevice(m_hCellEnd.data(), m_dCellEnd, nullptr, int(sizeof(uint) * m_numGridCells));
    uint maxCellSize = 0;
    for (uint i = 0; i < m_numGridCells; i++)
        if (m_hCellStart[i] != 0xffffffffu) maxCellSize = std::max(maxCellSize, m_hCellEnd[i] - m_hCellStart[i]);
    printf("maximum particles per cell = %u\n", maxCellSize);
}

void ParticleSystem::dumpSolventGrid() {
    if (!m_solventEnabled) return;
    copyArrayFromDevice(m_hCellStart.data(), m_dCellStart, nullptr, int(sizeof(uint) * m_numSolventCells));
    copyArrayFromDevice(m_hCellEnd.data(), m_dCellEnd, nullptr, int(sizeof(uint) * m_numSolventCells));
    std::vector<float> com(4 * size_t(m_numSolventCells));
    copyArrayFromDevice(com.data(), m_dSolventCellCOM, nullptr, int(com.size() * sizeof(float)));
    uint maxCellSize = 0;
    for (uint i = 0; i < m_numSolventCells; i++) {
        if (m_hCellStart[i] == 0xffffffffu) continue;
        const uint c = m_hCellEnd[i] - m_hCellStart[i];
        if (c > maxCellSize) {
            printf("[%u] %u CoM = {%f %f} \n", i, c, com[4 * size_t(i)], com[4 * size_t(i) + 1]);
            maxCellSize = c;
        }
    }
    printf("maximum particles per solvent cell = %u\n", maxCellSize);
}

void ParticleSystem::dumpParticles(uint start, uint count) {
    getArray(POSITION);
    getArray(VELOCITY);
    for (uint i = start; i < start + count && i < m_numParticles + m_numSolvent; i++) {
        const float *P = &m_hPos[4 * size_t(i)], *V = &m_hVel[4 * size_t(i)];
        printf("pos: (%.4f, %.4f, %.4f, %.4f)\n", P[0], P[1], P[2], P[3]);
        printf("vel: (%.4f, %.4f, %.4f, %.4f)\n", V[0], V[1], V[2], V[3]);
    }
}

// particleSystem.cpp:481-532: N, a JSON parameter line, then "A|B x y z vx vy vz"
void ParticleSystem::writeOutputs(const std::string &fileName, int iteration, float timeDelta) {
    const size_t N = size_t(m_numParticles) + m_numSolvent;
    getArray(POSITION);
    getArray(VELOCITY);
    std::ofstream fout(fileName, std::ios::out | std::ios::app);
    fout << N << std::endl;
    const SimParams &p = m_params;
    fout << "{"
         << "\"iteration\":" << iteration << ","
         << "\"dt\":" << timeDelta << ","
         << "\"time\":" << iteration * timeDelta << ","
         << "\"num_particles\":" << p.numParticles << ","
         << "\"num_filaments\":" << p.numFilaments << ","
         << "\"filament_size\":" << p.filamentSize << ","
         << "\"bond_spring_coef\":" << p.bondSpringK << ","
         << "\"bond_spring_length\":" << p.bondSpringL << ","
         << "\"bond_bending_coef\":" << p.bondBendingK << ","
         << "\"activity\":" << p.activity << ","
         << "\"kbT\":" << p.kbT << ","
         << "\"gamma\":" << p.gamma << ","
         << "\"particle_radius\":" << p.particleRadius << ","
         << "\"particle_mass\":" << p.particleMass << ","
         << "\"num_cells\":" << p.numCells << ","
         << "\"grid_size\":" << "[" << p.gridSize.x << "," << p.gridSize.y << "," << p.gridSize.z << "],"
         << "\"box_size\":" << "[" << p.boxSize.x << "," << p.boxSize.y << "," << p.boxSize.z << "],"
         << "\"cell_size\":" << "[" << p.cellSize.x << "," << p.cellSize.y << "," << p.cellSize.z << "],"
         << "\"boundaries\":" << "[" << int(p.boundaryX) << "," << int(p.boundaryY) << "," << int(p.boundaryZ)
         << "],"
         << "\"reverse_probability\":" << p.reverseProbability << ","
         << "\"collision_spring_coef\":" << p.spring << ","
         << "\"collision_damping_coef\":" << p.damping << ","
         << "\"collision_shear_coef\":" << p.shear << ","
         << "\"collision_attraction_coef\":" << p.attraction << ","
         << "\"boundary_damping_coef\":" << p.boundaryDamping << "}" << std::endl;
    for (size_t i = 0; i < N; i++) {
        const char *type = i < m_numParticles ? "A" : "B";
        fout << type << " " << m_hPos[i * 4] << " " << m_hPos[i * 4 + 1] << " " << m_hPos[i * 4 + 2] << " "
             << m_hVel[i * 4] << " " << m_hVel[i * 4 + 1] << " " << m_hVel[i * 4 + 2] << std::endl;
    }
}

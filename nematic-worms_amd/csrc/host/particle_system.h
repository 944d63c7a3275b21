// particle_system.h -- host orchestrator of the worm time step.
//
// Restates the reference's ParticleSystem (simulation/particleSystem.h:24-209,
// simulation/particleSystem.cpp): same constructor, defaults, setters, array
// access, per-step call order and .xyzv output, driving the device only
// through the C-ABI of include/nw_abi.h.  Additions: construction from an
// NWSystemConfig (box/cell/SRD/physics variants of the benchmark workloads),
// seeded initial conditions, zero-initialised force arrays.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "nw_abi.h"
#include "nw_system.h"

class ParticleSystem {
public:
    // particleSystem.cpp:39-111
    ParticleSystem(uint numFilaments, uint filamentSize, nw_uint3 gridSize, bool srdSolvent = false);
    explicit ParticleSystem(const NWSystemConfig &cfg);
    ~ParticleSystem();

    enum ParticleArray { POSITION, VELOCITY, TANGENT, FORCE, FORCE_OLD };

    void update(float deltaTime);           // particleSystem.cpp:249-339
    void reset();                           // particleSystem.cpp:583-678
    void resetRandom(int mode, float kT);   // NW_INIT_RANDOM / NW_INIT_LATTICE_RANDOM

    float *getArray(ParticleArray array);   // host copy of N (+ solvent for pos/vel) records
    void setArray(ParticleArray array, const float *data, int start, int count);
    float *deviceArray(ParticleArray array);

    int getNumParticles() const { return int(m_numParticles); }
    uint getNumSolvent() const { return m_numSolvent; }
    uint getNumCellsAlloc() const { return m_numCellsAlloc; }
    uint64_t getStep() const { return m_updateCount; }

    void dumpGrid();
    void dumpSolventGrid();
    void dumpParticles(uint start, uint count);
    void writeOutputs(const std::string &fileName, int iteration, float timeDelta);
    // nematic / polar order and speed along the tangent of the worms (on the device)
    void observables(NWObservables *out);
    // the library's autotuner on a copy of this system's state (fused steps only)
    int autotune(float deltaTime, const char *cachePath, int useCache, NWTuning *out);

    // setters / getters (particleSystem.h:53-138)
    void setDamping(float x) { m_params.gamma = x; }
    void setTemperature(float x) { m_params.kbT = x; m_solvent.kbT = x; }
    void setGravity(float x) { m_params.gravity = nw_float3{0.0f, x, 0.0f}; }
    void setActivity(float x) { m_params.activity = x; }
    void setReverseProbability(float x) { m_params.reverseProbability = x; }
    void setCollideSpring(float x) { m_params.spring = x; }
    void setCollideDamping(float x) { m_params.damping = x; }
    void setCollideShear(float x) { m_params.shear = x; }
    void setCollideAttraction(float x) { m_params.attraction = x; }
    void setBondSpringConstant(float x) { m_params.bondSpringK = x; }
    void setBondSpringLength(float x) { m_params.bondSpringL = x; }
    void setBondBendingConstant(float x) { m_params.bondBendingK = x; }
    void setBoundaries(BoundaryType x, BoundaryType y, BoundaryType z) {
        m_params.boundaryX = x;
        m_params.boundaryY = y;
        m_params.boundaryZ = z;
    }
    float getParticleRadius() const { return m_params.particleRadius; }
    nw_uint3 getGridSize() const { return m_params.gridSize; }
    nw_float3 getOrigin() const { return m_params.origin; }
    nw_float3 getCellSize() const { return m_params.cellSize; }

    SimParams &params() { return m_params; }
    SolventParams &solvent() { return m_solvent; }
    NWExtParams &ext() { return m_ext; }
    uint64_t rngOffset() const { return m_rngOffset; }
    void setRngOffset(uint64_t o) { m_rngOffset = o; }
    uint64_t seed() const { return m_seed; }
    void setPerCallSequence(bool on) { m_perCall = on; }
    // the fused entry points when their sort plans fit, else the per-call sequence
    bool fusedActive() const;

private:
    void _defaults(uint numFilaments, uint filamentSize, nw_uint3 gridSize, bool srd);
    void _initSolventParams(const NWSystemConfig *cfg);
    void _solventUpdate(float dt);
    void _updateFused(float deltaTime);
    void _srdShift();
    void _initialize();
    void _finalize();
    float _frand();  // host uniform in [0, 1] (stands in for rand()/RAND_MAX)
    void _upload(float *dev, const std::vector<float> &host, size_t count4);
    // the solvent kept in SRD-cell order (every m_solventSortEvery collisions);
    // m_dSolventId maps each solvent slot to its original index
    void _solventOrderReset();
    void _maybeSortSolvent();
    void _sortSolvent();

    bool m_bInitialized = false;
    uint64_t m_updateCount = 0;
    uint m_numParticles = 0, m_numFilaments = 0, m_filamentSize = 0;

    std::vector<float> m_hPos, m_hVel, m_hTangent, m_hForce;
    std::vector<uint> m_hCellStart, m_hCellEnd;

    float *m_dPos = nullptr, *m_dVel = nullptr, *m_dTangent = nullptr, *m_dForce = nullptr, *m_dForceOld = nullptr;
    float *m_dRandom = nullptr, *m_dSortedPos = nullptr, *m_dSortedVel = nullptr, *m_dSortedTangent = nullptr;
    uint *m_dGridParticleHash = nullptr, *m_dGridParticleIndex = nullptr, *m_dCellStart = nullptr,
         *m_dCellEnd = nullptr;
    float *m_dSolventCellCOM = nullptr;

    SimParams m_params{};
    SolventParams m_solvent{};
    NWExtParams m_ext{};
    nw_uint3 m_gridSize{};
    uint m_numGridCells = 0;
    uint m_numCellsAlloc = 0;

    bool m_solventEnabled = false;
    uint m_numSolvent = 0, m_numSolventCells = 0, m_solventUpdateCount = 0, m_solventUpdateGap = 5;
    bool m_shiftCentred = false;
    bool m_langevinFdt = false;
    bool m_perCall = false;     // reference per-call sequence instead of the fused entry points
    uint m_solventSortEvery = 0;    // collisions between solvent re-sorts (0: never)
    bool m_solventPermuted = false; // the device solvent is not in original order
    // the fused steps stored no tangents (the pair term reads none): the
    // device array is rebuilt from the positions when it is read
    bool m_tangentStale = false;
    void _materializeTangents();
    uint *m_dSolventId = nullptr;
    std::vector<uint> m_hSolventId;

    uint64_t m_seed = 0;
    uint64_t m_rngOffset = 0;   // device generator offset owned by this system
    uint64_t m_hostDraws = 0;   // host uniform stream position
};

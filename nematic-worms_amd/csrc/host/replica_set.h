// replica_set.h -- many independent worm systems stepped together through the
// batched fused step of the C-ABI (nwBatchMdStep), or one replica per launch
// sequence (nwFusedMdStep per replica) for comparison.
#pragma once

#include <cstdint>
#include <vector>

#include "nw_abi.h"
#include "nw_system.h"

class ReplicaSet {
public:
    ReplicaSet(const NWReplicaSpec *specs, uint32_t numReplicas, const NWReplicaPhysics &phys);
    ~ReplicaSet();
    ReplicaSet(const ReplicaSet &) = delete;
    ReplicaSet &operator=(const ReplicaSet &) = delete;

    void reset(int mode, float kT);
    void update(float dt, uint32_t steps, int form);
    void sync();

    uint32_t numReplicas() const { return uint32_t(m_reps.size()); }
    uint64_t numParticles(int replica) const;
    void getArray(int which, int replica, float *host);
    void setArray(int which, int replica, const float *host);
    void params(int replica, SimParams *p, NWExtParams *e) const;
    void rngState(int replica, uint64_t *seed, uint64_t *offset) const;
    void observables(int replica, NWObservables *out);

private:
    float *_array(int which) const;
    void _uploadBatch(float dt);

    NWReplicaPhysics m_phys;
    std::vector<NWReplicaParams> m_reps;  // offset = the replica's generator position now
    std::vector<uint64_t> m_off;          // first bead of every replica (+ total)
    uint64_t m_N = 0;
    float *m_dPos = nullptr, *m_dVel = nullptr, *m_dForce = nullptr, *m_dForceOld = nullptr;
    float *m_dTangent = nullptr, *m_dSortedPos = nullptr, *m_dSortedTangent = nullptr;
    uint *m_dIndex = nullptr;
    bool m_batchCurrent = false;  // the library's batch holds these replicas at their offsets
    float m_batchDt = 0.0f;
    uint64_t m_batchSteps = 0;    // batched steps since the upload
};

// nw_abi.hip -- the extern "C" surface of libnwamd.so (declared in
// include/nw_abi.h).  Mirrors simulation/particleSystem_cuda.cu:36-371: same
// names, arguments, data layout and error behaviour, with the work done by the
// HIP kernels of nw_kernels.hip / nw_sort.hip on one library stream.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <map>
#include <string>
#include <vector>

#include "nw_abi.h"
#include "nw_internal.hpp"
#include "nw_bucket.hpp"
#include "nw_observe.hpp"
#include "nw_adjoint.hpp"

namespace {

using namespace nw;

struct Timer {
    bool enabled = false;
    uint32_t stride = 1, seen = 0;  // time every stride-th launch
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
    double totalMs = 0.0;
    uint64_t launches = 0;
};

struct State {
    bool initialized = false;
    int device = 0;
    hipStream_t stream = nullptr;
    bool ownStream = false;
    SimParams params{};
    SolventParams solvent{};
    NWExtParams ext{};
    uint64_t seed = 0;
    uint64_t offset = 0;
    int errorMode = NW_ERRORS_EXIT;
    int lastError = 0;
    std::string lastErrorString;
    uint32_t lastHashBound = 0;   // cell count of the most recent calcHash
    const uint32_t *lastHashPtr = nullptr;
    bool lastHashSrd = false;     // timer family of the sort / reorder that follow
    SortScratch sort;
    BucketScratch bucket;
    double *obsPartial = nullptr;  // observe_rows x kObsSums, then kObsSums totals
    uint64_t obsCap = 0;
    // batched replicas
    struct Batch {
        std::vector<BatchRep> reps;
        std::vector<uint32_t> beadOff, cellOff;
        BatchRep *dev = nullptr;
        uint64_t *stepCtr = nullptr;
        uint32_t R = 0, N = 0, cells = 0, waves = 0, blocks = 0;
        int mode = 1;
        bool tan = false, unit = true;
        uint64_t steps = 0;
        BucketScratch scratch;
        std::vector<SortRep> sort;  // every replica's own sort plan
        SortRep *sortDev = nullptr;
        BatchSortPlan sortPlan;
        // the step as graphs: [0] force -> forceOld, [1] the buffers' roles
        // exchanged (double-buffered forces, when no replica reverses)
        hipGraphExec_t exec[2] = {nullptr, nullptr};
        hipGraph_t graph[2] = {nullptr, nullptr};
        std::vector<uintptr_t> key;  // arguments the graphs were captured with
        bool swap = false;
    } batch;
    // adjoint of the fused MD step: checkpoints of the last nwAdjointForward
    struct Adjoint {
        bool ready = false;
        KP kp{};
        float dt = 0.f, gamma = 0.f, kT = 0.f, lgvStd = 0.f;
        int langevin = 0;
        uint32_t W = 0, N = 0, C = 0, steps = 0, K = 0;
        uint64_t seed = 0, endOffset = 0;
        std::vector<uint64_t> ckptOffset;  // generator offset at every checkpoint
        // checkpoints x (2 NT + 2 N): pos, vel (beads then solvent), force, forceOld
        float4 *ckpt = nullptr;
        // SRD collisions of the run (nwAdjointForwardSrd): the solvent follows
        // the beads, a collision after steps 0, gap, 2 gap, ...
        bool srd = false;
        uint32_t Ns = 0, NT = 0, Cs = 0, gap = 0;
        int rotBeads = 0;
        float cosA = 0.f, sinA = 0.f;
        std::vector<float> shifts;  // the grid shift of every collision, 3 floats each
        // a segment's collisions: cell of every particle, cell order, CSR
        // offsets, (u, sign) per cell; the full state while recomputing
        uint32_t *collHash = nullptr, *collIds = nullptr, *collOff = nullptr;
        float4 *collCom = nullptr, *cellSum = nullptr, *wPos = nullptr, *wVel = nullptr;
        uint64_t capCollHash = 0, capCollIds = 0, capCollOff = 0, capCollCom = 0, capCellSum = 0, capWPos = 0,
                 capWVel = 0;
        // one segment: states after every step and the cell lists of every step
        float4 *segPos = nullptr, *segVel = nullptr, *segSPos = nullptr, *work = nullptr, *lam = nullptr;
        uint32_t *segIds = nullptr, *segOff = nullptr, *segCell = nullptr;
        double *partial = nullptr, *grad = nullptr;
        uint64_t capCkpt = 0, capPos = 0, capVel = 0, capSPos = 0, capIds = 0, capOff = 0, capCell = 0,
                 capWork = 0, capLam = 0, capPartial = 0;
    } adj;
    // SRD solvent of the batch (nwBatchSetSolvent)
    struct BatchSolvent {
        std::vector<SrdRep> reps;
        SrdRep *dev = nullptr;
        std::vector<uint32_t> solvOff, cellOff;  // R + 1 each: first solvent particle, first SRD cell
        uint32_t R = 0, Ns = 0, cells = 0;
        float cosA = 0.f, sinA = 0.f;
        int rotBeads = 0;
        BucketScratch sort;  // the SRD cell sort of the whole batch
        float *shift = nullptr;  // nwBatchSrdStep: the replicas' grid shifts on the device
        uint64_t capShift = 0;
    } bsol;
    // adjoint of a batched run: checkpoints of the last nwBatchAdjointForward
    struct BatchAdjoint {
        bool ready = false, srd = false;
        uint32_t steps = 0, K = 0, gap = 0, N = 0, NT = 0, C = 0, Cs = 0, R = 0, waves = 0;
        uint64_t ctr0 = 0;  // the batch's step count before the run
        float dt = 0.f;
        float4 *ckpt = nullptr;   // checkpoints x (2 NT + 2 N): pos, vel, force, forceOld
        float *shifts = nullptr;  // collisions x R x 3 grid shifts
        float4 *segPos = nullptr, *segVel = nullptr, *segSPos = nullptr, *work = nullptr, *lam = nullptr;
        float4 *wPos = nullptr, *wVel = nullptr, *collCom = nullptr, *cellSum = nullptr;
        uint32_t *segIds = nullptr, *segOff = nullptr, *hash = nullptr, *sCell = nullptr;
        uint32_t *collHash = nullptr, *collIds = nullptr, *collOff = nullptr;
        double *partial = nullptr, *grad = nullptr;
        uint64_t capCkpt = 0, capShifts = 0, capPos = 0, capVel = 0, capSPos = 0, capWork = 0, capLam = 0,
                 capWPos = 0, capWVel = 0, capCollCom = 0, capCellSum = 0, capIds = 0, capOff = 0, capHash = 0,
                 capSCell = 0, capCollHash = 0, capCollIds = 0, capCollOff = 0, capPartial = 0, capGrad = 0;
    } badj;
    // nwSortSolvent: keys, order, and the reordered solvent before it is copied back
    uint32_t *solKey = nullptr, *solIdx = nullptr, *solId = nullptr;
    float4 *solPos = nullptr, *solVel = nullptr;
    uint64_t capSolKey = 0, capSolIdx = 0, capSolId = 0, capSolPos = 0, capSolVel = 0;
    std::map<std::string, Timer> timers;
};

State &S() {
    static State s;
    return s;
}

// Records an error; in exit mode prints and terminates like helper_cuda.h's check().
bool fail(const char *file, int line, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    State &s = S();
    s.lastError = code ? code : -1;
    s.lastErrorString = buf;
    if (s.errorMode == NW_ERRORS_EXIT) {
        fprintf(stderr, "%s:%d: %s\n", file, line, buf);
        exit(EXIT_FAILURE);
    }
    return false;
}

#define NW_HIP(call)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) {                                                                      \
            fail(__FILE__, __LINE__, int(e_), "HIP error %d (%s) in %s", int(e_), hipGetErrorString(e_), \
                 #call);                                                                             \
            return;                                                                                  \
        }                                                                                            \
    } while (0)

#define NW_HIP_R(call, ret)                                                                          \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) {                                                                      \
            fail(__FILE__, __LINE__, int(e_), "HIP error %d (%s) in %s", int(e_), hipGetErrorString(e_), \
                 #call);                                                                             \
            return ret;                                                                              \
        }                                                                                            \
    } while (0)

#define NW_REQUIRE(cond, ...)                                  \
    do {                                                       \
        if (!(cond)) {                                         \
            fail(__FILE__, __LINE__, -22, __VA_ARGS__);        \
            return;                                            \
        }                                                      \
    } while (0)

#define NW_REQUIRE_R(cond, ret, ...)                           \
    do {                                                       \
        if (!(cond)) {                                         \
            fail(__FILE__, __LINE__, -22, __VA_ARGS__);        \
            return ret;                                        \
        }                                                      \
    } while (0)

}  // namespace

// ---------------------------------------------------------------------------
// Scratch of the allocation-free entry point (nwFusedMdStepNoAlloc): one
// array per device, held until the process ends.  The host-side name of the
// array is not a device address: host code uses only the address
// hipGetSymbolAddress returns for the current device.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kNoAllocMaxDevices = 64;
void *g_noallocScratch[kNoAllocMaxDevices] = {};

// the address of `scratch` on the current device, resolved once per device;
// false (and *out NULL) when the device or the symbol cannot be resolved
bool noalloc_scratch(void **out, int *device, hipError_t *err) {
    *out = nullptr;
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess && (dev < 0 || dev >= kNoAllocMaxDevices)) e = hipErrorInvalidDevice;
    if (e == hipSuccess && !g_noallocScratch[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e == hipSuccess && ptr == nullptr) e = hipErrorInvalidSymbol;
        if (e == hipSuccess) g_noallocScratch[dev] = ptr;
    }
    *device = dev;
    *err = e;
    if (e != hipSuccess) return false;
    *out = g_noallocScratch[dev];
    return true;
}

bool ensure_init() {
    State &s = S();
    if (s.initialized) return true;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        fail(__FILE__, __LINE__, -19, "no HIP device available (hipGetDeviceCount: %s)", hipGetErrorString(e));
        return false;
    }
    if (s.device >= count) s.device = 0;
    if ((e = hipSetDevice(s.device)) != hipSuccess) {
        fail(__FILE__, __LINE__, int(e), "hipSetDevice(%d): %s", s.device, hipGetErrorString(e));
        return false;
    }
    if ((e = hipStreamCreate(&s.stream)) != hipSuccess) {
        fail(__FILE__, __LINE__, int(e), "hipStreamCreate: %s", hipGetErrorString(e));
        return false;
    }
    s.ownStream = true;
    s.initialized = true;
    // resolve the allocation-free entry point's scratch on this device now
    // (set-up time); a failure is reported by that entry point when called
    void *ptr = nullptr;
    int dev = 0;
    hipError_t se = hipSuccess;
    (void)noalloc_scratch(&ptr, &dev, &se);
    return true;
}

#define NW_INIT()                  \
    do {                           \
        if (!ensure_init()) return; \
    } while (0)

// the multiplier of div_exact for a divisor n: ceil(2^32 / n), below 2^32
// for n >= 2 (Granlund & Montgomery 1994); 0 for n = 1 (and the invalid 0)
static uint32_t div_magic_for(uint64_t n) {
    if (n < 2) return 0u;
    const uint64_t two32 = 1ull << 32;
    return uint32_t((two32 + n - 1) / n);
}

KP make_kp_for(const SimParams &p, const SolventParams &sp, const NWExtParams &ext) {
    KP kp{};
    kp.p = p;
    kp.s = sp;
    kp.e = ext;
    kp.fdivMagic = div_magic_for(p.filamentSize);
    kp.gxMagic = div_magic_for(p.gridSize.x);
    kp.layerMagic = div_magic_for(uint64_t(p.gridSize.x) * p.gridSize.y);
    return kp;
}

KP make_kp() { return make_kp_for(S().params, S().solvent, S().ext); }

// A timed family's launches between two events on the stream; a family of
// a single launch (single = true) hands the events to that launch's dispatch
// packet instead, which times the kernel alone.
struct TimedScope {
    Timer *t = nullptr;
    hipStream_t st;
    bool single = false;
    std::pair<hipEvent_t, hipEvent_t> ev{};
    TimedScope(const char *name, hipStream_t stream, bool singleLaunch = false) : st(stream), single(singleLaunch) {
        auto it = S().timers.find(name);
        if (it == S().timers.end() || !it->second.enabled) return;
        if (it->second.seen++ % it->second.stride != 0) return;
        t = &it->second;
        if (!t->pool.empty()) {
            ev = t->pool.back();
            t->pool.pop_back();
        } else {
            (void)hipEventCreate(&ev.first);
            (void)hipEventCreate(&ev.second);
        }
        if (single) {
            LaunchTiming &lt = launch_timing();
            lt.start = ev.first;
            lt.stop = ev.second;
            lt.used = false;
        } else {
            (void)hipEventRecord(ev.first, st);
        }
    }
    ~TimedScope() {
        if (!t) return;
        if (single) {
            LaunchTiming &lt = launch_timing();
            const bool used = lt.used;
            lt = LaunchTiming{};
            if (!used) {  // nothing launched (an empty system)
                t->pool.push_back(ev);
                return;
            }
        } else {
            (void)hipEventRecord(ev.second, st);
        }
        t->pending.push_back(ev);
        t->launches++;
    }
};

uint64_t blocks_for(uint64_t n) { return (n + 3) / 4; }

}  // namespace

namespace nw {
LaunchTiming &launch_timing() {
    static thread_local LaunchTiming lt;
    return lt;
}
KP current_kp() { return make_kp(); }
}  // namespace nw

extern "C" {

// ---------------------------------------------------------------------------
// init / RNG / memory
// ---------------------------------------------------------------------------
void cudaInit(int argc, char **argv) {
    State &s = S();
    for (int i = 1; i < argc && argv; ++i) {
        const char *a = argv[i];
        while (*a == '-') ++a;
        if (strncasecmp(a, "device=", 7) == 0) s.device = atoi(a + 7);
    }
    if (!ensure_init()) {
        // reference: "No CUDA Capable devices found, exiting..." + exit(EXIT_SUCCESS)
        return;
    }
    s.seed = 0;
    s.offset = 0;
}

void randomizeUniform(float *device, uint n) {
    NW_INIT();
    State &s = S();
    NW_HIP(launch_randomize(device, n, 0, 0.f, 1.f, s.seed, s.offset, s.stream));
    s.offset += blocks_for(n);
}

void randomizeGaussian(float *device, uint n, float mean, float stddev) {
    NW_INIT();
    State &s = S();
    NW_HIP(launch_randomize(device, n, 1, mean, stddev, s.seed, s.offset, s.stream));
    s.offset += blocks_for(n);
}

void allocateArray(void **devPtr, size_t size) {
    NW_INIT();
    NW_HIP(hipMalloc(devPtr, size));
}

void freeArray(void *devPtr) {
    NW_INIT();
    NW_HIP(hipFree(devPtr));
}

void threadSync(void) {
    NW_INIT();
    NW_HIP(hipDeviceSynchronize());
}

void copyArrayFromDevice(void *host, const void *device, struct cudaGraphicsResource **, int size) {
    NW_INIT();
    State &s = S();
    if (size <= 0) return;
    NW_HIP(hipMemcpyAsync(host, device, size_t(size), hipMemcpyDeviceToHost, s.stream));
    NW_HIP(hipStreamSynchronize(s.stream));
}

void copyArrayToDevice(void *device, const void *host, int offset, int size) {
    NW_INIT();
    State &s = S();
    if (size <= 0) return;
    NW_HIP(hipMemcpyAsync(static_cast<char *>(device) + offset, host, size_t(size), hipMemcpyHostToDevice, s.stream));
    NW_HIP(hipStreamSynchronize(s.stream));
}

void setParameters(SimParams *hostParams) {
    NW_REQUIRE(hostParams != nullptr, "setParameters: NULL");
    S().params = *hostParams;
}

void setSolventParameters(SolventParams *solventParams) {
    NW_REQUIRE(solventParams != nullptr, "setSolventParameters: NULL");
    S().solvent = *solventParams;
}

uint iDivUp(uint a, uint b) { return (a % b != 0) ? (a / b + 1) : (a / b); }

void computeGridSize(uint n, uint blockSize, uint *numBlocks, uint *numThreads) {
    *numThreads = n < blockSize ? n : blockSize;
    *numBlocks = *numThreads ? iDivUp(n, *numThreads) : 0;
}

// ---------------------------------------------------------------------------
// the time step
// ---------------------------------------------------------------------------
void integrateFilaments(float *pos, float *vel, float *f, float *f_old, float deltaTime, uint numParticles) {
    NW_INIT();
    State &s = S();
    TimedScope ts("integrate", s.stream);
    NW_HIP(launch_integrate_filaments(reinterpret_cast<float4 *>(pos), reinterpret_cast<float4 *>(vel),
                                      reinterpret_cast<float4 *>(f), reinterpret_cast<float4 *>(f_old), deltaTime,
                                      numParticles, make_kp(), s.stream));
}

void integrateSystem(float *pos, float *vel, float *f, float *f_old, float deltaTime, uint numParticles) {
    integrateFilaments(pos, vel, f, f_old, deltaTime, numParticles);
}

void integrateSolvent(float *pos, float *vel, float deltaTime, uint startIndex, uint numSolvent) {
    NW_INIT();
    State &s = S();
    TimedScope ts("srd_stream", s.stream);
    NW_HIP(launch_integrate_solvent(reinterpret_cast<float4 *>(pos), reinterpret_cast<float4 *>(vel), deltaTime,
                                    startIndex, numSolvent, make_kp(), s.stream));
}

void calcHash(uint *gridParticleHash, uint *gridParticleIndex, float *pos, int numParticles, bool srd) {
    NW_INIT();
    State &s = S();
    NW_REQUIRE(numParticles >= 0, "calcHash: negative particle count %d", numParticles);
    TimedScope ts(srd ? "srd_hash" : "hash", s.stream);
    NW_HIP(launch_calc_hash(gridParticleHash, gridParticleIndex, reinterpret_cast<const float4 *>(pos),
                            uint32_t(numParticles), srd, make_kp(), s.stream));
    // every key calcHash writes lies below the product of the grid's sides
    // (numCells may disagree with it; the sort must never clamp a key)
    const nw_uint3 g = srd ? s.solvent.gridSize : s.params.gridSize;
    const uint64_t bound = uint64_t(g.x) * g.y * g.z;
    s.lastHashBound = bound > 0 && bound <= 0xffffffffull ? uint32_t(bound) : 0u;
    s.lastHashPtr = gridParticleHash;
    s.lastHashSrd = srd;
}

void sortParticles(uint *dGridParticleHash, uint *dGridParticleIndex, uint numParticles) {
    NW_INIT();
    State &s = S();
    TimedScope ts(s.lastHashSrd ? "srd_sort" : "sort", s.stream);
    if (s.lastHashPtr == dGridParticleHash && s.lastHashBound > 0) {
        // calcHash wrote these keys: every one below the cell count
        const hipError_t e = bucket_sort_pairs(dGridParticleHash, dGridParticleIndex, numParticles, s.lastHashBound,
                                               s.bucket, s.stream);
        if (e == hipErrorNotSupported)
            NW_HIP(counting_sort_pairs(dGridParticleHash, dGridParticleIndex, numParticles, s.lastHashBound, s.sort,
                                       s.stream));
        else
            NW_HIP(e);
    } else {
        NW_HIP(radix_sort_pairs(dGridParticleHash, dGridParticleIndex, numParticles, 32, s.sort, s.stream));
    }
}

void nwSortParticlesBounded(uint *dGridParticleHash, uint *dGridParticleIndex, uint numParticles, uint keyBound) {
    NW_INIT();
    State &s = S();
    NW_REQUIRE(keyBound > 0, "nwSortParticlesBounded: keyBound must be > 0");
    TimedScope ts("sort", s.stream);
    NW_HIP(counting_sort_pairs(dGridParticleHash, dGridParticleIndex, numParticles, keyBound, s.sort, s.stream));
}

void nwRadixSortPairs(uint *dGridParticleHash, uint *dGridParticleIndex, uint numParticles, uint keyBits) {
    NW_INIT();
    State &s = S();
    NW_REQUIRE(keyBits >= 1 && keyBits <= 32, "nwRadixSortPairs: keyBits %u outside 1..32", keyBits);
    TimedScope ts("sort", s.stream);
    NW_HIP(radix_sort_pairs(dGridParticleHash, dGridParticleIndex, numParticles, keyBits, s.sort, s.stream));
}

void reorderDataAndFindCellStart(uint *cellStart, uint *cellEnd, float *sortedPos, float *sortedVel,
                                 float *sortedTangent, uint *gridParticleHash, uint *gridParticleIndex, float *pos,
                                 float *vel, float *tangent, uint numParticles, uint numCells) {
    NW_INIT();
    State &s = S();
    TimedScope ts(s.lastHashSrd ? "srd_reorder" : "reorder", s.stream);
    NW_HIP(launch_reorder(cellStart, cellEnd, reinterpret_cast<float4 *>(sortedPos),
                          reinterpret_cast<float4 *>(sortedVel), reinterpret_cast<float4 *>(sortedTangent),
                          gridParticleHash, gridParticleIndex, reinterpret_cast<const float4 *>(pos),
                          reinterpret_cast<const float4 *>(vel), reinterpret_cast<const float4 *>(tangent),
                          numParticles, numCells, s.stream));
}

void collideFilaments(float *force, float *sortedPos, float *sortedVel, float *sortedTangent,
                      uint *gridParticleIndex, uint *cellStart, uint *cellEnd, uint numParticles, uint numCells) {
    NW_INIT();
    State &s = S();
    const KP kp = make_kp();
    NW_REQUIRE(kp.p.filamentSize > 0, "collideFilaments: filamentSize is 0 (call setParameters first)");
    NW_REQUIRE(kp.p.gridSize.x > 0 && kp.p.gridSize.y > 0 && kp.p.gridSize.z > 0,
               "collideFilaments: empty grid (call setParameters first)");
    NW_REQUIRE(uint64_t(kp.p.gridSize.x) * kp.p.gridSize.y * kp.p.gridSize.z <= numCells,
               "collideFilaments: grid of %u x %u x %u cells exceeds numCells %u", kp.p.gridSize.x,
               kp.p.gridSize.y, kp.p.gridSize.z, numCells);
    const bool needVel = kp.e.pairPotential == NW_PAIR_REF_DEM;
    const bool needTan = kp.e.pairPotential == NW_PAIR_REF_DEM || kp.p.activity != 0.0f;
    NW_REQUIRE(!needVel || sortedVel != nullptr, "collideFilaments: sortedVel is NULL");
    NW_REQUIRE(!needTan || sortedTangent != nullptr, "collideFilaments: sortedTangent is NULL");
    TimedScope ts("collide", s.stream);
    NW_HIP(launch_collide_filaments(reinterpret_cast<float4 *>(force), reinterpret_cast<const float4 *>(sortedPos),
                                    reinterpret_cast<const float4 *>(sortedVel),
                                    reinterpret_cast<const float4 *>(sortedTangent), gridParticleIndex, cellStart,
                                    cellEnd, numParticles, numCells, kp, s.stream));
}

void collide(float *force, float *sortedPos, float *sortedVel, float *sortedTangent, uint *gridParticleIndex,
             uint *cellStart, uint *cellEnd, uint numParticles, uint numCells) {
    collideFilaments(force, sortedPos, sortedVel, sortedTangent, gridParticleIndex, cellStart, cellEnd, numParticles,
                     numCells);
}

void filamentForces(float *force, float *tangent, float *pos, float numFilaments) {
    NW_INIT();
    State &s = S();
    NW_REQUIRE(numFilaments >= 0.0f && numFilaments < 4294967296.0f, "filamentForces: bad worm count %g",
               double(numFilaments));
    TimedScope ts("filament", s.stream);
    NW_HIP(launch_filament_forces(reinterpret_cast<float4 *>(force), reinterpret_cast<float4 *>(tangent),
                                  reinterpret_cast<const float4 *>(pos), uint32_t(numFilaments), make_kp(),
                                  s.stream));
}

void langevinThermostat(float *force, float *vel, float *random, float gamma, float kbT, uint numParticles) {
    (void)random;
    NW_INIT();
    State &s = S();
    const float stddev = sqrtf(2.0f * gamma * kbT);
    TimedScope ts("langevin", s.stream);
    NW_HIP(launch_langevin(reinterpret_cast<float4 *>(force), reinterpret_cast<const float4 *>(vel), stddev,
                           numParticles, s.seed, s.offset, make_kp(), s.stream));
    s.offset += numParticles;  // == blocks_for(4 * numParticles)
}

void reverseFilaments(float *force, float *forceOld, float *vel, float *pos, float *uniform, uint numFilaments) {
    (void)uniform;
    NW_INIT();
    State &s = S();
    TimedScope ts("reverse", s.stream);
    NW_HIP(launch_reverse_filaments(reinterpret_cast<float4 *>(force), reinterpret_cast<float4 *>(forceOld),
                                    reinterpret_cast<float4 *>(vel), reinterpret_cast<float4 *>(pos), numFilaments,
                                    s.seed, s.offset, make_kp(), s.stream));
    s.offset += blocks_for(numFilaments);
}

void solventCellCentorOfMomentum(float *solventCellCOM, float *sortedVel, uint *cellStart, uint *cellEnd,
                                 uint *gridParticleIndex, uint numCells) {
    (void)gridParticleIndex;
    NW_INIT();
    State &s = S();
    TimedScope ts("srd_com", s.stream);
    NW_HIP(launch_srd_com(reinterpret_cast<float4 *>(solventCellCOM), reinterpret_cast<const float4 *>(sortedVel),
                          cellStart, cellEnd, numCells, s.stream));
}

void solventIsokineticThermostat(float *vel, float *sortedVel, uint *cellStart, uint *cellEnd,
                                 uint *gridParticleIndex, uint numCells) {
    NW_INIT();
    State &s = S();
    TimedScope ts("srd_thermo", s.stream);
    NW_HIP(launch_srd_isokinetic(reinterpret_cast<float4 *>(vel), reinterpret_cast<float4 *>(sortedVel), cellStart,
                                 cellEnd, gridParticleIndex, numCells, make_kp(), s.stream));
}

void collideSolvent(float *pos, float *vel, float *solventCellCOM, float *random, uint numSolventCells,
                    uint numParticles) {
    (void)random;
    NW_INIT();
    State &s = S();
    // cos/sin of the rotation angle once, on the host, in double precision
    const double a = double(s.solvent.alpha);
    const float cosA = float(std::cos(a)), sinA = float(std::sin(a));
    TimedScope ts("srd_collide", s.stream);
    NW_HIP(launch_collide_solvent(reinterpret_cast<const float4 *>(pos), reinterpret_cast<float4 *>(vel),
                                  reinterpret_cast<const float4 *>(solventCellCOM), numSolventCells, numParticles,
                                  cosA, sinA, s.seed, s.offset, make_kp(), s.stream));
    s.offset += blocks_for(numSolventCells);
}

// ---------------------------------------------------------------------------
// fused step
// ---------------------------------------------------------------------------
// The fused MD step of nwFusedMdStep; offOut / cellOut (NULL: the library's
// scratch) receive the CSR cell offsets and the cell of every sorted slot.
// fixed: the scratch to use as it is (no allocation, no timing events), else
// the library's own, grown as needed.
static bool fused_md(const KP &kp, float *pos, float *vel, float *force, float *forceOld, float *tangent,
                     uint *gridParticleHash, uint *gridParticleIndex, float *sortedPos, float *sortedVel,
                     float *sortedTangent, float deltaTime, float langevinGamma, float langevinKbT, int langevin,
                     uint numFilaments, uint32_t *offOut, uint32_t *cellOut, const float *posIn = nullptr,
                     const float *velIn = nullptr, bool swapForces = false, const BucketScratch *fixed = nullptr) {
    State &s = S();
    const uint32_t n = kp.p.filamentSize;
    NW_REQUIRE_R(n >= 1 && n <= 64, false, "nwFusedMdStep: worms of %u beads (1..64 supported)", n);
    const uint64_t cells = uint64_t(kp.p.gridSize.x) * kp.p.gridSize.y * kp.p.gridSize.z;
    NW_REQUIRE_R(cells > 0 && cells <= kp.p.numCells, false, "nwFusedMdStep: grid does not fit numCells");
    const uint64_t N = uint64_t(numFilaments) * n;
    NW_REQUIRE_R(N < (1ull << 31), false, "nwFusedMdStep: too many beads");
    const bool needVel = kp.e.pairPotential == NW_PAIR_REF_DEM;
    const bool needTan = needVel || kp.p.activity != 0.0f;
    NW_REQUIRE_R(!needVel || sortedVel != nullptr, false, "nwFusedMdStep: sortedVel is NULL");
    NW_REQUIRE_R(!needTan || sortedTangent != nullptr, false, "nwFusedMdStep: sortedTangent is NULL");
    NW_REQUIRE_R(!needTan || tangent != nullptr, false,
                 "nwFusedMdStep: tangent is NULL but the pair term reads tangents");
    if (N == 0) return true;
    const BucketPlan bp = fused_plan(N, cells, false);
    NW_REQUIRE_R(!*bucket_plan_problem(bp), false, "nwFusedMdStep: %s", bucket_plan_problem(bp));
    if (!fixed) NW_HIP_R(ensure_bucket_scratch(s.bucket, bp), false);
    BucketScratch base = fixed ? *fixed : s.bucket;
    BucketScratch sc = base;
    if (offOut) sc.off = offOut;
    if (cellOut) sc.sortedCell = cellOut;
    const char *const noTimer = "";  // no timer carries this name

    NW_REQUIRE_R(!swapForces || kp.p.reverseProbability == 0.0f, false,
                 "nwFusedMdStepSwapForces: no polarity reversal (it moves the old force)");
    MdFrontArgs a;
    a.swapForces = swapForces ? 1 : 0;
    a.pos = reinterpret_cast<float4 *>(pos);
    a.vel = reinterpret_cast<float4 *>(vel);
    a.force = reinterpret_cast<float4 *>(force);
    a.forceOld = reinterpret_cast<float4 *>(forceOld);
    a.tangent = reinterpret_cast<float4 *>(tangent);
    a.posIn = reinterpret_cast<const float4 *>(posIn);
    a.velIn = reinterpret_cast<const float4 *>(velIn);
    a.hash = gridParticleHash ? gridParticleHash : base.hash;
    a.dt = deltaTime;
    a.W = numFilaments;
    a.seed = s.seed;
    a.offRev = s.offset;
    if (kp.p.reverseProbability > 0.0f) s.offset += blocks_for(numFilaments);
    a.offLgv = s.offset;
    a.langevin = langevin ? 1 : 0;
    a.lgvStd = sqrtf(2.0f * langevinGamma * langevinKbT);
    if (langevin) s.offset += N;
    {
        TimedScope ts(fixed ? noTimer : "md_front", s.stream, true);
        NW_HIP_R(launch_md_front(a, kp, s.stream), false);
    }
    {
        TimedScope ts(fixed ? noTimer : "md_tiles", s.stream, true);
        NW_HIP_R(launch_hash_tiles(a.hash, bp, base, s.stream), false);
    }
    {
        TimedScope ts(fixed ? noTimer : "md_sort", s.stream, true);
        NW_HIP_R(launch_md_cells(bp, sc, gridParticleIndex, a.pos, needVel ? a.vel : nullptr,
                               needTan ? a.tangent : nullptr, reinterpret_cast<float4 *>(sortedPos),
                               needVel ? reinterpret_cast<float4 *>(sortedVel) : nullptr,
                               needTan ? reinterpret_cast<float4 *>(sortedTangent) : nullptr, s.stream), false);
    }
    {
        TimedScope ts(fixed ? noTimer : "md_collide", s.stream, true);
        NW_HIP_R(launch_collide_off(swapForces ? a.forceOld : a.force, reinterpret_cast<const float4 *>(sortedPos),
                                  reinterpret_cast<const float4 *>(sortedVel),
                                  reinterpret_cast<const float4 *>(sortedTangent), gridParticleIndex,
                                  sc.off, sc.sortedCell, uint32_t(N), kp, s.stream), false);
    }
    s.lastHashPtr = nullptr;
    return true;
}

void nwFusedMdStep(float *pos, float *vel, float *force, float *forceOld, float *tangent, uint *gridParticleHash,
                   uint *gridParticleIndex, float *sortedPos, float *sortedVel, float *sortedTangent,
                   float deltaTime, float langevinGamma, float langevinKbT, int langevin, uint numFilaments) {
    NW_INIT();
    (void)fused_md(make_kp(), pos, vel, force, forceOld, tangent, gridParticleHash, gridParticleIndex, sortedPos,
                   sortedVel, sortedTangent, deltaTime, langevinGamma, langevinKbT, langevin, numFilaments, nullptr,
                   nullptr);
}

void nwFusedMdStepSwapForces(float *pos, float *vel, float *force, float *forceOld, float *tangent,
                             uint *gridParticleHash, uint *gridParticleIndex, float *sortedPos, float *sortedVel,
                             float *sortedTangent, float deltaTime, float langevinGamma, float langevinKbT,
                             int langevin, uint numFilaments) {
    NW_INIT();
    (void)fused_md(make_kp(), pos, vel, force, forceOld, tangent, gridParticleHash, gridParticleIndex, sortedPos,
                   sortedVel, sortedTangent, deltaTime, langevinGamma, langevinKbT, langevin, numFilaments, nullptr,
                   nullptr, nullptr, nullptr, true);
}

void nwFilamentTangents(float *tangent, const float *pos, uint numFilaments) {
    NW_INIT();
    State &s = S();
    NW_REQUIRE(tangent != nullptr && pos != nullptr, "nwFilamentTangents: NULL array");
    const KP kp = make_kp();
    NW_REQUIRE(kp.p.filamentSize >= 1 && kp.p.filamentSize <= 64, "nwFilamentTangents: worms of %u beads (1..64)",
               kp.p.filamentSize);
    NW_HIP(launch_filament_tangents(reinterpret_cast<const float4 *>(pos), reinterpret_cast<float4 *>(tangent),
                                    numFilaments, kp, s.stream));
}

void nwFusedMdStepNoAlloc(float *pos, float *vel, float *force, float *forceOld, float *tangent,
                          uint *gridParticleHash, uint *gridParticleIndex, float *sortedPos, float *sortedVel,
                          float *sortedTangent, float deltaTime, float langevinGamma, float langevinKbT, int langevin,
                          uint numFilaments) {
    NW_INIT();
    void *base = nullptr;
    int dev = -1;
    hipError_t e = hipSuccess;
    if (!noalloc_scratch(&base, &dev, &e)) {
        fail(__FILE__, __LINE__, int(e), "nwFusedMdStepNoAlloc: no address of the scratch array on device %d: %s", dev,
             hipGetErrorString(e));
        return;
    }
    const KP kp = make_kp();
    const uint64_t N = uint64_t(numFilaments) * kp.p.filamentSize;
    const uint64_t cells = uint64_t(kp.p.gridSize.x) * kp.p.gridSize.y * kp.p.gridSize.z;
    NW_REQUIRE(N < (1ull << 31) && cells > 0 && cells < (1ull << 32),
               "nwFusedMdStepNoAlloc: %llu beads over %llu cells", (unsigned long long)N, (unsigned long long)cells);
    // the sort's arrays carved from the scratch, 256-byte aligned: the cell
    // of every bead (unless the caller takes it), the packed pairs, the tile
    // x bucket prefix matrix, the cell of every sorted slot, the CSR offsets
    const BucketPlan bp = fused_plan(N, cells, false);
    NW_REQUIRE(!*bucket_plan_problem(bp), "nwFusedMdStepNoAlloc: %s", bucket_plan_problem(bp));
    auto words = [](uint64_t n) { return (n * 4 + 255) & ~uint64_t(255); };
    const uint64_t need = (gridParticleHash ? 0 : words(N)) + words(uint64_t(bp.tiles) * bp.tileItems) +
                          words(uint64_t(bp.tiles) * (bp.buckets + 1)) + words(N) + words(cells + 2);
    NW_REQUIRE(need <= kNoAllocScratchBytes,
               "nwFusedMdStepNoAlloc: %llu beads over %llu cells need %llu bytes of scratch; the fixed scratch "
               "array holds %llu (use nwFusedMdStep)",
               (unsigned long long)N, (unsigned long long)cells, (unsigned long long)need,
               (unsigned long long)kNoAllocScratchBytes);
    unsigned char *at = static_cast<unsigned char *>(base);
    auto take = [&at, &words](uint64_t n) {
        uint32_t *p = reinterpret_cast<uint32_t *>(at);
        at += words(n);
        return p;
    };
    BucketScratch sc{};
    sc.hash = gridParticleHash ? nullptr : take(N);
    sc.pairs = take(uint64_t(bp.tiles) * bp.tileItems);
    sc.P = take(uint64_t(bp.tiles) * (bp.buckets + 1));
    sc.sortedCell = take(N);
    sc.off = take(cells + 2);
    (void)fused_md(kp, pos, vel, force, forceOld, tangent, gridParticleHash, gridParticleIndex, sortedPos, sortedVel,
                   sortedTangent, deltaTime, langevinGamma, langevinKbT, langevin, numFilaments, nullptr, nullptr,
                   nullptr, nullptr, false, &sc);
}

// The fused SRD collision with the parameters kp; offOut (may be NULL: the
// library's scratch) receives the CSR offsets of the cell order.
static bool fused_srd(const KP &kp, float *pos, float *vel, float *solventCellCOM, uint *gridParticleHash,
                      uint *gridParticleIndex, uint32_t *offOut, float deltaTime, uint numParticles, uint numSolvent,
                      int isokinetic) {
    State &s = S();
    const uint64_t NT = uint64_t(numParticles) + numSolvent;
    NW_REQUIRE_R(NT < (1ull << 31), false, "nwFusedSrdStep: too many particles");
    const uint64_t cells = uint64_t(kp.s.gridSize.x) * kp.s.gridSize.y * kp.s.gridSize.z;
    NW_REQUIRE_R(cells > 0 && cells == kp.s.numCells, false, "nwFusedSrdStep: SRD grid does not match numCells");
    NW_REQUIRE_R(gridParticleIndex != nullptr && solventCellCOM != nullptr, false,
                 "nwFusedSrdStep: NULL output buffer");
    if (NT == 0) return true;
    const BucketPlan bp = fused_plan(NT, cells, true);
    NW_REQUIRE_R(!*bucket_plan_problem(bp), false, "nwFusedSrdStep: %s", bucket_plan_problem(bp));
    NW_HIP_R(ensure_bucket_scratch(s.bucket, bp), false);
    BucketScratch sc = s.bucket;
    if (offOut) sc.off = offOut;
    const uint64_t off = s.offset;
    s.offset += blocks_for(cells);
    const double alpha = double(kp.s.alpha);
    const float cosA = float(std::cos(alpha)), sinA = float(std::sin(alpha));
    float4 *P = reinterpret_cast<float4 *>(pos), *V = reinterpret_cast<float4 *>(vel);
    float4 *com = reinterpret_cast<float4 *>(solventCellCOM);
    uint32_t *hash = gridParticleHash ? gridParticleHash : s.bucket.hash;
    const int iso = isokinetic ? 1 : 0;
    {
        TimedScope ts("srdf_tiles", s.stream, true);
        NW_HIP_R(launch_srd_tiles(P, hash, bp, s.bucket, kp, s.stream), false);
    }
    {
        TimedScope ts("srdf_cells", s.stream, true);
        NW_HIP_R(launch_srd_cells(bp, sc, gridParticleIndex, V, com, iso, s.seed, off, kp, s.stream), false);
    }
    {
        TimedScope ts("srdf_rotate", s.stream, true);
        NW_HIP_R(launch_srd_rotate(hash, com, s.bucket.lam, P, V, uint32_t(NT), numParticles, deltaTime, cosA, sinA,
                                   iso, kp, s.stream),
                 false);
    }
    s.lastHashPtr = nullptr;
    return true;
}

void nwFusedSrdStep(float *pos, float *vel, float *solventCellCOM, uint *gridParticleHash, uint *gridParticleIndex,
                    float deltaTime, uint numParticles, uint numSolvent, int isokinetic) {
    NW_INIT();
    (void)fused_srd(make_kp(), pos, vel, solventCellCOM, gridParticleHash, gridParticleIndex, nullptr, deltaTime,
                    numParticles, numSolvent, isokinetic);
}

// ---------------------------------------------------------------------------
// additions
// ---------------------------------------------------------------------------
void nwSetErrorMode(int mode) { S().errorMode = mode == NW_ERRORS_RETURN ? NW_ERRORS_RETURN : NW_ERRORS_EXIT; }

int nwGetLastError(void) {
    const int e = S().lastError;
    S().lastError = 0;
    return e;
}

const char *nwGetLastErrorString(void) { return S().lastErrorString.c_str(); }

int nwDeviceAvailable(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return 0;
    const int mode = S().errorMode;
    S().errorMode = NW_ERRORS_RETURN;
    const bool ok = ensure_init();
    S().errorMode = mode;
    return ok ? 1 : 0;
}

void nwRngSetSeed(uint64_t seed, uint64_t offset) {
    S().seed = seed;
    S().offset = offset;
}

void nwRngGetState(uint64_t *seed, uint64_t *offset) {
    if (seed) *seed = S().seed;
    if (offset) *offset = S().offset;
}

void nwPhiloxBlock(uint64_t seed, uint64_t block, uint32_t out[4]) {
    const u4 b = philox_block(seed, block);
    out[0] = b.x;
    out[1] = b.y;
    out[2] = b.z;
    out[3] = b.w;
}

void nwSetExtParameters(const NWExtParams *ext) {
    NW_REQUIRE(ext != nullptr, "nwSetExtParameters: NULL");
    S().ext = *ext;
}

void nwGetExtParameters(NWExtParams *ext) {
    if (ext) *ext = S().ext;
}

void *nwGetStream(void) { return S().stream; }

void nwSetStream(void *stream) { S().stream = static_cast<hipStream_t>(stream); }

double nwSymEig3Max(const double a[6]) {
    const double a11 = a[0], a12 = a[1], a13 = a[2], a22 = a[3], a23 = a[4], a33 = a[5];
    const double p1 = a12 * a12 + a13 * a13 + a23 * a23;
    if (p1 == 0.0) return std::fmax(a11, std::fmax(a22, a33));
    const double q = (a11 + a22 + a33) / 3.0;
    const double p2 = (a11 - q) * (a11 - q) + (a22 - q) * (a22 - q) + (a33 - q) * (a33 - q) + 2.0 * p1;
    const double p = std::sqrt(p2 / 6.0);
    const double b11 = (a11 - q) / p, b22 = (a22 - q) / p, b33 = (a33 - q) / p;
    const double b12 = a12 / p, b13 = a13 / p, b23 = a23 / p;
    const double det = b11 * (b22 * b33 - b23 * b23) - b12 * (b12 * b33 - b23 * b13) + b13 * (b12 * b23 - b22 * b13);
    const double r = std::fmin(1.0, std::fmax(-1.0, det / 2.0));
    return q + 2.0 * p * std::cos(std::acos(r) / 3.0);
}

static void observe_kp(const KP &kp, const float *pos, const float *vel, uint numFilaments, NWObservables *out);

void nwObservables(const float *pos, const float *vel, uint numFilaments, NWObservables *out) {
    NW_INIT();
    NW_REQUIRE(out != nullptr && pos != nullptr && vel != nullptr, "nwObservables: NULL argument");
    observe_kp(make_kp(), pos, vel, numFilaments, out);
}

// the observables of the worms of kp's SimParams
static void observe_kp(const KP &kp, const float *pos, const float *vel, uint numFilaments, NWObservables *out) {
    State &s = S();
    const uint32_t n = kp.p.filamentSize;
    NW_REQUIRE(n >= 1 && n <= 64, "nwObservables: worms of %u beads (1..64 supported)", n);
    std::memset(out, 0, sizeof(*out));
    const uint64_t rows = observe_rows(numFilaments, n);
    if (rows == 0) return;
    const uint64_t need = (rows + 1) * kObsSums;
    if (need > s.obsCap) {
        if (s.obsPartial) (void)hipFree(s.obsPartial);
        s.obsPartial = nullptr;
        s.obsCap = 0;
        NW_HIP(hipMalloc(&s.obsPartial, need * sizeof(double)));
        s.obsCap = need;
    }
    double *total = s.obsPartial + rows * kObsSums;
    NW_HIP(launch_observe(reinterpret_cast<const float4 *>(pos), reinterpret_cast<const float4 *>(vel), numFilaments,
                          s.obsPartial, total, kp, s.stream));
    double sums[kObsSums];
    NW_HIP(hipMemcpyAsync(sums, total, sizeof(sums), hipMemcpyDeviceToHost, s.stream));
    NW_HIP(hipStreamSynchronize(s.stream));
    static_assert(kObsSums == NW_OBS_SUMS, "NWObservables.sums layout");
    std::memcpy(out->sums, sums, sizeof(sums));
    const double nb = sums[kObsBeads], nbo = sums[kObsBonds];
    out->beads = nb;
    out->bonds = nbo;
    if (nb > 0) {
        const double Q[6] = {1.5 * sums[kObsTxx] / nb - 0.5, 1.5 * sums[kObsTxy] / nb, 1.5 * sums[kObsTxz] / nb,
                             1.5 * sums[kObsTyy] / nb - 0.5, 1.5 * sums[kObsTyz] / nb, 1.5 * sums[kObsTzz] / nb - 0.5};
        out->nematicS = nwSymEig3Max(Q);
        const double mx = sums[kObsTx] / nb, my = sums[kObsTy] / nb, mz = sums[kObsTz] / nb;
        out->polarP = std::sqrt(mx * mx + my * my + mz * mz);
        out->polarSpeed = sums[kObsVdotT] / nb;
        out->meanV2 = sums[kObsV2] / nb;
    }
    if (nbo > 0) {
        const double Q[6] = {1.5 * sums[kObsBxx] / nbo - 0.5, 1.5 * sums[kObsBxy] / nbo, 1.5 * sums[kObsBxz] / nbo,
                             1.5 * sums[kObsByy] / nbo - 0.5, 1.5 * sums[kObsByz] / nbo,
                             1.5 * sums[kObsBzz] / nbo - 0.5};
        out->bondS = nwSymEig3Max(Q);
    }
}

// ---------------------------------------------------------------------------
// batched replicas
// ---------------------------------------------------------------------------
static void batch_drop_graph(State::Batch &b) {
    for (int k = 0; k < 2; ++k) {
        if (b.exec[k]) (void)hipGraphExecDestroy(b.exec[k]);
        if (b.graph[k]) (void)hipGraphDestroy(b.graph[k]);
        b.exec[k] = nullptr;
        b.graph[k] = nullptr;
    }
    b.key.clear();
}

void nwBatchSetReplicas(const NWReplicaParams *reps, uint numReplicas) {
    NW_INIT();
    State &s = S();
    State::Batch &b = s.batch;
    NW_REQUIRE(numReplicas == 0 || reps != nullptr, "nwBatchSetReplicas: NULL replicas");
    batch_drop_graph(b);
    s.bsol.R = 0;  // a new batch has no solvent until nwBatchSetSolvent
    s.badj.ready = false;
    b.reps.assign(numReplicas, BatchRep{});
    b.beadOff.assign(numReplicas + 1, 0u);
    b.cellOff.assign(numReplicas + 1, 0u);
    uint64_t N = 0, C = 0, waves = 0, blocks = 0;
    for (uint r = 0; r < numReplicas; ++r) {
        const NWReplicaParams &q = reps[r];
        const SimParams &p = q.params;
        const uint32_t n = p.filamentSize, W = p.numFilaments;
        NW_REQUIRE(n >= 1 && n <= 64, "nwBatchSetReplicas: replica %u has worms of %u beads (1..64)", r, n);
        const bool oneLayer = q.ext.stencil == NW_STENCIL_2D || p.gridSize.z == 1;
        NW_REQUIRE(oneLayer && p.gridSize.x >= 3 && p.gridSize.y >= 1,
                   "nwBatchSetReplicas: replica %u needs a one-layer stencil and gridSize.x >= 3", r);
        NW_REQUIRE(q.ext.pairPotential == reps[0].ext.pairPotential,
                   "nwBatchSetReplicas: replica %u: one pair potential per batch", r);
        const uint64_t cells = uint64_t(p.gridSize.x) * p.gridSize.y * p.gridSize.z;
        BatchRep &d = b.reps[r];
        d.kp = make_kp_for(p, SolventParams{}, q.ext);
        d.kp.p.numParticles = W * n;
        d.beadOffset = uint32_t(N);
        d.cellOffset = uint32_t(C);
        d.waveOffset = uint32_t(waves);
        d.blockOffset = uint32_t(blocks);
        d.seed = q.seed;
        d.rngBase = q.offset;
        d.revBlocks = p.reverseProbability > 0.0f ? uint32_t(blocks_for(W)) : 0u;
        d.rngStride = d.revBlocks + (q.langevin ? uint64_t(W) * n : 0u);
        d.lgvStd = sqrtf(2.0f * q.langevinGamma * q.langevinKbT);
        d.lgvGamma = q.langevinGamma;
        d.langevin = q.langevin ? 1 : 0;
        d.numBeads = W * n;
        N += uint64_t(W) * n;
        C += cells;
        waves += (uint64_t(W) + 64u / n - 1) / (64u / n);
        blocks += (uint64_t(W) * n + kPairBlock - 1) / kPairBlock;
        NW_REQUIRE(N < (1ull << 31) && C < (1ull << 32) - 2, "nwBatchSetReplicas: batch too large");
        b.beadOff[r + 1] = uint32_t(N);
        b.cellOff[r + 1] = uint32_t(C);
        const bool wca = q.ext.pairPotential == NW_PAIR_WCA;
        if (!wca || p.activity != 0.0f) b.tan = true;
        if (!(q.ext.stencil == NW_STENCIL_2D && p.gridSize.y >= 3)) b.unit = false;
    }
    b.R = numReplicas;
    b.N = uint32_t(N);
    b.swap = true;
    for (uint r = 0; r < numReplicas; ++r) b.swap = b.swap && b.reps[r].revBlocks == 0;
    b.cells = uint32_t(C);
    b.waves = uint32_t(waves);
    b.blocks = uint32_t(blocks);
    b.mode = numReplicas && reps[0].ext.pairPotential == NW_PAIR_WCA ? 1 : 0;
    if (b.mode == 0) b.tan = true;
    if (numReplicas == 0) b.tan = false, b.unit = true;
    b.steps = 0;
    if (b.dev) (void)hipFree(b.dev);
    b.dev = nullptr;
    if (numReplicas) {
        // the records, then their wavefront and workgroup offsets packed (the
        // kernels' replica lookups bisect those)
        static_assert(sizeof(BatchRep) % alignof(uint32_t) == 0, "offsets follow the records");
        const size_t recBytes = sizeof(BatchRep) * numReplicas;
        std::vector<unsigned char> table(recBytes + 2 * sizeof(uint32_t) * numReplicas);
        memcpy(table.data(), b.reps.data(), recBytes);
        uint32_t *off = reinterpret_cast<uint32_t *>(table.data() + recBytes);
        for (uint r = 0; r < numReplicas; ++r) {
            off[r] = b.reps[r].waveOffset;
            off[numReplicas + r] = b.reps[r].blockOffset;
        }
        NW_HIP(hipMalloc(&b.dev, table.size()));
        NW_HIP(hipMemcpy(b.dev, table.data(), table.size(), hipMemcpyHostToDevice));
    }
    if (!b.stepCtr) NW_HIP(hipMalloc(&b.stepCtr, sizeof(uint64_t)));
    NW_HIP(hipMemset(b.stepCtr, 0, sizeof(uint64_t)));
    b.sort.assign(numReplicas, SortRep{});
    b.sortPlan = BatchSortPlan{};
    if (b.sortDev) (void)hipFree(b.sortDev);
    b.sortDev = nullptr;
    if (N) {
        const char *why = make_sort_reps(b.beadOff.data(), b.cellOff.data(), numReplicas, b.sort.data(), b.sortPlan);
        NW_REQUIRE(!*why, "nwBatchSetReplicas: %s", why);
        NW_HIP(hipMalloc(&b.sortDev, sizeof(SortRep) * numReplicas));
        NW_HIP(hipMemcpy(b.sortDev, b.sort.data(), sizeof(SortRep) * numReplicas, hipMemcpyHostToDevice));
        NW_HIP(ensure_batch_sort_scratch(b.scratch, b.sortPlan));
    }
}

void nwBatchLayout(uint *beadOffset, uint *cellOffset) {
    const State::Batch &b = S().batch;
    for (uint r = 0; r <= b.R; ++r) {
        if (beadOffset) beadOffset[r] = b.beadOff.empty() ? 0u : b.beadOff[r];
        if (cellOffset) cellOffset[r] = b.cellOff.empty() ? 0u : b.cellOff[r];
    }
}

void nwBatchRngState(uint replica, uint64_t *offset) {
    const State::Batch &b = S().batch;
    NW_REQUIRE(replica < b.R, "nwBatchRngState: replica %u of %u", replica, b.R);
    if (offset) *offset = b.reps[replica].rngBase + b.steps * b.reps[replica].rngStride;
}

// the launches of one batched step
// (timed: each kernel may carry the events of its timer family -- never
// while the stream is being captured into a graph)
static hipError_t batch_step(State::Batch &b, const BatchArrays &a, const float4 *vel, const float4 *tan,
                             float4 *sVel, float4 *sTan, float dt, hipStream_t st, bool timed = false) {
    hipError_t e;
    {
        TimedScope ts(timed ? "batch_front" : "", st, true);
        if ((e = launch_md_front_batch(a, b.dev, b.R, b.waves, b.stepCtr, dt, st)) != hipSuccess) return e;
    }
    {
        TimedScope ts(timed ? "batch_tiles" : "", st, true);
        if ((e = launch_hash_tiles_batch(a.hash, b.sortDev, b.R, b.sortPlan, b.scratch, st)) != hipSuccess) return e;
    }
    {
        TimedScope ts(timed ? "batch_sort" : "", st, true);
        if ((e = launch_md_cells_batch(b.sortDev, b.R, b.sortPlan, b.scratch, a.ids, a.pos, vel, tan, a.sPos, sVel,
                                       sTan, st)) != hipSuccess)
            return e;
    }
    TimedScope ts(timed ? "batch_collide" : "", st, true);
    return launch_collide_batch(a, b.dev, b.R, b.blocks, b.mode, b.tan, b.unit, b.stepCtr, st);
}

void nwBatchMdStep(float *pos, float *vel, float *force, float *forceOld, float *tangent, uint *gridParticleIndex,
                   float *sortedPos, float *sortedVel, float *sortedTangent, float deltaTime, uint steps,
                   int useGraph) {
    NW_INIT();
    State &s = S();
    State::Batch &b = s.batch;
    if (b.N == 0 || steps == 0) return;
    NW_REQUIRE(pos && vel && force && forceOld && tangent && gridParticleIndex && sortedPos,
               "nwBatchMdStep: NULL array");
    NW_REQUIRE(b.mode == 1 || sortedVel != nullptr, "nwBatchMdStep: sortedVel is NULL");
    NW_REQUIRE(!b.tan || sortedTangent != nullptr, "nwBatchMdStep: sortedTangent is NULL");
    BatchArrays a{};
    a.pos = reinterpret_cast<float4 *>(pos);
    a.vel = reinterpret_cast<float4 *>(vel);
    a.force = reinterpret_cast<float4 *>(force);
    a.forceOld = reinterpret_cast<float4 *>(forceOld);
    a.tangent = reinterpret_cast<float4 *>(tangent);
    a.swapForces = b.swap ? 1 : 0;
    a.hash = b.scratch.hash;
    a.ids = gridParticleIndex;
    a.off = b.scratch.off;
    a.sCell = b.scratch.sortedCell;
    a.sPos = reinterpret_cast<float4 *>(sortedPos);
    a.sVel = b.mode == 0 ? reinterpret_cast<float4 *>(sortedVel) : nullptr;
    a.sTan = b.tan ? reinterpret_cast<float4 *>(sortedTangent) : nullptr;
    const float4 *gv = b.mode == 0 ? a.vel : nullptr;
    const float4 *gt = b.tan ? a.tangent : nullptr;
    uint32_t dtBits;
    std::memcpy(&dtBits, &deltaTime, 4);
    const std::vector<uintptr_t> key = {uintptr_t(pos), uintptr_t(vel), uintptr_t(force), uintptr_t(forceOld),
                                        uintptr_t(tangent), uintptr_t(gridParticleIndex), uintptr_t(sortedPos),
                                        uintptr_t(sortedVel), uintptr_t(sortedTangent), uintptr_t(dtBits)};
    // double-buffered forces: phase 0 reads force and writes the new force
    // into forceOld's buffer, phase 1 the other way round; after an odd
    // number of steps the two buffers exchange contents once
    BatchArrays ar[2] = {a, a};
    ar[1].force = a.forceOld;
    ar[1].forceOld = a.force;
    int ph = 0;
    uint done = 0;
    if (useGraph && (!b.exec[0] || b.key != key)) {
        // one direct step first (scratch and kernel attributes are set up
        // outside the capture), then capture the step (both phases)
        batch_drop_graph(b);
        {
            NW_HIP(batch_step(b, ar[0], gv, gt, a.sVel, a.sTan, deltaTime, s.stream, true));
        }
        ++done;
        ++b.steps;
        if (b.swap) ph = 1;
        if (done < steps) {
            hipError_t e = hipSuccess;
            for (int k = 0; k < (b.swap ? 2 : 1) && e == hipSuccess; ++k) {
                if ((e = hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal)) != hipSuccess) break;
                const hipError_t e1 = batch_step(b, ar[k], gv, gt, a.sVel, a.sTan, deltaTime, s.stream);
                hipGraph_t g = nullptr;
                const hipError_t e2 = hipStreamEndCapture(s.stream, &g);
                e = e1 != hipSuccess ? e1 : e2;
                if (e != hipSuccess) break;
                b.graph[k] = g;
                e = hipGraphInstantiate(&b.exec[k], g, nullptr, nullptr, 0);
            }
            if (e != hipSuccess) {
                // as in the replay loop below: the contents go back to the
                // caller's roles before the error is reported
                batch_drop_graph(b);
                if (ph == 1) (void)launch_swap4(a.force, a.forceOld, b.N, s.stream);
                fail(__FILE__, __LINE__, int(e), "nwBatchMdStep: graph capture failed: HIP error %d (%s)", int(e),
                     hipGetErrorString(e));
                return;
            }
            b.key = key;
        }
    }
    for (; done < steps; ++done) {
        hipError_t e = hipSuccess;
        if (useGraph) {
            TimedScope ts("batch_step", s.stream);
            e = hipGraphLaunch(b.exec[ph], s.stream);
        } else {
            e = batch_step(b, ar[ph], gv, gt, a.sVel, a.sTan, deltaTime, s.stream, true);
        }
        if (e != hipSuccess) {
            // the buffers' roles are exchanged after an odd number of steps:
            // exchange their contents back so force / forceOld keep their
            // meaning for the caller, then report the failing step
            if (ph == 1) (void)launch_swap4(a.force, a.forceOld, b.N, s.stream);
            fail(__FILE__, __LINE__, int(e), "nwBatchMdStep: step %u of %u failed: HIP error %d (%s)", done, steps,
                 int(e), hipGetErrorString(e));
            return;
        }
        ++b.steps;
        if (b.swap) ph ^= 1;
    }
    if (ph == 1) NW_HIP(launch_swap4(a.force, a.forceOld, b.N, s.stream));
}

// ---------------------------------------------------------------------------
// adjoint of the fused MD step
// ---------------------------------------------------------------------------
extern "C++" {
template <typename T>
static bool grow(T *&p, uint64_t &cap, uint64_t need) {
    if (need <= cap) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(reinterpret_cast<void **>(&p), need * sizeof(T)) != hipSuccess) return false;
    cap = need;
    return true;
}
}

static bool copy_d2d(void *dst, const void *src, size_t bytes, hipStream_t st) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st) == hipSuccess;
}

static uint32_t collisions_in(uint32_t steps, uint32_t gap) { return gap ? (steps + gap - 1) / gap : 0; }

// the SolventParams of collision c of the run
static KP collision_kp(const State::Adjoint &A, uint32_t c) {
    KP kp = A.kp;
    kp.s.randOffset.x = A.shifts[3 * size_t(c)];
    kp.s.randOffset.y = A.shifts[3 * size_t(c) + 1];
    kp.s.randOffset.z = A.shifts[3 * size_t(c) + 2];
    return kp;
}

// The forward run of nwAdjointForward / nwAdjointForwardSrd (srd: numSolvent,
// gap and shifts set in A by the caller, A.srd true).
static void adjoint_forward(float *pos, float *vel, float *force, float *forceOld, float *tangent,
                            uint numFilaments, float deltaTime, float langevinGamma, float langevinKbT, int langevin,
                            uint steps, uint checkpointEvery) {
    State &s = S();
    State::Adjoint &A = s.adj;
    const KP kp = make_kp();
    const SimParams &p = kp.p;
    NW_REQUIRE(kp.e.pairPotential == NW_PAIR_WCA && kp.e.bondedMode == NW_BONDED_CHAIN,
               "nwAdjointForward: WCA pairs with chain bonds only");
    NW_REQUIRE(p.activity == 0.0f && p.reverseProbability == 0.0f,
               "nwAdjointForward: no extensile pair term and no reversal");
    NW_REQUIRE((kp.e.stencil == NW_STENCIL_2D || p.gridSize.z == 1) && p.gridSize.x >= 3,
               "nwAdjointForward: one-layer stencil and gridSize.x >= 3 only");
    NW_REQUIRE(p.boundaryX == PERIODIC && p.boundaryY == PERIODIC && p.boundaryZ == PERIODIC,
               "nwAdjointForward: periodic boxes only");
    NW_REQUIRE(!langevin || langevinGamma == p.gamma,
               "nwAdjointForward: the Langevin gamma must be SimParams.gamma (one gamma to differentiate)");
    NW_REQUIRE(checkpointEvery >= 1 && steps >= 1, "nwAdjointForward: steps and checkpointEvery >= 1");
    const uint32_t n = p.filamentSize;
    NW_REQUIRE(n >= 1 && n <= 64, "nwAdjointForward: worms of 1..64 beads");
    const uint64_t N = uint64_t(numFilaments) * n;
    const uint64_t C = uint64_t(p.gridSize.x) * p.gridSize.y * p.gridSize.z;
    if (A.srd) {
        NW_REQUIRE(!langevin, "nwAdjointForwardSrd: no Langevin thermostat with the solvent");
        NW_REQUIRE(kp.s.gridSize.x * kp.s.gridSize.y * kp.s.gridSize.z == kp.s.numCells && kp.s.numCells > 0,
                   "nwAdjointForwardSrd: SRD grid does not match numCells");
        NW_REQUIRE(N + A.Ns < (1ull << 31), "nwAdjointForwardSrd: too many particles");
        const double alpha = double(kp.s.alpha);
        A.cosA = float(std::cos(alpha));
        A.sinA = float(std::sin(alpha));
        A.rotBeads = kp.e.srdCollideFilaments != 0 ? 1 : 0;
        A.Cs = kp.s.numCells;
    }
    A.NT = uint32_t(N + (A.srd ? A.Ns : 0));
    A.kp = kp;
    A.dt = deltaTime;
    A.gamma = langevinGamma;
    A.kT = langevinKbT;
    A.lgvStd = sqrtf(2.0f * langevinGamma * langevinKbT);
    A.langevin = langevin ? 1 : 0;
    A.W = numFilaments;
    A.N = uint32_t(N);
    A.C = uint32_t(C);
    A.steps = steps;
    A.K = checkpointEvery;
    A.seed = s.seed;
    const uint32_t nck = (steps + checkpointEvery - 1) / checkpointEvery;
    const uint64_t NT = A.NT, stride = 2 * NT + 2 * N;
    NW_REQUIRE(grow(A.ckpt, A.capCkpt, uint64_t(nck) * stride) && grow(A.segSPos, A.capSPos, N) &&
                   grow(A.segIds, A.capIds, NT),
               "nwAdjointForward: out of device memory");
    if (A.srd)
        NW_REQUIRE(grow(A.collCom, A.capCollCom, A.Cs) && grow(A.collHash, A.capCollHash, NT),
                   "nwAdjointForwardSrd: out of device memory");
    A.ckptOffset.assign(nck, 0);
    const size_t bytesT = NT * sizeof(float4), bytes = N * sizeof(float4);
    for (uint32_t k = 0; k < steps; ++k) {
        if (k % checkpointEvery == 0) {
            const uint32_t c = k / checkpointEvery;
            float4 *ck = A.ckpt + uint64_t(c) * stride;
            NW_REQUIRE(copy_d2d(ck, pos, bytesT, s.stream) && copy_d2d(ck + NT, vel, bytesT, s.stream) &&
                           copy_d2d(ck + 2 * NT, force, bytes, s.stream) &&
                           copy_d2d(ck + 2 * NT + N, forceOld, bytes, s.stream),
                       "nwAdjointForward: checkpoint copy failed");
            A.ckptOffset[c] = s.offset;
        }
        // sorted positions and ids go to the segment store (scratch here)
        if (!fused_md(kp, pos, vel, force, forceOld, tangent, nullptr, A.segIds, reinterpret_cast<float *>(A.segSPos),
                      nullptr, nullptr, deltaTime, langevinGamma, langevinKbT, langevin, numFilaments, nullptr,
                      nullptr))
            return;
        if (A.srd && k % A.gap == 0 &&
            !fused_srd(collision_kp(A, k / A.gap), pos, vel, reinterpret_cast<float *>(A.collCom), A.collHash,
                       A.segIds, nullptr, deltaTime * float(A.gap), uint32_t(N), A.Ns, 0))
            return;
    }
    A.endOffset = s.offset;
    A.ready = true;
}

void nwAdjointForward(float *pos, float *vel, float *force, float *forceOld, float *tangent, uint numFilaments,
                      float deltaTime, float langevinGamma, float langevinKbT, int langevin, uint steps,
                      uint checkpointEvery) {
    NW_INIT();
    State::Adjoint &A = S().adj;
    A.ready = false;
    A.srd = false;
    A.Ns = A.gap = 0;
    A.shifts.clear();
    adjoint_forward(pos, vel, force, forceOld, tangent, numFilaments, deltaTime, langevinGamma, langevinKbT, langevin,
                    steps, checkpointEvery);
}

void nwAdjointForwardSrd(float *pos, float *vel, float *force, float *forceOld, float *tangent, uint numFilaments,
                         uint numSolvent, float deltaTime, uint srdGap, const float *srdShifts, uint steps,
                         uint checkpointEvery) {
    NW_INIT();
    State::Adjoint &A = S().adj;
    A.ready = false;
    NW_REQUIRE(srdGap >= 1 && srdShifts != nullptr, "nwAdjointForwardSrd: srdGap >= 1 and the shifts are required");
    A.srd = true;
    A.Ns = numSolvent;
    A.gap = srdGap;
    A.shifts.assign(srdShifts, srdShifts + 3 * size_t(collisions_in(steps, srdGap)));
    adjoint_forward(pos, vel, force, forceOld, tangent, numFilaments, deltaTime, S().params.gamma, 0.0f, 0, steps,
                    checkpointEvery);
}

static bool ensure_segment(State::Adjoint &A) {
    const uint64_t N = A.N, K = A.K, NT = A.NT;
    if (!(grow(A.segPos, A.capPos, (K + 1) * N) && grow(A.segVel, A.capVel, (K + 1) * N) &&
          grow(A.segSPos, A.capSPos, K * N) && grow(A.segIds, A.capIds, std::max(K * N, NT)) &&
          grow(A.segOff, A.capOff, K * (uint64_t(A.C) + 2)) && grow(A.segCell, A.capCell, K * N) &&
          grow(A.work, A.capWork, 3 * N) && grow(A.lam, A.capLam, 4 * N + NT)))
        return false;
    if (!A.srd) return true;
    // at most this many collisions in a segment of K steps
    const uint64_t cpS = (K + A.gap - 1) / A.gap + 1;
    return grow(A.collHash, A.capCollHash, cpS * NT) && grow(A.collIds, A.capCollIds, cpS * NT) &&
           grow(A.collOff, A.capCollOff, cpS * (uint64_t(A.Cs) + 2)) &&
           grow(A.collCom, A.capCollCom, cpS * A.Cs) && grow(A.cellSum, A.capCellSum, A.Cs) &&
           grow(A.wPos, A.capWPos, NT) && grow(A.wVel, A.capWVel, NT);
}

void nwAdjointBackward(const float *dLdPos, const float *dLdVel, double grad[4]) {
    NW_INIT();
    State &s = S();
    State::Adjoint &A = s.adj;
    NW_REQUIRE(A.ready, "nwAdjointBackward: no nwAdjointForward run to differentiate");
    NW_REQUIRE(dLdPos != nullptr && grad != nullptr, "nwAdjointBackward: NULL argument");
    const uint64_t N = A.N, K = A.K;
    NW_REQUIRE(ensure_segment(A), "nwAdjointBackward: out of device memory");
    const uint64_t rowsPerStep = adj_step_blocks(A.W, A.kp.p.filamentSize);
    NW_REQUIRE(grow(A.partial, A.capPartial, K * rowsPerStep * kAdjParams), "nwAdjointBackward: out of memory");
    if (!A.grad) NW_HIP(hipMalloc(&A.grad, kAdjParams * sizeof(double)));
    NW_HIP(hipMemsetAsync(A.grad, 0, kAdjParams * sizeof(double), s.stream));
    const uint64_t NT = A.NT, stride = 2 * NT + 2 * N;
    const size_t bytes = N * sizeof(float4);
    // lv covers every particle (the solvent's adjoints follow the beads')
    float4 *lr = A.lam, *lF[2] = {A.lam + N, A.lam + 2 * N}, *lG = A.lam + 3 * N, *lv = A.lam + 4 * N;
    NW_HIP(hipMemcpyAsync(lr, dLdPos, bytes, hipMemcpyDeviceToDevice, s.stream));
    NW_HIP(hipMemsetAsync(lv, 0, NT * sizeof(float4), s.stream));
    if (dLdVel) NW_HIP(hipMemcpyAsync(lv, dLdVel, bytes, hipMemcpyDeviceToDevice, s.stream));
    NW_HIP(hipMemsetAsync(lF[0], 0, bytes, s.stream));
    NW_HIP(hipMemsetAsync(lG, 0, bytes, s.stream));
    int cur = 0;
    float4 *F = A.work, *G = A.work + N, *T = A.work + 2 * N;
    const uint32_t nck = (A.steps + A.K - 1) / A.K;
    std::vector<uint64_t> lgvOff(K);
    std::vector<int> collSlot(K);  // the collision store slot of each step of the segment, -1: none
    for (int c = int(nck) - 1; c >= 0; --c) {
        const uint32_t s0 = uint32_t(c) * A.K, len = std::min<uint32_t>(A.K, A.steps - s0);
        const float4 *ck = A.ckpt + uint64_t(c) * stride;
        NW_HIP(hipMemcpyAsync(A.segPos, ck, bytes, hipMemcpyDeviceToDevice, s.stream));
        NW_HIP(hipMemcpyAsync(A.segVel, ck + NT, bytes, hipMemcpyDeviceToDevice, s.stream));
        NW_HIP(hipMemcpyAsync(F, ck + 2 * NT, bytes, hipMemcpyDeviceToDevice, s.stream));
        NW_HIP(hipMemcpyAsync(G, ck + 2 * NT + N, bytes, hipMemcpyDeviceToDevice, s.stream));
        if (A.srd) {
            NW_HIP(hipMemcpyAsync(A.wPos, ck, NT * sizeof(float4), hipMemcpyDeviceToDevice, s.stream));
            NW_HIP(hipMemcpyAsync(A.wVel, ck + NT, NT * sizeof(float4), hipMemcpyDeviceToDevice, s.stream));
        }
        // recompute the segment, keeping every step's state and cell lists
        s.offset = A.ckptOffset[c];
        int slots = 0;
        for (uint32_t j = 0; j < len; ++j) {
            // step j reads slot j and writes slot j + 1
            float4 *pj = A.segPos + uint64_t(j + 1) * N, *vj = A.segVel + uint64_t(j + 1) * N;
            lgvOff[j] = s.offset;
            collSlot[j] = -1;
            if (A.srd) {
                // the full state steps in place; the beads' copies fill the slots
                if (!fused_md(A.kp, reinterpret_cast<float *>(A.wPos), reinterpret_cast<float *>(A.wVel),
                              reinterpret_cast<float *>(F), reinterpret_cast<float *>(G),
                              reinterpret_cast<float *>(T), nullptr, A.segIds + uint64_t(j) * N,
                              reinterpret_cast<float *>(A.segSPos + uint64_t(j) * N), nullptr, nullptr, A.dt, A.gamma,
                              A.kT, A.langevin, A.W, A.segOff + uint64_t(j) * (A.C + 2),
                              A.segCell + uint64_t(j) * N, nullptr, nullptr))
                    return;
                NW_HIP(hipMemcpyAsync(pj, A.wPos, bytes, hipMemcpyDeviceToDevice, s.stream));
                NW_HIP(hipMemcpyAsync(vj, A.wVel, bytes, hipMemcpyDeviceToDevice, s.stream));
                const uint32_t k = s0 + j;
                if (k % A.gap == 0) {
                    const int q = slots++;
                    collSlot[j] = q;
                    if (!fused_srd(collision_kp(A, k / A.gap), reinterpret_cast<float *>(A.wPos),
                                   reinterpret_cast<float *>(A.wVel),
                                   reinterpret_cast<float *>(A.collCom + uint64_t(q) * A.Cs),
                                   A.collHash + uint64_t(q) * NT, A.collIds + uint64_t(q) * NT,
                                   A.collOff + uint64_t(q) * (A.Cs + 2), A.dt * float(A.gap), A.N, A.Ns, 0))
                        return;
                }
                continue;
            }
            if (!fused_md(A.kp, reinterpret_cast<float *>(pj), reinterpret_cast<float *>(vj),
                          reinterpret_cast<float *>(F), reinterpret_cast<float *>(G), reinterpret_cast<float *>(T),
                          nullptr, A.segIds + uint64_t(j) * N, reinterpret_cast<float *>(A.segSPos + uint64_t(j) * N),
                          nullptr, nullptr, A.dt, A.gamma, A.kT, A.langevin, A.W,
                          A.segOff + uint64_t(j) * (A.C + 2), A.segCell + uint64_t(j) * N,
                          reinterpret_cast<const float *>(A.segPos + uint64_t(j) * N),
                          reinterpret_cast<const float *>(A.segVel + uint64_t(j) * N)))
                return;
        }
        // the adjoint through the segment, last step first
        for (int j = int(len) - 1; j >= 0; --j) {
            if (collSlot[j] >= 0) {
                // the step's collision came after its MD part: undo it first
                const uint64_t q = uint64_t(collSlot[j]);
                NW_HIP(launch_srd_adj(A.collIds + q * NT, A.collOff + q * (A.Cs + 2), A.collHash + q * NT,
                                      A.collCom + q * A.Cs, A.wVel, lv, A.cellSum, A.N, uint32_t(NT), A.Cs, A.cosA,
                                      A.sinA, A.rotBeads, s.stream));
            }
            double *part = A.partial + uint64_t(j) * rowsPerStep * kAdjParams;
            AdjStepArgs as;
            as.pos = A.segPos + uint64_t(j + 1) * N;
            as.vel = A.segVel + uint64_t(j + 1) * N;
            as.sPos = A.segSPos + uint64_t(j) * N;
            as.ids = A.segIds + uint64_t(j) * N;
            as.off = A.segOff + uint64_t(j) * (A.C + 2);
            as.lF = lF[cur];
            as.lFout = lF[1 - cur];
            as.lr = lr;
            as.lv = lv;
            as.lG = lG;
            as.partial = part;
            as.W = A.W;
            as.langevin = A.langevin;
            as.lgvStd = A.lgvStd;
            as.gamma = A.gamma;
            as.dt = A.dt;
            as.seed = A.seed;
            as.offLgv = lgvOff[j];
            {
                TimedScope ts("adj_step", s.stream, true);
                NW_HIP(launch_adj_step(as, A.kp, s.stream));
            }
            cur = 1 - cur;
        }
        NW_HIP(launch_adj_total(A.partial, uint64_t(len) * rowsPerStep, A.grad, s.stream));
    }
    s.offset = A.endOffset;  // the generator continues after the forward run
    NW_HIP(hipMemcpyAsync(grad, A.grad, kAdjParams * sizeof(double), hipMemcpyDeviceToHost, s.stream));
    NW_HIP(hipStreamSynchronize(s.stream));
}

// ---------------------------------------------------------------------------
// SRD solvent and adjoint of batched replicas
// ---------------------------------------------------------------------------
void nwBatchSetSolvent(const NWReplicaSolvent *solvent, uint numReplicas) {
    NW_INIT();
    State &s = S();
    State::Batch &b = s.batch;
    State::BatchSolvent &q = s.bsol;
    s.badj.ready = false;
    q.R = q.Ns = q.cells = 0;
    q.reps.clear();
    if (q.dev) (void)hipFree(q.dev);
    q.dev = nullptr;
    if (numReplicas == 0) return;
    NW_REQUIRE(solvent != nullptr && numReplicas == b.R,
               "nwBatchSetSolvent: one solvent per replica of the current batch (%u)", b.R);
    q.reps.assign(numReplicas, SrdRep{});
    q.solvOff.assign(numReplicas + 1, b.N);
    q.cellOff.assign(numReplicas + 1, 0u);
    uint64_t NT = b.N, C = 0;
    const float alpha = solvent[0].solvent.alpha;
    const int rot = b.reps[0].kp.e.srdCollideFilaments != 0 ? 1 : 0;
    for (uint r = 0; r < numReplicas; ++r) {
        const SolventParams &sp = solvent[r].solvent;
        const uint64_t cells = uint64_t(sp.gridSize.x) * sp.gridSize.y * sp.gridSize.z;
        NW_REQUIRE(cells > 0 && cells == sp.numCells, "nwBatchSetSolvent: replica %u: SRD grid does not match numCells",
                   r);
        NW_REQUIRE(sp.alpha == alpha, "nwBatchSetSolvent: replica %u: one rotation angle per batch", r);
        NW_REQUIRE((b.reps[r].kp.e.srdCollideFilaments != 0 ? 1 : 0) == rot,
                   "nwBatchSetSolvent: replica %u: one srdCollideFilaments per batch", r);
        SrdRep &d = q.reps[r];
        d.kp = make_kp_for(b.reps[r].kp.p, sp, b.reps[r].kp.e);
        d.beadOffset = b.beadOff[r];
        d.numBeads = b.reps[r].numBeads;
        d.solvOffset = uint32_t(NT);
        d.numSolvent = solvent[r].numSolvent;
        d.cellOffset = uint32_t(C);
        d.cells = uint32_t(cells);
        d.seed = b.reps[r].seed;
        d.rngBase = b.reps[r].rngBase;
        d.cellBlocks = blocks_for(cells);
        NT += solvent[r].numSolvent;
        C += cells;
        NW_REQUIRE(NT < (1ull << 31) && C < (1ull << 32) - 2, "nwBatchSetSolvent: batch too large");
        q.solvOff[r + 1] = uint32_t(NT);
        q.cellOff[r + 1] = uint32_t(C);
    }
    const BucketPlan bp = make_bucket_plan(NT, C, 0);
    NW_REQUIRE(!*bucket_plan_problem(bp), "nwBatchSetSolvent: %s", bucket_plan_problem(bp));
    q.Ns = uint32_t(NT - b.N);
    q.cells = uint32_t(C);
    q.cosA = float(std::cos(double(alpha)));
    q.sinA = float(std::sin(double(alpha)));
    q.rotBeads = rot;
    NW_HIP(hipMalloc(&q.dev, sizeof(SrdRep) * numReplicas));
    NW_HIP(hipMemcpy(q.dev, q.reps.data(), sizeof(SrdRep) * numReplicas, hipMemcpyHostToDevice));
    q.R = numReplicas;
}

void nwBatchSolventLayout(uint *solventOffset, uint *srdCellOffset) {
    const State::BatchSolvent &q = S().bsol;
    for (uint r = 0; r <= q.R; ++r) {
        if (solventOffset) solventOffset[r] = q.solvOff.empty() ? 0u : q.solvOff[r];
        if (srdCellOffset) srdCellOffset[r] = q.cellOff.empty() ? 0u : q.cellOff[r];
    }
}

// One batched MD step writing its cell lists into a (ids, off, sCell, sPos),
// so a replayed step keeps them: the WCA chain worms of the adjoint (no
// velocity or tangent gathers).
static hipError_t batch_step_into(State::Batch &b, const BatchArrays &a, float dt, hipStream_t st) {
    BucketScratch sc = b.scratch;
    sc.off = a.off;
    sc.sortedCell = a.sCell;
    hipError_t e;
    if ((e = launch_md_front_batch(a, b.dev, b.R, b.waves, b.stepCtr, dt, st)) != hipSuccess) return e;
    if ((e = launch_hash_tiles_batch(a.hash, b.sortDev, b.R, b.sortPlan, b.scratch, st)) != hipSuccess) return e;
    if ((e = launch_md_cells_batch(b.sortDev, b.R, b.sortPlan, sc, a.ids, a.pos, nullptr, nullptr, a.sPos, nullptr,
                                   nullptr, st)) != hipSuccess)
        return e;
    return launch_collide_batch(a, b.dev, b.R, b.blocks, b.mode, b.tan, b.unit, b.stepCtr, st);
}

// Collision c of every replica of the batch over the full arrays (all beads,
// then the solvent): cell of every particle into hash, the cell order into
// ids, its CSR offsets into offOut, (u, sign) per cell into com.
static bool batch_srd_collision(State &s, uint32_t c, float4 *P, float4 *V, uint32_t *hash, uint32_t *ids,
                                uint32_t *offOut, float4 *com, const float *shift, float dtSrd) {
    State::Batch &b = s.batch;
    State::BatchSolvent &q = s.bsol;
    const uint32_t NT = b.N + q.Ns;
    const BucketPlan bp = make_bucket_plan(NT, q.cells, 0);
    NW_HIP_R(ensure_bucket_scratch(q.sort, bp), false);
    BucketScratch sc = q.sort;
    if (offOut) sc.off = offOut;
    NW_HIP_R(launch_srd_hash_batch(P, hash, q.dev, q.R, b.N, NT, shift, s.stream), false);
    NW_HIP_R(launch_hash_tiles(hash, bp, q.sort, s.stream, true), false);
    NW_HIP_R(launch_srd_cells_batch(bp, sc, ids, V, com, q.dev, q.R, c, s.stream), false);
    NW_HIP_R(launch_srd_rotate_batch(hash, com, P, V, q.dev, q.R, b.N, NT, dtSrd, q.cosA, q.sinA, q.rotBeads,
                                     s.stream),
             false);
    return true;
}

void nwBatchSrdStep(float *pos, float *vel, float *solventCellCOM, uint *gridParticleHash, uint *gridParticleIndex,
                    float deltaTime, const float *shifts, uint collision) {
    NW_INIT();
    State &s = S();
    State::Batch &b = s.batch;
    State::BatchSolvent &q = s.bsol;
    NW_REQUIRE(b.R > 0 && q.R == b.R, "nwBatchSrdStep: no batch solvent (nwBatchSetReplicas, nwBatchSetSolvent)");
    NW_REQUIRE(pos && vel && solventCellCOM && gridParticleHash && gridParticleIndex && shifts,
               "nwBatchSrdStep: NULL argument");
    NW_REQUIRE(grow(q.shift, q.capShift, uint64_t(b.R) * 3), "nwBatchSrdStep: out of device memory");
    // a synchronous copy: the previous collision's kernels may still read the buffer
    NW_HIP(hipStreamSynchronize(s.stream));
    NW_HIP(hipMemcpy(q.shift, shifts, size_t(b.R) * 3 * sizeof(float), hipMemcpyHostToDevice));
    (void)batch_srd_collision(s, collision, reinterpret_cast<float4 *>(pos), reinterpret_cast<float4 *>(vel),
                              gridParticleHash, gridParticleIndex, nullptr, reinterpret_cast<float4 *>(solventCellCOM),
                              q.shift, deltaTime);
}

void nwBatchAdjointForward(float *pos, float *vel, float *force, float *forceOld, float *tangent, float deltaTime,
                           uint srdGap, const float *srdShifts, uint steps, uint checkpointEvery) {
    NW_INIT();
    State &s = S();
    State::Batch &b = s.batch;
    State::BatchSolvent &q = s.bsol;
    State::BatchAdjoint &A = s.badj;
    A.ready = false;
    NW_REQUIRE(b.R > 0 && b.N > 0, "nwBatchAdjointForward: no batch (nwBatchSetReplicas)");
    NW_REQUIRE(pos && vel && force && forceOld && tangent, "nwBatchAdjointForward: NULL array");
    NW_REQUIRE(checkpointEvery >= 1 && steps >= 1, "nwBatchAdjointForward: steps and checkpointEvery >= 1");
    NW_REQUIRE(b.mode == 1 && !b.tan, "nwBatchAdjointForward: WCA pairs without the extensile term only");
    for (uint32_t r = 0; r < b.R; ++r) {
        const KP &kp = b.reps[r].kp;
        NW_REQUIRE(kp.e.bondedMode == NW_BONDED_CHAIN && kp.p.reverseProbability == 0.0f,
                   "nwBatchAdjointForward: replica %u: chain bonds and no reversal only", r);
        NW_REQUIRE(kp.p.boundaryX == PERIODIC && kp.p.boundaryY == PERIODIC && kp.p.boundaryZ == PERIODIC,
                   "nwBatchAdjointForward: replica %u: periodic boxes only", r);
        NW_REQUIRE(!b.reps[r].langevin || b.reps[r].lgvGamma == kp.p.gamma,
                   "nwBatchAdjointForward: replica %u: the Langevin gamma must be SimParams.gamma", r);
    }
    const bool srd = srdGap > 0;
    if (srd) {
        NW_REQUIRE(q.R == b.R, "nwBatchAdjointForward: SRD collisions need the batch's solvent (nwBatchSetSolvent)");
        NW_REQUIRE(srdShifts != nullptr, "nwBatchAdjointForward: NULL grid shifts");
        for (uint32_t r = 0; r < b.R; ++r)
            NW_REQUIRE(b.reps[r].rngStride == 0,
                       "nwBatchAdjointForward: replica %u: no Langevin thermostat with the solvent", r);
    }
    const uint64_t N = b.N, NT = N + (srd ? q.Ns : 0), C = b.cells, Cs = srd ? q.cells : 0;
    A.srd = srd;
    A.gap = srdGap;
    A.steps = steps;
    A.K = checkpointEvery;
    A.dt = deltaTime;
    A.N = b.N;
    A.NT = uint32_t(NT);
    A.C = b.cells;
    A.Cs = uint32_t(Cs);
    A.R = b.R;
    A.waves = b.waves;
    A.ctr0 = b.steps;
    const uint32_t nck = (steps + checkpointEvery - 1) / checkpointEvery;
    const uint64_t stride = 2 * NT + 2 * N;
    NW_REQUIRE(grow(A.ckpt, A.capCkpt, uint64_t(nck) * stride) && grow(A.hash, A.capHash, NT) &&
                   grow(A.sCell, A.capSCell, N) && grow(A.segIds, A.capIds, N) && grow(A.segSPos, A.capSPos, N) &&
                   grow(A.segOff, A.capOff, C + 2),
               "nwBatchAdjointForward: out of device memory");
    if (srd) {
        const uint64_t nColl = collisions_in(steps, srdGap);
        NW_REQUIRE(grow(A.shifts, A.capShifts, nColl * b.R * 3) && grow(A.collHash, A.capCollHash, NT) &&
                       grow(A.collIds, A.capCollIds, NT) && grow(A.collOff, A.capCollOff, Cs + 2) &&
                       grow(A.collCom, A.capCollCom, Cs),
                   "nwBatchAdjointForward: out of device memory");
        NW_HIP(hipMemcpy(A.shifts, srdShifts, nColl * b.R * 3 * sizeof(float), hipMemcpyHostToDevice));
    }
    BatchArrays a{};
    a.pos = reinterpret_cast<float4 *>(pos);
    a.vel = reinterpret_cast<float4 *>(vel);
    a.force = reinterpret_cast<float4 *>(force);
    a.forceOld = reinterpret_cast<float4 *>(forceOld);
    a.tangent = reinterpret_cast<float4 *>(tangent);
    a.hash = A.hash;
    a.ids = A.segIds;
    a.off = A.segOff;
    a.sCell = A.sCell;
    a.sPos = A.segSPos;
    const size_t bytesT = NT * sizeof(float4), bytes = N * sizeof(float4);
    for (uint32_t k = 0; k < steps; ++k) {
        if (k % checkpointEvery == 0) {
            float4 *ck = A.ckpt + uint64_t(k / checkpointEvery) * stride;
            NW_REQUIRE(copy_d2d(ck, pos, bytesT, s.stream) && copy_d2d(ck + NT, vel, bytesT, s.stream) &&
                           copy_d2d(ck + 2 * NT, force, bytes, s.stream) &&
                           copy_d2d(ck + 2 * NT + N, forceOld, bytes, s.stream),
                       "nwBatchAdjointForward: checkpoint copy failed");
        }
        {
            TimedScope ts("batch_adj_fwd_md", s.stream);
            NW_HIP(batch_step_into(b, a, deltaTime, s.stream));
        }
        ++b.steps;
        if (srd && k % srdGap == 0) {
            TimedScope ts("batch_adj_fwd_srd", s.stream);
            const uint32_t c = k / srdGap;
            if (!batch_srd_collision(s, c, a.pos, a.vel, A.collHash, A.collIds, A.collOff, A.collCom,
                                     A.shifts + uint64_t(c) * b.R * 3, deltaTime * float(srdGap)))
                return;
        }
    }
    A.ready = true;
}

static bool ensure_batch_segment(State::BatchAdjoint &A) {
    const uint64_t N = A.N, K = A.K, NT = A.NT;
    if (!(grow(A.segPos, A.capPos, (K + 1) * N) && grow(A.segVel, A.capVel, (K + 1) * N) &&
          grow(A.segSPos, A.capSPos, K * N) && grow(A.segIds, A.capIds, K * N) &&
          grow(A.segOff, A.capOff, K * (uint64_t(A.C) + 2)) && grow(A.work, A.capWork, 3 * N) &&
          grow(A.lam, A.capLam, 4 * N + NT) && grow(A.wPos, A.capWPos, NT) && grow(A.wVel, A.capWVel, NT) &&
          grow(A.partial, A.capPartial, uint64_t(A.waves) * kAdjParams) &&
          grow(A.grad, A.capGrad, uint64_t(A.R) * kAdjParams)))
        return false;
    if (!A.srd) return true;
    const uint64_t cpS = (K + A.gap - 1) / A.gap + 1;  // collisions in a segment, at most
    return grow(A.collHash, A.capCollHash, cpS * NT) && grow(A.collIds, A.capCollIds, cpS * NT) &&
           grow(A.collOff, A.capCollOff, cpS * (uint64_t(A.Cs) + 2)) &&
           grow(A.collCom, A.capCollCom, cpS * A.Cs) && grow(A.cellSum, A.capCellSum, A.Cs);
}

void nwBatchAdjointBackward(const float *dLdPos, const float *dLdVel, double *grad) {
    NW_INIT();
    State &s = S();
    State::Batch &b = s.batch;
    State::BatchSolvent &q = s.bsol;
    State::BatchAdjoint &A = s.badj;
    NW_REQUIRE(A.ready, "nwBatchAdjointBackward: no nwBatchAdjointForward run to differentiate");
    NW_REQUIRE(dLdPos != nullptr && grad != nullptr, "nwBatchAdjointBackward: NULL argument");
    NW_REQUIRE(b.R == A.R && b.N == A.N, "nwBatchAdjointBackward: the batch changed since the forward run");
    NW_REQUIRE(ensure_batch_segment(A), "nwBatchAdjointBackward: out of device memory");
    const uint64_t N = A.N, NT = A.NT, K = A.K, C = A.C, stride = 2 * NT + 2 * N;
    const size_t bytes = N * sizeof(float4);
    float4 *lr = A.lam, *lF[2] = {A.lam + N, A.lam + 2 * N}, *lG = A.lam + 3 * N, *lv = A.lam + 4 * N;
    NW_HIP(hipMemcpyAsync(lr, dLdPos, bytes, hipMemcpyDeviceToDevice, s.stream));
    NW_HIP(hipMemsetAsync(lv, 0, NT * sizeof(float4), s.stream));
    if (dLdVel) NW_HIP(hipMemcpyAsync(lv, dLdVel, bytes, hipMemcpyDeviceToDevice, s.stream));
    NW_HIP(hipMemsetAsync(lF[0], 0, bytes, s.stream));
    NW_HIP(hipMemsetAsync(lG, 0, bytes, s.stream));
    NW_HIP(hipMemsetAsync(A.partial, 0, uint64_t(A.waves) * kAdjParams * sizeof(double), s.stream));
    int cur = 0;
    float4 *F = A.work, *G = A.work + N, *T = A.work + 2 * N;
    const uint32_t nck = (A.steps + A.K - 1) / A.K;
    std::vector<int> collSlot(K);
    // the step velocities are read only for the Langevin partials
    bool lgv = false;
    for (uint32_t r = 0; r < b.R; ++r) lgv = lgv || b.reps[r].langevin != 0;
    for (int c = int(nck) - 1; c >= 0; --c) {
        const uint32_t s0 = uint32_t(c) * A.K, len = std::min<uint32_t>(A.K, A.steps - s0);
        const float4 *ck = A.ckpt + uint64_t(c) * stride;
        NW_HIP(hipMemcpyAsync(A.wPos, ck, NT * sizeof(float4), hipMemcpyDeviceToDevice, s.stream));
        NW_HIP(hipMemcpyAsync(A.wVel, ck + NT, NT * sizeof(float4), hipMemcpyDeviceToDevice, s.stream));
        NW_HIP(hipMemcpyAsync(F, ck + 2 * NT, bytes, hipMemcpyDeviceToDevice, s.stream));
        NW_HIP(hipMemcpyAsync(G, ck + 2 * NT + N, bytes, hipMemcpyDeviceToDevice, s.stream));
        NW_HIP(launch_set_step(b.stepCtr, A.ctr0 + s0, s.stream));
        // recompute the segment: the full state in place, every step's bead
        // state and cell lists and every collision's cells kept
        int slots = 0;
        for (uint32_t j = 0; j < len; ++j) {
            BatchArrays a{};
            a.pos = A.wPos;
            a.vel = A.wVel;
            a.force = F;
            a.forceOld = G;
            a.tangent = T;
            a.hash = A.hash;
            a.ids = A.segIds + uint64_t(j) * N;
            a.off = A.segOff + uint64_t(j) * (C + 2);
            a.sCell = A.sCell;
            a.sPos = A.segSPos + uint64_t(j) * N;
            NW_HIP(batch_step_into(b, a, A.dt, s.stream));
            NW_HIP(hipMemcpyAsync(A.segPos + uint64_t(j + 1) * N, A.wPos, bytes, hipMemcpyDeviceToDevice, s.stream));
            if (lgv)
                NW_HIP(hipMemcpyAsync(A.segVel + uint64_t(j + 1) * N, A.wVel, bytes, hipMemcpyDeviceToDevice,
                                      s.stream));
            collSlot[j] = -1;
            const uint32_t k = s0 + j;
            if (A.srd && k % A.gap == 0) {
                const int sl = slots++;
                collSlot[j] = sl;
                const uint64_t u = uint64_t(sl);
                if (!batch_srd_collision(s, k / A.gap, A.wPos, A.wVel, A.collHash + u * NT, A.collIds + u * NT,
                                         A.collOff + u * (A.Cs + 2), A.collCom + u * A.Cs,
                                         A.shifts + uint64_t(k / A.gap) * A.R * 3, A.dt * float(A.gap)))
                    return;
            }
        }
        // the adjoint through the segment, last step first
        for (int j = int(len) - 1; j >= 0; --j) {
            if (collSlot[j] >= 0) {
                const uint64_t u = uint64_t(collSlot[j]);
                NW_HIP(launch_srd_adj(A.collIds + u * NT, A.collOff + u * (A.Cs + 2), A.collHash + u * NT,
                                      A.collCom + u * A.Cs, A.wVel, lv, A.cellSum, A.N, uint32_t(NT), A.Cs, q.cosA,
                                      q.sinA, q.rotBeads, s.stream));
            }
            AdjStepArgs as{};
            as.pos = A.segPos + uint64_t(j + 1) * N;
            as.vel = A.segVel + uint64_t(j + 1) * N;
            as.sPos = A.segSPos + uint64_t(j) * N;
            as.ids = A.segIds + uint64_t(j) * N;
            as.off = A.segOff + uint64_t(j) * (C + 2);
            as.lF = lF[cur];
            as.lFout = lF[1 - cur];
            as.lr = lr;
            as.lv = lv;
            as.lG = lG;
            as.partial = A.partial;
            as.dt = A.dt;
            {
                TimedScope ts("badj_step", s.stream, true);
                NW_HIP(launch_adj_step_batch(as, b.dev, b.R, A.waves, A.ctr0 + s0 + uint32_t(j), s.stream));
            }
            cur = 1 - cur;
        }
    }
    NW_HIP(launch_set_step(b.stepCtr, A.ctr0 + A.steps, s.stream));  // where the forward run left it
    NW_HIP(launch_adj_total_batch(A.partial, b.dev, b.R, A.waves, A.grad, s.stream));
    NW_HIP(hipMemcpyAsync(grad, A.grad, uint64_t(A.R) * kAdjParams * sizeof(double), hipMemcpyDeviceToHost,
                          s.stream));
    NW_HIP(hipStreamSynchronize(s.stream));
}

void nwBatchPolarSpeedGrad(const float *pos, const float *vel, float *dPdPos, float *dPdVel, double *P) {
    NW_INIT();
    State &s = S();
    State::Batch &b = s.batch;
    NW_REQUIRE(pos != nullptr && vel != nullptr && P != nullptr, "nwBatchPolarSpeedGrad: NULL argument");
    for (uint32_t r = 0; r < b.R; ++r) {
        const BatchRep &rep = b.reps[r];
        const size_t o = size_t(rep.beadOffset) * 4;
        NWObservables obs;
        observe_kp(rep.kp, pos + o, vel + o, rep.kp.p.numFilaments, &obs);
        P[r] = obs.polarSpeed;
        if (!dPdPos) continue;
        if (obs.beads > 0) {
            NW_HIP(launch_polar_speed_grad(reinterpret_cast<const float4 *>(pos + o),
                                           reinterpret_cast<const float4 *>(vel + o), rep.kp.p.numFilaments,
                                           1.0 / obs.beads, reinterpret_cast<float4 *>(dPdPos + o),
                                           dPdVel ? reinterpret_cast<float4 *>(dPdVel + o) : nullptr, rep.kp,
                                           s.stream));
        } else {
            NW_HIP(hipMemsetAsync(dPdPos + o, 0, size_t(rep.numBeads) * 16, s.stream));
            if (dPdVel) NW_HIP(hipMemsetAsync(dPdVel + o, 0, size_t(rep.numBeads) * 16, s.stream));
        }
    }
    NW_HIP(hipStreamSynchronize(s.stream));
}

// unit eigenvector of the symmetric 3x3 a (a11 a12 a13 a22 a23 a33) for the
// eigenvalue lam: the longest cross product of two rows of a - lam I
static void sym_eigvec3(const double a[6], double lam, double e[3]) {
    const double m[3][3] = {{a[0] - lam, a[1], a[2]}, {a[1], a[3] - lam, a[4]}, {a[2], a[4], a[5] - lam}};
    double best = -1.0;
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j) {
            const double c[3] = {m[i][1] * m[j][2] - m[i][2] * m[j][1], m[i][2] * m[j][0] - m[i][0] * m[j][2],
                                 m[i][0] * m[j][1] - m[i][1] * m[j][0]};
            const double l = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
            if (l > best) {
                best = l;
                e[0] = c[0];
                e[1] = c[1];
                e[2] = c[2];
            }
        }
    const double l = std::sqrt(best);
    if (l > 0) {
        e[0] /= l;
        e[1] /= l;
        e[2] /= l;
    } else {
        e[0] = 1.0;
        e[1] = e[2] = 0.0;
    }
}

double nwBondOrderGrad(const float *pos, uint numFilaments, float *dSdPos) {
    if (!ensure_init()) return 0.0;
    State &s = S();
    NW_REQUIRE_R(pos != nullptr, 0.0, "nwBondOrderGrad: NULL positions");
    NWObservables o;
    nwObservables(pos, pos, numFilaments, &o);
    if (o.bonds <= 0) return 0.0;
    const double nb = o.bonds;
    const double Q[6] = {1.5 * o.sums[kObsBxx] / nb - 0.5, 1.5 * o.sums[kObsBxy] / nb, 1.5 * o.sums[kObsBxz] / nb,
                         1.5 * o.sums[kObsByy] / nb - 0.5, 1.5 * o.sums[kObsByz] / nb, 1.5 * o.sums[kObsBzz] / nb - 0.5};
    const double S = nwSymEig3Max(Q);
    if (dSdPos) {
        double e[3];
        sym_eigvec3(Q, S, e);
        NW_HIP_R(launch_bond_order_grad(reinterpret_cast<const float4 *>(pos), numFilaments, e, 3.0 / nb,
                                        reinterpret_cast<float4 *>(dSdPos), make_kp(), s.stream),
                 S);
        NW_HIP_R(hipStreamSynchronize(s.stream), S);
    }
    return S;
}

void nwSortSolvent(float *pos, float *vel, uint *solventId, uint numParticles, uint numSolvent) {
    NW_INIT();
    State &s = S();
    NW_REQUIRE(pos && vel, "nwSortSolvent: NULL positions or velocities");
    if (numSolvent == 0) return;
    const KP kp = make_kp();
    const uint64_t cells = uint64_t(kp.s.gridSize.x) * kp.s.gridSize.y * kp.s.gridSize.z;
    NW_REQUIRE(cells > 0 && cells == kp.s.numCells, "nwSortSolvent: SRD grid does not match numCells");
    NW_REQUIRE(uint64_t(numParticles) + numSolvent < (1ull << 31), "nwSortSolvent: too many particles");
    const uint32_t N = numParticles, Ns = numSolvent;
    NW_REQUIRE(grow(s.solKey, s.capSolKey, Ns) && grow(s.solIdx, s.capSolIdx, Ns) && grow(s.solId, s.capSolId, Ns) &&
                   grow(s.solPos, s.capSolPos, Ns) && grow(s.solVel, s.capSolVel, Ns),
               "nwSortSolvent: out of device memory");
    float4 *P = reinterpret_cast<float4 *>(pos), *V = reinterpret_cast<float4 *>(vel);
    TimedScope ts("srd_resort", s.stream);
    NW_HIP(launch_solvent_keys(P, N, Ns, s.solKey, s.solIdx, kp, s.stream));
    // stable by cell: solIdx becomes the new order
    const hipError_t e = bucket_sort_pairs(s.solKey, s.solIdx, Ns, uint32_t(cells), s.bucket, s.stream);
    if (e == hipErrorNotSupported)
        NW_HIP(counting_sort_pairs(s.solKey, s.solIdx, Ns, uint32_t(cells), s.sort, s.stream));
    else
        NW_HIP(e);
    NW_HIP(launch_solvent_gather(P, V, solventId, s.solIdx, N, Ns, s.solPos, s.solVel, s.solId, s.stream));
    const size_t b4 = size_t(Ns) * sizeof(float4);
    NW_HIP(hipMemcpyAsync(P + N, s.solPos, b4, hipMemcpyDeviceToDevice, s.stream));
    NW_HIP(hipMemcpyAsync(V + N, s.solVel, b4, hipMemcpyDeviceToDevice, s.stream));
    if (solventId)
        NW_HIP(hipMemcpyAsync(solventId, s.solId, size_t(Ns) * 4, hipMemcpyDeviceToDevice, s.stream));
    s.lastHashPtr = nullptr;
}

double nwPolarSpeedGrad(const float *pos, const float *vel, uint numFilaments, float *dPdPos, float *dPdVel) {
    if (!ensure_init()) return 0.0;
    State &s = S();
    NW_REQUIRE_R(pos != nullptr && vel != nullptr, 0.0, "nwPolarSpeedGrad: NULL positions or velocities");
    NWObservables o;
    nwObservables(pos, vel, numFilaments, &o);
    if (o.beads <= 0) return 0.0;
    if (dPdPos)
        NW_HIP_R(launch_polar_speed_grad(reinterpret_cast<const float4 *>(pos), reinterpret_cast<const float4 *>(vel),
                                         numFilaments, 1.0 / o.beads, reinterpret_cast<float4 *>(dPdPos),
                                         reinterpret_cast<float4 *>(dPdVel), make_kp(), s.stream),
                 o.polarSpeed);
    NW_HIP_R(hipStreamSynchronize(s.stream), o.polarSpeed);
    return o.polarSpeed;
}

void nwSetTuning(const NWTuning *t) {
    NW_REQUIRE(t != nullptr, "nwSetTuning: NULL tuning");
    auto oneOf = [](uint32_t v, std::initializer_list<uint32_t> ok) {
        if (v == 0) return true;
        for (uint32_t x : ok)
            if (v == x) return true;
        return false;
    };
    NW_REQUIRE(oneOf(t->mdTileItems, {1024, 2048, 4096, 8192, 16384}), "nwSetTuning: mdTileItems %u",
               t->mdTileItems);
    NW_REQUIRE(oneOf(t->srdTileItems, {4096, 8192, 16384}), "nwSetTuning: srdTileItems %u", t->srdTileItems);
    NW_REQUIRE(t->mdMembers == 0 || (t->mdMembers >= 64 && t->mdMembers <= 2048), "nwSetTuning: mdMembers %u",
               t->mdMembers);
    NW_REQUIRE(t->srdMembers == 0 || (t->srdMembers >= 64 && t->srdMembers <= 2048), "nwSetTuning: srdMembers %u",
               t->srdMembers);
    NW_REQUIRE(oneOf(t->pairBlock, {128, 256, 512}), "nwSetTuning: pairBlock %u", t->pairBlock);
    NW_REQUIRE(oneOf(t->frontBlock, {64, 128, 256, 512, 1024}), "nwSetTuning: frontBlock %u", t->frontBlock);
    NW_REQUIRE(t->pairStage <= 2048, "nwSetTuning: pairStage %u", t->pairStage);
    Tuning &g = tuning();
    g.mdTileItems = t->mdTileItems;
    g.mdMembers = t->mdMembers;
    g.srdTileItems = t->srdTileItems;
    g.srdMembers = t->srdMembers;
    g.pairBlock = t->pairBlock;
    g.frontBlock = t->frontBlock;
    g.pairStage = t->pairStage;
}

void nwGetTuning(NWTuning *t) {
    if (!t) return;
    *t = NWTuning{};
    const Tuning &g = tuning();
    t->mdTileItems = g.mdTileItems;
    t->mdMembers = g.mdMembers;
    t->srdTileItems = g.srdTileItems;
    t->srdMembers = g.srdMembers;
    t->pairBlock = g.pairBlock;
    t->frontBlock = g.frontBlock;
    t->pairStage = g.pairStage;
}

int nwFusedPlanInfo(uint numParticles, uint numCells, uint filamentSize, uint out[4]) {
    const BucketPlan bp = fused_plan(numParticles, numCells, filamentSize == 0);
    if (out) {
        out[0] = bp.shift;
        out[1] = bp.buckets;
        out[2] = bp.tileItems;
        out[3] = bp.tiles;
    }
    if (filamentSize > 64 || numCells == 0) return 0;
    return *bucket_plan_problem(bp) ? 0 : 1;
}

int nwFusedPlanFits(uint numParticles, uint numCells, uint filamentSize) {
    if (filamentSize > 64 || numCells == 0) return 0;
    return *bucket_plan_problem(fused_plan(numParticles, numCells, filamentSize == 0)) ? 0 : 1;
}

uint nwGetDeviceFlags(int clear) {
    if (!ensure_init()) return 0;
    State &s = S();
    uint32_t v = 0;
    uint32_t *d = device_flags();
    NW_HIP_R(hipStreamSynchronize(s.stream), 0);
    NW_HIP_R(hipMemcpy(&v, d, 4, hipMemcpyDeviceToHost), 0);
    if (clear) {
        const uint32_t z = 0;
        NW_HIP_R(hipMemcpy(d, &z, 4, hipMemcpyHostToDevice), 0);
    }
    return v;
}

void nwTimerEnable(const char *kernel, int enable) {
    if (!kernel) return;
    Timer &t = S().timers[kernel];
    t.enabled = enable > 0;
    if (enable > 0) {
        t.stride = uint32_t(enable);
        t.seen = 0;
        t.totalMs = 0.0;
        t.launches = 0;
    }
}

void nwTimerRead(const char *kernel, double *totalMs, uint64_t *launches) {
    if (!kernel) return;
    Timer &t = S().timers[kernel];
    for (auto &ev : t.pending) {
        (void)hipEventSynchronize(ev.second);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ev.first, ev.second);
        t.totalMs += ms;
        t.pool.push_back(ev);
    }
    t.pending.clear();
    if (totalMs) *totalMs = t.totalMs;
    if (launches) *launches = t.launches;
}

}  // extern "C"

// nw_tune.hip -- the autotuner of the fused steps' launch shapes
// (nwAutotuneFused, include/nw_abi.h).  It is a client of the library's own
// C-ABI: on a copy of the caller's state it runs nwFusedMdStep (and
// nwFusedSrdStep) under each candidate NWTuning, reads the per-kernel timers
// (HIP events on the kernels' dispatch packets) and keeps the fastest choice
// of every knob.  The knobs act on different kernels, so they are searched
// one kernel at a time: the front's workgroup size on the front, the MD
// sort's tile and bucket sizes jointly on its two levels, the pair kernel's
// workgroup size on the pair kernel, the SRD sort's tile and bucket sizes on
// the collision's two sort levels.  Choices are cached in a text file, one
// line per device and system size.
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "nw_abi.h"
#include "nw_internal.hpp"

namespace {

// a candidate within this fraction of the built-in choice's time keeps the
// built-in choice (timer noise is about 1%)
constexpr double kKeepDefaultWithin = 0.01;

struct Copy {
    void *p = nullptr;
    ~Copy() {
        if (p) (void)hipFree(p);
    }
};

bool alloc_copy(Copy &c, const void *src, size_t bytes, hipStream_t st) {
    if (hipMalloc(&c.p, bytes ? bytes : 16) != hipSuccess) return false;
    if (src && bytes) return hipMemcpyAsync(c.p, src, bytes, hipMemcpyDeviceToDevice, st) == hipSuccess;
    return hipMemsetAsync(c.p, 0, bytes ? bytes : 16, st) == hipSuccess;
}

std::string device_key() {
    int dev = 0;
    hipDeviceProp_t prop{};
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
    std::string k = std::string(prop.gcnArchName) + "/" + prop.name + "/" + std::to_string(prop.multiProcessorCount);
    for (char &c : k)
        if (c == ' ' || c == '\t') c = '_';
    return k;
}

std::string cache_file(const char *path) {
    if (path) return path;
    if (const char *e = std::getenv("NW_TUNE_CACHE")) return e;
    const char *home = std::getenv("HOME");
    return std::string(home ? home : ".") + "/.cache/nwamd/tuning.txt";
}

void make_parent_dirs(const std::string &file) {
    for (size_t i = 1; i < file.size(); ++i)
        if (file[i] == '/') (void)mkdir(file.substr(0, i).c_str(), 0755);
}

std::string size_key(uint64_t N, uint64_t C, uint32_t n, uint64_t Ns, uint64_t Cs) {
    std::ostringstream o;
    o << N << ' ' << C << ' ' << n << ' ' << Ns << ' ' << Cs;
    return o.str();
}

bool cache_lookup(const std::string &file, const std::string &dev, const std::string &size, NWTuning &t) {
    std::ifstream in(file);
    std::string line;
    bool found = false;
    while (std::getline(in, line)) {  // the last line of a key wins
        const size_t a = line.find('\t'), b = line.rfind('\t');
        if (a == std::string::npos || b == a) continue;
        if (line.substr(0, a) != dev || line.substr(a + 1, b - a - 1) != size) continue;
        NWTuning c{};
        std::istringstream v(line.substr(b + 1));
        if (v >> c.mdTileItems >> c.mdMembers >> c.srdTileItems >> c.srdMembers >> c.pairBlock >> c.frontBlock >>
            c.pairStage) {
            t = c;
            found = true;
        }
    }
    return found;
}

void cache_store(const std::string &file, const std::string &dev, const std::string &size, const NWTuning &t) {
    make_parent_dirs(file);
    std::ofstream out(file, std::ios::app);
    if (!out) return;  // an unwritable cache only costs the next run its timing
    out << dev << '\t' << size << '\t' << t.mdTileItems << ' ' << t.mdMembers << ' ' << t.srdTileItems << ' '
        << t.srdMembers << ' ' << t.pairBlock << ' ' << t.frontBlock << ' ' << t.pairStage << '\n';
}

double avg_ms(const char *fam) {
    double ms = 0.0;
    uint64_t n = 0;
    nwTimerRead(fam, &ms, &n);
    return n ? ms / double(n) : 1e30;
}

}  // namespace

extern "C" int nwAutotuneFused(float *pos, float *vel, float *force, float *forceOld, float *tangent,
                               uint numFilaments, uint numSolvent, float deltaTime, float langevinGamma,
                               float langevinKbT, int langevin, float srdDeltaTime, int isokinetic, uint reps,
                               const char *cachePath, int useCache, NWTuning *out) {
    if (!nwDeviceAvailable()) return 0;
    const nw::KP kp = nw::current_kp();
    const uint32_t n = kp.p.filamentSize;
    const uint64_t N = uint64_t(numFilaments) * n, NT = N + numSolvent;
    const uint64_t C = uint64_t(kp.p.gridSize.x) * kp.p.gridSize.y * kp.p.gridSize.z;
    const uint64_t Cs = numSolvent ? uint64_t(kp.s.gridSize.x) * kp.s.gridSize.y * kp.s.gridSize.z : 0;
    if (!pos || !vel || !force || !forceOld || !tangent || N == 0 || n < 1 || n > 64 || C == 0 ||
        NT >= (1ull << 31) || (numSolvent && Cs == 0))
        return 0;
    if (reps == 0) reps = 5;
    const std::string file = cache_file(cachePath), dev = device_key(), size = size_key(N, C, n, numSolvent, Cs);
    NWTuning best{};
    if (useCache && !file.empty() && cache_lookup(file, dev, size, best)) {
        nwSetTuning(&best);
        if (nwGetLastError() != 0) return 0;
        if (out) *out = best;
        return 2;
    }
    hipStream_t st = static_cast<hipStream_t>(nwGetStream());
    // the copies the candidates run on
    Copy P, V, F, G, T, H, I, SP, SV, ST, COM;
    const size_t b4N = size_t(N) * 16, b4T = size_t(NT) * 16;
    if (!alloc_copy(P, pos, b4T, st) || !alloc_copy(V, vel, b4T, st) || !alloc_copy(F, force, b4N, st) ||
        !alloc_copy(G, forceOld, b4N, st) || !alloc_copy(T, tangent, b4N, st) ||
        !alloc_copy(H, nullptr, size_t(NT) * 4, st) || !alloc_copy(I, nullptr, size_t(NT) * 4, st) ||
        !alloc_copy(SP, nullptr, b4N, st) || !alloc_copy(SV, nullptr, b4N, st) || !alloc_copy(ST, nullptr, b4N, st) ||
        !alloc_copy(COM, nullptr, size_t(Cs ? Cs : 1) * 16, st))
        return 0;
    uint64_t seed = 0, offset = 0;
    nwRngGetState(&seed, &offset);
    NWTuning saved{};
    nwGetTuning(&saved);

    auto fP = static_cast<float *>(P.p), fV = static_cast<float *>(V.p);
    auto md_step = [&]() {
        nwFusedMdStep(fP, fV, static_cast<float *>(F.p), static_cast<float *>(G.p), static_cast<float *>(T.p),
                      static_cast<uint *>(H.p), static_cast<uint *>(I.p), static_cast<float *>(SP.p),
                      static_cast<float *>(SV.p), static_cast<float *>(ST.p), deltaTime, langevinGamma, langevinKbT,
                      langevin, numFilaments);
    };
    auto srd_step = [&]() {
        nwFusedSrdStep(fP, fV, static_cast<float *>(COM.p), static_cast<uint *>(H.p), static_cast<uint *>(I.p),
                       srdDeltaTime, uint(N), numSolvent, isokinetic);
    };
    // the summed average duration of `fams` over reps steps under tuning t
    // (one untimed step first: first launches, LDS grants); negative when
    // the library refused the setting or a step failed
    auto measure = [&](const NWTuning &t, std::initializer_list<const char *> fams, bool srd) {
        nwSetTuning(&t);
        if (nwGetLastError() != 0) return -1.0;
        if (srd) srd_step();
        else md_step();
        for (const char *f : fams) nwTimerEnable(f, 1);
        for (uint k = 0; k < reps; ++k) {
            if (srd) srd_step();
            else md_step();
        }
        double tot = 0.0;
        for (const char *f : fams) {
            tot += avg_ms(f);
            nwTimerEnable(f, 0);
        }
        return nwGetLastError() == 0 ? tot : -1.0;
    };
    // the plan a candidate asks for is the plan the step runs (else skipped:
    // it fell back to the built-in plan)
    auto plan_is = [&](uint64_t np, uint64_t cells, uint32_t fs, uint32_t items) {
        uint out4[4] = {0, 0, 0, 0};
        nwFusedPlanInfo(uint(np), uint(cells), fs, out4);
        return out4[2] == items;
    };
    // one knob (or a pair) of `best` over its candidates
    auto search = [&](std::vector<NWTuning> cands, std::initializer_list<const char *> fams, bool srd) {
        const double base = measure(best, fams, srd);
        double bestT = base;
        NWTuning pick = best;
        for (const NWTuning &c : cands) {
            const double t = measure(c, fams, srd);
            if (t >= 0.0 && t < bestT) {
                bestT = t;
                pick = c;
            }
        }
        if (base >= 0.0 && bestT > base * (1.0 - kKeepDefaultWithin)) pick = best;
        best = pick;
        nwSetTuning(&best);
    };

    best = NWTuning{};
    {
        std::vector<NWTuning> c;
        for (uint32_t fb : {64u, 128u, 512u, 1024u}) {
            NWTuning t = best;
            t.frontBlock = fb;
            c.push_back(t);
        }
        search(c, {"md_front"}, false);
    }
    {
        std::vector<NWTuning> c;
        for (uint32_t items : {1024u, 2048u, 4096u, 8192u})
            for (uint32_t mem : {256u, 512u, 1024u, 1536u, 2048u}) {
                NWTuning t = best;
                t.mdTileItems = items;
                t.mdMembers = mem;
                nwSetTuning(&t);
                if (nwGetLastError() == 0 && plan_is(N, C, n, items)) c.push_back(t);
            }
        nwSetTuning(&best);
        search(c, {"md_tiles", "md_sort"}, false);
    }
    {
        // the pair workgroup and its stage (0: pairBlock + 128 per row)
        std::vector<NWTuning> c;
        for (uint32_t pb : {128u, 256u, 512u})
            for (uint32_t stage : {0u, 2u * pb}) {
                NWTuning t = best;
                t.pairBlock = pb;
                t.pairStage = stage;
                c.push_back(t);
            }
        search(c, {"md_collide"}, false);
    }
    if (numSolvent) {
        std::vector<NWTuning> c;
        for (uint32_t items : {4096u, 8192u, 16384u})
            for (uint32_t mem : {768u, 1536u, 2048u}) {
                NWTuning t = best;
                t.srdTileItems = items;
                t.srdMembers = mem;
                nwSetTuning(&t);
                if (nwGetLastError() == 0 && plan_is(NT, Cs, 0, items)) c.push_back(t);
            }
        nwSetTuning(&best);
        search(c, {"srdf_tiles", "srdf_cells"}, true);
    }
    (void)hipStreamSynchronize(st);
    nwRngSetSeed(seed, offset);
    if (nwGetLastError() != 0) {
        nwSetTuning(&saved);
        return 0;
    }
    nwSetTuning(&best);
    if (!file.empty()) cache_store(file, dev, size, best);
    if (out) *out = best;
    return 1;
}

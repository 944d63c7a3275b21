// main.cpp -- command-line driver `nw`, restating simulation/main.cpp:22-168:
// same flags (-name=value, names matched case-insensitively), defaults, step
// loop and .xyzv output.  Throughput is printed per elapsed iteration (the
// reference divides by the total iteration count, main.cpp:48).  Extra flags:
// -seed=N, -wca (WCA excluded volume + chain bonds + drive), -drive=F,
// -solventSort=K (keep the SRD solvent in cell order on the device, re-sorted
// every K collisions; frames stay in original particle order).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <strings.h>

#include "particle_system.h"

namespace {

const char *strip_dashes(const char *a) {
    while (*a == '-') ++a;
    return a;
}

// flag present with exactly this name (before any '=')
bool has_flag(int argc, char **argv, const char *name) {
    const size_t len = strlen(name);
    for (int i = 1; i < argc; ++i) {
        const char *a = strip_dashes(argv[i]);
        const char *eq = strchr(a, '=');
        const size_t alen = eq ? size_t(eq - a) : strlen(a);
        if (alen == len && strncasecmp(a, name, len) == 0) return true;
    }
    return false;
}

// value of the last argument whose name starts with `name` ("-name=value" or "-namevalue")
const char *flag_value(int argc, char **argv, const char *name) {
    const size_t len = strlen(name);
    const char *v = nullptr;
    for (int i = 1; i < argc; ++i) {
        const char *a = strip_dashes(argv[i]);
        if (strncasecmp(a, name, len) == 0) v = a[len] == '=' ? a + len + 1 : a + len;
    }
    return v;
}

int flag_int(int argc, char **argv, const char *name) {
    const char *v = flag_value(argc, argv, name);
    return v ? atoi(v) : 0;
}

float flag_float(int argc, char **argv, const char *name) {
    const char *v = flag_value(argc, argv, name);
    return v ? float(atof(v)) : 0.0f;
}

}  // namespace

int main(int argc, char *argv[]) {
    cudaInit(argc, argv);

    // simulation/main.cpp:22-40
    std::string outFile = "nw.xyzv";
    uint numFilaments = 6000, filamentSize = 35;
    float timestep = 0.002f;
    int printRate = 5000, iterations = 5000000;
    float kBend = 200.f, kBond = 57.146436f, hardness = kBond, activity = 0.01f, reverse = 0.0005f, kbT = 2.0f,
          drag = 0.01f;
    BoundaryType boundX = PERIODIC, boundY = WALL_NO_SLIP, boundZ = WALL;
    nw_uint3 gridSize{1024, 128, 1};
    bool srdSolvent = false;
    uint64_t seed = 0;

    // simulation/main.cpp:54-119
    if (has_flag(argc, argv, "iterations")) iterations = flag_int(argc, argv, "iterations");
    if (has_flag(argc, argv, "numFilaments")) numFilaments = uint(flag_int(argc, argv, "numFilaments"));
    if (has_flag(argc, argv, "filamentSize")) filamentSize = uint(flag_int(argc, argv, "filamentSize"));
    if (has_flag(argc, argv, "dt")) timestep = flag_float(argc, argv, "dt");
    if (has_flag(argc, argv, "printRate")) printRate = int(flag_float(argc, argv, "printRate"));
    if (has_flag(argc, argv, "gamma")) drag = flag_float(argc, argv, "gamma");
    if (has_flag(argc, argv, "kbT")) kbT = flag_float(argc, argv, "kbT");
    if (has_flag(argc, argv, "kBend")) kBend = flag_float(argc, argv, "kBend");
    if (has_flag(argc, argv, "kBond")) kBond = flag_float(argc, argv, "kBond");
    if (has_flag(argc, argv, "activity")) activity = flag_float(argc, argv, "activity");
    if (has_flag(argc, argv, "reverse")) reverse = flag_float(argc, argv, "reverse");
    if (has_flag(argc, argv, "hardness")) hardness = flag_float(argc, argv, "hardness");
    if (has_flag(argc, argv, "gridX")) gridSize.x = uint(flag_int(argc, argv, "gridX"));
    if (has_flag(argc, argv, "gridY")) gridSize.y = uint(flag_int(argc, argv, "gridY"));
    if (has_flag(argc, argv, "gridZ")) gridSize.z = uint(flag_int(argc, argv, "gridZ"));
    if (has_flag(argc, argv, "boundX")) boundX = BoundaryType(flag_int(argc, argv, "boundX"));
    if (has_flag(argc, argv, "boundY")) boundY = BoundaryType(flag_int(argc, argv, "boundY"));
    if (has_flag(argc, argv, "boundZ")) boundZ = BoundaryType(flag_int(argc, argv, "boundZ"));
    if (has_flag(argc, argv, "srd")) srdSolvent = true;
    if (has_flag(argc, argv, "o")) {
        const char *v = flag_value(argc, argv, "o");
        if (v && *v) outFile = v;
    }
    if (has_flag(argc, argv, "seed")) seed = uint64_t(atoll(flag_value(argc, argv, "seed")));
    if (printRate <= 0) printRate = 1;

    NWSystemConfig cfg{};
    cfg.numFilaments = numFilaments;
    cfg.filamentSize = filamentSize;
    cfg.gridSize = gridSize;
    cfg.srd = srdSolvent;
    cfg.seed = seed;
    if (has_flag(argc, argv, "solventSort")) cfg.solventSortEvery = uint32_t(flag_int(argc, argv, "solventSort"));
    if (has_flag(argc, argv, "wca")) {
        cfg.ext.pairPotential = NW_PAIR_WCA;
        cfg.ext.bondedMode = NW_BONDED_CHAIN;
        cfg.ext.wcaEpsilon = 1.0f;
        cfg.ext.wcaSigma = 1.0f;
        cfg.ext.wcaCutoff = 1.12246204830937f;
        cfg.ext.driveForce = has_flag(argc, argv, "drive") ? flag_float(argc, argv, "drive") : 1.0f;
    }

    ParticleSystem *psystem = new ParticleSystem(cfg);
    psystem->setBoundaries(boundX, boundY, boundZ);
    psystem->setDamping(drag);
    psystem->setTemperature(kbT);
    psystem->setBondBendingConstant(kBend);
    psystem->setBondSpringConstant(kBond);
    psystem->setActivity(activity);
    psystem->setReverseProbability(reverse);
    psystem->setCollideSpring(hardness);

    psystem->reset();
    psystem->writeOutputs(outFile, 0, timestep);

    threadSync();
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 1; i <= iterations; ++i) {
        psystem->update(timestep);
        if (i % printRate == 0) {
            psystem->writeOutputs(outFile, i, timestep);
            const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            const double perIter = s / i;
            printf("[%i / %i] Throughput = %.4f KParticles/s, Time = %.5f s, Size = %u particles, "
                   "NumDevsUsed = %u, Workgroup = %u\n",
                   i, iterations, 1.0e-3 * psystem->getNumParticles() / perIter, perIter,
                   psystem->getNumParticles(), 1u, 0u);
            fflush(stdout);
        }
    }
    threadSync();
    delete psystem;
    return 0;
}

/*
 * nw_abi.h -- the C-ABI of libnwamd.so.
 *
 * Part 1 reproduces, symbol for symbol and argument for argument, the extern
 * "C" surface that the reference's host (simulation/particleSystem.cpp,
 * simulation/main.cpp) links against: simulation/particleSystem.cuh:12-112,
 * implemented upstream in simulation/particleSystem_cuda.cu:36-371.  A host
 * built against the reference header links against this library unchanged.
 * Every pointer is a device pointer obtained from allocateArray(); float*
 * arrays are float4 records (pos.w = radius, vel.w = mass).
 *
 * Part 2 is additive (prefix nw): RNG seeding, error reporting, stream
 * selection, the extension parameter block, kernel timing and sorting with an
 * explicit key bound.
 *
 * Error convention (reference: common/helper_cuda.h:566-598): by default a
 * failing call prints "file:line: message" and calls exit(EXIT_FAILURE).  After
 * nwSetErrorMode(NW_ERRORS_RETURN) a failing call records the error (read with
 * nwGetLastError / nwGetLastErrorString) and returns without launching.
 *
 * Threading: like the reference, the library keeps process-global state (the
 * parameter blocks, the RNG generator, one stream) and is not re-entrant.
 */
#ifndef NW_ABI_H
#define NW_ABI_H

#include <stddef.h>
#include <stdint.h>
#include "nw_params.h"

#ifdef __cplusplus
extern "C" {
#endif

/* opaque; the argument is accepted and ignored as in the reference
   (simulation/particleSystem_cuda.cu:86-90) */
struct cudaGraphicsResource;

/* ======================= Part 1: reference surface ======================= */

/* particleSystem.cuh:14 / particleSystem_cuda.cu:40-54.  Selects the device
   (-device=N on the command line, else device 0), creates the library stream
   and seeds the Philox generator with its default seed. */
void cudaInit(int argc, char **argv);

/* particleSystem.cuh:15 / particleSystem_cuda.cu:56-59.  n uniforms in (0, 1]
   from the global counter-based generator; the generator advances. */
void randomizeUniform(float *device, uint n);

/* particleSystem.cuh:16 / particleSystem_cuda.cu:61-64.  n normals N(mean, stddev). */
void randomizeGaussian(float *device, uint n, float mean, float stddev);

/* particleSystem.cuh:18 declares int, particleSystem_cuda.cu:66 defines size_t;
   size_t is used here, as in the reference's definition.  A host compiled
   against the reference header passes an int in the 64-bit size register,
   whose upper half the x86-64 SysV ABI leaves unspecified -- the same
   mismatch the reference has between its header and its definition.  GCC
   and Clang form an int argument with 32-bit instructions (which clear the
   upper half) when it comes from int arithmetic, as every size in
   particleSystem.cpp does; an int narrowed from a 64-bit value may keep the
   old upper half, and the allocation is then larger than asked (never
   smaller) or fails with an error.  tests/test_gpu_int_callers.py calls
   this entry point from C through the reference's int prototype. */
void allocateArray(void **devPtr, size_t size);
/* particleSystem.cuh:19 / particleSystem_cuda.cu:71-74 */
void freeArray(void *devPtr);
/* particleSystem.cuh:21 / particleSystem_cuda.cu:76-79: waits for all device work */
void threadSync(void);

/* particleSystem.cuh:23 / particleSystem_cuda.cu:86-90 */
void copyArrayFromDevice(void *host, const void *device,
                         struct cudaGraphicsResource **cuda_vbo_resource, int size);
/* particleSystem.cuh:24 / particleSystem_cuda.cu:81-84 (offset in bytes) */
void copyArrayToDevice(void *device, const void *host, int offset, int size);

/* particleSystem.cuh:26-27 / particleSystem_cuda.cu:92-102.  The blocks are
   copied by value into library state and travel to the kernels as kernel
   arguments (no constant-memory upload, no host synchronisation). */
void setParameters(SimParams *hostParams);
void setSolventParameters(SolventParams *solventParams);

/* particleSystem.cuh:29-34 / particleSystem_cuda.cu:117-133
   (functor particles_kernel_impl.cuh:71-142) */
void integrateFilaments(float *pos, float *vel, float *f, float *f_old,
                        float deltaTime, uint numParticles);

/* particleSystem.cuh:36-40 / particleSystem_cuda.cu:136-149
   (functor particles_kernel_impl.cuh:145-205) */
void integrateSolvent(float *pos, float *vel, float deltaTime,
                      uint startIndex, uint numSolvent);

/* particleSystem.cuh:42-46 / particleSystem_cuda.cu:151-169
   (kernel particles_kernel_impl.cuh:312-342).  srd=false: MD grid of
   SimParams; srd=true: shifted SRD grid of SolventParams. */
#ifdef __cplusplus
void calcHash(uint *gridParticleHash, uint *gridParticleIndex, float *pos,
              int numParticles, bool srd = false);
#else
void calcHash(uint *gridParticleHash, uint *gridParticleIndex, float *pos,
              int numParticles, nw_bool srd);
#endif

/* particleSystem.cuh:48-59 / particleSystem_cuda.cu:171-205
   (kernel particles_kernel_impl.cuh:346-424).  cellStart[0..numCells) is set
   to 0xffffffff first; cellEnd of empty cells is left untouched.  vel/sortedVel
   and tangent/sortedTangent may be NULL. */
void reorderDataAndFindCellStart(uint *cellStart, uint *cellEnd,
                                 float *sortedPos, float *sortedVel, float *sortedTangent,
                                 uint *gridParticleHash, uint *gridParticleIndex,
                                 float *pos, float *vel, float *tangent,
                                 uint numParticles, uint numCells);

/* particleSystem.cuh:61-69 / particleSystem_cuda.cu:207-235
   (kernels particles_kernel_impl.cuh:427-556, 694-744).  Adds the pair forces
   of every bead to force[gridParticleIndex[i]]. */
void collideFilaments(float *force, float *sortedPos, float *sortedVel, float *sortedTangent,
                      uint *gridParticleIndex, uint *cellStart, uint *cellEnd,
                      uint numParticles, uint numCells);

/* particleSystem.cuh:71-76 / particleSystem_cuda.cu:237-256
   (kernel particles_kernel_impl.cuh:640-691).  Draws one uniform per SRD cell
   from the global generator (in registers; `random` is not written) and
   rotates v - u_cell by +-alpha. */
void collideSolvent(float *pos, float *vel, float *solventCellCOM, float *random,
                    uint numSolventCells, uint numParticles);

/* particleSystem.cuh:78-81 / particleSystem_cuda.cu:258-274
   (kernel particles_kernel_impl.cuh:751-818).  numFilaments is a float in the
   reference signature and stays one here. */
void filamentForces(float *force, float *tangent, float *pos, float numFilaments);

/* particleSystem.cuh:83 / particleSystem_cuda.cu:277-282.  Stable ascending
   sort of (hash -> index) pairs, in place. */
void sortParticles(uint *dGridParticleHash, uint *dGridParticleIndex, uint numParticles);

/* particleSystem.cuh:85-90 / particleSystem_cuda.cu:284-306
   (kernel particles_kernel_impl.cuh:820-857).  One uniform per worm from the
   global generator (in registers; `uniform` is not written). */
void reverseFilaments(float *force, float *forceOld, float *vel, float *pos,
                      float *uniform, uint numFilaments);

/* particleSystem.cuh:92-97 / particleSystem_cuda.cu:308-328
   (kernel particles_kernel_impl.cuh:859-874).  f += -gamma v + xi with
   xi ~ N(0, 2 gamma kbT) per component, drawn in registers from the global
   generator exactly as randomizeGaussian(random, 4*numParticles, 0, sigma)
   would have written them; `random` is not written. */
void langevinThermostat(float *force, float *vel, float *random,
                        float gamma, float kbT, uint numParticles);

/* particleSystem.cuh:99-104 / particleSystem_cuda.cu:331-349
   (kernel particles_kernel_impl.cuh:559-594).  The misspelling is the ABI. */
void solventCellCentorOfMomentum(float *solventCellCOM, float *sortedVel,
                                 uint *cellStart, uint *cellEnd,
                                 uint *gridParticleIndex, uint numCells);

/* particleSystem.cuh:106-111 / particleSystem_cuda.cu:351-369
   (kernel particles_kernel_impl.cuh:597-637) */
void solventIsokineticThermostat(float *vel, float *sortedVel,
                                 uint *cellStart, uint *cellEnd,
                                 uint *gridParticleIndex, uint numCells);

/* particleSystem_cuda.cu:105-115 (inside the reference's extern "C" block).
   computeGridSize takes C++ references upstream; pointers are the same ABI. */
uint iDivUp(uint a, uint b);
void computeGridSize(uint n, uint blockSize, uint *numBlocks, uint *numThreads);

/* The names used for this path in the original particles sample, kept as
   aliases: integrateSystem == integrateFilaments, collide == collideFilaments. */
void integrateSystem(float *pos, float *vel, float *f, float *f_old,
                     float deltaTime, uint numParticles);
void collide(float *force, float *sortedPos, float *sortedVel, float *sortedTangent,
             uint *gridParticleIndex, uint *cellStart, uint *cellEnd,
             uint numParticles, uint numCells);

/* ======================= Part 2: additions ======================= */

enum { NW_ERRORS_EXIT = 0, NW_ERRORS_RETURN = 1 };

/* Error mode (NW_ERRORS_EXIT default, like helper_cuda.h). */
void nwSetErrorMode(int mode);
/* 0 if no error since the last call; clears the recorded error. */
int nwGetLastError(void);
const char *nwGetLastErrorString(void);
/* 1 if a HIP device is usable (initialises the library on first success). */
int nwDeviceAvailable(void);

/* Philox4x32-10 generator state: key = seed, counter offset in 128-bit blocks. */
void nwRngSetSeed(uint64_t seed, uint64_t offset);
void nwRngGetState(uint64_t *seed, uint64_t *offset);

/* Host evaluation of the generator: block b of key `seed` (the four 32-bit
   words every device draw is built from).  For host-side draws of the driver
   (initial conditions, the SRD grid shift). */
void nwPhiloxBlock(uint64_t seed, uint64_t block, uint32_t out[4]);

/* Extension parameters (physics variants); zero-filled block = reference. */
void nwSetExtParameters(const NWExtParams *ext);
void nwGetExtParameters(NWExtParams *ext);

/* Library stream: a hipStream_t created with default flags and priority by
   cudaInit.  nwSetStream replaces it (NULL = legacy null stream). */
void *nwGetStream(void);
void nwSetStream(void *stream);

/* Stable (hash -> index) sort when every key is known to be < keyBound
   (counting sort with keyBound bins).  Keys >= keyBound are still sorted
   correctly, on a slow path, and set nwGetDeviceFlags() bit 0. */
void nwSortParticlesBounded(uint *dGridParticleHash, uint *dGridParticleIndex,
                            uint numParticles, uint keyBound);
/* Stable LSD radix sort over the low `keyBits` bits of the keys (1..32);
   keys must fit in keyBits. */
void nwRadixSortPairs(uint *dGridParticleHash, uint *dGridParticleIndex,
                      uint numParticles, uint keyBits);

/* Fused MD step: the computation of integrateFilaments, reverseFilaments (when
   SimParams.reverseProbability > 0), langevinThermostat (when `langevin`),
   filamentForces, calcHash, sortParticles, reorderDataAndFindCellStart and
   collideFilaments -- in the order ParticleSystem::update issues them
   (simulation/particleSystem.cpp:257-330) -- in four launches, for worms of
   1..64 beads.  No counterpart in the reference.  Draws advance the generator
   exactly as the separate calls do.  On return gridParticleIndex holds the
   bead ids in stable cell order, sortedPos (and sortedVel / sortedTangent,
   which may be NULL when the WCA pair term without activity does not read
   them) the gathered records, gridParticleHash (may be NULL) the cell of
   every bead in bead order; the reference's cellStart/cellEnd tables are not
   produced.  tangent may be NULL when the pair term reads no tangents (the
   WCA term without activity): the step then stores none (16 B per bead
   less), and nwFilamentTangents gives them later, bit for bit.  Fails (NW error -22) when the grid has more than ~67M cells or
   the beads x cells product exceeds the sort's tile x bucket table
   (nwFusedPlanFits). */
void nwFusedMdStep(float *pos, float *vel, float *force, float *forceOld, float *tangent,
                   uint *gridParticleHash, uint *gridParticleIndex,
                   float *sortedPos, float *sortedVel, float *sortedTangent,
                   float deltaTime, float langevinGamma, float langevinKbT, int langevin,
                   uint numFilaments);

/* nwFusedMdStep with the force arrays double-buffered (no counterpart in
   the reference): the new force is written to forceOld's buffer and force's
   buffer, left untouched, is the new old force -- the caller swaps the two
   pointers after the call, and the step reads and writes 16 B per bead less.
   Same results as nwFusedMdStep after the swap when force.w equals
   forceOld.w (both zero from a cleared start); no polarity reversal (it
   permutes the old force). */
void nwFusedMdStepSwapForces(float *pos, float *vel, float *force, float *forceOld, float *tangent,
                             uint *gridParticleHash, uint *gridParticleIndex, float *sortedPos, float *sortedVel,
                             float *sortedTangent, float deltaTime, float langevinGamma, float langevinKbT,
                             int langevin, uint numFilaments);

/* The bead tangents filamentForces writes (particles_kernel_impl.cuh:751-818:
   unit(r[p+1] - r[p-1]), one-sided at the worm ends, none for worms of
   fewer than 3 beads, no tail tangent for 3) from the positions alone,
   without forces: equal bit for bit to what a fused step stores, for a
   caller that passed no tangent array to its fused steps.  No counterpart
   in the reference.  Worms of 1..64 beads (SimParams.filamentSize). */
void nwFilamentTangents(float *tangent, const float *pos, uint numFilaments);

/* nwFusedMdStep for callers that cannot allocate device memory -- code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
   No counterpart in the reference.  Same arguments and results as
   nwFusedMdStep (bit for bit), but the cell sort's scratch (the cell of
   every bead when gridParticleHash is NULL, the sort pairs, the tile x
   bucket prefix matrix, the cell of every sorted slot, the CSR offsets)
   lives in a fixed 32 MiB device array of the library, one per device, held
   until the process ends; the call allocates nothing and records no timing
   events.  Systems whose scratch exceeds the array -- about 1.5 million
   beads at one bead per cell -- fail with NW error -22 and a message giving
   the bytes needed, before any launch.  Every call on a device shares that
   device's array, so calls that overlap on one device (two streams, two
   threads) race; a call fully overwrites what it reads, so nothing carries
   over from the previous call. */
void nwFusedMdStepNoAlloc(float *pos, float *vel, float *force, float *forceOld, float *tangent,
                          uint *gridParticleHash, uint *gridParticleIndex, float *sortedPos, float *sortedVel,
                          float *sortedTangent, float deltaTime, float langevinGamma, float langevinKbT, int langevin,
                          uint numFilaments);

/* Fused SRD collision step: calcHash(srd) + sortParticles +
   [solventIsokineticThermostat] + solventCellCentorOfMomentum + collideSolvent
   + integrateSolvent (simulation/particleSystem.cpp:352-410) in three launches
   over the numParticles beads followed by numSolvent solvent particles, with
   the SolventParams of the last setSolventParameters.  No counterpart in the
   reference.  solventCellCOM receives (centre-of-mass velocity, rotation sign)
   per SRD cell, gridParticleIndex the particle ids in stable cell order,
   gridParticleHash (may be NULL) the SRD cell of every particle in particle
   order. */
void nwFusedSrdStep(float *pos, float *vel, float *solventCellCOM, uint *gridParticleHash,
                    uint *gridParticleIndex, float deltaTime, uint numParticles, uint numSolvent,
                    int isokinetic);

/* 1 when the fused steps' cell sort can plan numParticles particles over
   numCells cells (filamentSize 1..64: the MD step's tiles of whole worms;
   0: the SRD step), else 0.  No device work. */
int nwFusedPlanFits(uint numParticles, uint numCells, uint filamentSize);
/* The plan of that cell sort (no counterpart in the reference; no device
   work): out[0] log2 of the cells per bucket, out[1] buckets, out[2]
   particles per first-level tile, out[3] tiles.  Returns what
   nwFusedPlanFits returns. */
int nwFusedPlanInfo(uint numParticles, uint numCells, uint filamentSize, uint out[4]);

/* ------------------------------------------------------------------------
   Launch shapes and the autotuner (no counterpart in the reference, which
   launches every kernel with computeGridSize(n, 256),
   simulation/particleSystem_cuda.cu:111-115).  A zero field takes the
   built-in choice.  Every setting gives the same results bit for bit --
   the sort is stable, every pair sum runs in stencil order -- only the
   time differs.
   ------------------------------------------------------------------------ */
typedef struct NWTuning {
    uint32_t mdTileItems;   /* MD cell sort, first level: beads per tile (1024, 2048, 4096, 8192, 16384) */
    uint32_t mdMembers;     /* MD cell sort, second level: bucket population aimed at (64..2048) */
    uint32_t srdTileItems;  /* SRD cell sort, first level: particles per tile (4096, 8192, 16384) */
    uint32_t srdMembers;    /* SRD cell sort, second level (64..2048) */
    uint32_t pairBlock;     /* sorted beads per pair-kernel workgroup (128, 256; 512 for the WCA term) */
    uint32_t frontBlock;    /* threads per workgroup of the fused front (64, 128, 256, 512, 1024) */
    uint32_t pairStage;     /* beads staged in LDS per neighbour row of a pair workgroup, on average (the
                               three rows share one pool): pairBlock + 128 or 2 pairBlock where that fits
                               the workgroup's LDS (0: pairBlock + 128) */
    uint32_t reserved[9];
} NWTuning;
/* Applies to the fused single-system steps (nwFusedMdStep and its variants,
   nwFusedSrdStep) and to nwFusedPlanInfo; values outside the lists above
   fail with NW error -22 and leave the tuning unchanged.  A tile / bucket
   pair whose plan does not fit a system falls back to the built-in plan. */
void nwSetTuning(const NWTuning *tuning);
void nwGetTuning(NWTuning *tuning);

/* The autotuner: times the candidate launch shapes of nwFusedMdStep on a
   copy of the given state (pos / vel hold the numFilaments worms' beads,
   followed by numSolvent solvent particles) -- the front's workgroup size,
   the MD sort's tile and bucket sizes jointly, the pair kernel's workgroup
   size, each over `reps` steps with HIP events on the kernels' dispatch
   packets -- and, when numSolvent > 0, those of nwFusedSrdStep (the SRD
   sort's tile and bucket sizes) with the SolventParams of the last
   setSolventParameters.  The fastest choice of each is set (nwSetTuning)
   and written to out.  The state itself is not touched; the generator is
   left where it was.  cachePath (NULL: $NW_TUNE_CACHE, else
   $HOME/.cache/nwamd/tuning.txt; "": no cache) holds one line per device
   and system size (beads, MD cells, beads per worm, solvent, SRD cells):
   with useCache a recorded line is applied without timing.  Returns 2 when
   the choice came from the cache, 1 when it was timed, 0 on an error. */
int nwAutotuneFused(float *pos, float *vel, float *force, float *forceOld, float *tangent, uint numFilaments,
                    uint numSolvent, float deltaTime, float langevinGamma, float langevinKbT, int langevin,
                    float srdDeltaTime, int isokinetic, uint reps, const char *cachePath, int useCache,
                    NWTuning *out);

/* ------------------------------------------------------------------------
   Batched replicas: many independent worm systems of uneven size stepped by
   one launch of each fused kernel per step (no counterpart in the
   reference).  Replica r's beads occupy [beadOffset[r], beadOffset[r+1]) of
   the concatenated float4 arrays, its cells [cellOffset[r], ...) of one
   concatenated cell space; each keeps its own SimParams, NWExtParams and
   Philox stream, so a replica's trajectory is the one nwFusedMdStep gives it
   alone (same seed and offset), bit for bit.
   ------------------------------------------------------------------------ */
typedef struct NWReplicaParams {
    SimParams params;     /* box, grid (one layer, gridSize.x >= 3), worms, force constants */
    NWExtParams ext;      /* pair potential and stencil must agree over the batch */
    uint64_t seed;        /* the replica's Philox key */
    uint64_t offset;      /* its generator offset before the first batched step */
    float langevinGamma;  /* noise sigma = sqrt(2 gamma kT), as nwFusedMdStep */
    float langevinKbT;
    int langevin;
    uint32_t reserved[5];
} NWReplicaParams;

/* Replaces the batch (one per process) and resets its step count. */
void nwBatchSetReplicas(const NWReplicaParams *reps, uint numReplicas);
/* numReplicas + 1 bead and cell offsets of the current batch. */
void nwBatchLayout(uint *beadOffset, uint *cellOffset);
/* `steps` fused MD steps of every replica over the concatenated arrays
   (sortedVel / sortedTangent may be NULL as in nwFusedMdStep).  useGraph:
   after one direct step, the step is captured once as a hipGraph (per
   argument set) and replayed. */
void nwBatchMdStep(float *pos, float *vel, float *force, float *forceOld, float *tangent,
                   uint *gridParticleIndex, float *sortedPos, float *sortedVel, float *sortedTangent,
                   float deltaTime, uint steps, int useGraph);
/* Generator offset of one replica after the batched steps so far. */
void nwBatchRngState(uint replica, uint64_t *offset);

/* SRD solvent of the batch (no counterpart in the reference): replica r's
   numSolvent solvent particles follow all beads of the batch, in replica
   order, and its SRD cells follow the previous replicas' cells (offsets:
   nwBatchSolventLayout).  solvent.randOffset is not read (each collision
   brings its grid shifts).  One rotation angle and one srdCollideFilaments
   over the batch.  Replaced by the next nwBatchSetReplicas. */
typedef struct NWReplicaSolvent {
    SolventParams solvent;  /* the replica's SRD grid, cell size, alpha */
    uint numSolvent;
    uint reserved[3];
} NWReplicaSolvent;
void nwBatchSetSolvent(const NWReplicaSolvent *solvent, uint numReplicas);
/* numReplicas + 1 entries each: first solvent particle (counted from the
   first bead) and first SRD cell of every replica. */
void nwBatchSolventLayout(uint *solventOffset, uint *srdCellOffset);

/* ------------------------------------------------------------------------
   Differentiation through the fused MD step (no counterpart in the
   reference): reverse-mode gradients of a loss of the final configuration
   with respect to theta = (driveForce, bondBendingK, wcaEpsilon, gamma) --
   the gamma of both the drag and the noise amplitude sqrt(2 gamma kT).
   Dry worms: WCA pairs, chain bonds, bending, drive, Langevin; periodic box,
   one-layer stencil, no extensile term, no reversal.
   ------------------------------------------------------------------------ */
/* `steps` nwFusedMdStep steps from (pos, vel, force, forceOld), in place,
   with the SimParams / NWExtParams of the last setParameters /
   nwSetExtParameters; the state and generator offset are checkpointed every
   checkpointEvery steps for nwAdjointBackward.  langevinGamma must equal
   SimParams.gamma when langevin is on. */
void nwAdjointForward(float *pos, float *vel, float *force, float *forceOld, float *tangent, uint numFilaments,
                      float deltaTime, float langevinGamma, float langevinKbT, int langevin, uint steps,
                      uint checkpointEvery);
/* nwAdjointForward with SRD collisions (no counterpart in the reference):
   pos / vel hold the numFilaments worms' beads followed by numSolvent solvent
   particles, as nwFusedSrdStep; after steps 0, srdGap, 2 srdGap, ... of the
   run the fused SRD collision (time step deltaTime * srdGap, the
   SolventParams of the last setSolventParameters with the grid shift
   srdShifts[3c .. 3c+2] for collision c, ceil(steps / srdGap) of them)
   follows the MD step, as ParticleSystem::update orders them
   (simulation/particleSystem.cpp:249-339).  No Langevin thermostat (the
   reference applies it only without solvent, particleSystem.cpp:276); no
   isokinetic rescaling.  nwAdjointBackward then differentiates through the
   collisions with their cells and rotation signs held fixed. */
void nwAdjointForwardSrd(float *pos, float *vel, float *force, float *forceOld, float *tangent, uint numFilaments,
                         uint numSolvent, float deltaTime, uint srdGap, const float *srdShifts, uint steps,
                         uint checkpointEvery);
/* The adjoint of the last nwAdjointForward run, backwards from dLdPos /
   dLdVel (adjoints of the final positions / velocities, device float4 per
   bead; dLdVel may be NULL): every segment is recomputed from its checkpoint
   with the Philox noise replayed, then swept backwards.  grad[4] receives
   d loss / d theta.  The generator is left where the forward run left it.
   Synchronises. */
void nwAdjointBackward(const float *dLdPos, const float *dLdVel, double grad[4]);
/* One SRD collision of every replica of the batch (nwFusedSrdStep over all
   replicas in one launch per kernel; no counterpart in the reference): pos /
   vel hold all beads, then the solvent of nwBatchSetSolvent; shifts
   (host, 3 per replica) the replicas' grid shifts; replica r draws its
   rotation signs from its stream at offset + collision * ceil(cells / 4)
   (cells: its SRD cells), as nwFusedSrdStep does after nwRngSetSeed(seed,
   that offset).  solventCellCOM receives (u, sign) per SRD cell of the
   concatenated cell space, gridParticleHash the cell of every particle,
   gridParticleIndex the particles in cell order.  No isokinetic rescaling. */
void nwBatchSrdStep(float *pos, float *vel, float *solventCellCOM, uint *gridParticleHash, uint *gridParticleIndex,
                    float deltaTime, const float *shifts, uint collision);

/* Differentiation through batched steps (no counterpart in the reference):
   `steps` batched MD steps of the current batch (WCA chain worms, no
   extensile term, no reversal, periodic boxes) over pos / vel (all beads,
   then, with srdGap > 0, the solvent of nwBatchSetSolvent), force,
   forceOld, tangent (beads); with srdGap > 0, after steps 0, srdGap, ... the
   SRD collision of every replica (time step deltaTime * srdGap, grid shift
   srdShifts[(c * numReplicas + r) * 3 ..] for collision c of replica r, its
   signs drawn from the replica's stream at offset + c * ceil(cells / 4)):
   replica r's trajectory is the one nwAdjointForward / nwAdjointForwardSrd
   gives it alone from the same seed and offset, bit for bit.  Checkpoints
   every checkpointEvery steps.  With the solvent: no Langevin thermostat. */
void nwBatchAdjointForward(float *pos, float *vel, float *force, float *forceOld, float *tangent, float deltaTime,
                           uint srdGap, const float *srdShifts, uint steps, uint checkpointEvery);
/* The adjoint of the last nwBatchAdjointForward run from dLdPos / dLdVel
   (beads, device float4; dLdVel may be NULL; the solvent's adjoints start
   at zero): grad[4 r + k] = d loss / d theta_k of replica r's own
   (driveForce, bondBendingK, wcaEpsilon, gamma).  Collisions are held fixed
   (cells and signs).  Synchronises. */
void nwBatchAdjointBackward(const float *dLdPos, const float *dLdVel, double *grad);
/* Polar speed P[r] of every replica and, when dPdPos is not NULL, dP_r /
   dpos (and dPdVel: dP_r / dvel) of its beads.  Synchronises. */
void nwBatchPolarSpeedGrad(const float *pos, const float *vel, float *dPdPos, float *dPdVel, double *P);

/* Bond-tangent nematic order S of the worms (NWObservables.bondS) and, when
   dSdPos is not NULL, dS/dpos (device float4 per bead, w = 0).
   Synchronises. */
double nwBondOrderGrad(const float *pos, uint numFilaments, float *dSdPos);

/* Reorders the numSolvent solvent particles that follow the numParticles
   beads in pos / vel by their SRD cell (the grid of the last
   setSolventParameters, unshifted), stably; solventId (device, numSolvent;
   may be NULL) is permuted alike, so slot k keeps the original index of the
   particle now in it.  SRD collisions are invariant under the order up to
   the rounding of the cell sums; a cell-ordered solvent turns their random
   accesses into near-sequential ones.  No counterpart in the reference (a
   data-layout step of the host driver, include/nw_system.h). */
void nwSortSolvent(float *pos, float *vel, uint *solventId, uint numParticles, uint numSolvent);

/* Polar speed P = <v . t> of the worms (NWObservables.polarSpeed) and, when
   dPdPos is not NULL, dP/dpos and (dPdVel not NULL) dP/dvel (device float4
   per bead, w = 0).  Synchronises. */
double nwPolarSpeedGrad(const float *pos, const float *vel, uint numFilaments, float *dPdPos, float *dPdVel);

/* Order parameters of a worm configuration (the reference's analysis,
   analysis/analyze_correlations.py:13-38, analysis/afa/utils.py:35-49, on
   the device).  Bead tangent t = unit(r[p+1] - r[p-1]), one-sided at the worm
   ends, bond tangent b = unit(r[p+1] - r[p]), both under the minimum image.
   No counterpart in the reference's C-ABI. */
#define NW_OBS_SUMS 19
typedef struct NWObservables {
    double nematicS;    /* largest eigenvalue of <1.5 t t - 0.5 I> over bead tangents */
    double polarP;      /* |<t>| */
    double bondS;       /* largest eigenvalue of <1.5 b b - 0.5 I> over bond tangents */
    double polarSpeed;  /* <v . t>: mean speed along the tangent */
    double meanV2;      /* <|v|^2> */
    double beads;       /* beads with a tangent (worms of >= 2 beads) */
    double bonds;
    /* FP64 sums: t t (xx xy xz yy yz zz), t (x y z), beads, b b (6), bonds,
       v . t, |v|^2 */
    double sums[NW_OBS_SUMS];
} NWObservables;

/* Observables of numFilaments worms of SimParams.filamentSize beads (pos,
   vel: float4 per bead, worm-contiguous) with the box of the last
   setParameters.  Synchronises the library stream.  Deterministic: a fixed
   reduction order. */
void nwObservables(const float *pos, const float *vel, uint numFilaments, NWObservables *out);

/* Largest eigenvalue of the symmetric 3x3 matrix (a11 a12 a13 a22 a23 a33),
   FP64, closed form (the trigonometric solution of the characteristic
   cubic). */
double nwSymEig3Max(const double a[6]);

/* Device-side anomaly flags (bit 0: sort key above bound).  Synchronises. */
uint nwGetDeviceFlags(int clear);

/* Kernel timing with HIP events on the library stream while enabled:
   enable = 0 stops, enable = k >= 1 (re)starts timing every k-th launch of
   the named kernel family.  The per-call families record an event pair
   around their launches (a few microseconds of stream time each; sampling
   keeps that out of most steps); each fused family is one kernel, and its
   events ride on that kernel's dispatch packet (hipExtLaunchKernelGGL), so
   they time its execution alone.  Names of the per-call entry points:
   "integrate", "langevin", "filament", "hash", "sort", "reorder", "collide",
   "srd_hash", "srd_sort", "srd_reorder", "srd_com", "srd_thermo",
   "srd_collide", "srd_stream", "reverse"; of the fused steps: "md_front",
   "md_tiles" (the sort's first level), "md_sort", "md_collide",
   "srdf_tiles", "srdf_cells", "srdf_rotate"; of the batched step's kernels
   (launched directly, not from a replayed graph): "batch_front",
   "batch_tiles", "batch_sort", "batch_collide", and "batch_step" around a
   graph replay; of the adjoint sweeps: "adj_step", "badj_step"; the solvent
   re-sort nwSortSolvent: "srd_resort". */
void nwTimerEnable(const char *kernel, int enable);
/* Synchronises; returns total milliseconds and launch count since enable. */
void nwTimerRead(const char *kernel, double *totalMs, uint64_t *launches);

#ifdef __cplusplus
}
#endif
#endif

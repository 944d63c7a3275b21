"""Benchmark of the worm time step on one MI355X.

Workload (default "worms_srd_800k"): 40,000 worms x 20 beads = 800,000 beads in
a 2D periodic box L = 1000 with 5,000,000 SRD solvent particles (5 per unit
cell), WCA excluded volume, harmonic bonds k = 57 r0 = 0.8, bending kappa = 5,
drive f_a = 1 along the bond tangent, SRD collision every 10 steps (90 deg,
random sign per cell, grid shift in [-0.5, 0.5), beads included), dt = 1e-3.
A step is one ParticleSystem::update() through the C-ABI; inputs are resident
in HBM.  The system first relaxes from its lattice start for --relax-steps
untimed steps (the order parameters of the start and of the timed state are
reported).  value = beads x steps / wall time of the timed region, the
window charged its exact share of the solvent re-sort that the host driver
runs once every 20 collisions (the re-sorts the window held replaced by
steps / 200 re-sorts at the average re-sort time).

Also reported: per-kernel average durations (HIP events on the library
stream), the roofline of the dominant kernel (algorithmic bytes per launch /
average duration vs 8 TB/s) and the oracle's FP32 OpenMP step on the host
cores over a bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nematic-worms_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "bead-steps/s & %HBM roofline: plain vs batched; adjoint/forward cost"

# Algorithmic bytes per launch of each kernel family: every distinct input
# element read once and every output written once (float4 = 16 B, uint = 4 B).
# N beads, NT = beads + solvent, C = MD cells, Cs = SRD cells.
ALG_BYTES = {
    # per-call path (the reference's entry points)
    "integrate": lambda w: 128 * w["N"],                       # R pos vel f f_old, W the same
    "langevin": lambda w: 48 * w["N"],                         # R vel f, W f (noise in registers)
    "filament": lambda w: 64 * w["N"],                         # R pos, RMW force, W tangent
    "hash": lambda w: 24 * w["N"],                             # R pos, W hash index
    "sort": lambda w: 16 * w["N"] + 8 * w["C"],                # R/W keys+values, bin counts
    "reorder": lambda w: (40 if w["wca"] else 104) * w["N"] + 12 * w["C"],  # R hash index, gather pos (vel tan)
    "collide": lambda w: (52 if w["wca"] else 84) * w["N"] + 8 * w["C"],  # R sPos sIdx (+sVel sTan), RMW f
    "srd_hash": lambda w: 24 * w["NT"],
    "srd_sort": lambda w: 16 * w["NT"] + 8 * w["Cs"],
    "srd_reorder": lambda w: 72 * w["NT"] + 12 * w["Cs"],      # R hash index, gather pos vel
    "srd_com": lambda w: 16 * w["NT"] + 24 * w["Cs"],          # R sVel, cell bounds, W com
    "srd_thermo": lambda w: 68 * w["NT"] + 8 * w["Cs"],
    "srd_collide": lambda w: 48 * w["NT"] + 16 * w["Cs"],      # R pos, RMW vel, R com
    "srd_stream": lambda w: 64 * w["Ns"],                      # RMW pos vel of the solvent
    # fused path (nwFusedMdStep / nwFusedSrdStep); Pm / Ps = bytes of the
    # sort's tile x bucket prefix matrix (written by the tile pass, read once)
    # R pos vel f f_old, W pos vel f (+ f_old unless the host driver swaps
    # the force buffers: no reversal), W cell, W tangent when the pair term
    # reads tangents (the host driver otherwise rebuilds them on demand)
    "md_front": lambda w: ((132 if w["swap"] else 148) - (0 if w["tan_store"] else 16)) * w["N"],
    "md_tiles": lambda w: 8 * w["N"] + w["Pm"],                # R cell, W pair (4 B packed)
    "md_sort": lambda w: 44 * w["N"] + 4 * w["C"] + w["Pm"],   # R pairs, W ids cells, gather pos, W offsets
    "md_collide": lambda w: (56 if w["wca"] else 88) * w["N"] + 4 * w["C"],  # R sPos sIdx sCell (+sVel sTan), RMW f
    "srdf_tiles": lambda w: 24 * w["NT"] + w["Ps"],            # R pos, W cell, W pair
    "srdf_cells": lambda w: 24 * w["NT"] + 20 * w["Cs"] + w["Ps"],  # R pairs, W ids, gather vel, W offsets com
    # R cell, RMW vel (+pos), R com; beads only when rotated or rescaled
    "srdf_rotate": lambda w: 36 * w["Nrot"] + 32 * w["Ns"] + 16 * w["Cs"],
    # batched step over all replicas (WCA, no tangent gathers, swapped force
    # buffers): the fused kernels' bytes over the concatenated beads and cells
    "batch_front": lambda w: 132 * w["N"],
    "batch_tiles": lambda w: 8 * w["N"] + w["Pm"],
    "batch_sort": lambda w: 44 * w["N"] + 4 * w["C"] + w["Pm"],
    "batch_collide": lambda w: 56 * w["N"] + 4 * w["C"],
    # one adjoint step: R r', own lF, sorted r' and ids of the cell lists,
    # RMW lr lv lG, W lF of the previous force (+ R v' with the Langevin
    # thermostat), R the CSR offsets
    "adj_step": lambda w: (180 if w["lgv"] else 164) * w["N"] + 4 * w["C"],
    "badj_step": lambda w: (180 if w["lgv"] else 164) * w["N"] + 4 * w["C"],
}


# families that are one kernel (the roofline is quoted for a single kernel)
SINGLE_KERNEL = {"integrate", "langevin", "filament", "hash", "collide", "srd_com", "srd_thermo",
                 "srd_hash", "srd_collide", "srd_stream", "md_front", "md_tiles", "md_sort", "md_collide",
                 "srdf_tiles", "srdf_cells", "srdf_rotate"}


# kernel behind each single-kernel family of the fused path (the names that
# rocprofv3 reports), for the HBM traffic measured by its counters
FAMILY_KERNEL = {"md_sort": "k_bucket_cells<false, 256>", "srdf_cells": "k_bucket_cells<true, 256>",
                 "srdf_rotate": "k_srd_rotate"}


def dispatched_kernel(family, w):
    """The kernel this run dispatches for a single-kernel family of the fused
    path (the tile kernels follow the sort's plan of this run)."""
    tun = w.get("tuning") or {}
    if family == "md_front":
        return f"k_md_front<{tun.get('frontBlock') or 256}>"
    if family == "md_collide":
        # pair_stage: pairBlock + 128 beads per row run, or 2 pairBlock when pairStage asks for more
        pb = tun.get('pairBlock') or 256
        stage = 2 * pb if (tun.get('pairStage') or 0) > pb + 128 else pb + 128
        return f"k_collide_rows<1, false, true, {pb}, {stage}>"
    if family == "md_tiles":
        items = w["md_plan"][2]
        shapes = {1024: "4, 256", 2048: "8, 256", 4096: "4, 1024", 8192: "8, 1024", 16384: "16, 1024"}
        return f"k_tiles_staged<false, {shapes[items]}>" if items in shapes else None
    if family == "srdf_tiles":
        items = w["srd_plan"][2]
        shapes = {4096: "4, 1024", 8192: "16, 512", 16384: "16, 1024"}
        return f"k_tiles_staged<true, {shapes[items]}>" if items in shapes else None
    return FAMILY_KERNEL.get(family)


def measured_traffic(workload, family, w):
    """HBM bytes per dispatch of the family's kernel from the committed
    counter summary of this workload (profiles/<workload>_traffic.json,
    written by python -m nwamd.profsum traffic) -- taken from that file, not
    measured in this run -- and the file it came from; (None, None) when the
    file does not list the kernel this run dispatches."""
    path = os.path.join(ROOT, "profiles", f"{workload}_traffic.json")
    kern = dispatched_kernel(family, w)
    if not kern or not os.path.exists(path):
        return None, None
    b = json.load(open(path)).get("bytes_per_dispatch", {}).get(kern)
    if not b:
        return None, None
    return int(b), f"profiles/{workload}_traffic.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; committed)"


def build_workload(name, per_call=False, solvent_sort_every=0):
    from nwamd import abi, system
    if name == "worms_srd_800k":
        cfg = system.config_srd_800k(solvent_sort_every=solvent_sort_every)
    elif name == "worms_dry_4k":
        cfg = system.config_dry_4k()
    else:
        raise SystemExit(f"unknown workload {name}")
    cfg.perCallSequence = 1 if per_call else 0
    s = system.System(cfg)
    system.apply_worm_physics(s, kappa=5.0, gamma=2.0, kT=0.1)
    s.reset(system.NW_INIT_LATTICE_RANDOM if cfg.srd else system.NW_INIT_RANDOM, 0.1)
    p = s.params
    w = dict(N=s.n, Ns=s.ns, NT=s.n + s.ns, C=p.numCells, Cs=s.solvent.numCells if cfg.srd else 0,
             wca=cfg.ext.pairPotential == 1, swap=p.reverseProbability == 0.0,
             tan_store=cfg.ext.pairPotential != 1 or p.activity != 0.0)
    rot_beads = bool(cfg.ext.srdCollideFilaments) or cfg.ext.srdThermostat == abi.NW_SRD_THERMO_ISOKINETIC
    w["Nrot"] = w["NT"] if rot_beads else w["Ns"]
    w["md_plan"] = fused_plan(w["N"], w["C"], cfg.filamentSize)
    w["Pm"] = 4 * w["md_plan"][3] * (w["md_plan"][1] + 1)
    w["srd_plan"] = fused_plan(w["NT"], w["Cs"], 0) if cfg.srd else (0, 0, 0, 0)
    w["Ps"] = 4 * w["srd_plan"][3] * (w["srd_plan"][1] + 1)
    return s, cfg, w


def fused_plan(n, cells, filament_size):
    """(log2 cells per bucket, buckets, particles per tile, tiles) of the
    fused cell sort for n particles over `cells` (nwFusedPlanInfo)"""
    import ctypes
    from nwamd import abi
    out = (ctypes.c_uint * 4)()
    abi.load().nwFusedPlanInfo(n, cells, filament_size, out)
    return tuple(out)


def batched_form(args):
    """256 independent replicas (worms U[500, 8000] x U{10..30} beads, area
    fraction 0.3, f_a log-uniform [0.1, 10], kappa U[1, 20], Langevin gamma 2,
    kT 0.1, dt 1e-3): aggregate bead-steps/s of the batched step replayed as a
    hipGraph, and of the same replicas one launch sequence per replica."""
    from nwamd import replicas, system
    specs = replicas.replica_specs_config2(args.replicas)
    rs = replicas.ReplicaSet(specs, replicas.worm_physics(gamma=2.0, kT=0.1, fdt=True))
    rs.reset(system.NW_INIT_LATTICE_RANDOM, 0.1)
    N = rs.n()
    out = dict(workload=f"{args.replicas} replicas", replicas=args.replicas, beads=N)
    graph_form = replicas.NW_REPLICAS_LAUNCHES if args.batched_no_graph else replicas.NW_REPLICAS_GRAPH
    for key, form, steps in (("graph", graph_form, args.batched_steps),
                             ("one_per_launch", replicas.NW_REPLICAS_ONE_PER_LAUNCH, args.one_per_launch_steps)):
        if steps <= 0:
            continue
        rs.update(args.dt, max(2, steps // 10), form)  # warm-up (the graph is captured here)
        rs.sync()
        t0 = time.perf_counter()
        rs.update(args.dt, steps, form)
        rs.sync()
        el = time.perf_counter() - t0
        out[key] = dict(value=round(N * steps / el, 1), unit="bead-steps/s", steps=steps,
                        ms_per_step=round(el / steps * 1e3, 4))
    if "graph" in out and "one_per_launch" in out:
        out["speedup"] = round(out["graph"]["value"] / out["one_per_launch"]["value"], 2)
    # the kernels the graph replays, launched directly for a few steps with
    # events on their dispatch packets: per-kernel durations, the roofline of
    # the longest
    import ctypes

    import numpy as np
    from nwamd import abi
    lib = abi.load()
    fams = ["batch_front", "batch_tiles", "batch_sort", "batch_collide"]
    timers(lib, fams, True)
    rs.update(args.dt, 10, replicas.NW_REPLICAS_LAUNCHES)
    rs.sync()
    R = rs.R
    bo, co = np.zeros(R + 1, np.uint32), np.zeros(R + 1, np.uint32)
    lib.nwBatchLayout(bo.ctypes.data_as(ctypes.c_void_p), co.ctypes.data_as(ctypes.c_void_p))
    wb = dict(N=N, C=int(co[-1]), Pm=batch_matrix_bytes(bo, co))
    per = {f: abi.timer_read(lib, f) for f in fams}
    out["kernels"] = {f: dict(avg_us=round(ms / max(c, 1) * 1e3, 2), launches=c,
                              gbs=round(ALG_BYTES[f](wb) / (ms / max(c, 1) * 1e-3) / 1e9, 1) if c else None)
                      for f, (ms, c) in per.items()}
    dom = max(fams, key=lambda f: per[f][0])
    out["roofline"] = roofline_of(lib, dom, wb, 1)
    timers(lib, fams, False)
    o = rs.observables(0)
    out["replica0_S"] = round(o["nematicS"], 4)
    rs.close()
    return out


def adjoint_form(args):
    """Dry worms, 2,000 x 16 beads in a periodic 200 x 200 box (WCA + bonds
    k 57 r0 0.8 + bending kappa 5 + drive 1 + Langevin gamma 2 kT 0.1 with
    FDT noise, dt 1e-3): `steps` forward steps alone,
    then the same number differentiated -- forward with checkpoints every
    100 steps, loss = final bond-tangent nematic order S, adjoint sweep with
    recompute and replayed noise -- giving d S / d (f_a, kappa, epsilon, gamma).
    Starts from rows of straight worms (lattice start)."""
    import ctypes
    from nwamd import abi, system
    from nwamd.abi import DeviceArray
    lib = abi.load()
    c = system.NWSystemConfig()
    c.numFilaments, c.filamentSize = 2000, 16
    G = int(200.0 / 2.0 ** (1.0 / 6.0))
    c.gridSize = abi.UInt3(G, G, 1)
    c.boxSize = abi.Float3(200.0, 200.0, 200.0 / G)
    c.langevinFdt = 1
    c.seed = 4242
    c.ext = system.wca_ext(stencil=abi.NW_STENCIL_2D, srd_beads=0)
    s = system.System(c)
    system.apply_worm_physics(s, kappa=5.0, gamma=2.0, kT=0.1)
    # 32,000 beads in 200 x 200 fill 63 % of the area: rows of straight
    # worms with random polarity (rejection sampling would not finish)
    s.reset(system.NW_INIT_LATTICE_RANDOM, 0.1)
    N, W, steps, K, dt = s.n, c.numFilaments, args.adjoint_steps, 100, args.dt
    s.update(dt, 20)
    s.sync()
    t0 = time.perf_counter()
    s.update(dt, steps)
    s.sync()
    t_fwd = time.perf_counter() - t0
    p = s.params
    lib.setParameters(ctypes.byref(p))
    lib.nwSetExtParameters(ctypes.byref(c.ext))
    seed, off = s.rng_state()
    lib.nwRngSetSeed(seed, off)
    arr = [ctypes.c_void_p(s.device(k)) for k in (system.POSITION, system.VELOCITY, system.FORCE, system.FORCE_OLD,
                                                  system.TANGENT)]
    dS = DeviceArray(lib, 16 * N)
    g = (ctypes.c_double * 4)()
    lib.threadSync()
    lib.nwTimerEnable(b"adj_step", 10)
    t0 = time.perf_counter()
    lib.nwAdjointForward(*arr, W, dt, p.gamma, p.kbT / dt, 1, steps, K)
    S = lib.nwBondOrderGrad(arr[0], W, dS.p)
    t1 = time.perf_counter()
    lib.nwAdjointBackward(dS.p, None, g)
    t2 = time.perf_counter()
    roof = roofline_of(lib, "adj_step", dict(N=N, C=G * G, lgv=True), 10)
    lib.nwTimerEnable(b"adj_step", 0)
    s.close()
    return dict(workload="2000 worms x 16 beads, L 200, dry", beads=N, steps=steps, checkpoint_every=K,
                forward_ms=round(t_fwd * 1e3, 2), forward_value=round(N * steps / t_fwd, 1),
                checkpointed_forward_ms=round((t1 - t0) * 1e3, 2), backward_ms=round((t2 - t1) * 1e3, 2),
                adjoint_over_forward=round((t2 - t0) / t_fwd, 3), backward_over_forward=round((t2 - t1) / t_fwd, 3),
                value=round(N * steps / (t2 - t0), 1), unit="bead-steps/s (forward + adjoint)",
                loss_S=round(S, 6), grad=dict(drive=g[0], kappa=g[1], epsilon=g[2], gamma=g[3]), roofline=roof)


def batch_adjoint_form(args):
    """Batched replicas differentiated through SRD collisions: R replicas of
    1,000 worms x 12 beads in periodic 120 x 120 boxes (WCA + bonds k 57 r0
    0.8 + bending kappa 5 + drive f_a uniform in [0.5, 5] per replica, dt
    1e-3, no Langevin thermostat: the solvent takes its place), each with SRD
    solvent at 3 per unit cell (43,200 particles at uniform positions, ids in
    cell order, Maxwellian kT 0.1, mass 1; beads join the collisions), a collision every 10 steps (90 degrees,
    random sign per cell, random grid shift per replica and collision).
    `steps` batched steps with checkpoints every 50, loss = sum over replicas
    of (polar speed - 0.5)^2, adjoint sweep with the collisions' cells and
    signs held fixed, giving d loss / d f_a of every replica.  Starts from
    rows of straight worms (lattice start)."""
    import ctypes
    import math

    import numpy as np
    from nwamd import abi, replicas, system
    from nwamd.abi import DeviceArray
    lib = abi.load()
    R, steps, K, gap, dt, kT = args.batch_adjoint_replicas, args.batch_adjoint_steps, 50, 10, args.dt, 0.1
    rng = np.random.default_rng(31337)
    specs = (replicas.NWReplicaSpec * R)()
    for r in range(R):
        specs[r].numFilaments, specs[r].filamentSize, specs[r].boxSize = 1000, 12, 120.0
        specs[r].driveForce = float(rng.uniform(0.5, 5.0))
        specs[r].bendingK, specs[r].activity = 5.0, 0.0
        specs[r].seed = int(rng.integers(1, 2 ** 62))
    rs = replicas.ReplicaSet(specs, replicas.worm_physics(gamma=2.0, kT=kT, fdt=True))
    rs.reset(system.NW_INIT_LATTICE_RANDOM, kT)
    reps = (abi.NWReplicaParams * R)()
    sol = (abi.NWReplicaSolvent * R)()
    beads, bvel, spos, svel = [], [], [], []
    for r in range(R):
        p, e = rs.params(r)
        e.srdCollideFilaments, e.srdThermostat = 1, abi.NW_SRD_THERMO_NONE
        reps[r].params, reps[r].ext = p, e
        reps[r].seed, reps[r].offset = rs.rng_state(r)
        reps[r].langevin = 0
        beads.append(rs.get(system.POSITION, r))
        bvel.append(rs.get(system.VELOCITY, r))
        sp = abi.SolventParams()
        sp.enabled = True
        sp.gridSize = abi.UInt3(120, 120, 1)
        sp.numCells = 120 * 120
        sp.cellSize = abi.Float3(1.0, 1.0, p.boxSize.z)
        sp.alpha, sp.kbT, sp.particleMass = math.pi / 2, kT, 1.0
        Ns = 3 * sp.numCells
        sp.numParticles = Ns
        sol[r].solvent, sol[r].numSolvent = sp, Ns
        q = np.zeros((Ns, 4), np.float32)
        q[:, :2] = rng.random((Ns, 2)) * 120.0 + p.origin.x
        q[:, 2] = p.origin.z + 0.5 * p.boxSize.z
        q[:, 3] = 0.5
        # ids in SRD-cell order (the order the host driver's solvent re-sort
        # keeps); 1,000 steps move the solvent ~0.03 cells, so it stays so
        cell = np.floor(q[:, 1] - p.origin.y) * 120 + np.floor(q[:, 0] - p.origin.x)
        q = q[np.argsort(cell, kind="stable")]
        v = np.zeros((Ns, 4), np.float32)
        v[:, :2] = rng.normal(0.0, math.sqrt(kT), (Ns, 2))
        v[:, 3] = 1.0
        spos.append(q)
        svel.append(v)
    rs.close()
    lib.nwBatchSetReplicas(reps, R)
    lib.nwBatchSetSolvent(sol, R)
    N = sum(len(b) for b in beads)
    Nsol = sum(len(q) for q in spos)
    P = DeviceArray.from_numpy(lib, np.concatenate(beads + spos))
    V = DeviceArray.from_numpy(lib, np.concatenate(bvel + svel))
    zero = np.zeros((N, 4), np.float32)
    F, G, T = (DeviceArray.from_numpy(lib, zero) for _ in range(3))
    dP, dV = DeviceArray(lib, 16 * N), DeviceArray(lib, 16 * N)
    nColl = (steps + gap - 1) // gap
    sh = np.zeros((nColl, R, 3), np.float32)
    sh[:, :, :2] = rng.random((nColl, R, 2)) - 0.5
    sh = np.ascontiguousarray(sh.reshape(-1))
    shp = sh.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    g = (ctypes.c_double * (4 * R))()
    Pr = (ctypes.c_double * R)()
    # warm-up: a short differentiated run (allocations, first launches)
    lib.nwBatchAdjointForward(P.p, V.p, F.p, G.p, T.p, dt, gap, shp, 2 * gap, gap)
    lib.nwBatchPolarSpeedGrad(P.p, V.p, dP.p, dV.p, Pr)
    lib.nwBatchAdjointBackward(dP.p, dV.p, g)
    lib.threadSync()
    lib.nwTimerEnable(b"badj_step", 10)
    t0 = time.perf_counter()
    lib.nwBatchAdjointForward(P.p, V.p, F.p, G.p, T.p, dt, gap, shp, steps, K)
    lib.threadSync()
    t1 = time.perf_counter()
    lib.nwBatchPolarSpeedGrad(P.p, V.p, dP.p, dV.p, Pr)
    lib.nwBatchAdjointBackward(dP.p, dV.p, g)
    t2 = time.perf_counter()
    co = np.zeros(R + 1, np.uint32)
    lib.nwBatchLayout(None, co.ctypes.data_as(ctypes.c_void_p))
    roof = roofline_of(lib, "badj_step", dict(N=N, C=int(co[-1]), lgv=False), 10)
    lib.nwTimerEnable(b"badj_step", 0)
    Pv = np.array(Pr[:R])
    dfa = 2.0 * (Pv - 0.5) * np.array(g[:])[0::4]  # d loss / d f_a of every replica
    return dict(workload=f"{R} replicas x 1000 worms x 12 beads, L 120, SRD 3 per cell", replicas=R, beads=N,
                solvent=Nsol, steps=steps, checkpoint_every=K, srd_gap=gap,
                forward_ms=round((t1 - t0) * 1e3, 2), backward_ms=round((t2 - t1) * 1e3, 2),
                backward_over_forward=round((t2 - t1) / (t1 - t0), 3),
                adjoint_over_forward=round((t2 - t0) / (t1 - t0), 3),
                value=round(N * steps / (t2 - t0), 1), unit="bead-steps/s (forward + adjoint)",
                loss=round(float(((Pv - 0.5) ** 2).sum()), 6),
                dloss_dfa=dict(replica0=float(dfa[0]), replica1=float(dfa[1]), mean=float(dfa.mean()),
                               min=float(dfa.min()), max=float(dfa.max())), roofline=roof)


def roofline_of(lib, family, w, stride):
    """roofline of one timed single-kernel family: its algorithmic bytes per
    launch over its average duration (HIP events on its dispatch packets)"""
    from nwamd import abi
    ms, cnt = abi.timer_read(lib, family)
    avg_s = ms / max(cnt, 1) * 1e-3
    alg = ALG_BYTES[family](w)
    achieved = alg / avg_s / 1e9 if avg_s > 0 else 0.0
    return dict(bound="hbm", kernel=family, achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None, alg_bytes_per_launch=int(alg),
                avg_launch_us=round(avg_s * 1e6, 2), launches_timed=cnt, timed_every=stride)


def batch_matrix_bytes(bead_off, cell_off):
    """bytes of the batched sort's prefix matrices: every replica sorted
    against its own plan (tiles of 4096 beads, buckets of 2^S cells aiming
    at 512 members; make_sort_reps in the library)"""
    tot = 0
    for r in range(len(bead_off) - 1):
        n, c = int(bead_off[r + 1] - bead_off[r]), int(cell_off[r + 1] - cell_off[r])
        S = 0
        while S < 11 and (1 << (S + 1)) * n / c <= 512.0:
            S += 1
        tot += 4 * ((n + 4095) // 4096) * (((c + (1 << S) - 1) >> S) + 1)
    return tot


def timers(lib, names, enable):
    for n in names:
        lib.nwTimerEnable(n.encode(), 1 if enable else 0)


def cpu_baseline(s, cfg, dt, steps):
    """The oracle's FP32 step (OpenMP) on a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    from nwamd import system
    seed, offset = s.rng_state()
    cs = orc.CpuSystem(s.params, s.solvent, cfg.ext, s.get(system.POSITION), s.get(system.VELOCITY),
                       s.get(system.TANGENT), s.get(system.FORCE), s.get(system.FORCE_OLD), num_solvent=s.ns,
                       srd=bool(cfg.srd), srd_gap=cfg.srdGap or 5, dt=dt, seed=seed, offset=offset,
                       step=s.step_count, langevin_fdt=bool(cfg.langevinFdt),
                       shift_centred=bool(cfg.srdShiftCentred))
    cs.step(1)  # warm-up (page faults, thread pool)
    t0 = time.perf_counter()
    cores = cs.step(steps)
    el = time.perf_counter() - t0
    return dict(value=s.n * steps / el, unit="bead-steps/s", cores=int(cores), kind="port",
                sample=f"{steps} steps of the full workload ({s.n} beads + {s.ns} solvent, "
                       f"{(steps + 9) // 10 if cfg.srd else 0} SRD collisions), oracle FP32 OpenMP step",
                seconds=round(el, 3))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="worms_srd_800k")
    ap.add_argument("--dt", type=float, default=1e-3)
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--replicas", type=int, default=256, help="replicas of the batched form")
    ap.add_argument("--batched-steps", type=int, default=200, help="timed steps of the batched form (0: skip)")
    ap.add_argument("--batched-no-graph", action="store_true",
                    help="launch the batched step's kernels directly instead of replaying the captured graph "
                         "(for counter passes of a profiler)")
    ap.add_argument("--one-per-launch-steps", type=int, default=20,
                    help="timed steps of the replicas one launch sequence each (0: skip)")
    ap.add_argument("--adjoint-steps", type=int, default=2000,
                    help="steps of the differentiated dry-worm run (0: skip)")
    ap.add_argument("--batch-adjoint-steps", type=int, default=1000,
                    help="steps of the batched run differentiated through SRD collisions (0: skip)")
    ap.add_argument("--batch-adjoint-replicas", type=int, default=64)
    ap.add_argument("--solvent-sort-every", type=int, default=20,
                    help="collisions between re-sorts of the solvent by SRD cell on the device "
                         "(the host driver's data layout; 0: the reference's particle order)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="at the end, one elementwise kernel reading and writing this many MiB "
                         "(a known-size reference for HBM byte counters)")
    ap.add_argument("--gather-probe-mb", type=int, default=0,
                    help="at the end, one gather of this many MiB of 16-byte rows, whole 128-byte lines in random "
                         "order, each line read once (a known-size reference for counters of gather kernels)")
    ap.add_argument("--profile-steps", type=int, default=20,
                    help="untimed steps after the warm-up that rank the kernel families")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in launch shapes (default: the library's autotuner picks them on a copy "
                         "of the relaxed state before the warm-up)")
    ap.add_argument("--relax-steps", type=int, default=20000,
                    help="untimed steps from the lattice start before the warm-up, so the timed state is a "
                         "relaxed one (its order parameters are reported)")
    ap.add_argument("--shuffled-steps", type=int, default=200,
                    help="timed steps of the same state with its worms permuted in memory (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--per-call-steps", type=int, default=100,
                    help="steps of the reference per-call sequence timed beside (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(a batch of replicas would shard across GPUs by replica, without communication)",
              file=sys.stderr)
        sys.exit(2)

    from nwamd import abi
    lib = abi.load()
    if not lib.nwDeviceAvailable():
        raise SystemExit("bench.py: no HIP device")
    lib.cudaInit(0, None)
    s, cfg, w = build_workload(args.workload, solvent_sort_every=args.solvent_sort_every)
    dt = args.dt
    names = list(ALG_BYTES)

    # relaxation from the lattice start (untimed), timing every solvent
    # re-sort it contains
    obs0 = s.observables()
    resort_every = (cfg.srdGap or 5) * args.solvent_sort_every if cfg.srd and args.solvent_sort_every else 0
    if resort_every:
        lib.nwTimerEnable(b"srd_resort", 1)
    t0 = time.perf_counter()
    s.update(dt, args.relax_steps)
    s.sync()
    relax_s = time.perf_counter() - t0
    resort_ms, resorts = abi.timer_read(lib, "srd_resort") if resort_every else (0.0, 0)
    lib.nwTimerEnable(b"srd_resort", 0)
    obs1 = s.observables()
    state = dict(relax_steps=args.relax_steps, relax_seconds=round(relax_s, 2),
                 start=dict(nematicS=round(obs0["nematicS"], 4), polarP=round(obs0["polarP"], 4),
                            polarSpeed=round(obs0["polarSpeed"], 4)),
                 timed=dict(nematicS=round(obs1["nematicS"], 4), polarP=round(obs1["polarP"], 4),
                            polarSpeed=round(obs1["polarSpeed"], 4)))

    # launch shapes of the fused kernels: the library's autotuner, timed on a
    # copy of the relaxed state (or from its cache file)
    tuning, tuned = None, "built-in"
    if not args.no_autotune:
        t0 = time.perf_counter()
        rc, tuning = s.autotune(dt)
        tuned = {1: "timed", 2: "cache"}.get(rc, "failed")
        tuning["tuner_seconds"] = round(time.perf_counter() - t0, 3)
        w["tuning"] = tuning
        w["md_plan"] = fused_plan(w["N"], w["C"], cfg.filamentSize)
        w["Pm"] = 4 * w["md_plan"][3] * (w["md_plan"][1] + 1)
        if cfg.srd:
            w["srd_plan"] = fused_plan(w["NT"], w["Cs"], 0)
            w["Ps"] = 4 * w["srd_plan"][3] * (w["srd_plan"][1] + 1)

    # untimed warm-up, then (also untimed) every kernel family timed over a
    # few steps to find the dominant one
    s.update(dt, args.warmup)
    s.sync()
    timers(lib, names, True)
    s.update(dt, args.profile_steps)
    s.sync()
    per = {}
    for n in names:
        ms, cnt = abi.timer_read(lib, n)
        if cnt:
            per[n] = dict(ms_per_launch=ms / cnt, launches=cnt, ms_total=ms)
    timers(lib, names, False)
    single = [k for k in per if k in SINGLE_KERNEL]
    dominant = max(single, key=lambda k: per[k]["ms_total"]) if single else "md_collide"

    # timed region: only the dominant kernel carries events, on every 10th
    # launch when it runs every step (an event pair stalls the stream ~5 us)
    stride = 10 if per.get(dominant, {}).get("launches", 0) >= args.profile_steps else 1
    lib.nwTimerEnable(dominant.encode(), stride)
    if resort_every:
        lib.nwTimerEnable(b"srd_resort", 1)  # counts the re-sorts inside the window
    s.sync()
    t0 = time.perf_counter()
    s.update(dt, args.steps)
    s.sync()
    el = time.perf_counter() - t0
    dom_ms, dom_n = abi.timer_read(lib, dominant)
    lib.nwTimerEnable(dominant.encode(), 0)
    in_window = abi.timer_read(lib, "srd_resort")[1] if resort_every else 0
    lib.nwTimerEnable(b"srd_resort", 0)

    # the solvent re-sort runs once every resort_every steps: the window is
    # charged its share exactly (the re-sorts it held replaced by the
    # expected number, at the average re-sort time of the relaxation)
    per_resort_s = resort_ms / resorts * 1e-3 if resorts else 0.0
    el_charged = el + ((args.steps / resort_every - in_window) * per_resort_s if resort_every else 0.0)
    resort = (dict(every_steps=resort_every, avg_ms=round(per_resort_s * 1e3, 3), in_window=in_window,
                   expected_in_window=round(args.steps / resort_every, 3), window_wall_s=round(el, 6))
              if resort_every else None)
    value = w["N"] * args.steps / el_charged
    alg = ALG_BYTES[dominant](w)
    avg_s = dom_ms / max(dom_n, 1) * 1e-3
    achieved = alg / avg_s / 1e9 if avg_s > 0 else 0.0
    traffic, traffic_source = measured_traffic(args.workload, dominant, w)
    roof = dict(bound="hbm", kernel=dominant, achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, traffic_source=traffic_source,
                alg_bytes_per_launch=int(alg),
                avg_launch_us=round(avg_s * 1e6, 2), launches_timed=dom_n, timed_every=stride)
    kernels = {k: dict(avg_us=round(v["ms_per_launch"] * 1e3, 2), launches=v["launches"],
                       gbs=round(ALG_BYTES[k](w) / (v["ms_per_launch"] * 1e-3) / 1e9, 1))
               for k, v in per.items()}

    # the same system with its worms in a random order in memory (the
    # physics is unchanged; bead order no longer follows position, so the
    # sort's gathers and the front's stores lose their locality)
    shuffled = None
    if args.shuffled_steps > 0:
        import numpy as np
        from nwamd import system as nsys
        W, n = cfg.numFilaments, cfg.filamentSize
        perm = np.random.default_rng(7).permutation(W)
        for which in (nsys.POSITION, nsys.VELOCITY, nsys.TANGENT, nsys.FORCE, nsys.FORCE_OLD):
            a = s.get(which)
            beads = a[:W * n].reshape(W, n, 4)[perm].reshape(W * n, 4)
            s.set(which, np.concatenate([beads, a[W * n:]]))
        s.update(dt, args.warmup)
        s.sync()
        t0 = time.perf_counter()
        s.update(dt, args.shuffled_steps)
        s.sync()
        el3 = time.perf_counter() - t0
        shuffled = dict(value=round(w["N"] * args.shuffled_steps / el3, 1), unit="bead-steps/s",
                        steps=args.shuffled_steps, ms_per_step=round(el3 / args.shuffled_steps * 1e3, 4),
                        form="the timed state with its worms permuted at random in memory (raw window)")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(s, cfg, dt, args.cpu_steps)
    s.close()

    per_call = None
    if args.per_call_steps > 0:
        s2, _, _ = build_workload(args.workload, per_call=True, solvent_sort_every=args.solvent_sort_every)
        s2.update(dt, args.warmup)
        s2.sync()
        t0 = time.perf_counter()
        s2.update(dt, args.per_call_steps)
        s2.sync()
        el2 = time.perf_counter() - t0
        per_call = dict(value=round(w["N"] * args.per_call_steps / el2, 1), unit="bead-steps/s",
                        steps=args.per_call_steps, ms_per_step=round(el2 / args.per_call_steps * 1e3, 4),
                        form="the reference's per-call C-ABI sequence (ParticleSystem::update, 10+ launches/step)")
        s2.close()

    batched = batched_form(args) if args.batched_steps > 0 else None
    adjoint = adjoint_form(args) if args.adjoint_steps > 0 else None
    batch_adjoint = batch_adjoint_form(args) if args.batch_adjoint_steps > 0 else None

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "bead-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(el_charged / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "beads": w["N"], "solvent": w["Ns"], "md_cells": w["C"],
                   "srd_cells": w["Cs"], "dt": dt,
                   "solvent_sort_every": args.solvent_sort_every,
                   "form": "plain: one replica, ParticleSystem::update through the fused C-ABI entry points",
                   "md_plan": list(w["md_plan"]), "srd_plan": list(w["srd_plan"]),
                   "tuning": tuning, "tuning_source": tuned},
        "state": state, "solvent_resort": resort, "shuffled_worm_order": shuffled,
        "roofline": roof, "cpu_baseline": cpu, "per_call_sequence": per_call, "batched": batched, "adjoint": adjoint,
        "batched_adjoint_srd": batch_adjoint,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)
    if args.copy_probe_mb > 0:
        import torch
        a = torch.ones(args.copy_probe_mb * 2 ** 20 // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        torch.neg(a, out=b)
        torch.cuda.synchronize()
        print(f"copy probe: neg over {a.numel()} floats, {a.numel() * 4} B read and {a.numel() * 4} B written",
              flush=True)
    if args.gather_probe_mb > 0:
        import torch
        rows = args.gather_probe_mb * 2 ** 20 // 16
        lines = rows // 8
        table = torch.ones((lines * 8, 4), dtype=torch.float32, device="cuda")
        g = torch.Generator(device="cuda")
        g.manual_seed(5)
        perm = torch.randperm(lines, device="cuda", generator=g)
        idx = (perm[:, None] * 8 + torch.arange(8, device="cuda")[None, :]).reshape(-1)
        torch.cuda.synchronize()
        out = torch.index_select(table, 0, idx)
        torch.cuda.synchronize()
        print(f"gather probe: index_select of {idx.numel()} rows of 16 B in {lines} whole 128 B lines, random line "
              f"order: {table.numel() * 4} B of rows + {idx.numel() * 8} B of indices read, {out.numel() * 4} B "
              f"written", flush=True)


if __name__ == "__main__":
    main()

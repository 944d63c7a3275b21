"""The batched forms checked against the oracle replica by replica.

* Batched replicas shaped like the 256-replica configuration (worms
  U[500, 2000] x U{10..30} beads, box at area fraction 0.3, f_a
  log-uniform in [0.1, 10], kappa U[1, 20], own seeds, rejection-sampled
  start), stepped with the hipGraph: every replica against the oracle's
  whole FP32 step (orc.CpuSystem, the reference's ParticleSystem::update,
  simulation/particleSystem.cpp:249-339) started from the same state and
  Philox offset.
* Batched SRD replicas shaped like the differentiated 64-replica
  configuration (collision every 10 steps, 90 degrees, beads in the
  collisions, loss = sum over replicas of (polar speed - 0.5)^2): every
  replica's d loss / d f_a against the FP64 reference adjoint
  (oracle/adjoint_ref.py) and FP64 central differences.
* At the full sizes of both configurations: properties that hold at any
  size (every replica of the graph-replayed batch equal bit for bit to the
  same replica stepped alone; determinism and independence of the
  checkpoint spacing of the batched gradients; finite states)."""
import ctypes
import math

import numpy as np
import pytest

import config_cases as cc
import orc
from adjoint_ref import Model, Srd
from nwamd import abi, replicas, system
from nwamd.abi import DeviceArray
from nwamd.replicas import NW_REPLICAS_GRAPH, NW_REPLICAS_ONE_PER_LAUNCH, ReplicaSet, worm_physics
from nwtest import force_tol_check

pytestmark = pytest.mark.gpu

DT = 1e-3


def config2_specs(R, seed, wmin=500, wmax=2000):
    rng = np.random.default_rng(seed)
    specs = (replicas.NWReplicaSpec * R)()
    for r in range(R):
        s = specs[r]
        s.numFilaments = int(rng.integers(wmin, wmax + 1))
        s.filamentSize = int(rng.integers(10, 31))
        s.boxSize = replicas.box_for(s.numFilaments * s.filamentSize)
        s.driveForce = float(np.exp(rng.uniform(np.log(0.1), np.log(10.0))))
        s.bendingK = float(rng.uniform(1.0, 20.0))
        s.activity = 0.0
        s.seed = int(rng.integers(1, 2 ** 62))
    return specs


def periodic_diff(a, b, L):
    d = (np.asarray(a, np.float64) - np.asarray(b, np.float64))
    return d - L * np.round(d / L)


def test_config2_shaped_replicas_match_the_oracle(nwlib):
    R = 8
    specs = config2_specs(R, seed=2202)
    rs = ReplicaSet(specs, worm_physics(gamma=2.0, kT=0.1, fdt=True))
    rs.reset(system.NW_INIT_RANDOM, 0.1)
    cpus = []
    for r in range(R):
        p, e = rs.params(r)
        seed, off = rs.rng_state(r)
        cpus.append(orc.CpuSystem(p, abi.SolventParams(), e, rs.get(system.POSITION, r),
                                  rs.get(system.VELOCITY, r), rs.get(system.TANGENT, r), rs.get(system.FORCE, r),
                                  rs.get(system.FORCE_OLD, r), num_solvent=0, srd=False, srd_gap=10, dt=DT,
                                  seed=seed, offset=off, langevin_fdt=True))
    sizes = [(specs[r].numFilaments, specs[r].filamentSize) for r in range(R)]
    # the first step (a direct one): forces from identical positions
    rs.update(DT, 1, NW_REPLICAS_GRAPH)
    rs.sync()
    for r, cpu in enumerate(cpus):
        cpu.step(1)
        assert np.array_equal(rs.get(system.POSITION, r), cpu.get("pos")), (r, sizes[r])
        force_tol_check(rs.get(system.FORCE, r), cpu.get("force"), name=f"replica {r} {sizes[r]} force")
    # then the captured graph, replayed
    K = 24
    rs.update(DT, K, NW_REPLICAS_GRAPH)
    rs.sync()
    for r, cpu in enumerate(cpus):
        cpu.step(K)
        p, _ = rs.params(r)
        L = np.array([p.boxSize.x, p.boxSize.y, p.boxSize.z])
        gp, gv = rs.get(system.POSITION, r), rs.get(system.VELOCITY, r)
        dp = np.abs(periodic_diff(gp[:, :3], cpu.get("pos")[:, :3], L)).max()
        dv = np.abs(gv[:, :3].astype(np.float64) - cpu.get("vel")[:, :3]).max()
        assert dp < 1e-4 and dv < 1e-3, (r, sizes[r], dp, dv)
        # the generator advanced as the oracle's: one block of four normals per bead and step
        assert rs.rng_state(r)[1] == cpu.c.offset, r
    rs.close()


def test_config2_full_size_graph_equals_each_replica_alone(nwlib):
    """256 replicas of the configuration (21.3M beads): the graph-replayed
    batch against the same replicas one launch sequence each, every replica
    bit for bit, states finite."""
    specs = replicas.replica_specs_config2(256)
    phys = worm_physics(gamma=2.0, kT=0.1, fdt=True)
    a, b = ReplicaSet(specs, phys), ReplicaSet(specs, phys)
    for rs in (a, b):
        rs.reset(system.NW_INIT_LATTICE_RANDOM, 0.1)
    a.update(DT, 1, NW_REPLICAS_GRAPH)
    a.update(DT, 4, NW_REPLICAS_GRAPH)
    b.update(DT, 5, NW_REPLICAS_ONE_PER_LAUNCH)
    a.sync()
    b.sync()
    for r in range(a.R):
        for w in range(5):
            x, y = a.get(w, r), b.get(w, r)
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (r, w)
            if w == 0:
                assert np.isfinite(x).all(), r
        assert a.rng_state(r) == b.rng_state(r)
    a.close()
    b.close()


# ---- batched + adjoint with SRD ------------------------------------------------
SRD_GAP = 10
C4_WORMS = [6, 8, 5, 7]
C4_FA = [0.7, 2.3, 4.1, 1.4]
C4_SEEDS = [(901, 0), (902, 17), (903, 5), (904, 123)]
C4_N, C4_L = 12, 12.0


def c4_inputs(k):
    p, e, s, bp, bv, sp, sv = cc.srd_replica(C4_WORMS[k], C4_N, C4_L, C4_FA[k], seed=40 + k)
    return p, e, s, bp, bv, sp, sv


def c4_shifts(k, steps, s):
    cell = (s.cellSize.x, s.cellSize.y, s.cellSize.z)
    return np.array([orc.srd_shift(C4_SEEDS[k][0], c, True, cell)
                     for c in range((steps + SRD_GAP - 1) // SRD_GAP)], np.float32)


def batched_polar_speed_gradient(lib, ins, seeds, shifts, steps, K):
    """nwBatchAdjointForward + nwBatchPolarSpeedGrad + nwBatchAdjointBackward:
    per replica (P, d P / d theta)"""
    R = len(ins)
    reps = (abi.NWReplicaParams * R)()
    sol = (abi.NWReplicaSolvent * R)()
    for k, (p, e, s, *_rest) in enumerate(ins):
        reps[k].params, reps[k].ext = p, e
        reps[k].seed, reps[k].offset = seeds[k]
        reps[k].langevin = 0
        sol[k].solvent, sol[k].numSolvent = s, s.numParticles
    lib.nwBatchSetReplicas(reps, R)
    lib.nwBatchSetSolvent(sol, R)
    P0 = np.concatenate([x[3] for x in ins] + [x[5] for x in ins])
    V0 = np.concatenate([x[4] for x in ins] + [x[6] for x in ins])
    N = sum(len(x[3]) for x in ins)
    zero = np.zeros((N, 4), np.float32)
    d = [DeviceArray.from_numpy(lib, a) for a in (P0, V0, zero, zero, zero)]
    sh = np.ascontiguousarray(np.stack(shifts, axis=1).reshape(-1))  # collision-major, then replica
    lib.nwBatchAdjointForward(d[0].p, d[1].p, d[2].p, d[3].p, d[4].p, DT, SRD_GAP,
                              sh.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), steps, K)
    dP, dV = DeviceArray(lib, 16 * N), DeviceArray(lib, 16 * N)
    Pr = (ctypes.c_double * R)()
    lib.nwBatchPolarSpeedGrad(d[0].p, d[1].p, dP.p, dV.p, Pr)
    g = (ctypes.c_double * (4 * R))()
    lib.nwBatchAdjointBackward(dP.p, dV.p, g)
    state = (d[0].download(np.float32, P0.shape), d[1].download(np.float32, V0.shape))
    return np.array(Pr[:]), np.array(g[:]).reshape(R, 4), state


def c4_model(k, steps, fa=None):
    p, e, s, *_ = c4_inputs(k)
    theta = (C4_FA[k] if fa is None else fa, 5.0, 1.0, 2.0)
    srd = Srd(s, s.numParticles, SRD_GAP, c4_shifts(k, steps, s), True)
    return Model(p, e, p.numFilaments, DT, 0.0, False, C4_SEEDS[k][0], C4_SEEDS[k][1], srd=srd), theta


def c4_loss64(k, steps, fa):
    m, theta = c4_model(k, steps, fa)
    p, e, s, bp, bv, sp, sv = c4_inputs(k)
    zero = np.zeros((len(bp), 4), np.float32)
    rf, vf, *_ = m.run(np.concatenate([bp, sp]), np.concatenate([bv, sv]), zero, zero, theta, steps)
    P, _, _ = cc.polar_speed_and_grad(p, rf, vf, p.numFilaments)
    return (P - 0.5) ** 2


def test_config4_shaped_replica_gradients_match_fp64_reference_and_differences(nwlib):
    """Four uneven SRD replicas (5..8 worms x 12 beads, 12 x 12 unit SRD
    cells with 3 solvent particles each, own f_a, seeds and grid shifts), 40
    steps with a collision every 10, checkpoints every 10: d loss / d f_a of
    every replica from the batched adjoint against the FP64 reference
    gradient (2e-3) and FP64 central differences with h = 1e-4 f_a (1e-3)."""
    steps, K = 40, 10
    R = len(C4_WORMS)
    ins = [c4_inputs(k) for k in range(R)]
    shifts = [c4_shifts(k, steps, ins[k][2]) for k in range(R)]
    Pg, g, (fP, fV) = batched_polar_speed_gradient(nwlib, ins, C4_SEEDS, shifts, steps, K)
    dl_gpu = 2.0 * (Pg - 0.5) * g[:, 0]
    bo = np.concatenate([[0], np.cumsum([len(x[3]) for x in ins])])
    for k in range(R):
        p, e, s, bp, bv, sp, sv = ins[k]
        m, theta = c4_model(k, steps)
        zero = np.zeros((len(bp), 4), np.float32)
        P0, V0 = np.concatenate([bp, sp]), np.concatenate([bv, sv])
        rf, vf, *_ = m.run(P0, V0, zero, zero, theta, steps)
        # the FP32 batched trajectory tracks the FP64 one
        np.testing.assert_allclose(fP[bo[k]:bo[k + 1], :3], rf[:m.N], atol=5e-5, err_msg=f"replica {k}")
        P64, dPdr, dPdv = cc.polar_speed_and_grad(p, rf, vf, p.numFilaments)
        assert abs(Pg[k] - P64) < 1e-4, (k, Pg[k], P64)
        ref = m.gradient(P0, V0, zero, zero, theta, steps, dPdr, dPdv)
        dl_ref = 2.0 * (P64 - 0.5) * ref[0]
        h = 1e-4 * C4_FA[k]
        dl_fd = (c4_loss64(k, steps, C4_FA[k] + h) - c4_loss64(k, steps, C4_FA[k] - h)) / (2 * h)
        assert dl_fd != 0.0
        assert abs(dl_gpu[k] - dl_ref) <= 2e-3 * abs(dl_ref), (k, dl_gpu[k], dl_ref)
        assert abs(dl_gpu[k] - dl_fd) <= 1e-3 * abs(dl_fd), (k, dl_gpu[k], dl_fd)
        # the other parameters of the replica against the reference too
        np.testing.assert_allclose(g[k, :3], ref[:3], rtol=2e-3, atol=2e-4 * np.abs(ref).max())
        assert g[k, 3] == 0.0


def test_config4_full_size_gradients_deterministic_and_checkpoint_independent(nwlib):
    """64 replicas x 1,000 worms x 12 beads in 120 x 120 boxes with 43,200
    solvent particles each (768,000 beads + 2,764,800 solvent), 100 steps:
    two runs give the same gradients bit for bit, a different checkpoint
    spacing gives them to the FP64 summation order, all finite."""
    R, steps = 64, 100
    rng = np.random.default_rng(64)
    fa = rng.uniform(0.5, 5.0, R)
    ins = [cc.srd_replica(1000, 12, 120.0, float(fa[k]), seed=1000 + k) for k in range(R)]
    seeds = [(int(rng.integers(1, 2 ** 62)), 0) for _ in range(R)]
    shifts = []
    for k in range(R):
        sh = np.zeros(((steps + SRD_GAP - 1) // SRD_GAP, 3), np.float32)
        sh[:, :2] = rng.random((len(sh), 2)) - 0.5
        shifts.append(sh)
    P1, g1, (fP, fV) = batched_polar_speed_gradient(nwlib, ins, seeds, shifts, steps, 50)
    P2, g2, _ = batched_polar_speed_gradient(nwlib, ins, seeds, shifts, steps, 50)
    P3, g3, _ = batched_polar_speed_gradient(nwlib, ins, seeds, shifts, steps, 20)
    assert np.isfinite(fP).all() and np.isfinite(fV).all()
    assert np.all(np.isfinite(g1)) and np.all(np.abs(g1[:, 0]) > 0)
    assert np.array_equal(P1, P2) and np.array_equal(g1, g2)
    assert np.array_equal(P1, P3)
    np.testing.assert_allclose(g3, g1, rtol=1e-9, atol=1e-12 * np.abs(g1).max())
    assert np.all(g1[:, 3] == 0.0)
    # replicas with larger drive move faster along their tangents
    assert np.corrcoef(fa, P1)[0, 1] > 0.8

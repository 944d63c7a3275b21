"""Differentiation through batched steps (nwBatchSetSolvent,
nwBatchAdjointForward / nwBatchAdjointBackward, nwBatchPolarSpeedGrad):
every replica of a batch of uneven replicas (3..5 worms, own drive,
stiffness, WCA epsilon, seed and SRD grid shifts) against the same replica
alone through nwAdjointForwardSrd / nwAdjointForward / nwPolarSpeedGrad,
which the FP64 reference and central differences pin: trajectories bit for
bit, gradients to the FP64 summation order."""
import ctypes

import numpy as np
import pytest

import adjoint_case as case
import orc
from nwamd import abi
from nwamd.abi import DeviceArray

pytestmark = pytest.mark.gpu

THETAS = [(1.3, 4.0, 1.0, 2.0), (0.6, 6.0, 1.2, 2.0), (2.1, 3.0, 0.8, 2.0)]
WORMS = [5, 3, 4]
SEEDS = [(77, 1000), (78, 5), (1337, 123456)]


def replica_inputs(k, srd):
    """params, ext, solvent (or None), beads' pos/vel, solvent pos/vel, loss weights"""
    W, n = WORMS[k], 8
    if srd:
        p, e, s = case.srd_system(THETAS[k])
        P, V, _, w, wv = case.srd_initial_state(seed=6 + k)
        N0 = len(w)
        bp, bv, sp, sv = P[:W * n], V[:W * n], P[N0:], V[N0:]
    else:
        p, e = case.system(THETAS[k])
        s = None
        bp, bv, _, w = case.initial_state(seed=5 + k)
        wv = np.random.default_rng(20 + k).normal(0, 1, w.shape).astype(np.float32)
        wv[:, 3] = 0
        bp, bv = bp[:W * n], bv[:W * n]
        sp = sv = np.zeros((0, 4), np.float32)
    p.numFilaments = W
    p.numParticles = W * n
    return p, e, s, bp, bv, sp, sv, w[:W * n], wv[:W * n]


def shifts_of(k, steps, s):
    cell = (s.cellSize.x, s.cellSize.y, s.cellSize.z)
    return np.array([orc.srd_shift(SEEDS[k][0], c, True, cell)
                     for c in range((steps + case.SRD_GAP - 1) // case.SRD_GAP)], np.float32)


def alone(lib, k, srd, steps, K):
    p, e, s, bp, bv, sp, sv, w, wv = replica_inputs(k, srd)
    lib.setParameters(ctypes.byref(p))
    lib.nwSetExtParameters(ctypes.byref(e))
    if srd:
        lib.setSolventParameters(ctypes.byref(s))
    lib.nwRngSetSeed(*SEEDS[k])
    P, V = np.concatenate([bp, sp]), np.concatenate([bv, sv])
    zero = np.zeros((len(bp), 4), np.float32)
    d = [DeviceArray.from_numpy(lib, a) for a in (P, V, zero, zero, zero)]
    if srd:
        sh = np.ascontiguousarray(shifts_of(k, steps, s).reshape(-1))
        lib.nwAdjointForwardSrd(d[0].p, d[1].p, d[2].p, d[3].p, d[4].p, p.numFilaments, len(sp), case.DT,
                                case.SRD_GAP, sh.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), steps, K)
    else:
        lib.nwAdjointForward(d[0].p, d[1].p, d[2].p, d[3].p, d[4].p, p.numFilaments, case.DT, p.gamma,
                             case.KT / case.DT, 1, steps, K)
    final = (d[0].download(np.float32, (len(P), 4)), d[1].download(np.float32, (len(P), 4)))
    dW, dWv = DeviceArray.from_numpy(lib, w), DeviceArray.from_numpy(lib, wv)
    g = (ctypes.c_double * 4)()
    lib.nwAdjointBackward(dW.p, dWv.p, g)
    return final, np.array(g)


def batched(lib, srd, steps, K):
    R = len(THETAS)
    ins = [replica_inputs(k, srd) for k in range(R)]
    reps = (abi.NWReplicaParams * R)()
    for k, (p, e, *_rest) in enumerate(ins):
        reps[k].params, reps[k].ext = p, e
        reps[k].seed, reps[k].offset = SEEDS[k]
        reps[k].langevinGamma, reps[k].langevinKbT = p.gamma, case.KT / case.DT
        reps[k].langevin = 0 if srd else 1
    lib.nwBatchSetReplicas(reps, R)
    bo = np.zeros(R + 1, np.uint32)
    lib.nwBatchLayout(bo.ctypes.data, None)
    beads = [x[3] for x in ins]
    assert list(np.diff(bo)) == [len(b) for b in beads]
    so = np.zeros(R + 1, np.uint32)
    shifts = None
    if srd:
        sol = (abi.NWReplicaSolvent * R)()
        for k, x in enumerate(ins):
            sol[k].solvent, sol[k].numSolvent = x[2], len(x[5])
        lib.nwBatchSetSolvent(sol, R)
        lib.nwBatchSolventLayout(so.ctypes.data, None)
        assert so[0] == bo[-1] and list(np.diff(so)) == [len(x[5]) for x in ins]
        per = [shifts_of(k, steps, ins[k][2]) for k in range(R)]
        shifts = np.ascontiguousarray(np.stack(per, axis=1).reshape(-1))  # collision-major, then replica
    P = np.concatenate([x[3] for x in ins] + [x[5] for x in ins])
    V = np.concatenate([x[4] for x in ins] + [x[6] for x in ins])
    N = int(bo[-1])
    zero = np.zeros((N, 4), np.float32)
    d = [DeviceArray.from_numpy(lib, a) for a in (P, V, zero, zero, zero)]
    sp = shifts.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if srd else None
    lib.nwBatchAdjointForward(d[0].p, d[1].p, d[2].p, d[3].p, d[4].p, case.DT, case.SRD_GAP if srd else 0, sp,
                              steps, K)
    fP, fV = d[0].download(np.float32, P.shape), d[1].download(np.float32, V.shape)
    dW = DeviceArray.from_numpy(lib, np.concatenate([x[7] for x in ins]))
    dWv = DeviceArray.from_numpy(lib, np.concatenate([x[8] for x in ins]))
    g = (ctypes.c_double * (4 * R))()
    lib.nwBatchAdjointBackward(dW.p, dWv.p, g)
    out = []
    for k in range(R):
        rows = np.r_[bo[k]:bo[k + 1], so[k]:so[k + 1]] if srd else np.r_[bo[k]:bo[k + 1]]
        out.append(((fP[rows], fV[rows]), np.array(g[4 * k:4 * k + 4])))
    return out


@pytest.mark.parametrize("srd", [True, False], ids=["srd", "dry_langevin"])
def test_batched_adjoint_equals_every_replica_alone(nwlib, srd):
    steps, K = 23, 7
    got = batched(nwlib, srd, steps, K)
    for k, ((bP, bV), bg) in enumerate(got):
        (aP, aV), ag = alone(nwlib, k, srd, steps, K)
        assert np.array_equal(bP.view(np.uint32), aP.view(np.uint32)), k
        assert np.array_equal(bV.view(np.uint32), aV.view(np.uint32)), k
        assert np.all(np.abs(ag[:3]) > 0)
        np.testing.assert_allclose(bg, ag, rtol=1e-9, atol=1e-12 * np.abs(ag).max())
        if srd:
            assert bg[3] == 0.0


def test_batched_polar_speed_equals_every_replica_alone(nwlib):
    lib = nwlib
    R = len(THETAS)
    ins = [replica_inputs(k, False) for k in range(R)]
    reps = (abi.NWReplicaParams * R)()
    for k, (p, e, *_rest) in enumerate(ins):
        reps[k].params, reps[k].ext = p, e
        reps[k].seed, reps[k].offset = SEEDS[k]
    lib.nwBatchSetReplicas(reps, R)
    bo = np.zeros(R + 1, np.uint32)
    lib.nwBatchLayout(bo.ctypes.data, None)
    P = np.concatenate([x[3] for x in ins])
    V = np.concatenate([x[4] for x in ins])
    dP, dV = DeviceArray.from_numpy(lib, P), DeviceArray.from_numpy(lib, V)
    gP, gV = DeviceArray(lib, 16 * len(P)), DeviceArray(lib, 16 * len(P))
    Pr = (ctypes.c_double * R)()
    lib.nwBatchPolarSpeedGrad(dP.p, dV.p, gP.p, gV.p, Pr)
    G, GV = gP.download(np.float32, P.shape), gV.download(np.float32, P.shape)
    for k, x in enumerate(ins):
        lib.setParameters(ctypes.byref(x[0]))
        a, b = DeviceArray.from_numpy(lib, x[3]), DeviceArray.from_numpy(lib, x[4])
        ga, gb = DeviceArray(lib, 16 * len(x[3])), DeviceArray(lib, 16 * len(x[3]))
        assert lib.nwPolarSpeedGrad(a.p, b.p, x[0].numFilaments, ga.p, gb.p) == Pr[k]
        sl = slice(bo[k], bo[k + 1])
        assert np.array_equal(G[sl], ga.download(np.float32, x[3].shape))
        assert np.array_equal(GV[sl], gb.download(np.float32, x[3].shape))


def test_batched_srd_step_equals_every_replica_alone(nwlib):
    """nwBatchSrdStep (collision 2 of the replicas' streams) against
    nwFusedSrdStep on each replica alone with the matching generator offset:
    positions, velocities, cell momenta and signs, cells and cell order bit
    for bit."""
    lib = nwlib
    R = len(THETAS)
    ins = [replica_inputs(k, True) for k in range(R)]
    reps = (abi.NWReplicaParams * R)()
    sol = (abi.NWReplicaSolvent * R)()
    for k, x in enumerate(ins):
        reps[k].params, reps[k].ext = x[0], x[1]
        reps[k].seed, reps[k].offset = SEEDS[k]
        sol[k].solvent, sol[k].numSolvent = x[2], len(x[5])
    lib.nwBatchSetReplicas(reps, R)
    lib.nwBatchSetSolvent(sol, R)
    bo, so, co = (np.zeros(R + 1, np.uint32) for _ in range(3))
    lib.nwBatchLayout(bo.ctypes.data, None)
    lib.nwBatchSolventLayout(so.ctypes.data, co.ctypes.data)
    P = np.concatenate([x[3] for x in ins] + [x[5] for x in ins])
    V = np.concatenate([x[4] for x in ins] + [x[6] for x in ins])
    NT, C = len(P), int(co[-1])
    shifts = np.ascontiguousarray(np.stack([shifts_of(k, 15, ins[k][2])[2] for k in range(R)]).reshape(-1))
    dt = float(np.float32(case.DT) * np.float32(case.SRD_GAP))
    collision = 2
    dP, dV = DeviceArray.from_numpy(lib, P), DeviceArray.from_numpy(lib, V)
    com, h, ids = DeviceArray(lib, 16 * C), DeviceArray(lib, 4 * NT), DeviceArray(lib, 4 * NT)
    lib.nwBatchSrdStep(dP.p, dV.p, com.p, h.p, ids.p, dt, shifts.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                       collision)
    bP, bV = dP.download(np.float32, P.shape), dV.download(np.float32, V.shape)
    bC, bH, bI = com.download(np.float32, (C, 4)), h.download(np.uint32, (NT,)), ids.download(np.uint32, (NT,))
    slot0 = 0
    for k, x in enumerate(ins):
        p, e, s = x[0], x[1], x[2]
        N, Ns, Cs = len(x[3]), len(x[5]), int(co[k + 1] - co[k])
        s.randOffset = abi.Float3(*(float(v) for v in shifts[3 * k:3 * k + 3]))
        lib.setParameters(ctypes.byref(p))
        lib.setSolventParameters(ctypes.byref(s))
        lib.nwSetExtParameters(ctypes.byref(e))
        lib.nwRngSetSeed(SEEDS[k][0], SEEDS[k][1] + collision * ((Cs + 3) // 4))
        aP = DeviceArray.from_numpy(lib, np.concatenate([x[3], x[5]]))
        aV = DeviceArray.from_numpy(lib, np.concatenate([x[4], x[6]]))
        aC, aH, aI = DeviceArray(lib, 16 * Cs), DeviceArray(lib, 4 * (N + Ns)), DeviceArray(lib, 4 * (N + Ns))
        lib.nwFusedSrdStep(aP.p, aV.p, aC.p, aH.p, aI.p, dt, N, Ns, 0)
        rows = np.r_[bo[k]:bo[k + 1], so[k]:so[k + 1]]
        assert np.array_equal(bP[rows].view(np.uint32), aP.download(np.float32, (N + Ns, 4)).view(np.uint32)), k
        assert np.array_equal(bV[rows].view(np.uint32), aV.download(np.float32, (N + Ns, 4)).view(np.uint32)), k
        assert np.array_equal(bC[co[k]:co[k + 1]].view(np.uint32), aC.download(np.float32, (Cs, 4)).view(np.uint32))
        assert np.array_equal(bH[rows] - co[k], aH.download(np.uint32, (N + Ns,)))
        g = bI[slot0:slot0 + N + Ns].astype(np.int64)
        local = np.where(g < bo[-1], g - bo[k], N + g - so[k])
        assert np.array_equal(local, aI.download(np.uint32, (N + Ns,)))
        slot0 += N + Ns

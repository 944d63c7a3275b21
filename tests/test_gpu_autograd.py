"""The torch.autograd binding (nwamd.autograd): gradients through WormRun
equal nwAdjointBackward's, and the bond-order loss S(final positions) -- the
nematic-order objective -- differentiates to FP64 central differences."""
import ctypes

import numpy as np
import pytest

import adjoint_case as case
from adjoint_ref import Model
from test_gpu_adjoint import _bond_order_S, fp64_model, gpu_gradient

torch = pytest.importorskip("torch")
from nwamd.autograd import WormRun, bond_order  # noqa: E402

pytestmark = pytest.mark.gpu


def run_and_theta(steps=40, K=7):
    p, e = case.system()
    run = WormRun(p, e, p.numFilaments, case.DT, case.KT / case.DT, case.SEED, case.OFFSET, steps, K)
    theta = torch.tensor(case.THETA, dtype=torch.float64, requires_grad=True)
    return run, theta


def test_autograd_matches_the_adjoint_abi(nwlib):
    pos, vel, _, w = case.initial_state()
    run, theta = run_and_theta()
    P = run(theta, torch.from_numpy(pos).cuda(), torch.from_numpy(vel).cuda())
    loss = (P.double() * torch.from_numpy(w).double().cuda()).sum()
    loss.backward()
    (direct,), _ = gpu_gradient(nwlib, 40, 7)
    np.testing.assert_array_equal(theta.grad.numpy(), direct)


def test_bond_order_objective_gradient(nwlib):
    pos, vel, zero, _ = case.initial_state()
    run, theta = run_and_theta()
    P = run(theta, torch.from_numpy(pos).cuda(), torch.from_numpy(vel).cuda())
    S = bond_order(P, run)
    S.backward()
    g = theta.grad.numpy()
    p, _ = case.system()
    fd = np.zeros(4)
    for k in range(4):
        h = 1e-4 * abs(case.THETA[k])
        vals = []
        for sgn in (1, -1):
            t = list(case.THETA)
            t[k] += sgn * h
            rf, *_ = fp64_model(t).run(pos, vel, zero, zero, t, 40)
            vals.append(_bond_order_S(p, rf, p.numFilaments))
        fd[k] = (vals[0] - vals[1]) / (2 * h)
    print(f"S-loss gradient {g} fd {fd} rel err {np.abs(g - fd) / np.abs(fd)}", flush=True)
    np.testing.assert_allclose(g, fd, rtol=1e-3, atol=1e-4 * np.abs(fd).max())


def _replica_srd_setup():
    import test_gpu_batch_adjoint as tb
    from nwamd import abi
    R = len(tb.THETAS)
    ins = [tb.replica_inputs(k, True) for k in range(R)]
    reps = (abi.NWReplicaParams * R)()
    sol = (abi.NWReplicaSolvent * R)()
    for k, x in enumerate(ins):
        reps[k].params, reps[k].ext = x[0], x[1]
        reps[k].seed, reps[k].offset = tb.SEEDS[k]
        sol[k].solvent, sol[k].numSolvent = x[2], len(x[5])
    steps = 23
    shifts = np.stack([tb.shifts_of(k, steps, ins[k][2]) for k in range(R)], axis=1)
    P0 = np.concatenate([x[3] for x in ins] + [x[5] for x in ins])
    V0 = np.concatenate([x[4] for x in ins] + [x[6] for x in ins])
    return tb, ins, reps, sol, shifts, steps, P0, V0


def test_replica_srd_autograd_matches_the_batch_abi(nwlib):
    """ReplicaSrdRun: d loss / d f_a of every replica equals
    nwBatchAdjointBackward's, for a loss on the beads' final positions and
    velocities."""
    from nwamd.autograd import ReplicaSrdRun
    tb, ins, reps, sol, shifts, steps, P0, V0 = _replica_srd_setup()
    run = ReplicaSrdRun(reps, sol, case.DT, case.SRD_GAP, shifts, steps, 7)
    fa = torch.tensor([t[0] for t in tb.THETAS], dtype=torch.float64, requires_grad=True)
    pos, vel = run(fa, torch.from_numpy(P0).cuda(), torch.from_numpy(V0).cuda())
    N = run.bead_off[-1]
    w = torch.from_numpy(np.concatenate([x[7] for x in ins])).cuda().double()
    wv = torch.from_numpy(np.concatenate([x[8] for x in ins])).cuda().double()
    loss = (pos[:N].double() * w).sum() + (vel[:N].double() * wv).sum()
    loss.backward()
    direct = tb.batched(nwlib, True, steps, 7)
    np.testing.assert_array_equal(fa.grad.numpy(), np.array([g[0] for _, g in direct]))
    for k, ((dP, dV), _) in enumerate(direct):
        rows = np.r_[run.bead_off[k]:run.bead_off[k + 1]]
        assert np.array_equal(pos[rows].detach().cpu().numpy(), dP[:len(rows)])


def test_replica_srd_polar_speed_loss(nwlib):
    """loss = sum over replicas of (polar speed - 0.5)^2 through the torch
    binding against nwBatchPolarSpeedGrad + nwBatchAdjointBackward scaled by
    2 (P - 0.5) per replica."""
    import ctypes as C
    from nwamd.abi import DeviceArray
    from nwamd.autograd import ReplicaSrdRun, polar_speed
    tb, ins, reps, sol, shifts, steps, P0, V0 = _replica_srd_setup()
    run = ReplicaSrdRun(reps, sol, case.DT, case.SRD_GAP, shifts, steps, 7)
    fa = torch.tensor([t[0] for t in tb.THETAS], dtype=torch.float64, requires_grad=True)
    pos, vel = run(fa, torch.from_numpy(P0).cuda(), torch.from_numpy(V0).cuda())
    P = polar_speed(pos, vel, run)
    ((P - 0.5) ** 2).sum().backward()
    # the same by hand through the C-ABI (the batch and its run are still the last ones)
    lib = nwlib
    N = run.bead_off[-1]
    dPp, dPv = DeviceArray(lib, 16 * len(P0)), DeviceArray(lib, 16 * len(P0))
    pd, vd = DeviceArray.from_numpy(lib, pos.detach().cpu().numpy()), DeviceArray.from_numpy(lib, vel.detach().cpu().numpy())
    Pr = (C.c_double * run.R)()
    lib.nwBatchPolarSpeedGrad(pd.p, vd.p, dPp.p, dPv.p, Pr)
    assert list(Pr) == P.detach().tolist()
    g = (C.c_double * (4 * run.R))()
    lib.nwBatchAdjointBackward(dPp.p, dPv.p, g)
    want = 2.0 * (np.array(Pr[:]) - 0.5) * np.array(g[:])[0::4]
    np.testing.assert_allclose(fa.grad.numpy(), want, rtol=1e-5)
    assert np.all(np.abs(want) > 0)

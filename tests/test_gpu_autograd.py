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
    np.testing.assert_allclose(g, fd, rtol=5e-3, atol=5e-3 * np.abs(fd).max())

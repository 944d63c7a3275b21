"""The adjoint of the fused MD step on the GPU (nwAdjointForward /
nwAdjointBackward, FP32) against the FP64 reference gradient and FP64
central finite differences (h = 1e-4 relative), on the small dry-worm system
of adjoint_case.py; the bond-order loss gradient against FP64 differences."""
import ctypes

import numpy as np
import pytest

import adjoint_case as case
from adjoint_ref import Model
from nwamd import abi
from nwamd.abi import DeviceArray
from nwtest import random_worms, sim_params

pytestmark = pytest.mark.gpu


def fp64_model(theta=case.THETA):
    p, e = case.system(theta)
    return Model(p, e, p.numFilaments, case.DT, case.KT / case.DT, True, case.SEED, case.OFFSET)


def gpu_gradient(lib, steps, K, theta=case.THETA, twice=False):
    p, e = case.system(theta)
    lib.setParameters(ctypes.byref(p))
    lib.nwSetExtParameters(ctypes.byref(e))
    lib.nwRngSetSeed(case.SEED, case.OFFSET)
    pos, vel, zero, w = case.initial_state()
    d = [DeviceArray.from_numpy(lib, a) for a in (pos, vel, zero, zero, zero)]
    dW = DeviceArray.from_numpy(lib, w)
    W = p.numFilaments
    lib.nwAdjointForward(d[0].p, d[1].p, d[2].p, d[3].p, d[4].p, W, case.DT, theta[3], case.KT / case.DT, 1,
                         steps, K)
    seed, off_fwd = ctypes.c_uint64(), ctypes.c_uint64()
    lib.nwRngGetState(ctypes.byref(seed), ctypes.byref(off_fwd))
    g = (ctypes.c_double * 4)()
    lib.nwAdjointBackward(dW.p, None, g)
    out = [np.array(g)]
    if twice:
        lib.nwAdjointBackward(dW.p, None, g)
        out.append(np.array(g))
    off_bwd = ctypes.c_uint64()
    lib.nwRngGetState(ctypes.byref(seed), ctypes.byref(off_bwd))
    assert off_bwd.value == off_fwd.value == case.OFFSET + steps * p.numFilaments * p.filamentSize
    return out, d[0].download(np.float32, (W * p.filamentSize, 4))


def test_adjoint_matches_fp64_reference_and_differences(nwlib):
    steps = 40
    (g, g2), final = gpu_gradient(nwlib, steps, 7, twice=True)
    assert np.array_equal(g, g2)  # deterministic
    m = fp64_model()
    pos, vel, zero, w = case.initial_state()
    rf, *_ = m.run(pos, vel, zero, zero, case.THETA, steps)
    np.testing.assert_allclose(final[:, :3], rf, atol=2e-5)  # FP32 trajectory tracks FP64
    ref = m.gradient(pos, vel, zero, zero, case.THETA, steps, w)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(g, ref, rtol=2e-3, atol=2e-4 * scale)
    # central differences of the FP64 forward, h = 1e-4 relative
    fd = np.zeros(4)
    for k in range(4):
        h = 1e-4 * abs(case.THETA[k])
        tp, tm = list(case.THETA), list(case.THETA)
        tp[k] += h
        tm[k] -= h
        fd[k] = (fp64_model(tp).loss_linear(pos, vel, zero, zero, tp, steps, w) -
                 fp64_model(tm).loss_linear(pos, vel, zero, zero, tm, steps, w)) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-3, atol=2e-4 * scale)


def test_adjoint_is_independent_of_the_checkpoint_spacing(nwlib):
    """Segments are recomputed bit for bit, so the spacing only regroups the
    FP64 parameter sums."""
    (a,), _ = gpu_gradient(nwlib, 24, 1)
    (b,), _ = gpu_gradient(nwlib, 24, 5)
    (c,), _ = gpu_gradient(nwlib, 24, 24)
    np.testing.assert_allclose(a, b, rtol=1e-12)
    np.testing.assert_allclose(a, c, rtol=1e-12)


def _bond_order_S(p, pos, W):
    n = p.filamentSize
    r = np.asarray(pos, np.float64)[:, :3].reshape(W, n, 3)
    L = np.array([p.boxSize.x, p.boxSize.y, p.boxSize.z])
    d = r[:, 1:] - r[:, :-1]
    d -= L * np.round(d / L)
    b = (d / np.linalg.norm(d, axis=-1, keepdims=True)).reshape(-1, 3)
    Q = 1.5 * (b[:, :, None] * b[:, None, :]).mean(0) - 0.5 * np.eye(3)
    return np.linalg.eigvalsh(Q)[-1]


def test_bond_order_gradient_matches_differences(nwlib):
    rng = np.random.default_rng(4)
    p = sim_params(grid=(30, 30, 1), cell=1.0, W=40, n=6)
    pos, _, _ = random_worms(p, rng, twod=True, kink=0.6)
    lib = nwlib
    lib.setParameters(ctypes.byref(p))
    dP = DeviceArray.from_numpy(lib, pos)
    dG = DeviceArray(lib, 16 * len(pos))
    S = lib.nwBondOrderGrad(dP.p, p.numFilaments, dG.p)
    g = dG.download(np.float32, (len(pos), 4))
    assert S == pytest.approx(_bond_order_S(p, pos, p.numFilaments), abs=1e-9)
    pick = rng.choice(len(pos), 12, replace=False)
    base = pos.astype(np.float64)
    for i in pick:
        for a in range(2):
            h = 1e-4
            pp, pm = base.copy(), base.copy()
            pp[i, a] += h
            pm[i, a] -= h
            fd = (_bond_order_S(p, pp, p.numFilaments) - _bond_order_S(p, pm, p.numFilaments)) / (2 * h)
            assert g[i, a] == pytest.approx(fd, rel=2e-3, abs=2e-6), (i, a)
    assert np.all(g[:, 3] == 0)

"""The adjoint of the fused MD step on the GPU (nwAdjointForward /
nwAdjointBackward, FP32) against the FP64 reference gradient and FP64
central finite differences (h = 1e-4 relative), on the small dry-worm system
of adjoint_case.py, dry and with SRD collisions (nwAdjointForwardSrd); the
bond-order and polar-speed loss gradients against FP64 differences."""
import ctypes

import numpy as np
import pytest

import adjoint_case as case
from adjoint_ref import Model, Srd
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


# ---- with SRD collisions ---------------------------------------------------
def srd_fp64_model(steps, theta=case.THETA):
    p, e, s = case.srd_system(theta)
    srd = Srd(s, s.numParticles, case.SRD_GAP, case.srd_shifts(steps), e.srdCollideFilaments)
    return Model(p, e, p.numFilaments, case.DT, 0.0, False, case.SEED, case.OFFSET, srd=srd)


def srd_setup(lib, theta=case.THETA):
    p, e, s = case.srd_system(theta)
    lib.setParameters(ctypes.byref(p))
    lib.setSolventParameters(ctypes.byref(s))
    lib.nwSetExtParameters(ctypes.byref(e))
    lib.nwRngSetSeed(case.SEED, case.OFFSET)
    pos, vel, zero, w, wv = case.srd_initial_state()
    d = [DeviceArray.from_numpy(lib, a) for a in (pos, vel, zero, zero, zero)]
    return p, s, d, (pos, vel, zero, w, wv)


def srd_gpu_gradient(lib, steps, K):
    p, s, d, (pos, vel, zero, w, wv) = srd_setup(lib)
    sh = np.ascontiguousarray(case.srd_shifts(steps).reshape(-1))
    lib.nwAdjointForwardSrd(d[0].p, d[1].p, d[2].p, d[3].p, d[4].p, p.numFilaments, s.numParticles, case.DT,
                            case.SRD_GAP, sh.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), steps, K)
    NT = len(pos)
    final = (d[0].download(np.float32, (NT, 4)), d[1].download(np.float32, (NT, 4)))
    dW, dWv = DeviceArray.from_numpy(lib, w), DeviceArray.from_numpy(lib, wv)
    g = (ctypes.c_double * 4)()
    lib.nwAdjointBackward(dW.p, dWv.p, g)
    return np.array(g), final


def test_srd_adjoint_forward_is_the_fused_step_sequence(nwlib):
    """nwAdjointForwardSrd runs nwFusedMdStep then, every gap steps,
    nwFusedSrdStep with the collision's grid shift: bit for bit."""
    lib = nwlib
    steps = 11
    _, (fp, fv) = srd_gpu_gradient(lib, steps, 4)
    p, s, d, (pos, vel, *_) = srd_setup(lib)
    N, NT, W = p.numParticles, len(pos), p.numFilaments
    sh = case.srd_shifts(steps)
    h, i = DeviceArray(lib, 4 * NT), DeviceArray(lib, 4 * NT)
    sP = DeviceArray(lib, 16 * N)
    com = DeviceArray(lib, 16 * s.numCells)
    for k in range(steps):
        lib.nwFusedMdStep(d[0].p, d[1].p, d[2].p, d[3].p, d[4].p, h.p, i.p, sP.p, None, None, case.DT, p.gamma,
                          0.0, 0, W)
        if k % case.SRD_GAP == 0:
            s.randOffset = abi.Float3(*(float(x) for x in sh[k // case.SRD_GAP]))
            lib.setSolventParameters(ctypes.byref(s))
            # the SRD time step in FP32, as ParticleSystem forms it (deltaTime * gap)
            dt_srd = float(np.float32(case.DT) * np.float32(case.SRD_GAP))
            lib.nwFusedSrdStep(d[0].p, d[1].p, com.p, None, i.p, dt_srd, N, s.numParticles, 0)
    assert np.array_equal(d[0].download(np.float32, (NT, 4)).view(np.uint32), fp.view(np.uint32))
    assert np.array_equal(d[1].download(np.float32, (NT, 4)).view(np.uint32), fv.view(np.uint32))


def test_srd_adjoint_matches_fp64_reference_and_differences(nwlib):
    """8 collisions in 40 steps; loss = sum of w . r + wv . v over the beads."""
    steps = 40
    g, (fp, fv) = srd_gpu_gradient(nwlib, steps, 7)
    m = srd_fp64_model(steps)
    pos, vel, zero, w, wv = case.srd_initial_state()
    rf, vf, *_ = m.run(pos, vel, zero, zero, case.THETA, steps)
    np.testing.assert_allclose(fp[:m.N, :3], rf[:m.N], atol=2e-5)  # FP32 trajectory tracks FP64
    np.testing.assert_allclose(fv[:, :3], vf, atol=2e-4)
    ref = m.gradient(pos, vel, zero, zero, case.THETA, steps, w, wv)
    scale = np.abs(ref).max()
    assert ref[3] == 0.0 and g[3] == 0.0  # gamma does not enter without the Langevin thermostat
    np.testing.assert_allclose(g, ref, rtol=2e-3, atol=2e-4 * scale)
    fd = np.zeros(4)
    for k in range(3):
        h = 1e-4 * abs(case.THETA[k])
        tp, tm = list(case.THETA), list(case.THETA)
        tp[k] += h
        tm[k] -= h
        fd[k] = (srd_fp64_model(steps, tp).loss_linear(pos, vel, zero, zero, tp, steps, w, wv) -
                 srd_fp64_model(steps, tm).loss_linear(pos, vel, zero, zero, tm, steps, w, wv)) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-3, atol=2e-4 * scale)


def test_srd_adjoint_is_independent_of_the_checkpoint_spacing(nwlib):
    a, _ = srd_gpu_gradient(nwlib, 24, 1)
    b, _ = srd_gpu_gradient(nwlib, 24, 5)
    c, _ = srd_gpu_gradient(nwlib, 24, 24)
    np.testing.assert_allclose(a, b, rtol=1e-12)
    np.testing.assert_allclose(a, c, rtol=1e-12)


# ---- polar speed ------------------------------------------------------------
def _polar_speed(p, pos, vel, W):
    n = p.filamentSize
    r = np.asarray(pos, np.float64)[:, :3].reshape(W, n, 3)
    v = np.asarray(vel, np.float64)[:, :3].reshape(W, n, 3)
    L = np.array([p.boxSize.x, p.boxSize.y, p.boxSize.z])
    a = np.concatenate([r[:, :1], r[:, :-2], r[:, -2:-1]], axis=1)
    b = np.concatenate([r[:, 1:2], r[:, 2:], r[:, -1:]], axis=1)
    d = b - a
    d -= L * np.round(d / L)
    t = d / np.linalg.norm(d, axis=-1, keepdims=True)
    return (v * t).sum(-1).mean()


def test_polar_speed_gradient_matches_differences(nwlib):
    rng = np.random.default_rng(8)
    p = sim_params(grid=(30, 30, 1), cell=1.0, W=40, n=6)
    pos, vel, _ = random_worms(p, rng, twod=True, kink=0.6, speed=0.5)
    lib = nwlib
    lib.setParameters(ctypes.byref(p))
    dP, dV = DeviceArray.from_numpy(lib, pos), DeviceArray.from_numpy(lib, vel)
    gP, gV = DeviceArray(lib, 16 * len(pos)), DeviceArray(lib, 16 * len(pos))
    P = lib.nwPolarSpeedGrad(dP.p, dV.p, p.numFilaments, gP.p, gV.p)
    obs = abi.NWObservables()
    lib.nwObservables(dP.p, dV.p, p.numFilaments, ctypes.byref(obs))
    assert P == obs.polarSpeed
    assert P == pytest.approx(_polar_speed(p, pos, vel, p.numFilaments), abs=1e-9)
    g, gv = gP.download(np.float32, (len(pos), 4)), gV.download(np.float32, (len(pos), 4))
    base_r, base_v = pos.astype(np.float64), vel.astype(np.float64)
    # every position of the first two worms (both ends and the interior), a sample of the rest
    pick = list(range(12)) + list(rng.choice(np.arange(12, len(pos)), 12, replace=False))
    for i in pick:
        for a in range(2):
            h = 1e-4
            pp, pm = base_r.copy(), base_r.copy()
            pp[i, a] += h
            pm[i, a] -= h
            fd = (_polar_speed(p, pp, base_v, p.numFilaments) - _polar_speed(p, pm, base_v, p.numFilaments)) / (2 * h)
            assert g[i, a] == pytest.approx(fd, rel=2e-3, abs=2e-7), (i, a)
            vp, vm = base_v.copy(), base_v.copy()
            vp[i, a] += h
            vm[i, a] -= h
            fdv = (_polar_speed(p, base_r, vp, p.numFilaments) - _polar_speed(p, base_r, vm, p.numFilaments)) / (2 * h)
            assert gv[i, a] == pytest.approx(fdv, rel=2e-3, abs=2e-7), (i, a)
    assert np.all(g[:, 3] == 0) and np.all(gv[:, 3] == 0)


# ---- the dry differentiation workload's shape ----------------------------------
def test_config3_shaped_gradient_matches_differences(nwlib):
    """50 worms x 16 beads in a 32 x 32 box (the bead density of 2,000 x 16
    beads in 200 x 200), rows of straight worms with random polarity, 200
    steps with checkpoints every 20, loss = the final bond-tangent nematic
    order S: d S / d (f_a, kappa, epsilon, gamma) of the FP32 adjoint against
    FP64 central differences of the FP64 forward (h = 1e-4 relative) at
    1e-3."""
    import config_cases as cc
    lib = nwlib
    p, e = cc.config3_system()
    pos, vel = cc.config3_state()
    W, N = p.numFilaments, len(pos)
    steps, K = 200, 20
    theta = cc.C3_THETA
    lib.setParameters(ctypes.byref(p))
    lib.nwSetExtParameters(ctypes.byref(e))
    lib.nwRngSetSeed(cc.C3_SEED, cc.C3_OFFSET)
    zero = np.zeros((N, 4), np.float32)
    d = [DeviceArray.from_numpy(lib, a) for a in (pos, vel, zero, zero, zero)]
    lib.nwAdjointForward(d[0].p, d[1].p, d[2].p, d[3].p, d[4].p, W, cc.DT, theta[3], cc.KT / cc.DT, 1, steps, K)
    dS = DeviceArray(lib, 16 * N)
    S = lib.nwBondOrderGrad(d[0].p, W, dS.p)
    g = (ctypes.c_double * 4)()
    lib.nwAdjointBackward(dS.p, None, g)
    g = np.array(g)

    def S64(t):
        m = Model(p, e, W, cc.DT, cc.KT / cc.DT, True, cc.C3_SEED, cc.C3_OFFSET)
        rf, *_ = m.run(pos, vel, zero, zero, t, steps)
        return cc.bond_order_S(p, rf, W)

    assert abs(S - S64(theta)) < 1e-5
    fd = np.zeros(4)
    for k in range(4):
        h = 1e-4 * abs(theta[k])
        tp, tm = list(theta), list(theta)
        tp[k] += h
        tm[k] -= h
        fd[k] = (S64(tp) - S64(tm)) / (2 * h)
    rel = np.abs(g - fd) / np.abs(fd)
    print(f"config-3 shaped: S {S:.6f} grad {g} fd {fd} rel err {rel}", flush=True)
    assert np.all(np.abs(fd) > 0)
    np.testing.assert_allclose(g, fd, rtol=1e-3, atol=1e-4 * np.abs(fd).max())

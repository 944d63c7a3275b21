"""The FP64 reference gradient (oracle/adjoint_ref.py) pinned by central
finite differences of its own forward map, on the small dry-worm system of
adjoint_case.py (bonds, bending, drive, WCA between rows, Langevin with
replayed noise)."""
import numpy as np
import pytest

import adjoint_case as case
from adjoint_ref import Model, Srd


def model(theta):
    p, e = case.system(theta)
    return Model(p, e, p.numFilaments, case.DT, case.KT / case.DT, True, case.SEED, case.OFFSET)


def test_reference_gradient_matches_central_differences():
    pos, vel, zero, w = case.initial_state()
    steps = 12
    g = model(case.THETA).gradient(pos, vel, zero, zero, case.THETA, steps, w)
    fd = np.zeros(4)
    for m in range(4):
        h = 1e-6 * abs(case.THETA[m])
        tp, tm = list(case.THETA), list(case.THETA)
        tp[m] += h
        tm[m] -= h
        fd[m] = (model(tp).loss_linear(pos, vel, zero, zero, tp, steps, w) -
                 model(tm).loss_linear(pos, vel, zero, zero, tm, steps, w)) / (2 * h)
    assert np.all(np.abs(g) > 0)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-7 * np.abs(fd).max())


def srd_model(theta, steps):
    p, e, s = case.srd_system(theta)
    srd = Srd(s, s.numParticles, case.SRD_GAP, case.srd_shifts(steps), e.srdCollideFilaments)
    return Model(p, e, p.numFilaments, case.DT, 0.0, False, case.SEED, case.OFFSET, srd=srd)


def test_reference_gradient_with_srd_matches_central_differences():
    """Collisions after steps 0, 5, 10 mix the beads' velocities with the
    solvent's; the loss weighs the final positions and velocities."""
    pos, vel, zero, w, wv = case.srd_initial_state()
    steps = 12
    g = srd_model(case.THETA, steps).gradient(pos, vel, zero, zero, case.THETA, steps, w, wv)
    fd = np.zeros(4)
    for m in range(3):  # no Langevin thermostat with the solvent: gamma does not enter
        h = 1e-6 * abs(case.THETA[m])
        tp, tm = list(case.THETA), list(case.THETA)
        tp[m] += h
        tm[m] -= h
        fd[m] = (srd_model(tp, steps).loss_linear(pos, vel, zero, zero, tp, steps, w, wv) -
                 srd_model(tm, steps).loss_linear(pos, vel, zero, zero, tm, steps, w, wv)) / (2 * h)
    assert np.all(np.abs(g[:3]) > 0) and g[3] == 0.0
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-7 * np.abs(fd).max())


def test_srd_collision_conserves_cell_momentum_and_speeds():
    """The collision map of the reference model: every cell's momentum is
    kept and every rotated particle keeps its speed relative to the cell."""
    pos, vel, zero, w, wv = case.srd_initial_state()
    m = srd_model(case.THETA, 1)
    N = m.N
    v = np.asarray(vel, np.float64)
    r = np.asarray(pos, np.float64)
    mass = v[:, 3]
    vb, vs, col, off = m.collide(0, r[:N, :3], v[:N, :3], r[N:, :3], v[N:, :3], mass, case.OFFSET)
    V0, V1 = v[:, :3], np.concatenate([vb, vs])
    for a in range(3):
        np.testing.assert_allclose(np.bincount(col.h, mass * V1[:, a], col.Cs),
                                   np.bincount(col.h, mass * V0[:, a], col.Cs), atol=1e-12)
    u = np.stack([np.bincount(col.h, mass * V0[:, a], col.Cs) for a in range(3)], 1) / np.maximum(col.M, 1e-300)[:, None]
    np.testing.assert_allclose(np.linalg.norm(V1 - u[col.h], axis=1), np.linalg.norm(V0 - u[col.h], axis=1),
                               rtol=1e-12)
    # the relative velocity turns by +-alpha (alpha: pi/2 rounded to FP32)
    d0, d1 = (V0 - u[col.h])[:, :2], (V1 - u[col.h])[:, :2]
    np.testing.assert_allclose((d0 * d1).sum(1), np.cos(np.float32(np.pi / 2)) * (d0 * d0).sum(1), atol=1e-14)
    assert off == case.OFFSET + (col.Cs + 3) // 4


def test_tree_pairs_equal_all_pairs_forces():
    """The reference model's real-valued forces take the WCA pairs from a
    periodic k-d tree; over all pairs (the complex-step path) the same forces
    result, including worms across the periodic faces."""
    import config_cases as cc
    p, e = cc.config3_system()
    pos, vel = cc.config3_state()
    pos = pos.copy()
    pos[:, :2] += np.float32(15.7)  # push rows across the faces, then wrap
    L = p.boxSize.x
    pos[:, :2] = (np.mod(pos[:, :2] + L / 2, L) - L / 2).astype(np.float32)
    m = Model(p, e, p.numFilaments, cc.DT, 0.0, False, 1, 0)
    r = pos[:, :3].astype(np.float64)
    z = np.zeros_like(r)
    f_tree = m.forces(r, vel[:, :3].astype(np.float64), z, cc.C3_THETA)
    f_all = m.forces(r.astype(np.complex128), vel[:, :3].astype(np.float64), z, cc.C3_THETA).real
    i, j = m.pairs(r)
    assert len(i) > 100  # neighbouring rows interact
    np.testing.assert_allclose(f_tree, f_all, rtol=1e-12, atol=1e-12)


def test_polar_speed_gradient_matches_differences():
    import config_cases as cc
    rng = np.random.default_rng(12)
    p, _ = cc.dry_system(7, 5, 12.0, (1.0, 5.0, 1.0, 2.0))
    pos, vel = cc.lattice_worms(p, rng, jitter=0.2, speed=0.5)
    pos[:, :2] += np.float32(5.0)  # worms across the faces
    P, dr, dv = cc.polar_speed_and_grad(p, pos, vel, p.numFilaments)
    base_r, base_v = pos.astype(np.float64), vel.astype(np.float64)
    h = 1e-6
    for i in range(len(pos)):
        for a in range(2):
            pp, pm = base_r.copy(), base_r.copy()
            pp[i, a] += h
            pm[i, a] -= h
            fd = (cc.polar_speed_and_grad(p, pp, base_v, p.numFilaments)[0] -
                  cc.polar_speed_and_grad(p, pm, base_v, p.numFilaments)[0]) / (2 * h)
            assert abs(dr[i, a] - fd) <= 1e-7 + 1e-6 * abs(fd), (i, a, dr[i, a], fd)
            vp, vm = base_v.copy(), base_v.copy()
            vp[i, a] += h
            vm[i, a] -= h
            fdv = (cc.polar_speed_and_grad(p, base_r, vp, p.numFilaments)[0] -
                   cc.polar_speed_and_grad(p, base_r, vm, p.numFilaments)[0]) / (2 * h)
            assert abs(dv[i, a] - fdv) <= 1e-9 + 1e-6 * abs(fdv)


def test_order_parameter_closed_forms():
    """Straight worms along x: S = 1 whatever their polarity; the polar speed
    of beads moving along their tangent at speed u is u, against it -u."""
    import config_cases as cc
    rng = np.random.default_rng(13)
    p, _ = cc.dry_system(20, 8, 30.0, (1.0, 5.0, 1.0, 2.0))
    pos, vel = cc.lattice_worms(p, rng, jitter=0.0, speed=0.0)
    assert cc.bond_order_S(p, pos, p.numFilaments) == pytest.approx(1.0, abs=1e-12)
    n = p.filamentSize
    r = pos[:, :3].reshape(p.numFilaments, n, 3).astype(np.float64)
    heading = np.sign(r[:, -1, 0] - r[:, 0, 0])
    vel = np.zeros_like(pos)
    vel[:, 0] = np.repeat(heading, n) * 0.7
    P, _, _ = cc.polar_speed_and_grad(p, pos, vel, p.numFilaments)
    assert P == pytest.approx(0.7, abs=1e-6)
    P2, _, _ = cc.polar_speed_and_grad(p, pos, -vel, p.numFilaments)
    assert P2 == pytest.approx(-0.7, abs=1e-6)

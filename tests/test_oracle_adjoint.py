"""The FP64 reference gradient (oracle/adjoint_ref.py) pinned by central
finite differences of its own forward map, on the small dry-worm system of
adjoint_case.py (bonds, bending, drive, WCA between rows, Langevin with
replayed noise)."""
import numpy as np

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

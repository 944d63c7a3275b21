"""The FP64 reference gradient (oracle/adjoint_ref.py) pinned by central
finite differences of its own forward map, on the small dry-worm system of
adjoint_case.py (bonds, bending, drive, WCA between rows, Langevin with
replayed noise)."""
import numpy as np

import adjoint_case as case
from adjoint_ref import Model


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

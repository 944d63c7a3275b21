"""The small dry-worm system the adjoint tests share: 5 worms x 8 beads in
rows 1.0 apart (WCA active between rows), bonds, bending, drive, Langevin
(fluctuation-dissipation noise), periodic 12 x 12 box."""
import numpy as np

from nwamd import abi
from nwtest import ext_params, sim_params

THETA = (1.3, 4.0, 1.0, 2.0)  # f_a, kappa, epsilon, gamma
DT, KT = 1e-3, 0.1
SEED, OFFSET = 77, 1000


def system(theta=THETA):
    fa, kappa, eps, gamma = theta
    p = sim_params(grid=(10, 10, 1), cell=1.2, W=5, n=8, bondSpringL=0.8, activity=0.0)
    p.bondSpringK, p.bondBendingK, p.gamma, p.kbT = 57.0, kappa, gamma, KT
    e = ext_params(pairPotential=abi.NW_PAIR_WCA, bondedMode=abi.NW_BONDED_CHAIN, driveForce=fa,
                   stencil=abi.NW_STENCIL_2D)
    e.wcaEpsilon = eps
    return p, e


def initial_state(seed=5):
    p, _ = system()
    rng = np.random.default_rng(seed)
    N = p.numFilaments * p.filamentSize
    pos = np.zeros((N, 4), np.float32)
    for w in range(p.numFilaments):
        for q in range(p.filamentSize):
            pos[w * 8 + q, :3] = (-3.0 + 0.8 * q, -2.0 + 1.0 * w, p.origin.z + 0.6)
    pos[:, :2] += rng.normal(0, 0.03, (N, 2)).astype(np.float32)
    pos[:, 3] = 0.5
    vel = np.zeros((N, 4), np.float32)
    vel[:, :2] = rng.normal(0, 0.3, (N, 2))
    vel[:, 3] = 1.0
    w = rng.normal(0, 1, (N, 4))
    w[:, 3] = 0
    return pos, vel, np.zeros((N, 4), np.float32), w.astype(np.float32)

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


# ---- with SRD solvent: the same worms in a 12 x 12 SRD grid of unit cells,
# 3 solvent particles per cell, a collision after steps 0, 5, 10, ...
SRD_GAP = 5
SRD_PER_CELL = 3


def srd_system(theta=THETA):
    from nwtest import solvent_params
    p, e = system(theta)
    e.srdCollideFilaments = 1
    s = solvent_params(p, (12, 12, 1), (1.0, 1.0, p.boxSize.z), alpha=np.pi / 2, kbT=KT,
                       ns=12 * 12 * SRD_PER_CELL)
    return p, e, s


def srd_initial_state(seed=6):
    """beads then solvent (pos, vel: w = radius, mass 1), forces of the beads,
    loss weights of the beads' positions and velocities"""
    pos, vel, zero, w = initial_state()
    p, _, s = srd_system()
    rng = np.random.default_rng(seed)
    Ns = s.numParticles
    sp = np.zeros((Ns, 4), np.float32)
    sp[:, :2] = rng.random((Ns, 2)) * p.boxSize.x + p.origin.x
    sp[:, 2] = pos[0, 2]
    sp[:, 3] = 0.5
    sv = np.zeros((Ns, 4), np.float32)
    sv[:, :2] = rng.normal(0, np.sqrt(KT), (Ns, 2))
    sv[:, 3] = 1.0
    wv = rng.normal(0, 1, w.shape)
    wv[:, 3] = 0
    return np.concatenate([pos, sp]), np.concatenate([vel, sv]), zero, w, wv.astype(np.float32)


def srd_shifts(steps):
    """the grid shift of every collision of a run (ParticleSystem's centred
    shift stream)"""
    import orc
    p, _, s = srd_system()
    cell = (s.cellSize.x, s.cellSize.y, s.cellSize.z)
    return np.array([orc.srd_shift(SEED, c, True, cell) for c in range((steps + SRD_GAP - 1) // SRD_GAP)],
                    np.float32)

"""Builders for the parity cases shaped like the benchmark configurations,
at sizes the CPU oracle finishes in seconds, and the FP64 losses they use.

* dry worms for differentiation (2,000 x 16 beads in L = 200 scaled down to
  50 x 16 beads in L = 32: the same bead density, rows of straight worms
  with random polarity, WCA + bonds + bending + drive + Langevin);
* batched SRD replicas for differentiation (1,000 worms x 12 beads in
  L = 120 with 3 solvent particles per unit cell, scaled down);
* the order-parameter losses in FP64 numpy: the bond-tangent nematic S and
  the polar speed <v . t> with its gradient (tangent definitions of
  analysis/afa/utils.py:35-49 under the minimum image)."""
import math

import numpy as np

from nwamd import abi
from nwtest import ext_params, sim_params, solvent_params

DT, KT = 1e-3, 0.1


def lattice_worms(p, rng, rows_gap=None, jitter=0.02, speed=0.3):
    """Straight worms of n beads (bond 0.8) along x in rows, two-dimensional,
    random polarity, small jitter; pos.w = 0.5, vel.w = 1."""
    W, n = p.numFilaments, p.filamentSize
    L = p.boxSize.x
    per_row = max(1, int(L // (0.8 * n + 0.8)))
    rows = -(-W // per_row)
    gap = rows_gap or L / rows
    pos = np.zeros((W * n, 4), np.float32)
    for w in range(W):
        row, col = divmod(w, per_row)
        x0 = p.origin.x + 0.4 + col * (L / per_row)
        y = p.origin.y + 0.5 * gap + row * gap
        flip = rng.random() < 0.5
        for q in range(n):
            qq = n - 1 - q if flip else q
            pos[w * n + qq, :3] = (x0 + 0.8 * q, y, p.origin.z + 0.5 * p.cellSize.z)
    pos[:, :2] += rng.normal(0, jitter, (W * n, 2))
    pos[:, 3] = 0.5
    vel = np.zeros((W * n, 4), np.float32)
    vel[:, :2] = rng.normal(0, speed, (W * n, 2))
    vel[:, 3] = 1.0
    return pos, vel


def dry_system(W, n, L, theta, gamma_on=True):
    """2D periodic box of side L, MD cells >= 2^(1/6); theta = (f_a, kappa, eps, gamma)."""
    fa, kappa, eps, gamma = theta
    G = int(L / 2.0 ** (1.0 / 6.0))
    p = sim_params(grid=(G, G, 1), cell=L / G, W=W, n=n, bondSpringL=0.8, activity=0.0)
    p.bondSpringK, p.bondBendingK, p.gamma, p.kbT = 57.0, kappa, gamma, KT
    e = ext_params(pairPotential=abi.NW_PAIR_WCA, bondedMode=abi.NW_BONDED_CHAIN, driveForce=fa,
                   stencil=abi.NW_STENCIL_2D)
    e.wcaEpsilon = eps
    return p, e


# ---- config-3 shape: 50 worms x 16 beads in L = 32 (config 3's density) -------
C3_W, C3_N, C3_L = 50, 16, 32.0
C3_THETA = (1.0, 5.0, 1.0, 2.0)
C3_SEED, C3_OFFSET = 4242, 0


def config3_system(theta=C3_THETA):
    return dry_system(C3_W, C3_N, C3_L, theta)


def config3_state(seed=3):
    p, _ = config3_system()
    return lattice_worms(p, np.random.default_rng(seed))


# ---- config-4 shape: SRD replicas, 3 solvent per unit cell --------------------
def srd_replica(W, n, L, fa, seed):
    """params, ext, solvent of one replica (unit SRD cells, 90 degrees, beads
    join the collisions, no Langevin thermostat) and its initial beads and
    solvent (uniform positions in SRD-cell order, Maxwellian kT)."""
    p, e = dry_system(W, n, L, (fa, 5.0, 1.0, 2.0))
    e.srdCollideFilaments, e.srdThermostat = 1, abi.NW_SRD_THERMO_NONE
    Gs = int(round(L))
    s = solvent_params(p, (Gs, Gs, 1), (L / Gs, L / Gs, p.boxSize.z), alpha=math.pi / 2, kbT=KT,
                       ns=3 * Gs * Gs)
    rng = np.random.default_rng(seed)
    bp, bv = lattice_worms(p, rng)
    Ns = s.numParticles
    sp = np.zeros((Ns, 4), np.float32)
    sp[:, :2] = rng.random((Ns, 2)) * L + p.origin.x
    sp[:, 2] = bp[0, 2]
    sp[:, 3] = 0.5
    cell = np.floor((sp[:, 1] - p.origin.y) / s.cellSize.y) * Gs + np.floor((sp[:, 0] - p.origin.x) / s.cellSize.x)
    sp = sp[np.argsort(cell, kind="stable")]
    sv = np.zeros((Ns, 4), np.float32)
    sv[:, :2] = rng.normal(0, math.sqrt(KT), (Ns, 2))
    sv[:, 3] = 1.0
    return p, e, s, bp, bv, sp, sv


# ---- FP64 losses ---------------------------------------------------------------
def _mi(d, L):
    return d - L * np.round(d / L)


def box_of(p):
    return np.array([p.boxSize.x, p.boxSize.y, p.boxSize.z], np.float64)


def bond_order_S(p, pos, W):
    """largest eigenvalue of <1.5 b b - 0.5 I> over the bond tangents b"""
    n = p.filamentSize
    r = np.asarray(pos, np.float64)[:W * n, :3].reshape(W, n, 3)
    d = _mi(r[:, 1:] - r[:, :-1], box_of(p))
    b = (d / np.linalg.norm(d, axis=-1, keepdims=True)).reshape(-1, 3)
    Q = 1.5 * (b[:, :, None] * b[:, None, :]).mean(0) - 0.5 * np.eye(3)
    return float(np.linalg.eigvalsh(Q)[-1])


def polar_speed_and_grad(p, pos, vel, W):
    """P = <v . t> over the beads, t = unit(r[q+1] - r[q-1]) (one-sided at
    the ends); returns P, dP/dr (N, 3), dP/dv (N, 3)"""
    n = p.filamentSize
    N = W * n
    r = np.asarray(pos, np.float64)[:N, :3].reshape(W, n, 3)
    v = np.asarray(vel, np.float64)[:N, :3].reshape(W, n, 3)
    hi = np.concatenate([r[:, 1:], r[:, -1:]], axis=1)   # r[q+1], r[n-1] at the tail
    lo = np.concatenate([r[:, :1], r[:, :-1]], axis=1)   # r[q-1], r[0] at the head
    d = _mi(hi - lo, box_of(p))
    ln = np.linalg.norm(d, axis=-1, keepdims=True)
    t = d / ln
    P = float((v * t).sum() / N)
    dPdv = t / N
    g = (v - t * (v * t).sum(-1, keepdims=True)) / ln / N   # dP / dd of every bead's tangent
    dPdr = np.zeros_like(r)
    dPdr[:, 1:] += g[:, :-1]       # d = r[q+1] - r[q-1] for interior beads; head: r[1] - r[0]
    dPdr[:, -1] += g[:, -1]        # tail: r[n-1] - r[n-2]
    dPdr[:, :-1] -= g[:, 1:]
    dPdr[:, 0] -= g[:, 0]
    return P, dPdr.reshape(N, 3), dPdv.reshape(N, 3)

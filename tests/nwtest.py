"""Shared builders for seeded worm systems used by the CPU and GPU tests."""
import math

import numpy as np

from nwamd import abi


def sim_params(grid=(16, 16, 4), cell=1.0, W=10, n=8, radius=0.5, box=None, **kw):
    """SimParams in the reference's convention (simulation/particleSystem.cpp:64-105):
    origin = -box/2, box = grid * cell unless given."""
    p = abi.SimParams()
    gx, gy, gz = grid
    cx, cy, cz = (cell, cell, cell) if np.isscalar(cell) else cell
    p.gridSize = abi.UInt3(gx, gy, gz)
    p.numCells = gx * gy * gz
    p.cellSize = abi.Float3(cx, cy, cz)
    if box is None:
        box = (gx * cx, gy * cy, gz * cz)
    p.boxSize = abi.Float3(*box)
    p.origin = abi.Float3(-box[0] / 2, -box[1] / 2, -box[2] / 2)
    p.particleRadius = radius
    p.particleMass = 1.0
    p.boundaryX = p.boundaryY = p.boundaryZ = abi.PERIODIC
    p.numFilaments = W
    p.filamentSize = n
    p.numParticles = W * n
    p.bondSpringK = 57.0
    p.bondSpringL = radius * 0.8
    p.bondBendingK = 0.7
    p.activity = 0.3
    p.reverseProbability = 0.0
    p.kbT = 1.0
    p.gamma = 0.5
    p.spring = 0.5
    p.damping = 0.02
    p.shear = 0.1
    p.attraction = 0.0
    p.boundaryDamping = -0.5
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def ext_params(**kw):
    e = abi.NWExtParams()
    e.wcaEpsilon = 1.0
    e.wcaSigma = 1.0
    e.wcaCutoff = 2.0 ** (1.0 / 6.0)
    for k, v in kw.items():
        setattr(e, k, v)
    return e


def solvent_params(p, srd_grid, srd_cell, alpha=2.26893, kbT=1.0, offset=(0.0, 0.0, 0.0), ns=0):
    s = abi.SolventParams()
    s.enabled = True
    s.gridSize = abi.UInt3(*srd_grid)
    s.numCells = srd_grid[0] * srd_grid[1] * srd_grid[2]
    s.cellSize = abi.Float3(*srd_cell)
    s.randOffset = abi.Float3(*offset)
    s.numParticles = ns
    s.alpha = alpha
    s.kbT = kbT
    s.particleMass = 1.0
    return s


def box_of(p):
    return np.array([p.boxSize.x, p.boxSize.y, p.boxSize.z], np.float64)


def origin_of(p):
    return np.array([p.origin.x, p.origin.y, p.origin.z], np.float64)


def wrap_into_box(p, xyz):
    o, L = origin_of(p), box_of(p)
    return (o + np.mod(xyz - o, L)).astype(np.float32)


def random_worms(p, rng, bond=None, kink=0.25, speed=0.3, twod=False):
    """W worms of n beads: random head, random heading, bond-length steps with
    random kinks; pos.w = radius, vel.w = mass.  Returns pos, vel, tangent (float4)."""
    W, n = p.numFilaments, p.filamentSize
    bond = p.bondSpringL if bond is None else bond
    o, L = origin_of(p), box_of(p)
    pos = np.zeros((W * n, 4), np.float32)
    for w in range(W):
        head = o + rng.random(3) * L
        if twod:
            head[2] = o[2] + 0.5 * p.cellSize.z
        th = rng.random() * 2 * math.pi
        ph = 0.0 if twod else (rng.random() - 0.5) * 0.8
        r = head.copy()
        for q in range(n):
            pos[w * n + q, :3] = r
            th += (rng.random() - 0.5) * 2 * kink
            if not twod:
                ph += (rng.random() - 0.5) * kink
            step = bond * (0.9 + 0.2 * rng.random())
            r = r + step * np.array([math.cos(th) * math.cos(ph), math.sin(th) * math.cos(ph), math.sin(ph)])
    pos[:, :3] = wrap_into_box(p, pos[:, :3])
    pos[:, 3] = p.particleRadius
    vel = np.zeros((W * n, 4), np.float32)
    vel[:, :3] = rng.normal(0, speed, (W * n, 3))
    if twod:
        vel[:, 2] = 0
    vel[:, 3] = p.particleMass
    tan = np.zeros((W * n, 4), np.float32)
    t = rng.normal(0, 1, (W * n, 3))
    if twod:
        t[:, 2] = 0
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    tan[:, :3] = t
    return pos, vel, tan


def random_forces(rng, n, scale=1.0):
    f = np.zeros((n, 4), np.float32)
    f[:, :3] = rng.normal(0, scale, (n, 3))
    return f


def rel_err(gpu, ref, scale):
    """max |gpu - ref| / scale, scale = per-row magnitude floor."""
    d = np.abs(np.asarray(gpu, np.float64) - np.asarray(ref, np.float64))
    return float(np.max(d / scale)) if d.size else 0.0


def force_tol_check(gpu, ref, rtol=1e-5, floor=None, name="force"):
    """Per-bead check of FP32 results against the FP64 oracle:
    |F_gpu - F_oracle| <= rtol * (|F_oracle| + floor), floor defaulting to the
    mean |F_oracle| of the system (a bead whose terms cancel to well below the
    typical force is judged against the typical force)."""
    g = np.asarray(gpu, np.float64)[:, :3]
    r = np.asarray(ref, np.float64)[:, :3]
    err = np.linalg.norm(g - r, axis=1)
    mag = np.linalg.norm(r, axis=1)
    if floor is None:
        floor = float(np.mean(mag)) if mag.size else 0.0
    bound = rtol * (mag + floor) + 1e-30
    worst = int(np.argmax(err / bound))
    assert np.all(err <= bound), (f"{name}: bead {worst} |err| {err[worst]:.3e} > bound {bound[worst]:.3e} "
                                  f"(gpu {g[worst]}, oracle {r[worst]})")
    return float(np.max(err / bound))


def lattice_system_2d(W, n, G, rng, cell=1.1236):
    """W worms of n beads in straight rows along x (random polarity, small
    jitter) in a 2D periodic box of G x G cells of `cell` (>= the WCA
    cutoff): SimParams with bonds k = 57, r0 = 0.8, kappa = 5, and the
    pos / vel float4 arrays (w = radius 0.5 / mass 1)."""
    L = G * cell
    p = sim_params(grid=(G, G, 1), cell=cell, W=W, n=n, bondSpringL=0.8, activity=0.0)
    p.bondSpringK, p.bondBendingK = 57.0, 5.0
    per_row = max(1, int(L // (0.8 * n + 1.0)))
    rows = -(-W // per_row)
    assert rows * 1.0 <= L, "rows closer than one bead diameter"
    pos = np.zeros((W * n, 4), np.float32)
    w = np.arange(W)
    row, col = w // per_row, w % per_row
    flip = rng.random(W) < 0.5
    q = np.arange(n)
    qq = np.where(flip[:, None], n - 1 - q[None, :], q[None, :])
    pos[:, 0] = (p.origin.x + 0.5 + (col * (L / per_row))[:, None] + 0.8 * qq).reshape(-1)
    pos[:, 1] = np.repeat(p.origin.y + (row + 0.5) * (L / rows), n)
    pos[:, :2] += rng.normal(0, 0.02, (W * n, 2))
    pos[:, 2] = p.origin.z + 0.5 * cell
    pos[:, 3] = 0.5
    vel = np.zeros_like(pos)
    vel[:, :2] = rng.normal(0, 0.3, (W * n, 2))
    vel[:, 3] = 1.0
    return p, pos, vel

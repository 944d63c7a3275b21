"""FP32 error of every force term against the FP64 oracle, one term at a time.

Each case builds a system in which every checked bead feels exactly one
instance of one term, so the per-bead error is the term's own error:

  bond       harmonic chain bond k (r - r0) (AVS/NWWormsKernel_InternalForce.h:37-59): two-bead worms
  drive      f_a along the bond (AVS/NWWormsKernel_DriveForce.h:10-40): two-bead worms, k = 0
  bending    -grad(-kappa cos theta) (particles_kernel_impl.cuh:751-818): the end beads of three-bead worms
  WCA        (AVS/NWDeviceFunctions.h:73-81): isolated pairs of one-bead worms
  Langevin   -gamma v + xi (particles_kernel_impl.cuh:859-874)
  DEM        spring + dashpot + shear (collideParticles, particles_kernel_impl.cuh:427-462): isolated
             overlapping pairs
  extensile  -A (t_i - t_j) / (2 d) (extensileForce, particles_kernel_impl.cuh:477-487): isolated pairs
             with anti-aligned tangents beyond contact

The bonded terms and WCA are checked through both the reference's entry
points (filamentForces, collideFilaments) and the fused step the benchmark
runs (nwFusedMdStep with dt = 0: positions unchanged, the front's bonded
forces plus the LDS pair kernel).  The asserted bar is 1e-5 relative per bead,
unfloored, on geometry without catastrophic cancellation inside the term
(a bond not within 0.05 of its rest length, a WCA pair not within 1% of the
cutoff where 2 (s/r)^6 - 1 -> 0, a sum of parts not cancelling to below a
tenth of their magnitudes).  Every case also prints the unfloored maximum
over all of its geometry (lines starting with TERM)."""
import ctypes
import math

import numpy as np
import pytest

import orc
from nwamd import abi
from nwamd.abi import DeviceArray
from nwtest import ext_params, sim_params

pytestmark = pytest.mark.gpu

RC = 2.0 ** (1.0 / 6.0)


def dev(lib, a):
    return DeviceArray.from_numpy(lib, a)


def rel_errors(g, o, rows):
    g = np.asarray(g, np.float64)[rows, :3]
    o = np.asarray(o, np.float64)[rows, :3]
    err = np.linalg.norm(g - o, axis=1)
    mag = np.linalg.norm(o, axis=1)
    return err, mag


def report(name, path, err, mag, keep, scale=None):
    rel = err / np.maximum(mag, 1e-300)
    line = (f"TERM {name:9s} {path:8s} terms {len(err):7d}  max rel {rel[keep].max():.2e} (asserted set, "
            f"{int(keep.sum())} terms)  max rel all {rel.max():.2e}  p99.9 {np.quantile(rel, 0.999):.2e}")
    if scale is not None:
        line += f"  max |err|/scale {np.max(err / scale):.2e}"
    print(line, flush=True)
    return float(rel[keep].max())


def box_params(L, W, n, cell=1.2, **kw):
    G = int(L / cell)
    return sim_params(grid=(G, G, 1), cell=L / G, W=W, n=n, **kw)


def lattice_sites(p, count, spacing, rng):
    """count sites on a square lattice of the given spacing inside the box,
    at the mid-plane, each shifted a little at random"""
    per = int(p.boxSize.x // spacing)
    assert per * per >= count
    k = np.arange(count)
    xy = np.stack([(k % per) * spacing, (k // per) * spacing], 1) + 0.5 * spacing + p.origin.x
    xy += rng.uniform(-0.3, 0.3, xy.shape)
    z = np.full((count, 1), p.origin.z + 0.5 * p.cellSize.z)
    return np.concatenate([xy, z], 1)


def worms_from(points, p):
    pos = np.zeros((len(points), 4), np.float32)
    pos[:, :3] = points
    pos[:, 3] = p.particleRadius
    return pos


def fused_forces(lib, p, e, pos, vel=None, tan=None):
    """nwFusedMdStep with dt = 0, no Langevin: the bonded and pair forces at pos"""
    N = len(pos)
    vel = np.zeros((N, 4), np.float32) if vel is None else vel
    vel[:, 3] = 1.0
    zero = np.zeros((N, 4), np.float32)
    lib.setParameters(ctypes.byref(p))
    lib.nwSetExtParameters(ctypes.byref(e))
    d = [dev(lib, a) for a in (pos, vel, zero, zero, zero if tan is None else tan)]
    h, i = DeviceArray(lib, 4 * N), DeviceArray(lib, 4 * N)
    sp, sv, st = (DeviceArray(lib, 16 * N) for _ in range(3))
    lib.nwFusedMdStep(d[0].p, d[1].p, d[2].p, d[3].p, d[4].p, h.p, i.p, sp.p, sv.p, st.p, 0.0, 0.0, 0.0, 0,
                      p.numFilaments)
    assert np.array_equal(d[0].download(np.float32, (N, 4)), pos)
    return d[2].download(np.float32, (N, 4))


def bonded_both(lib, p, e, pos):
    """the bonded forces from filamentForces and from the fused step, and the FP64 oracle's"""
    N = len(pos)
    zero = np.zeros((N, 4), np.float32)
    lib.setParameters(ctypes.byref(p))
    lib.nwSetExtParameters(ctypes.byref(e))
    df, dt_, dp = dev(lib, zero), dev(lib, zero), dev(lib, pos)
    lib.filamentForces(df.p, dt_.p, dp.p, float(p.numFilaments))
    per_call = df.download(np.float32, (N, 4))
    fused = fused_forces(lib, p, e, pos)
    o, _ = orc.filament_forces(p, e, zero, pos, zero, p.numFilaments)
    return per_call, fused, o


def chain_ext(**kw):
    return ext_params(pairPotential=abi.NW_PAIR_WCA, bondedMode=abi.NW_BONDED_CHAIN, stencil=abi.NW_STENCIL_2D, **kw)


def test_bond_term(nwlib):
    rng = np.random.default_rng(101)
    W = 6000
    p = box_params(400.0, W, 2, bondSpringL=0.8, activity=0.0)
    p.bondSpringK, p.bondBendingK = 57.0, 0.0
    e = chain_ext(driveForce=0.0)
    a = lattice_sites(p, W, 5.0, rng)
    ln = rng.uniform(0.55, 1.05, W)
    ln[: W // 10] = 0.8 + rng.uniform(-1e-3, 1e-3, W // 10)  # near the rest length: cancellation in r - r0
    th = rng.uniform(0, 2 * np.pi, W)
    b = a + np.stack([ln * np.cos(th), ln * np.sin(th), 0 * th], 1)
    pos = worms_from(np.stack([a, b], 1).reshape(-1, 3), p)
    per_call, fused, o = bonded_both(nwlib, p, e, pos)
    for path, g in (("per-call", per_call), ("fused", fused)):
        err, mag = rel_errors(g, o, np.arange(2 * W))
        lw = np.repeat(ln, 2)
        keep = np.abs(lw - 0.8) >= 0.05
        worst = report("bond", path, err, mag, keep, scale=p.bondSpringK * lw)
        assert worst <= 1e-5


def test_drive_term(nwlib):
    rng = np.random.default_rng(102)
    W = 6000
    p = box_params(400.0, W, 2, bondSpringL=0.8, activity=0.0)
    p.bondSpringK, p.bondBendingK = 0.0, 0.0
    e = chain_ext(driveForce=1.3)
    a = lattice_sites(p, W, 5.0, rng)
    ln = rng.uniform(0.5, 1.2, W)
    th = rng.uniform(0, 2 * np.pi, W)
    b = a + np.stack([ln * np.cos(th), ln * np.sin(th), 0 * th], 1)
    pos = worms_from(np.stack([a, b], 1).reshape(-1, 3), p)
    per_call, fused, o = bonded_both(nwlib, p, e, pos)
    heads = np.arange(0, 2 * W, 2)
    for path, g in (("per-call", per_call), ("fused", fused)):
        err, mag = rel_errors(g, o, heads)
        worst = report("drive", path, err, mag, np.ones(W, bool))
        assert worst <= 1e-5
        assert np.all(g[1::2, :3] == 0.0)  # the tail bead has no bond ahead of it


def test_bending_term(nwlib):
    rng = np.random.default_rng(103)
    W = 6000
    p = box_params(400.0, W, 3, bondSpringL=0.8, activity=0.0)
    p.bondSpringK, p.bondBendingK = 0.0, 5.0
    e = chain_ext(driveForce=0.0)
    a = lattice_sites(p, W, 5.0, rng)
    l1, l2 = rng.uniform(0.6, 1.0, W), rng.uniform(0.6, 1.0, W)
    th = rng.uniform(0, 2 * np.pi, W)
    bend = rng.uniform(0.05, 2.6, W) * rng.choice([-1, 1], W)  # angle between consecutive bonds
    b = a + np.stack([l1 * np.cos(th), l1 * np.sin(th), 0 * th], 1)
    c = b + np.stack([l2 * np.cos(th + bend), l2 * np.sin(th + bend), 0 * th], 1)
    pos = worms_from(np.stack([a, b, c], 1).reshape(-1, 3), p)
    per_call, fused, o = bonded_both(nwlib, p, e, pos)
    ends = np.sort(np.concatenate([np.arange(0, 3 * W, 3), np.arange(2, 3 * W, 3)]))
    bend_e = np.abs(np.repeat(bend, 2))
    for path, g in (("per-call", per_call), ("fused", fused)):
        err, mag = rel_errors(g, o, ends)
        # the end forces are kappa / (d_ij d_jk) times the part of one bond
        # perpendicular to the other: it vanishes as the worm straightens
        keep = bend_e >= 0.2
        worst = report("bending", path, err, mag, keep)
        assert worst <= 1e-5


def pair_points(p, count, dmin, dmax, rng, spacing=5.0):
    """isolated pairs: a on a lattice of the given spacing, b at distance U[dmin, dmax)"""
    a = lattice_sites(p, count, spacing, rng)
    d = rng.uniform(dmin, dmax, count)
    th = rng.uniform(0, 2 * np.pi, count)
    b = a + np.stack([d * np.cos(th), d * np.sin(th), 0 * th], 1)
    return np.stack([a, b], 1).reshape(-1, 3), d


def percall_collide(lib, p, e, pos, vel, tan):
    N = len(pos)
    h, i = orc.calc_hash(p, None, pos)
    h, i = orc.sort_pairs(h, i)
    cs, ce, sp, sv, st = orc.reorder(h, i, pos, vel, tan, p.numCells)
    zero = np.zeros((N, 4), np.float32)
    lib.setParameters(ctypes.byref(p))
    lib.nwSetExtParameters(ctypes.byref(e))
    d = {k: dev(lib, v) for k, v in dict(sp=sp, sv=sv, st=st, i=i, cs=cs, ce=ce, f=zero).items()}
    lib.collideFilaments(d["f"].p, d["sp"].p, d["sv"].p, d["st"].p, d["i"].p, d["cs"].p, d["ce"].p, N, p.numCells)
    o = orc.collide_filaments(p, e, sp, sv, st, i, cs, ce, zero)
    return d["f"].download(np.float32, (N, 4)), o


def test_wca_term(nwlib):
    rng = np.random.default_rng(104)
    Np = 6000
    p = box_params(400.0, 2 * Np, 1, cell=RC, activity=0.0)
    pts, d = pair_points(p, Np, 0.85, RC * (1 - 1e-5), rng)
    pos = worms_from(pts, p)
    vel = np.zeros_like(pos)
    vel[:, 3] = 1.0
    e = chain_ext(driveForce=0.0)
    g_pc, o = percall_collide(nwlib, p, e, pos, vel, np.zeros_like(pos))
    g_f = fused_forces(nwlib, p, e, pos)
    dd = np.repeat(d, 2)
    for path, g in (("per-call", g_pc), ("fused", g_f)):
        err, mag = rel_errors(g, o, np.arange(2 * Np))
        keep = dd <= 0.99 * RC
        scale = 24.0 / dd ** 7  # the repulsive part's magnitude
        worst = report("WCA", path, err, mag, keep, scale=scale)
        assert worst <= 1e-5


def test_dem_term(nwlib):
    rng = np.random.default_rng(105)
    Np = 6000
    p = box_params(400.0, 2 * Np, 1, cell=1.0, activity=0.0)
    p.spring, p.damping, p.shear = 0.5, 0.02, 0.1
    pts, d = pair_points(p, Np, 0.5, 0.98, rng)
    pos = worms_from(pts, p)
    vel = np.zeros_like(pos)
    vel[:, :2] = rng.normal(0, 1.0, (2 * Np, 2))
    vel[:, 3] = 1.0
    e = ext_params(pairPotential=abi.NW_PAIR_REF_DEM, stencil=abi.NW_STENCIL_2D)
    g, o = percall_collide(nwlib, p, e, pos, vel, np.zeros_like(pos))
    err, mag = rel_errors(g, o, np.arange(2 * Np))
    # scale: the sum of the three parts' magnitudes
    rv = (vel[1::2, :3] - vel[0::2, :3]).astype(np.float64)
    scale = np.repeat(p.spring * (1.0 - d) + (p.damping + p.shear) * np.linalg.norm(rv, axis=1), 2)
    keep = mag >= 0.1 * scale
    worst = report("DEM", "per-call", err, mag, keep, scale=scale)
    assert worst <= 1e-5


def test_extensile_term(nwlib):
    rng = np.random.default_rng(106)
    Np = 3000
    # the extensile term has no cutoff of its own: every bead of the 3 x 3
    # cells of size 2 around a bead takes part, so the pairs sit 10 apart
    p = box_params(600.0, 2 * Np, 1, cell=2.0, activity=0.7)
    pts, d = pair_points(p, Np, 1.05, 1.9, rng, spacing=10.0)
    pos = worms_from(pts, p)
    vel = np.zeros_like(pos)
    vel[:, 3] = 1.0
    th = rng.uniform(0, 2 * np.pi, Np)
    dth = np.pi + rng.uniform(-0.9, 0.9, Np)  # t_i . t_j < -0.5
    tan = np.zeros_like(pos)
    tan[0::2, 0], tan[0::2, 1] = np.cos(th), np.sin(th)
    tan[1::2, 0], tan[1::2, 1] = np.cos(th + dth), np.sin(th + dth)
    e = ext_params(pairPotential=abi.NW_PAIR_REF_DEM, stencil=abi.NW_STENCIL_2D)
    g, o = percall_collide(nwlib, p, e, pos, vel, tan)
    err, mag = rel_errors(g, o, np.arange(2 * Np))
    assert np.all(mag > 0)
    worst = report("extensile", "per-call", err, mag, np.ones(2 * Np, bool))
    assert worst <= 1e-5


def test_langevin_term(nwlib):
    rng = np.random.default_rng(107)
    n = 50_000
    p = sim_params(W=n, n=1, gamma=2.0)
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0, math.sqrt(0.1), (n, 3))
    zero = np.zeros((n, 4), np.float32)
    nwlib.setParameters(ctypes.byref(p))
    seed, off = 2024, 77
    nwlib.nwRngSetSeed(seed, off)
    df, dv = dev(nwlib, zero), dev(nwlib, vel)
    gamma, kT = 2.0, 0.1 / 1e-3
    nwlib.langevinThermostat(df.p, dv.p, None, gamma, kT, n)
    g = df.download(np.float32, (n, 4))
    sd = float(np.sqrt(np.float32(2.0) * np.float32(gamma) * np.float32(kT)))
    noise = orc.random_gaussian(seed, off, 4 * n, 0.0, sd).reshape(n, 4)
    o = orc.langevin(p, zero, vel, noise)
    err, mag = rel_errors(g, o, np.arange(n))
    scale = gamma * np.linalg.norm(vel[:, :3], axis=1) + np.linalg.norm(noise[:, :3], axis=1)
    keep = mag >= 0.1 * scale
    worst = report("Langevin", "per-call", err, mag, keep, scale=scale)
    assert worst <= 1e-5

"""Parity of every C-ABI entry point (include/nw_abi.h) against the oracle.

Integer stages (hash, sort, cell bounds, permutations, reversal, uniforms) are
compared bit for bit.  Floating-point stages are compared with the FP64
oracle at the 1e-5 relative bar (force_tol_check: per bead, relative to the
bead's force floored at the system's mean force).
"""
import ctypes

import numpy as np
import pytest

import orc
from nwamd import abi
from nwamd.abi import DeviceArray
from nwtest import (ext_params, force_tol_check, random_forces, random_worms, sim_params, solvent_params)

pytestmark = pytest.mark.gpu

F4 = 16


def dev(lib, a):
    return DeviceArray.from_numpy(lib, a)


def down4(d, n):
    return d.download(np.float32, (n, 4))


def downu(d, n):
    return d.download(np.uint32, (n,))


def set_params(lib, p, e=None, s=None):
    lib.setParameters(ctypes.byref(p))
    lib.nwSetExtParameters(ctypes.byref(e if e is not None else abi.NWExtParams()))
    if s is not None:
        lib.setSolventParameters(ctypes.byref(s))


def run_hash(lib, p, pos, srd=False, s=None):
    n = pos.shape[0]
    dpos = dev(lib, pos)
    dh, di = DeviceArray(lib, 4 * n), DeviceArray(lib, 4 * n)
    lib.calcHash(dh.p, di.p, dpos.p, n, srd)
    return downu(dh, n), downu(di, n), (dh, di, dpos)


# ----------------------------------------------------------------- hashing
@pytest.mark.parametrize("grid,cell,twod", [((16, 16, 4), 1.0, False), ((64, 64, 1), 1.5625, True),
                                            ((89, 37, 1), 1.1236, True), ((8, 8, 8), 1.0, False)])
def test_calc_hash_md_bit_exact(nwlib, grid, cell, twod):
    rng = np.random.default_rng(1)
    p = sim_params(grid=grid, cell=cell, W=300, n=10)
    pos, _, _ = random_worms(p, rng, twod=twod)
    # add exact cell-boundary positions
    pos[:8, 0] = p.origin.x + np.arange(8, dtype=np.float32) * np.float32(cell)
    set_params(nwlib, p)
    h, idx, _ = run_hash(nwlib, p, pos)
    oh, oi = orc.calc_hash(p, None, pos)
    assert np.array_equal(h, oh)
    assert np.array_equal(idx, np.arange(pos.shape[0], dtype=np.uint32))
    assert h.max() < p.numCells


def face_positions(p, cell, n_per_axis, outside=False):
    """Positions on and within two ulps of the cell faces o + k c along every
    axis (quotients at or next to integers: the exact-division fallback of the
    cell index), plus, with outside, faces up to two boxes away."""
    ks = np.arange(-n_per_axis if outside else 0, n_per_axis + 1, dtype=np.float32)
    pts = []
    for a, (o, L) in enumerate(zip((p.origin.x, p.origin.y, p.origin.z), (p.boxSize.x, p.boxSize.y, p.boxSize.z))):
        g = np.float32(o) + ks * np.float32(cell[a])
        if outside:
            g = np.concatenate([g, g + np.float32(L), g - np.float32(L)])
        g = np.concatenate([g, np.nextafter(g, np.float32(np.inf)), np.nextafter(g, np.float32(-np.inf)),
                            np.nextafter(np.nextafter(g, np.float32(np.inf)), np.float32(np.inf)),
                            np.nextafter(np.nextafter(g, np.float32(-np.inf)), np.float32(-np.inf))])
        pts.append(g.astype(np.float32))
    n = max(len(x) for x in pts)
    pos = np.zeros((n, 4), np.float32)
    for a in range(3):
        pos[:, a] = np.resize(pts[a], n)
    return pos


@pytest.mark.parametrize("grid,cell", [((89, 37, 3), 1.1236), ((890, 890, 1), 1.1235955), ((33, 7, 5), 0.7)])
def test_calc_hash_md_cell_faces_bit_exact(nwlib, grid, cell):
    p = sim_params(grid=grid, cell=cell, W=10, n=10)
    pos = face_positions(p, (cell, cell, cell), max(grid), outside=True)
    set_params(nwlib, p)
    h, _, _ = run_hash(nwlib, p, pos)
    oh, _ = orc.calc_hash(p, None, pos)
    assert np.array_equal(h, oh)
    assert h.max() < p.numCells


@pytest.mark.parametrize("srd_grid,offset", [((8, 8, 1), (0.3, 1.7, 0.0)), ((100, 60, 1), (-0.5, 0.49, 0.0)),
                                             ((8, 8, 2), (1.99, 0.0, 0.5))])
def test_calc_hash_srd_bit_exact(nwlib, srd_grid, offset):
    rng = np.random.default_rng(2)
    p = sim_params(grid=(16, 16, 4), cell=1.0, W=200, n=10) if srd_grid[2] > 1 else \
        sim_params(grid=(2 * srd_grid[0], 2 * srd_grid[1], 1), cell=1.0, W=200, n=10)
    cell = (p.boxSize.x / srd_grid[0], p.boxSize.y / srd_grid[1], p.boxSize.z / srd_grid[2])
    s = solvent_params(p, srd_grid, cell, offset=offset)
    pos, _, _ = random_worms(p, rng, twod=srd_grid[2] == 1)
    set_params(nwlib, p, s=s)
    h, _, _ = run_hash(nwlib, p, pos, srd=True)
    oh, _ = orc.calc_hash(p, s, pos, srd=True)
    assert np.array_equal(h, oh)
    assert h.max() < s.numCells
    # positions on the shifted SRD cell faces and within two ulps of them
    faces = face_positions(p, cell, max(srd_grid))
    for a, off in enumerate(offset):
        faces[:, a] -= np.float32(off)
    h, _, _ = run_hash(nwlib, p, faces, srd=True)
    oh, _ = orc.calc_hash(p, s, faces, srd=True)
    assert np.array_equal(h, oh)


# ----------------------------------------------------------------- sorting
@pytest.mark.parametrize("n,bound", [(0, 16), (1, 16), (2, 3), (2049, 7), (100_000, 4096), (300_000, 1_000_000)])
def test_sort_bounded_bit_exact(nwlib, n, bound):
    rng = np.random.default_rng(n + 3)
    keys = rng.integers(0, bound, n, dtype=np.uint32)
    vals = rng.permutation(n).astype(np.uint32)
    dk, dv = dev(nwlib, keys), dev(nwlib, vals)
    nwlib.nwSortParticlesBounded(dk.p, dv.p, n, bound)
    gk, gv = downu(dk, n), downu(dv, n)
    ok, ov = orc.sort_pairs(keys, vals)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(gk, ok) and np.array_equal(gv, ov)
    assert np.array_equal(gv, vals[order])
    assert nwlib.nwGetDeviceFlags(1) & 1 == 0


@pytest.mark.parametrize("order", ["sorted", "nearly_sorted", "reversed"])
def test_sort_particles_cell_ordered_input_bit_exact(nwlib, order):
    """calcHash + sortParticles (the two-level bucket sort) on particles
    already (or nearly) in cell order, as the solvent kept in cell order
    hands them over: a first-level tile then holds every member of the
    buckets it reaches in one segment (the case the second level's member
    table is filled for in parallel); reversed, the segments come in the
    opposite tile order."""
    rng = np.random.default_rng(11)
    p = sim_params(grid=(256, 256, 1), cell=1.0, W=25_000, n=12)
    n = p.numParticles
    pos = np.zeros((n, 4), np.float32)
    pos[:, 0] = rng.uniform(p.origin.x, p.origin.x + p.boxSize.x, n)
    pos[:, 1] = rng.uniform(p.origin.y, p.origin.y + p.boxSize.y, n)
    pos[:, 2] = p.origin.z + 0.5 * p.boxSize.z
    h0, _ = orc.calc_hash(p, None, pos)
    idx = np.argsort(h0, kind="stable")
    pos = np.ascontiguousarray(pos[idx[::-1] if order == "reversed" else idx])
    if order == "nearly_sorted":
        k = rng.choice(n, n // 100, replace=False)
        pos[k, 0] = rng.uniform(p.origin.x, p.origin.x + p.boxSize.x, k.size)
    oh, oi = orc.calc_hash(p, None, pos)
    if order == "sorted":
        assert np.all(np.diff(oh.astype(np.int64)) >= 0)
    set_params(nwlib, p)
    dpos = dev(nwlib, pos)
    dh, di = DeviceArray(nwlib, 4 * n), DeviceArray(nwlib, 4 * n)
    nwlib.calcHash(dh.p, di.p, dpos.p, n, False)
    nwlib.sortParticles(dh.p, di.p, n)
    ok, ov = orc.sort_pairs(oh, oi)
    assert np.array_equal(downu(dh, n), ok) and np.array_equal(downu(di, n), ov)


def test_sort_keys_above_bound_still_sorted(nwlib):
    rng = np.random.default_rng(5)
    n = 5000
    keys = rng.integers(0, 64, n, dtype=np.uint32)
    keys[::97] = rng.integers(64, 1 << 31, keys[::97].shape[0], dtype=np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = dev(nwlib, keys), dev(nwlib, vals)
    nwlib.nwSortParticlesBounded(dk.p, dv.p, n, 64)
    ok, ov = orc.sort_pairs(keys, vals)
    assert np.array_equal(downu(dk, n), ok) and np.array_equal(downu(dv, n), ov)
    assert nwlib.nwGetDeviceFlags(1) & 1 == 1


@pytest.mark.parametrize("n,bits", [(1, 4), (3000, 8), (70_001, 20), (200_000, 32)])
def test_radix_sort_bit_exact(nwlib, n, bits):
    rng = np.random.default_rng(bits)
    hi = (1 << bits) if bits < 32 else (1 << 32)
    keys = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
    keys[: n // 3] = keys[0]  # long run of equal keys: stability
    vals = rng.permutation(n).astype(np.uint32)
    dk, dv = dev(nwlib, keys), dev(nwlib, vals)
    nwlib.nwRadixSortPairs(dk.p, dv.p, n, bits)
    ok, ov = orc.sort_pairs(keys, vals)
    assert np.array_equal(downu(dk, n), ok) and np.array_equal(downu(dv, n), ov)


def test_sort_particles_abi_after_calchash(nwlib):
    rng = np.random.default_rng(6)
    p = sim_params(grid=(32, 32, 8), cell=1.0, W=2000, n=12)
    pos, _, _ = random_worms(p, rng)
    set_params(nwlib, p)
    n = pos.shape[0]
    dpos = dev(nwlib, pos)
    dh, di = DeviceArray(nwlib, 4 * n), DeviceArray(nwlib, 4 * n)
    nwlib.calcHash(dh.p, di.p, dpos.p, n, False)
    nwlib.sortParticles(dh.p, di.p, n)
    oh, oi = orc.calc_hash(p, None, pos)
    ok, ov = orc.sort_pairs(oh, oi)
    assert np.array_equal(downu(dh, n), ok) and np.array_equal(downu(di, n), ov)
    # without a preceding calcHash on this buffer the general radix path runs
    keys = rng.integers(0, 1 << 30, n, dtype=np.uint32)
    dk, dv = dev(nwlib, keys), dev(nwlib, np.arange(n, dtype=np.uint32))
    nwlib.sortParticles(dk.p, dv.p, n)
    ok, ov = orc.sort_pairs(keys, np.arange(n, dtype=np.uint32))
    assert np.array_equal(downu(dk, n), ok) and np.array_equal(downu(dv, n), ov)


# ----------------------------------------------------------------- reorder
@pytest.mark.parametrize("with_vt", [True, False])
def test_reorder_bit_exact(nwlib, with_vt):
    rng = np.random.default_rng(7)
    p = sim_params(grid=(16, 16, 4), cell=1.0, W=300, n=8)
    pos, vel, tan = random_worms(p, rng)
    n, C = pos.shape[0], p.numCells + 100  # extra cells stay empty
    oh, oi = orc.calc_hash(p, None, pos)
    oh, oi = orc.sort_pairs(oh, oi)
    sentinel = (np.arange(C, dtype=np.uint32) * 2654435761 % 1000003).astype(np.uint32)
    d = {k: dev(nwlib, a) for k, a in dict(h=oh, i=oi, pos=pos, vel=vel, tan=tan, ce=sentinel).items()}
    cs = DeviceArray(nwlib, 4 * C)
    sp, sv, st = (DeviceArray(nwlib, F4 * n) for _ in range(3))
    nwlib.reorderDataAndFindCellStart(cs.p, d["ce"].p, sp.p, sv.p if with_vt else None, st.p if with_vt else None,
                                      d["h"].p, d["i"].p, d["pos"].p, d["vel"].p if with_vt else None,
                                      d["tan"].p if with_vt else None, n, C)
    ocs, oce, osp, osv, ost = orc.reorder(oh, oi, pos, vel if with_vt else None, tan if with_vt else None, C,
                                          cell_end_init=sentinel)
    assert np.array_equal(downu(cs, C), ocs)
    assert np.array_equal(downu(d["ce"], C), oce)  # empty cells keep their old cellEnd
    assert np.array_equal(down4(sp, n).view(np.uint32), osp.view(np.uint32))
    if with_vt:
        assert np.array_equal(down4(sv, n).view(np.uint32), osv.view(np.uint32))
        assert np.array_equal(down4(st, n).view(np.uint32), ost.view(np.uint32))


# ------------------------------------------------------------- pair forces
def sorted_system(p, pos, vel, tan):
    h, i = orc.calc_hash(p, None, pos)
    h, i = orc.sort_pairs(h, i)
    cs, ce, sp, sv, st = orc.reorder(h, i, pos, vel, tan, p.numCells)
    return dict(h=h, i=i, cs=cs, ce=ce, sp=sp, sv=sv, st=st)


def run_collide(lib, p, e, ss, force):
    n = force.shape[0]
    d = {k: dev(lib, ss[k]) for k in ("sp", "sv", "st", "i", "cs", "ce")}
    df = dev(lib, force)
    set_params(lib, p, e)
    lib.collideFilaments(df.p, d["sp"].p, d["sv"].p, d["st"].p, d["i"].p, d["cs"].p, d["ce"].p, n, p.numCells)
    return down4(df, n)


@pytest.mark.parametrize("grid,twod", [((16, 16, 4), False), ((32, 32, 1), True), ((2, 16, 4), False)])
def test_collide_ref_dem(nwlib, grid, twod):
    rng = np.random.default_rng(11)
    p = sim_params(grid=grid, cell=1.0, W=int(grid[0] * grid[1] * grid[2] * 0.12), n=8)
    pos, vel, tan = random_worms(p, rng, twod=twod)
    force = random_forces(rng, pos.shape[0], 0.1)
    e = ext_params(pairPotential=abi.NW_PAIR_REF_DEM, stencil=abi.NW_STENCIL_REF27)
    ss = sorted_system(p, pos, vel, tan)
    g = run_collide(nwlib, p, e, ss, force)
    o = orc.collide_filaments(p, e, ss["sp"], ss["sv"], ss["st"], ss["i"], ss["cs"], ss["ce"], force)
    force_tol_check(g, o, name="collide REF")
    assert np.array_equal(g[:, 3].view(np.uint32), force[:, 3].view(np.uint32))


def test_collide_2d_grid_counts_layer_three_times(nwlib):
    """gridSize.z == 1: the reference's 27-offset stencil visits its only layer 3x."""
    rng = np.random.default_rng(12)
    p = sim_params(grid=(32, 32, 1), cell=1.0, W=120, n=8)
    pos, vel, tan = random_worms(p, rng, twod=True)
    zero = np.zeros((pos.shape[0], 4), np.float32)
    ss = sorted_system(p, pos, vel, tan)
    g27 = run_collide(nwlib, p, ext_params(stencil=abi.NW_STENCIL_REF27), ss, zero)
    g9 = run_collide(nwlib, p, ext_params(stencil=abi.NW_STENCIL_2D), ss, zero)
    force_tol_check(g27, 3.0 * g9.astype(np.float64), name="3x layer")


@pytest.mark.parametrize("grid,cell,activity", [((89, 89, 1), 1.1236, 0.0), ((64, 64, 1), 1.5625, 0.2),
                                                ((16, 16, 8), 1.2, 0.0)])
def test_collide_wca(nwlib, grid, cell, activity):
    rng = np.random.default_rng(13)
    twod = grid[2] == 1
    p = sim_params(grid=grid, cell=cell, W=int(grid[0] * grid[1] * grid[2] * 0.08), n=10, bondSpringL=0.8,
                   activity=activity)
    pos, vel, tan = random_worms(p, rng, twod=twod)
    force = random_forces(rng, pos.shape[0], 1.0)
    e = ext_params(pairPotential=abi.NW_PAIR_WCA, stencil=abi.NW_STENCIL_2D if twod else abi.NW_STENCIL_REF27)
    ss = sorted_system(p, pos, vel, tan)
    g = run_collide(nwlib, p, e, ss, force)
    o = orc.collide_filaments(p, e, ss["sp"], ss["sv"], ss["st"], ss["i"], ss["cs"], ss["ce"], force)
    force_tol_check(g, o, name="collide WCA")


def test_collide_alias_matches(nwlib):
    rng = np.random.default_rng(14)
    p = sim_params(grid=(16, 16, 4), W=150, n=8)
    pos, vel, tan = random_worms(p, rng)
    ss = sorted_system(p, pos, vel, tan)
    force = random_forces(rng, pos.shape[0])
    e = ext_params()
    a = run_collide(nwlib, p, e, ss, force)
    d = {k: dev(nwlib, ss[k]) for k in ("sp", "sv", "st", "i", "cs", "ce")}
    df = dev(nwlib, force)
    nwlib.collide(df.p, d["sp"].p, d["sv"].p, d["st"].p, d["i"].p, d["cs"].p, d["ce"].p, pos.shape[0], p.numCells)
    assert np.array_equal(down4(df, pos.shape[0]).view(np.uint32), a.view(np.uint32))


# ----------------------------------------------------------- bonded forces
@pytest.mark.parametrize("n,mode", [(3, 0), (4, 0), (8, 0), (20, 0), (64, 0), (65, 0), (2, 1), (20, 1), (30, 1),
                                    (70, 1)])
def test_filament_forces(nwlib, n, mode):
    rng = np.random.default_rng(20 + n)
    W = 257
    p = sim_params(grid=(32, 32, 1), cell=1.0, W=W, n=n, bondBendingK=5.0, bondSpringL=0.8)
    pos, vel, tan = random_worms(p, rng, twod=True, kink=0.4)
    force = random_forces(rng, W * n, 0.5)
    e = ext_params(bondedMode=mode, driveForce=1.0)
    set_params(nwlib, p, e)
    df, dt_, dp = dev(nwlib, force), dev(nwlib, tan), dev(nwlib, pos)
    nwlib.filamentForces(df.p, dt_.p, dp.p, float(W))
    g, gt = down4(df, W * n), down4(dt_, W * n)
    o, ot = orc.filament_forces(p, e, force, pos, tan, W)
    force_tol_check(g, o, name=f"bonded n={n} mode={mode}")
    # tangents are unit vectors: compare componentwise with an absolute bound
    assert np.max(np.abs(gt[:, :3] - ot[:, :3])) < 1e-6
    assert np.array_equal(gt[:, 3], ot[:, 3].astype(np.float32))


# ---------------------------------------------------------------- Langevin
def test_langevin_matches_oracle_noise(nwlib):
    rng = np.random.default_rng(30)
    n = 50_001
    p = sim_params(W=n, n=1, gamma=2.0)
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 1, (n, 3))
    force = random_forces(rng, n)
    set_params(nwlib, p)
    seed, off = 1337, 12345
    nwlib.nwRngSetSeed(seed, off)
    df, dv = dev(nwlib, force), dev(nwlib, vel)
    gamma_arg, kT = 0.7, 0.3
    nwlib.langevinThermostat(df.p, dv.p, None, gamma_arg, kT, n)
    g = down4(df, n)
    s_ = ctypes.c_uint64()
    o_ = ctypes.c_uint64()
    nwlib.nwRngGetState(ctypes.byref(s_), ctypes.byref(o_))
    assert o_.value == off + n
    sd = float(np.sqrt(np.float32(2.0) * np.float32(gamma_arg) * np.float32(kT)))
    noise = orc.random_gaussian(seed, off, 4 * n, 0.0, sd).reshape(n, 4)
    o = orc.langevin(p, force, vel, noise)
    force_tol_check(g, o, name="langevin")


def test_randomize_uniform_and_gaussian(nwlib):
    n = 100_003
    d = DeviceArray(nwlib, 4 * n)
    nwlib.nwRngSetSeed(77, 5)
    nwlib.randomizeUniform(d.p, n)
    u = d.download(np.float32, (n,))
    assert np.array_equal(u.view(np.uint32), orc.random_uniform(77, 5, n).view(np.uint32))
    assert u.min() > 0.0 and u.max() <= 1.0
    nwlib.randomizeGaussian(d.p, n, 0.5, 2.0)
    z = d.download(np.float32, (n,))
    oz = orc.random_gaussian(77, 5 + (n + 3) // 4, n, 0.5, 2.0)
    assert np.max(np.abs(z - oz)) < 2e-5 * max(1.0, np.max(np.abs(oz)))


# -------------------------------------------------------------- integrators
@pytest.mark.parametrize("bx,by,bz", [(0, 0, 0), (1, 2, 1), (2, 1, 0)])
def test_integrate_filaments(nwlib, bx, by, bz):
    rng = np.random.default_rng(40 + bx + 3 * by)
    p = sim_params(grid=(16, 16, 4), W=400, n=10, boundaryX=bx, boundaryY=by, boundaryZ=bz)
    pos, vel, _ = random_worms(p, rng)
    n = pos.shape[0]
    f, fo = random_forces(rng, n, 3.0), random_forces(rng, n, 3.0)
    f[:, 3] = 0.0
    fo[:, 3] = rng.random(n).astype(np.float32)
    # push some beads across the boundaries
    vel[:50, :3] = 400.0 * np.sign(vel[:50, :3])
    dt = 0.01
    set_params(nwlib, p)
    d = [dev(nwlib, a) for a in (pos, vel, f, fo)]
    nwlib.integrateFilaments(d[0].p, d[1].p, d[2].p, d[3].p, dt, n)
    g = [down4(x, n) for x in d]
    o = orc.integrate_filaments(p, pos, vel, f, fo, dt, prec="f32")
    # FP32 restatement with the same operation order: bitwise equal
    for a, b, name in zip(g, o, ("pos", "vel", "f", "f_old")):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name
    o64 = orc.integrate_filaments(p, pos, vel, f, fo, dt, prec="f64")
    L = np.array([p.boxSize.x, p.boxSize.y, p.boxSize.z])
    d = g[0][:, :3] - o64[0][:, :3]
    d -= L * np.round(d / L)  # a bead on a periodic face may wrap on one side only
    assert np.max(np.abs(d)) < 1e-5 * 16


def test_integrate_solvent(nwlib):
    rng = np.random.default_rng(45)
    p = sim_params(grid=(16, 16, 1), W=10, n=10)
    N, Ns = 100, 5000
    pos, vel, _ = random_worms(p, rng)
    sp = np.zeros((Ns, 4), np.float32)
    sp[:, :3] = (np.array([p.origin.x, p.origin.y, p.origin.z]) + rng.random((Ns, 3)) * 16 * [1, 1, 1])
    sv = np.zeros((Ns, 4), np.float32)
    sv[:, :3] = rng.normal(0, 5, (Ns, 3))
    sv[:, 3] = 1.0
    P, V = np.concatenate([pos, sp]), np.concatenate([vel, sv])
    set_params(nwlib, p)
    dp, dv = dev(nwlib, P), dev(nwlib, V)
    nwlib.integrateSolvent(dp.p, dv.p, 0.05, N, Ns)
    gp, gv = down4(dp, N + Ns), down4(dv, N + Ns)
    op, ov = orc.integrate_solvent(p, P, V, 0.05, N, Ns, prec="f32")
    assert np.array_equal(gp.view(np.uint32), op.view(np.uint32))
    assert np.array_equal(gv.view(np.uint32), ov.view(np.uint32))


# ---------------------------------------------------------------------- SRD
def srd_system(rng, N_w=20, n=10, Ns=20000, grid=(32, 32, 1), srd_grid=(16, 16, 1), offset=(0.4, 1.1, 0.0)):
    p = sim_params(grid=grid, cell=1.0, W=N_w, n=n)
    pos, vel, _ = random_worms(p, rng, twod=True)
    sp = np.zeros((Ns, 4), np.float32)
    o = np.array([p.origin.x, p.origin.y, p.origin.z])
    L = np.array([p.boxSize.x, p.boxSize.y, p.boxSize.z])
    sp[:, :3] = o + rng.random((Ns, 3)) * L
    sp[:, 2] = o[2] + 0.5
    sv = np.zeros((Ns, 4), np.float32)
    sv[:, :2] = rng.normal(0, 1, (Ns, 2))
    sv[:, 3] = 1.0
    P, V = np.concatenate([pos, sp]), np.concatenate([vel, sv])
    cell = (p.boxSize.x / srd_grid[0], p.boxSize.y / srd_grid[1], p.boxSize.z / srd_grid[2])
    s = solvent_params(p, srd_grid, cell, alpha=2.26893, kbT=1.3, offset=offset, ns=Ns)
    h, i = orc.calc_hash(p, s, P, srd=True)
    h, i = orc.sort_pairs(h, i)
    cs, ce, spos, svel, _ = orc.reorder(h, i, P, V, None, s.numCells)
    return p, s, P, V, dict(h=h, i=i, cs=cs, ce=ce, sp=spos, sv=svel)


def test_srd_com(nwlib):
    rng = np.random.default_rng(50)
    p, s, P, V, ss = srd_system(rng)
    C = s.numCells
    d = {k: dev(nwlib, ss[k]) for k in ("sv", "cs", "ce", "i")}
    com = DeviceArray(nwlib, F4 * C)
    nwlib.solventCellCentorOfMomentum(com.p, d["sv"].p, d["cs"].p, d["ce"].p, d["i"].p, C)
    g = down4(com, C)
    o32 = orc.srd_com(ss["sv"], ss["cs"], ss["ce"], C, prec="f32")
    assert np.array_equal(g.view(np.uint32), o32.view(np.uint32))
    o64 = orc.srd_com(ss["sv"], ss["cs"], ss["ce"], C)
    force_tol_check(g, o64, name="srd com")


def test_srd_isokinetic(nwlib):
    rng = np.random.default_rng(51)
    p, s, P, V, ss = srd_system(rng)
    C, n = s.numCells, P.shape[0]
    set_params(nwlib, p, s=s)
    d = {k: dev(nwlib, ss[k]) for k in ("sv", "cs", "ce", "i")}
    dv = dev(nwlib, V)
    nwlib.solventIsokineticThermostat(dv.p, d["sv"].p, d["cs"].p, d["ce"].p, d["i"].p, C)
    gv, gs = down4(dv, n), down4(d["sv"], n)
    ov, os_ = orc.srd_isokinetic(s, V, ss["sv"], ss["cs"], ss["ce"], ss["i"], C, prec="f32")
    assert np.array_equal(gv.view(np.uint32), ov.view(np.uint32))
    assert np.array_equal(gs.view(np.uint32), os_.view(np.uint32))


@pytest.mark.parametrize("beads_too", [0, 1])
def test_collide_solvent(nwlib, beads_too):
    rng = np.random.default_rng(52 + beads_too)
    p, s, P, V, ss = srd_system(rng)
    C, n = s.numCells, P.shape[0]
    e = ext_params(srdCollideFilaments=beads_too)
    set_params(nwlib, p, e, s)
    com = orc.srd_com(ss["sv"], ss["cs"], ss["ce"], C, prec="f32")
    dp, dv, dc = dev(nwlib, P), dev(nwlib, V), dev(nwlib, com)
    nwlib.nwRngSetSeed(9, 100)
    nwlib.collideSolvent(dp.p, dv.p, dc.p, None, C, n)
    g = down4(dv, n)
    uni = orc.random_uniform(9, 100, C)
    o = orc.collide_solvent(p, s, e, P, V, com, uni, C)
    force_tol_check(g, o, rtol=1e-6, name="srd collide")
    if not beads_too:
        assert np.array_equal(g[: p.numParticles].view(np.uint32), V[: p.numParticles].view(np.uint32))
    # per-cell momentum is conserved when every member rotates
    if beads_too:
        h, _ = orc.calc_hash(p, s, P, srd=True)
        m0 = np.zeros((C, 2))
        m1 = np.zeros((C, 2))
        np.add.at(m0, h, V[:, :2].astype(np.float64))
        np.add.at(m1, h, g[:, :2].astype(np.float64))
        assert np.max(np.abs(m1 - m0)) < 1e-4


# ---------------------------------------------------------------- reversal
def test_reverse_filaments_bit_exact(nwlib):
    rng = np.random.default_rng(60)
    W, n = 3001, 7
    p = sim_params(W=W, n=n, reverseProbability=0.3)
    pos, vel, _ = random_worms(p, rng)
    f, fo = random_forces(rng, W * n), random_forces(rng, W * n)
    set_params(nwlib, p)
    nwlib.nwRngSetSeed(4, 8)
    d = [dev(nwlib, a) for a in (f, fo, vel, pos)]
    nwlib.reverseFilaments(d[0].p, d[1].p, d[2].p, d[3].p, None, W)
    g = [down4(x, W * n) for x in d]
    uni = orc.random_uniform(4, 8, W)
    o = orc.reverse_filaments(p, uni, W, f, fo, vel, pos)
    for a, b in zip(g, o):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert 0.2 < np.mean(uni <= 0.3) < 0.4


# -------------------------------------------------- full-size integer stages
def test_full_size_hash_sort_reorder_800k(nwlib):
    """800,000 beads on a 890 x 890 grid (L = 1000, cell 1.1236): hash, sort and
    cell bounds bit-exact at the bench workload's size."""
    rng = np.random.default_rng(70)
    G = 890
    p = sim_params(grid=(G, G, 1), cell=1000.0 / G, W=40_000, n=20, box=(1000.0, 1000.0, 1000.0 / G),
                   bondSpringL=0.8)
    n = 800_000
    pos = np.zeros((n, 4), np.float32)
    pos[:, :2] = (rng.random((n, 2)) * 1000.0 - 500.0).astype(np.float32)
    pos[:, 2] = np.float32(p.origin.z + 0.5 * p.cellSize.z)
    set_params(nwlib, p)
    dpos = dev(nwlib, pos)
    dh, di = DeviceArray(nwlib, 4 * n), DeviceArray(nwlib, 4 * n)
    nwlib.calcHash(dh.p, di.p, dpos.p, n, False)
    nwlib.sortParticles(dh.p, di.p, n)
    C = p.numCells
    cs, ce, sp = DeviceArray(nwlib, 4 * C), DeviceArray(nwlib, 4 * C), DeviceArray(nwlib, F4 * n)
    ce.upload(np.zeros(C, np.uint32))
    nwlib.reorderDataAndFindCellStart(cs.p, ce.p, sp.p, None, None, dh.p, di.p, dpos.p, None, None, n, C)
    oh, oi = orc.calc_hash(p, None, pos)
    oh, oi = orc.sort_pairs(oh, oi)
    ocs, oce, osp, _, _ = orc.reorder(oh, oi, pos, None, None, C)
    assert np.array_equal(downu(dh, n), oh)
    assert np.array_equal(downu(di, n), oi)
    assert np.array_equal(downu(cs, C), ocs)
    assert np.array_equal(downu(ce, C), oce)
    assert np.array_equal(down4(sp, n).view(np.uint32), osp.view(np.uint32))


def test_hash_sort_sparse_grid_two_pass_level1(nwlib):
    """4,000,000 particles on an 8000 x 7000 grid (56M cells, 0.07 per cell):
    the bucket sort's plan has 27,344 buckets of 2048 cells and level-1 tiles
    of 65,536 particles, more than the LDS stage holds, so level 1 runs as its
    two-pass kernel (packed pairs written straight to memory); hash and sort
    bit-exact."""
    rng = np.random.default_rng(71)
    gx, gy = 8000, 7000
    p = sim_params(grid=(gx, gy, 1), cell=1.0, W=200_000, n=20, box=(float(gx), float(gy), 1.0))
    n = 4_000_000
    pos = np.zeros((n, 4), np.float32)
    pos[:, 0] = (rng.random(n) * gx - 0.5 * gx).astype(np.float32)
    pos[:, 1] = (rng.random(n) * gy - 0.5 * gy).astype(np.float32)
    pos[:, 2] = np.float32(p.origin.z + 0.5 * p.cellSize.z)
    set_params(nwlib, p)
    assert nwlib.nwFusedPlanFits(n, gx * gy, 20) == 1  # the bucket sort, not the counting-sort fallback
    dpos = dev(nwlib, pos)
    dh, di = DeviceArray(nwlib, 4 * n), DeviceArray(nwlib, 4 * n)
    nwlib.calcHash(dh.p, di.p, dpos.p, n, False)
    nwlib.sortParticles(dh.p, di.p, n)
    oh, oi = orc.calc_hash(p, None, pos)
    assert oh.max() >= gx * gy - 10_000  # the grid is covered to its last rows
    ok, ov = orc.sort_pairs(oh, oi)
    assert np.array_equal(downu(dh, n), ok)
    assert np.array_equal(downu(di, n), ov)

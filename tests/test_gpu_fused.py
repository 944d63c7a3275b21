"""The fused entry points (nwFusedMdStep, nwFusedSrdStep) against the
reference's per-call sequence on the GPU and against the oracle.

The per-call path is itself checked against the oracle (test_gpu_parity.py);
the fused path evaluates the same expressions in the same order, so positions,
velocities, old forces, tangents, the stable cell order and every SRD output
must agree bit for bit.  Pair forces are summed per row of three cells instead
of per cell, so they agree to the 1e-5 bar."""
import ctypes

import numpy as np
import pytest

import orc
from nwamd import abi
from nwamd.abi import DeviceArray
from nwtest import ext_params, force_tol_check, random_forces, random_worms, sim_params, solvent_params

pytestmark = pytest.mark.gpu


def dev(lib, a):
    return DeviceArray.from_numpy(lib, a)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def per_call_md(lib, p, e, st, dt, langevin, gamma, kT):
    """ParticleSystem::update's MD part through the reference entry points."""
    N, W, C = p.numParticles, p.numFilaments, p.numCells
    lib.integrateFilaments(st["pos"].p, st["vel"].p, st["f"].p, st["fo"].p, dt, N)
    if p.reverseProbability > 0:
        lib.reverseFilaments(st["f"].p, st["fo"].p, st["vel"].p, st["pos"].p, None, W)
    if langevin:
        lib.langevinThermostat(st["f"].p, st["vel"].p, None, gamma, kT, N)
    lib.filamentForces(st["f"].p, st["tan"].p, st["pos"].p, float(W))
    lib.calcHash(st["h"].p, st["i"].p, st["pos"].p, N, False)
    lib.sortParticles(st["h"].p, st["i"].p, N)
    lib.reorderDataAndFindCellStart(st["cs"].p, st["ce"].p, st["sp"].p, st["sv"].p, st["st"].p, st["h"].p,
                                    st["i"].p, st["pos"].p, st["vel"].p, st["tan"].p, N, C)
    lib.collideFilaments(st["f"].p, st["sp"].p, st["sv"].p, st["st"].p, st["i"].p, st["cs"].p, st["ce"].p, N, C)


def state(lib, p, pos, vel, tan, f, fo, extra=0):
    N, C = p.numParticles, p.numCells
    d = dict(pos=dev(lib, pos), vel=dev(lib, vel), tan=dev(lib, tan), f=dev(lib, f), fo=dev(lib, fo))
    for k in ("h", "i"):
        d[k] = DeviceArray(lib, 4 * (N + extra))
    d["cs"], d["ce"] = DeviceArray(lib, 4 * C), DeviceArray(lib, 4 * C)
    for k in ("sp", "sv", "st"):
        d[k] = DeviceArray(lib, 16 * (N + extra))
    return d


def down(d, name, n, dtype=np.float32, w=4):
    return d[name].download(dtype, (n, w) if w > 1 else (n,))


@pytest.mark.parametrize("mode,grid,n,rev,lgv", [
    ("ref", (16, 16, 4), 8, 0.0, True),
    ("ref2d", (32, 32, 1), 10, 0.3, True),
    ("wca", (89, 89, 1), 20, 0.2, True),
    ("wca", (89, 89, 1), 20, 0.0, False),
    ("wca3d", (12, 12, 6), 30, 0.1, True),
    ("wca", (40, 40, 1), 3, 0.5, False),
])
def test_fused_md_step_matches_per_call(nwlib, mode, grid, n, rev, lgv):
    rng = np.random.default_rng([len(mode), grid[0], grid[2], n, int(rev * 10), int(lgv)])
    twod = grid[2] == 1
    cell = 1.0 if mode.startswith("ref") else 1.1236
    W = max(4, int(grid[0] * grid[1] * grid[2] * (0.1 if mode.startswith("ref") else 0.06)))
    p = sim_params(grid=grid, cell=cell, W=W, n=n, bondSpringL=0.8 if mode != "ref" else 0.4,
                   reverseProbability=rev, bondBendingK=3.0, activity=0.2 if mode == "ref" else 0.0)
    if mode.startswith("wca"):
        e = ext_params(pairPotential=abi.NW_PAIR_WCA, bondedMode=abi.NW_BONDED_CHAIN, driveForce=1.0,
                       stencil=abi.NW_STENCIL_2D if twod else abi.NW_STENCIL_REF27)
    else:
        e = ext_params(stencil=abi.NW_STENCIL_REF27)
    pos, vel, tan = random_worms(p, rng, twod=twod)
    f, fo = random_forces(rng, W * n, 0.5), random_forces(rng, W * n, 0.5)
    f[:, 3] = 0
    compare_md_step(nwlib, p, e, pos, vel, tan, f, fo, lgv, mode)


@pytest.mark.parametrize("mode", ["ref2d", "wca"])
def test_fused_md_step_crowded_bucket(nwlib, mode):
    """More beads in one bucket of 1024 cells than the sort stages in LDS
    (2048): that bucket takes the global-memory path."""
    rng = np.random.default_rng([5, len(mode)])
    grid, n = (64, 64, 1), 10
    cell = 1.0 if mode == "ref2d" else 1.1236
    W = 300 if mode == "ref2d" else 220
    p = sim_params(grid=grid, cell=cell, W=W, n=n, bondSpringL=0.4 if mode == "ref2d" else 0.8,
                   bondBendingK=3.0, activity=0.0)
    if mode == "wca":
        e = ext_params(pairPotential=abi.NW_PAIR_WCA, bondedMode=abi.NW_BONDED_CHAIN, driveForce=1.0,
                       stencil=abi.NW_STENCIL_2D)
    else:
        e = ext_params(stencil=abi.NW_STENCIL_REF27)
    pos, vel, tan = random_worms(p, rng, twod=True)
    # squeeze every worm into cell rows 1..14 (the first bucket holds rows 0..15)
    y0 = p.origin.y
    pos[:, 1] = (y0 + cell * (1.0 + (pos[:, 1] - y0) * (13.5 / (64 * cell)))).astype(np.float32)
    f, fo = random_forces(rng, W * n, 0.5), random_forces(rng, W * n, 0.5)
    f[:, 3] = 0
    compare_md_step(nwlib, p, e, pos, vel, tan, f, fo, False, mode)


def compare_md_step(lib, p, e, pos, vel, tan, f, fo, lgv, mode):
    W, n = p.numFilaments, p.filamentSize
    lib.setParameters(ctypes.byref(p))
    lib.nwSetExtParameters(ctypes.byref(e))
    dt, gamma, kT = 0.002, 0.8, 0.4
    a = state(lib, p, pos, vel, tan, f, fo)
    b = state(lib, p, pos, vel, tan, f, fo)
    lib.nwRngSetSeed(21, 1000)
    per_call_md(lib, p, e, a, dt, lgv, gamma, kT)
    sa, oa = ctypes.c_uint64(), ctypes.c_uint64()
    lib.nwRngGetState(ctypes.byref(sa), ctypes.byref(oa))
    lib.nwRngSetSeed(21, 1000)
    lib.nwFusedMdStep(b["pos"].p, b["vel"].p, b["f"].p, b["fo"].p, b["tan"].p, b["h"].p, b["i"].p, b["sp"].p,
                      b["sv"].p, b["st"].p, dt, gamma, kT, int(lgv), W)
    sb, ob = ctypes.c_uint64(), ctypes.c_uint64()
    lib.nwRngGetState(ctypes.byref(sb), ctypes.byref(ob))
    assert oa.value == ob.value
    N = W * n
    for k in ("pos", "vel", "fo", "tan"):
        assert np.array_equal(bits(down(a, k, N)), bits(down(b, k, N))), k
    assert np.array_equal(down(a, "i", N, np.uint32, 1), down(b, "i", N, np.uint32, 1))
    assert np.array_equal(bits(down(a, "sp", N)), bits(down(b, "sp", N)))
    # the fused step leaves the cell of every bead in bead order
    ha, hb, ib = down(a, "h", N, np.uint32, 1), down(b, "h", N, np.uint32, 1), down(b, "i", N, np.uint32, 1)
    assert np.array_equal(hb[ib], ha)
    # pair forces: the oracle on the per-call path's sorted state
    fa, fb = down(a, "f", N), down(b, "f", N)
    force_tol_check(fb, fa.astype(np.float64), rtol=1e-5, name=f"fused vs per-call ({mode})")


@pytest.mark.parametrize("beads,iso,crowd", [(1, 0, 0), (0, 1, 0), (1, 1, 0), (1, 1, 1)])
def test_fused_srd_step_matches_per_call(nwlib, beads, iso, crowd):
    """crowd: the solvent fills an 8 x 8 patch, so one bucket of cells holds
    far more members than the sort stages in LDS (the global-memory path)."""
    rng = np.random.default_rng(7 + beads + 2 * iso + 4 * crowd)
    lib = nwlib
    p = sim_params(grid=(64, 64, 1), cell=1.0, W=100, n=10)
    pos, vel, _ = random_worms(p, rng, twod=True)
    Ns, N = 30000, p.numParticles
    sp = np.zeros((Ns, 4), np.float32)
    span = 8.0 if crowd else 64.0
    sp[:, :2] = (rng.random((Ns, 2)) * span - 32).astype(np.float32)
    sp[:, 2] = p.origin.z + 0.5
    sv = np.zeros((Ns, 4), np.float32)
    sv[:, :2] = rng.normal(0, 1, (Ns, 2))
    sv[:, 3] = 1.0
    P, V = np.concatenate([pos, sp]), np.concatenate([vel, sv])
    s = solvent_params(p, (50, 50, 1), (64 / 50, 64 / 50, 1.0), alpha=np.pi / 2, kbT=0.3,
                       offset=(-0.3, 0.41, 0.0), ns=Ns)
    e = ext_params(srdCollideFilaments=beads)
    lib.setParameters(ctypes.byref(p))
    lib.setSolventParameters(ctypes.byref(s))
    lib.nwSetExtParameters(ctypes.byref(e))
    NT, Cs, dt = N + Ns, s.numCells, 0.05
    res = []
    for fused in (False, True):
        dP, dV = dev(lib, P), dev(lib, V)
        h, i = DeviceArray(lib, 4 * NT), DeviceArray(lib, 4 * NT)
        com = DeviceArray(lib, 16 * Cs)
        lib.nwRngSetSeed(3, 77)
        if fused:
            lib.nwFusedSrdStep(dP.p, dV.p, com.p, h.p, i.p, dt, N, Ns, iso)
        else:
            cs, ce = DeviceArray(lib, 4 * Cs), DeviceArray(lib, 4 * Cs)
            sP, sV = DeviceArray(lib, 16 * NT), DeviceArray(lib, 16 * NT)
            lib.calcHash(h.p, i.p, dP.p, NT, True)
            lib.sortParticles(h.p, i.p, NT)
            lib.reorderDataAndFindCellStart(cs.p, ce.p, sP.p, sV.p, None, h.p, i.p, dP.p, dV.p, None, NT, Cs)
            if iso:
                lib.solventIsokineticThermostat(dV.p, sV.p, cs.p, ce.p, i.p, Cs)
            lib.solventCellCentorOfMomentum(com.p, sV.p, cs.p, ce.p, i.p, Cs)
            lib.collideSolvent(dP.p, dV.p, com.p, None, Cs, NT)
            lib.integrateSolvent(dP.p, dV.p, dt, N, Ns)
        res.append((dP.download(np.float32, (NT, 4)), dV.download(np.float32, (NT, 4)),
                    com.download(np.float32, (Cs, 4)), i.download(np.uint32, (NT,)), h.download(np.uint32, (NT,))))
    (pa, va, ca, ia, ha), (pb, vb, cb, ib, hb) = res
    assert np.array_equal(hb[ib], ha)  # fused: cells in particle order; per call: sorted
    assert np.array_equal(bits(pa), bits(pb))
    assert np.array_equal(bits(va), bits(vb))
    assert np.array_equal(bits(ca[:, :3]), bits(cb[:, :3]))
    assert np.array_equal(ia, ib)  # the stable cell order
    uni = orc.random_uniform(3, 77, Cs)
    assert np.array_equal(cb[:, 3], np.where(uni > 0.5, 1.0, -1.0).astype(np.float32))


def test_fused_step_repeats_bitwise(nwlib):
    """Atomic histogram slots are arbitrary, the results are not: two runs agree bit for bit."""
    rng = np.random.default_rng(99)
    lib = nwlib
    p = sim_params(grid=(64, 64, 1), cell=1.1236, W=800, n=20, bondSpringL=0.8, activity=0.0)
    e = ext_params(pairPotential=abi.NW_PAIR_WCA, bondedMode=abi.NW_BONDED_CHAIN, driveForce=1.0,
                   stencil=abi.NW_STENCIL_2D)
    pos, vel, tan = random_worms(p, rng, twod=True)
    f = np.zeros_like(pos)
    lib.setParameters(ctypes.byref(p))
    lib.nwSetExtParameters(ctypes.byref(e))
    out = []
    for _ in range(2):
        b = state(lib, p, pos, vel, tan, f, f)
        lib.nwRngSetSeed(5, 0)
        for _ in range(5):
            lib.nwFusedMdStep(b["pos"].p, b["vel"].p, b["f"].p, b["fo"].p, b["tan"].p, b["h"].p, b["i"].p,
                              b["sp"].p, None, None, 0.001, 2.0, 100.0, 1, p.numFilaments)
        out.append([down(b, k, p.numParticles) for k in ("pos", "vel", "f")])
    for x, y in zip(*out):
        assert np.array_equal(bits(x), bits(y))


@pytest.mark.parametrize("mode,activity", [("wca", 0.0), ("wca", 0.3), ("ref", 0.2)])
def test_lds_and_global_pair_kernels_agree_bitwise(nwlib, mode, activity):
    rng = np.random.default_rng([7, int(activity * 10), len(mode)])
    lib = nwlib
    grid = (120, 97, 1)
    cell = 1.1236 if mode == "wca" else 1.0
    p = sim_params(grid=grid, cell=cell, W=1200, n=10, bondSpringL=0.8 if mode == "wca" else 0.4,
                   activity=activity)
    pos, vel, tan = random_worms(p, rng, twod=True)
    f = np.zeros_like(pos)
    res = []
    for kern in (abi.NW_COLLIDE_GLOBAL, abi.NW_COLLIDE_LDS):
        if mode == "wca":
            e = ext_params(pairPotential=abi.NW_PAIR_WCA, bondedMode=abi.NW_BONDED_CHAIN, driveForce=1.0,
                           stencil=abi.NW_STENCIL_2D, collideKernel=kern)
        else:
            e = ext_params(stencil=abi.NW_STENCIL_REF27, collideKernel=kern)
        lib.setParameters(ctypes.byref(p))
        lib.nwSetExtParameters(ctypes.byref(e))
        b = state(lib, p, pos, vel, tan, f, f)
        lib.nwRngSetSeed(1, 0)
        for _ in range(3):
            lib.nwFusedMdStep(b["pos"].p, b["vel"].p, b["f"].p, b["fo"].p, b["tan"].p, b["h"].p, b["i"].p,
                              b["sp"].p, b["sv"].p, b["st"].p, 0.002, 1.0, 0.5, 1, p.numFilaments)
        res.append([down(b, k, p.numParticles) for k in ("pos", "vel", "f")])
    for x, y in zip(*res):
        assert np.array_equal(bits(x), bits(y))


def banded_rows(p, rng, dense=6):
    """Straight worms along x filling every row of a one-layer grid: one bead
    per cell in even rows, `dense` per cell in odd rows; each row shifted by
    a random amount, so worms cross the x faces, and every bead jittered."""
    gx, gy = p.gridSize.x, p.gridSize.y
    c, n = p.cellSize.x, p.filamentSize
    o = np.array([p.origin.x, p.origin.y, p.origin.z])
    rows = []
    for y in range(gy):
        per = dense if y % 2 else 1
        k = np.arange(gx * per)
        x = (k + 0.5 + 0.3 * (rng.random(len(k)) - 0.5)) * (c / per) + rng.random() * gx * c
        x = np.mod(x, gx * c)
        yy = (y + 0.5 + 0.4 * (rng.random(len(k)) - 0.5)) * c
        rows.append(np.stack([x, yy, np.full(len(k), 0.5 * p.cellSize.z)], 1))
    r = np.concatenate(rows) + o
    assert len(r) % n == 0
    pos = np.zeros((len(r), 4), np.float32)
    pos[:, :3] = r
    pos[:, 3] = p.particleRadius
    vel = np.zeros_like(pos)
    vel[:, :3] = 0.1 * (rng.random((len(r), 3)) - 0.5)
    vel[:, 2] = 0.0
    vel[:, 3] = 1.0
    return pos, vel


@pytest.mark.parametrize("pair_block", [0, 128, 512])
def test_lds_pair_kernel_overfull_stage_bitwise(nwlib, pair_block):
    """Rows of six beads per cell between rows of one: a workgroup on a
    sparse row needs more neighbour-row beads than its LDS pool holds (three
    runs of pairBlock + 128 on average), so some runs stay unstaged and their
    beads take the global walk.  Forces, positions and velocities equal the
    global kernel's bit for bit, for every pair workgroup size."""
    from nwamd.abi import NWTuning
    rng = np.random.default_rng([31, pair_block])
    lib = nwlib
    grid, n = (120, 40, 1), 10
    W = (120 * 20 + 120 * 6 * 20) // n
    p = sim_params(grid=grid, cell=1.1236, W=W, n=n, bondSpringL=0.8, activity=0.0)
    pos, vel = banded_rows(p, rng)
    tan = np.zeros_like(pos)
    f = np.zeros_like(pos)
    res = []
    lib.nwSetTuning(ctypes.byref(NWTuning(pairBlock=pair_block)))
    try:
        for kern in (abi.NW_COLLIDE_GLOBAL, abi.NW_COLLIDE_LDS):
            e = ext_params(pairPotential=abi.NW_PAIR_WCA, bondedMode=abi.NW_BONDED_CHAIN, driveForce=1.0,
                           stencil=abi.NW_STENCIL_2D, collideKernel=kern)
            lib.setParameters(ctypes.byref(p))
            lib.nwSetExtParameters(ctypes.byref(e))
            b = state(lib, p, pos, vel, tan, f, f)
            lib.nwRngSetSeed(3, 0)
            lib.nwFusedMdStep(b["pos"].p, b["vel"].p, b["f"].p, b["fo"].p, b["tan"].p, b["h"].p, b["i"].p,
                              b["sp"].p, None, None, 1e-4, 1.0, 0.5, 1, p.numFilaments)
            assert lib.nwGetLastError() == 0
            res.append([down(b, k, p.numParticles) for k in ("pos", "vel", "f")])
    finally:
        lib.nwSetTuning(ctypes.byref(NWTuning()))
    assert np.all(np.isfinite(res[0][2]))
    assert np.abs(res[0][2][:, :2]).max() > 1.0  # the dense rows push hard
    for x, y in zip(*res):
        assert np.array_equal(bits(x), bits(y))


@pytest.mark.parametrize("mode,lgv", [("wca", 0), ("wca", 1), ("ref", 1)])
def test_swap_forces_step_matches_fused_step(nwlib, mode, lgv):
    """nwFusedMdStepSwapForces with the caller swapping force / forceOld after
    every call against nwFusedMdStep: positions, velocities, tangents and the
    (logical) force and old force bit for bit over 3 steps."""
    rng = np.random.default_rng(17 + lgv)
    lib = nwlib
    wca = mode == "wca"
    p = sim_params(grid=(64, 64, 1), cell=1.1236 if wca else 1.0, W=300, n=10, bondSpringL=0.8 if wca else 0.4,
                   activity=0.0 if wca else 0.2)
    pos, vel, tan = random_worms(p, rng, twod=True)
    N = p.numParticles
    e = ext_params(pairPotential=abi.NW_PAIR_WCA if wca else abi.NW_PAIR_REF_DEM,
                   bondedMode=abi.NW_BONDED_CHAIN if wca else abi.NW_BONDED_REF, driveForce=0.7 if wca else 0.0,
                   stencil=abi.NW_STENCIL_2D)
    lib.setParameters(ctypes.byref(p))
    lib.nwSetExtParameters(ctypes.byref(e))
    f0 = np.zeros((N, 4), np.float32)
    res = []
    for swap in (False, True):
        d = dict(pos=dev(lib, pos), vel=dev(lib, vel), tan=dev(lib, tan))
        fa, fb = dev(lib, f0), dev(lib, f0)
        h, i = DeviceArray(lib, 4 * N), DeviceArray(lib, 4 * N)
        sP, sV, sT = (DeviceArray(lib, 16 * N) for _ in range(3))
        lib.nwRngSetSeed(4, 40)
        for _ in range(3):
            args = (d["pos"].p, d["vel"].p, fa.p, fb.p, d["tan"].p, h.p, i.p, sP.p, None if wca else sV.p,
                    None if wca else sT.p, 1e-3, p.gamma, 0.05, lgv, p.numFilaments)
            if swap:
                lib.nwFusedMdStepSwapForces(*args)
                fa, fb = fb, fa
            else:
                lib.nwFusedMdStep(*args)
        lib.threadSync()
        res.append([x.download(np.float32, (N, 4)) for x in (d["pos"], d["vel"], d["tan"], fa, fb)])
    for a, b in zip(*res):
        assert np.array_equal(bits(a), bits(b))


def test_fused_srd_step_8192_particle_tiles(nwlib):
    """3,003,000 particles at 7 per SRD cell (655 x 655 cells): the sort's plan
    takes first-level tiles of 8192 particles -- the SRD tile kernel of 512
    threads x 16 particles in nwFusedSrdStep, the 1024 x 8 kernel in the
    per-call sortParticles.  calcHash + sortParticles against the oracle's
    stable sort bit for bit; nwFusedSrdStep against the per-call SRD sequence
    bit for bit."""
    rng = np.random.default_rng(8192)
    lib = nwlib
    G = 655
    p = sim_params(grid=(G, G, 1), cell=1.0, W=100, n=10)
    pos, vel, _ = random_worms(p, rng, twod=True)
    N = p.numParticles
    Ns = 7 * G * G - N
    NT, Cs = N + Ns, G * G
    plan = (ctypes.c_uint * 4)()
    assert lib.nwFusedPlanInfo(NT, Cs, 0, plan) == 1
    assert plan[2] == 8192, list(plan)  # particles per first-level tile
    sp = np.zeros((Ns, 4), np.float32)
    sp[:, :2] = (rng.random((Ns, 2)) * G - G / 2).astype(np.float32)
    sp[:, 2] = p.origin.z + 0.5
    sv = np.zeros((Ns, 4), np.float32)
    sv[:, :2] = rng.normal(0, 1, (Ns, 2))
    sv[:, 3] = 1.0
    P, V = np.concatenate([pos, sp]), np.concatenate([vel, sv])
    s = solvent_params(p, (G, G, 1), (1.0, 1.0, 1.0), alpha=np.pi / 2, kbT=0.3, offset=(0.21, -0.37, 0.0), ns=Ns)
    e = ext_params(srdCollideFilaments=1)
    lib.setParameters(ctypes.byref(p))
    lib.setSolventParameters(ctypes.byref(s))
    lib.nwSetExtParameters(ctypes.byref(e))
    dt = 0.05
    # the per-call hash and sort against the oracle
    oh, oi = orc.calc_hash(p, s, P, srd=True)
    oh, oi = orc.sort_pairs(oh, oi)
    res = []
    for fused in (False, True):
        dP, dV = dev(lib, P), dev(lib, V)
        h, i = DeviceArray(lib, 4 * NT), DeviceArray(lib, 4 * NT)
        com = DeviceArray(lib, 16 * Cs)
        lib.nwRngSetSeed(5, 11)
        if fused:
            lib.nwFusedSrdStep(dP.p, dV.p, com.p, h.p, i.p, dt, N, Ns, 0)
        else:
            cs, ce = DeviceArray(lib, 4 * Cs), DeviceArray(lib, 4 * Cs)
            sP, sV = DeviceArray(lib, 16 * NT), DeviceArray(lib, 16 * NT)
            lib.calcHash(h.p, i.p, dP.p, NT, True)
            lib.sortParticles(h.p, i.p, NT)
            assert np.array_equal(h.download(np.uint32, (NT,)), oh)
            assert np.array_equal(i.download(np.uint32, (NT,)), oi)
            lib.reorderDataAndFindCellStart(cs.p, ce.p, sP.p, sV.p, None, h.p, i.p, dP.p, dV.p, None, NT, Cs)
            lib.solventCellCentorOfMomentum(com.p, sV.p, cs.p, ce.p, i.p, Cs)
            lib.collideSolvent(dP.p, dV.p, com.p, None, Cs, NT)
            lib.integrateSolvent(dP.p, dV.p, dt, N, Ns)
        res.append((dP.download(np.float32, (NT, 4)), dV.download(np.float32, (NT, 4)),
                    com.download(np.float32, (Cs, 4)), i.download(np.uint32, (NT,)), h.download(np.uint32, (NT,))))
    (pa, va, ca, ia, ha), (pb, vb, cb, ib, hb) = res
    assert np.array_equal(ib, oi)  # the fused step's stable cell order
    assert np.array_equal(hb[ib], ha)
    assert np.array_equal(bits(pa), bits(pb))
    assert np.array_equal(bits(va), bits(vb))
    assert np.array_equal(bits(ca[:, :3]), bits(cb[:, :3]))


@pytest.mark.parametrize("n", [2, 3, 10, 20])
def test_tangentless_step_and_tangents_from_positions(nwlib, n):
    """nwFusedMdStep with no tangent array (the WCA term reads none) gives
    every other output of the step with one bit for bit, and
    nwFilamentTangents on the new positions gives the tangents that step
    stored, bit for bit (none for n < 3, no tail tangent for n = 3)."""
    rng = np.random.default_rng(40 + n)
    lib = nwlib
    p = sim_params(grid=(48, 48, 1), cell=1.1236, W=180, n=n, bondSpringL=0.8, activity=0.0)
    e = ext_params(pairPotential=abi.NW_PAIR_WCA, bondedMode=abi.NW_BONDED_CHAIN, driveForce=1.0,
                   stencil=abi.NW_STENCIL_2D)
    pos, vel, _ = random_worms(p, rng, twod=True)
    N = p.numFilaments * n
    f, fo = random_forces(rng, N, 0.5), random_forces(rng, N, 0.5)
    f[:, 3] = fo[:, 3] = 0
    sentinel = np.full((N, 4), 7.0, np.float32)
    lib.setParameters(ctypes.byref(p))
    lib.nwSetExtParameters(ctypes.byref(e))
    outs = []
    for with_tan in (True, False):
        d = {k: dev(lib, a) for k, a in dict(pos=pos, vel=vel, f=f, fo=fo, tan=sentinel).items()}
        h, i, sp = DeviceArray(lib, 4 * N), DeviceArray(lib, 4 * N), DeviceArray(lib, 16 * N)
        lib.nwRngSetSeed(8, 80)
        lib.nwFusedMdStep(d["pos"].p, d["vel"].p, d["f"].p, d["fo"].p, d["tan"].p if with_tan else None, h.p, i.p,
                          sp.p, None, None, 2e-3, 2.0, 50.0, 1, p.numFilaments)
        outs.append({k: d[k].download(np.uint32, (N, 4)) for k in ("pos", "vel", "f", "fo", "tan")})
        outs[-1]["i"] = i.download(np.uint32, (N,))
        if not with_tan:
            assert np.array_equal(outs[-1]["tan"], sentinel.view(np.uint32))  # nothing stored
            lib.nwFilamentTangents(d["tan"].p, d["pos"].p, p.numFilaments)
            outs[-1]["tan"] = d["tan"].download(np.uint32, (N, 4))
    a, b = outs
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    if n < 3:
        assert np.array_equal(a["tan"], sentinel.view(np.uint32))

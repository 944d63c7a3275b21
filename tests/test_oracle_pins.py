"""Pins of the oracle (oracle/liborc.so) on the CPU.

The reference ships no tests, fixtures or golden vectors for this path and its
kernel bodies cannot be compiled here (they need CUDA headers), so every
oracle function is pinned by something independent of its own formula:
closed forms, finite differences of the potential energy, brute-force O(N^2)
sums, library routines (numpy stable argsort / searchsorted / bincount),
invariants (Newton's third law, momentum conservation, time reversibility,
equipartition) and Philox known-answer vectors from an independent
implementation (tests/golden/philox_kat.json).
"""
import ctypes
import json
import math
import os

import numpy as np
import pytest

import orc
from nwamd import abi
import nwtest
from nwtest import ext_params, random_worms, sim_params, solvent_params

HERE = os.path.dirname(os.path.abspath(__file__))


# ------------------------------------------------------------------ Philox
def test_philox_known_answers():
    kat = json.load(open(os.path.join(HERE, "golden", "philox_kat.json")))
    for v in kat["vectors"]:
        assert orc.philox(v["counter"], v["key"]).tolist() == v["out"]


def test_generator_layout_and_uniform_range():
    u = orc.random_uniform(1337, 40, 10)
    for e in range(10):
        blk = orc.philox([40 + e // 4, 0, 0, 0], [1337, 0])
        assert u[e] == np.float32((blk[e % 4] >> 8) * 2.0 ** -24 + 2.0 ** -24)
    big = orc.random_uniform(3, 0, 1 << 20)
    assert big.min() > 0.0 and big.max() <= 1.0
    assert abs(big.mean() - 0.5) < 2e-3


def test_gaussian_moments_and_cdf():
    z = orc.random_gaussian(99, 7, 400_000, 0.0, 1.0)
    assert abs(z.mean()) < 5e-3
    assert abs(z.std() - 1.0) < 5e-3
    assert abs(np.mean(z ** 4) - 3.0) < 0.05
    for x, cdf in ((-1.0, 0.158655), (0.0, 0.5), (1.5, 0.933193)):
        assert abs(np.mean(z < x) - cdf) < 3e-3
    z2 = orc.random_gaussian(99, 7, 8, 2.0, 3.0)
    assert np.allclose(z2, 2.0 + 3.0 * z[:8])


# -------------------------------------------------------------------- hash
def test_hash_cell_centres_and_faces():
    p = sim_params(grid=(8, 4, 2), cell=1.0, W=1, n=1)
    ix, iy, iz = np.meshgrid(np.arange(8), np.arange(4), np.arange(2), indexing="ij")
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
    pos = np.zeros((ix.size, 4), np.float32)
    pos[:, 0] = p.origin.x + ix + 0.5
    pos[:, 1] = p.origin.y + iy + 0.5
    pos[:, 2] = p.origin.z + iz + 0.5
    h, idx = orc.calc_hash(p, None, pos)
    assert np.array_equal(h, (iz * 4 + iy) * 8 + ix)
    assert np.array_equal(idx, np.arange(ix.size))
    # a position on a lower face belongs to the cell above it (floor)
    pos[:, 0] = p.origin.x + ix
    h, _ = orc.calc_hash(p, None, pos)
    assert np.array_equal(h % 8, ix)


def test_hash_wrap_equals_reference_mask_on_pow2_grids():
    """The reference wraps with `& (G - 1)` (particles_kernel_impl.cuh:254-256);
    the oracle's modulo must agree on power-of-two grids, also outside the box."""
    rng = np.random.default_rng(0)
    p = sim_params(grid=(16, 8, 4), cell=1.5625, W=1, n=1)
    pos = np.zeros((5000, 4), np.float32)
    L = np.array([p.boxSize.x, p.boxSize.y, p.boxSize.z])
    o = np.array([p.origin.x, p.origin.y, p.origin.z])
    pos[:, :3] = (o + (rng.random((5000, 3)) * 3 - 1) * L).astype(np.float32)
    h, _ = orc.calc_hash(p, None, pos)
    g = np.floor((pos[:, :3] - o.astype(np.float32)) / np.float32(1.5625)).astype(np.int64)
    gx, gy, gz = g[:, 0] & 15, g[:, 1] & 7, g[:, 2] & 3
    assert np.array_equal(h, (gz * 8 + gy) * 16 + gx)


def test_hash_2d_grid_and_srd_translation():
    p = sim_params(grid=(20, 10, 1), cell=1.0, W=1, n=1)
    s = solvent_params(p, (10, 5, 1), (2.0, 2.0, 1.0), offset=(0.0, 0.0, 0.0))
    pos = np.zeros((10 * 5, 4), np.float32)
    ix, iy = np.meshgrid(np.arange(10), np.arange(5), indexing="ij")
    pos[:, 0] = p.origin.x + 2 * ix.ravel() + 0.25
    pos[:, 1] = p.origin.y + 2 * iy.ravel() + 0.25
    h0, _ = orc.calc_hash(p, s, pos, srd=True)
    assert np.array_equal(h0, iy.ravel() * 10 + ix.ravel())
    # shifting the grid by one SRD cell moves every particle one cell up (periodic)
    s.randOffset = abi.Float3(2.0, 0.0, 0.0)
    h1, _ = orc.calc_hash(p, s, pos, srd=True)
    assert np.array_equal(h1, iy.ravel() * 10 + (ix.ravel() + 1) % 10)
    s.randOffset = abi.Float3(20.0, 10.0, 0.0)  # a whole box: no change
    h2, _ = orc.calc_hash(p, s, pos, srd=True)
    assert np.array_equal(h2, h0)
    hm, _ = orc.calc_hash(p, None, pos)
    assert hm.max() < 200


# -------------------------------------------------------------------- sort
@pytest.mark.parametrize("n,hi", [(0, 4), (1, 4), (17, 3), (5000, 64), (100_000, 1 << 20)])
def test_sort_matches_numpy_stable_argsort(n, hi):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, hi, n, dtype=np.uint32)
    vals = rng.integers(0, 1 << 31, n, dtype=np.uint32)
    k, v = orc.sort_pairs(keys, vals)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])


# ----------------------------------------------------------------- reorder
def test_reorder_bounds_match_searchsorted():
    rng = np.random.default_rng(3)
    C = 500
    keys = np.sort(rng.integers(0, C - 50, 3000).astype(np.uint32))
    idx = rng.permutation(3000).astype(np.uint32)
    pos = rng.random((3000, 4)).astype(np.float32)
    stale = np.full(C, 12345, np.uint32)
    cs, ce, sp, sv, st = orc.reorder(keys, idx, pos, pos, None, C, cell_end_init=stale)
    lo = np.searchsorted(keys, np.arange(C), "left")
    hi = np.searchsorted(keys, np.arange(C), "right")
    nonempty = hi > lo
    assert np.array_equal(cs[nonempty], lo[nonempty]) and np.array_equal(ce[nonempty], hi[nonempty])
    assert np.all(cs[~nonempty] == 0xFFFFFFFF) and np.all(ce[~nonempty] == 12345)
    assert np.array_equal(sp, pos[idx]) and np.array_equal(sv, pos[idx]) and st is None


# ------------------------------------------------------------ pair terms
def pair_system(rA, rB, worms=(0, 1), grid=(10, 10, 10), cell=2.0, **kw):
    """Two interacting beads in a 3D grid (multiplicity 1): worm 0 = beads 0, 1,
    worm 1 = beads 2, 3.  The two actors sit near the origin; the other two
    beads are parked three or more cells away from them and from each other."""
    p = sim_params(grid=grid, cell=cell, W=2, n=2, **kw)
    pos = np.zeros((4, 4), np.float32)
    pos[:, :3] = [[7.0, 7.0, 7.0], [7.0, 7.0, 7.0], [-7.0, -7.0, -7.0], [-7.0, -7.0, -7.0]]
    a, b = {(0, 0): (0, 1), (0, 1): (0, 2), (1, 1): (2, 3)}[worms]
    other = [k for k in range(4) if k not in (a, b)]
    pos[other[0], :3] = [7.0, 7.0, 7.0]
    pos[other[1], :3] = [-7.0, -7.0, -7.0]
    pos[a, :3], pos[b, :3] = rA, rB
    return p, pos, a, b


def oracle_pair_force(p, e, pos, vel=None, tan=None):
    n = pos.shape[0]
    vel = np.zeros_like(pos) if vel is None else vel
    tan = np.zeros_like(pos) if tan is None else tan
    h, i = orc.calc_hash(p, None, pos)
    h, i = orc.sort_pairs(h, i)
    cs, ce, sp, sv, st = orc.reorder(h, i, pos, vel, tan, p.numCells)
    return orc.collide_filaments(p, e, sp, sv, st, i, cs, ce, np.zeros((n, 4), np.float32))


def fd_grad(U, x, h=1e-5):
    g = np.zeros(3)
    for k in range(3):
        d = np.zeros(3)
        d[k] = h
        g[k] = (U(x + d) - U(x - d)) / (2 * h)
    return g


def test_bond_force_is_minus_grad_of_harmonic_energy():
    k, L = 57.0, 0.4
    rA, rB = np.array([0.1, 0.2, 0.05]), np.array([0.55, 0.31, -0.02])
    p, pos, a, b = pair_system(rA, rB, worms=(0, 0), spring=0.0, damping=0.0, shear=0.0, activity=0.0)
    F = oracle_pair_force(p, ext_params(), pos)
    U = lambda x: 0.5 * k * (np.linalg.norm(rB.astype(np.float32).astype(float) - x) - L) ** 2
    g = fd_grad(U, pos[a, :3].astype(float))
    assert np.allclose(F[a, :3], -g, rtol=1e-6, atol=1e-8)
    assert np.allclose(F[b, :3], g, rtol=1e-6, atol=1e-8)


def test_dem_spring_fd_and_dashpot_decomposition():
    rA, rB = np.array([0.0, 0.0, 0.0]), np.array([0.6, 0.3, 0.1])
    p, pos, a, b = pair_system(rA, rB, worms=(0, 1), spring=0.5, damping=0.02, shear=0.1, activity=0.0)
    e = ext_params()
    F0 = oracle_pair_force(p, e, pos)
    rb = pos[b, :3].astype(float)
    U = lambda x: 0.5 * 0.5 * (1.0 - np.linalg.norm(rb - x)) ** 2
    assert np.allclose(F0[a, :3], -fd_grad(U, pos[a, :3].astype(float)), rtol=1e-6, atol=1e-9)
    vel = np.zeros_like(pos)
    vel[b, :3] = [0.3, -0.7, 0.2]
    F1 = oracle_pair_force(p, e, pos, vel=vel)
    vel2 = vel.copy()
    vel2[b, :3] *= 2
    F2 = oracle_pair_force(p, e, pos, vel=vel2)
    dF1, dF2 = F1[a, :3] - F0[a, :3], F2[a, :3] - F0[a, :3]
    assert np.allclose(dF2, 2 * dF1, rtol=1e-6)  # linear in relative velocity
    nrm = (rb - pos[a, :3]) / np.linalg.norm(rb - pos[a, :3])
    vr = vel[b, :3].astype(float)
    vn = np.dot(vr, nrm) * nrm
    assert np.allclose(dF1, 0.02 * vn + (0.02 + 0.1) * (vr - vn), rtol=1e-6, atol=1e-10)
    # out of contact: nothing
    rB2 = np.array([1.2, 0.0, 0.0])
    p, pos, a, b = pair_system(rA, rB2, worms=(0, 1), activity=0.0)
    assert np.allclose(oracle_pair_force(p, e, pos)[a, :3], 0.0)


def test_wca_force_fd_cutoff_and_epsilon_scaling():
    sig, eps = 1.0, 1.3
    rc = 2 ** (1 / 6) * sig
    rA, rB = np.array([0.0, 0.0, 0.0]), np.array([0.7, 0.5, 0.2])
    p, pos, a, b = pair_system(rA, rB, worms=(0, 1), activity=0.0)
    e = ext_params(pairPotential=abi.NW_PAIR_WCA, wcaEpsilon=eps, wcaSigma=sig, wcaCutoff=rc)
    F = oracle_pair_force(p, e, pos)
    rb = pos[b, :3].astype(float)
    U = lambda x: 4 * eps * ((sig / np.linalg.norm(rb - x)) ** 12 - (sig / np.linalg.norm(rb - x)) ** 6)
    assert np.allclose(F[a, :3], -fd_grad(U, pos[a, :3].astype(float), 1e-6), rtol=1e-6)
    assert np.allclose(F[b, :3], -F[a, :3], rtol=1e-12)
    e2 = ext_params(pairPotential=abi.NW_PAIR_WCA, wcaEpsilon=2 * eps, wcaSigma=sig, wcaCutoff=rc)
    assert np.allclose(oracle_pair_force(p, e2, pos)[a, :3], 2 * F[a, :3], rtol=1e-12)
    # the force vanishes at r_c and beyond
    for r in (rc * (1 - 1e-7), rc * 1.0001, 1.5):
        p, pos, a, b = pair_system(rA, np.array([r, 0, 0]), worms=(0, 1), activity=0.0)
        assert np.linalg.norm(oracle_pair_force(p, e, pos)[a, :3]) < 1e-5
    # same worm: excluded
    p, pos, a, b = pair_system(rA, rB, worms=(1, 1), activity=0.0)
    assert np.allclose(oracle_pair_force(p, e, pos)[a, :3], 0.0)


def test_extensile_closed_form_and_threshold():
    A = 0.7
    rA, rB = np.array([0.0, 0.0, 0.0]), np.array([1.4, 0.2, 0.0])  # beyond DEM contact
    p, pos, a, b = pair_system(rA, rB, worms=(0, 1), activity=A)
    tan = np.zeros_like(pos)
    tA, tB = np.array([1.0, 0.0, 0.0]), np.array([-0.8, 0.6, 0.0])
    tan[a, :3], tan[b, :3] = tA, tB
    F = oracle_pair_force(p, ext_params(), pos, tan=tan)
    d = np.linalg.norm(pos[b, :3].astype(float) - pos[a, :3])
    assert np.allclose(F[a, :3], -A * (tA - tB) / (2 * d), rtol=1e-6)
    assert np.allclose(F[b, :3], -F[a, :3], rtol=1e-6)
    tan[b, :3] = [-0.4, math.sqrt(1 - 0.16), 0.0]  # dot = -0.4 > -0.5: no force
    assert np.allclose(oracle_pair_force(p, ext_params(), pos, tan=tan)[a, :3], 0.0)


def brute_force(p, e, pos, vel, tan):
    """O(N^2) sum of the reference pair terms over every pair whose cells are
    neighbours in the 27-cell stencil (3D grid, every axis >= 3 cells)."""
    n = pos.shape[0]
    fs = p.filamentSize
    L = np.array([p.boxSize.x, p.boxSize.y, p.boxSize.z])
    o = np.array([p.origin.x, p.origin.y, p.origin.z], np.float32)
    c = np.array([p.cellSize.x, p.cellSize.y, p.cellSize.z], np.float32)
    G = np.array([p.gridSize.x, p.gridSize.y, p.gridSize.z])
    cell = np.floor((pos[:, :3] - o) / c).astype(np.int64) % G
    r = pos[:, :3].astype(np.float64)
    F = np.zeros((n, 3))
    R2 = 2 * p.particleRadius
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            dc = (cell[j] - cell[i] + G) % G
            if not np.all((dc <= 1) | (dc >= G - 1)):
                continue
            d = r[j] - r[i]
            d -= L * np.round(d / L)
            dist = np.linalg.norm(d)
            same = i // fs == j // fs
            if e.pairPotential == abi.NW_PAIR_WCA:
                if not same and dist < e.wcaCutoff:
                    s6 = (e.wcaSigma / dist) ** 6
                    F[i] -= 24 * e.wcaEpsilon * s6 * (2 * s6 - 1) / dist ** 2 * d
            elif same and abs(i % fs - j % fs) == 1:
                F[i] += -p.bondSpringK * (p.bondSpringL - dist) * d / dist
            elif dist < R2:
                nrm = d / dist
                vr = vel[j, :3].astype(float) - vel[i, :3]
                vt = vr - np.dot(vr, nrm) * nrm
                F[i] += -p.spring * (R2 - dist) * nrm + p.damping * vr + p.shear * vt
            if not same and (e.pairPotential != abi.NW_PAIR_WCA or p.activity != 0):
                tA, tB = tan[i, :3].astype(float), tan[j, :3].astype(float)
                if np.dot(tA, tB) < -0.5:
                    F[i] += -p.activity * (tA - tB) / (2 * dist)
    return F


@pytest.mark.parametrize("mode", ["ref", "wca"])
def test_cell_list_equals_brute_force_and_newton3(mode):
    rng = np.random.default_rng(7 if mode == "ref" else 8)
    p = sim_params(grid=(5, 6, 4), cell=1.3, W=14, n=6, bondSpringL=0.5)
    pos, vel, tan = random_worms(p, rng, bond=0.5)
    e = ext_params(pairPotential=abi.NW_PAIR_WCA if mode == "wca" else abi.NW_PAIR_REF_DEM)
    F = oracle_pair_force(p, e, pos, vel, tan)[:, :3]
    B = brute_force(p, e, pos, vel, tan)
    assert np.allclose(F, B, rtol=1e-9, atol=1e-9)
    # every pair term is antisymmetric: the total vanishes
    assert np.abs(F.sum(axis=0)).max() < 1e-9 * max(1.0, np.abs(F).max())


def test_two_dimensional_grid_counts_its_layer_three_times():
    rng = np.random.default_rng(9)
    p = sim_params(grid=(16, 16, 1), cell=1.0, W=40, n=6)
    pos, vel, tan = random_worms(p, rng, twod=True)
    F27 = oracle_pair_force(p, ext_params(stencil=abi.NW_STENCIL_REF27), pos, vel, tan)
    F9 = oracle_pair_force(p, ext_params(stencil=abi.NW_STENCIL_2D), pos, vel, tan)
    assert np.allclose(F27, 3 * F9, rtol=1e-12, atol=1e-12)


# ----------------------------------------------------------------- bonded
def bending_energy(r, kappa, L):
    U = 0.0
    for q in range(len(r) - 2):
        a = r[q + 1] - r[q]
        b = r[q + 2] - r[q + 1]
        a -= L * np.round(a / L)
        b -= L * np.round(b / L)
        U += -kappa * np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b))
    return U


@pytest.mark.parametrize("n", [3, 5, 9])
def test_bending_is_minus_grad_of_cosine_energy(n):
    rng = np.random.default_rng(n)
    p = sim_params(grid=(8, 8, 8), cell=1.0, W=1, n=n, bondBendingK=2.5, bondSpringL=0.8)
    pos, vel, tan = random_worms(p, rng, kink=0.6)
    zero = np.zeros_like(pos)
    F, T = orc.filament_forces(p, ext_params(), zero, pos, tan, 1)
    L = np.array([p.boxSize.x, p.boxSize.y, p.boxSize.z])
    r = pos[:, :3].astype(np.float64)
    for q in range(n):
        def U(x, q=q):
            rr = r.copy()
            rr[q] = x
            return bending_energy(rr, 2.5, L)
        assert np.allclose(F[q, :3], -fd_grad(U, r[q], 1e-6), rtol=1e-5, atol=1e-7)
    assert np.abs(F[:, :3].sum(axis=0)).max() < 1e-12
    # zero torque about the worm's first bead (unwrapped chain)
    ru = np.cumsum(np.vstack([r[:1], np.diff(r, axis=0) - L * np.round(np.diff(r, axis=0) / L)]), axis=0)
    assert np.abs(np.cross(ru - ru[0], F[:, :3]).sum(axis=0)).max() < 1e-10
    # tangents: unit vectors along the chord of the neighbours
    nrm = np.linalg.norm(T[:, :3], axis=1)
    written = nrm > 0
    assert np.allclose(nrm[written], 1.0)


def test_straight_worm_feels_no_bending_and_chain_terms_fd():
    p = sim_params(grid=(8, 8, 8), cell=1.0, W=1, n=6, bondBendingK=4.0, bondSpringL=0.8, bondSpringK=57.0)
    pos = np.zeros((6, 4), np.float32)
    pos[:, 0] = np.arange(6) * 0.8 - 2.0
    zero = np.zeros_like(pos)
    F, T = orc.filament_forces(p, ext_params(), zero, pos, zero, 1)
    assert np.abs(F[:, :3]).max() < 1e-12
    assert np.allclose(T[:, :3], [[1.0, 0.0, 0.0]] * 6)
    # chain mode: bonds are -grad of sum k/2 (|d| - r0)^2, drive = f_a unit(r[p+1] - r[p]) for p < n-1
    rng = np.random.default_rng(4)
    pos[:, :3] += rng.normal(0, 0.1, (6, 3)).astype(np.float32)
    e = ext_params(bondedMode=abi.NW_BONDED_CHAIN, driveForce=0.0)
    p.bondBendingK = 0.0
    F, _ = orc.filament_forces(p, e, zero, pos, zero, 1)
    r = pos[:, :3].astype(np.float64)

    def Ub(x, q):
        rr = r.copy()
        rr[q] = x
        return sum(0.5 * 57.0 * (np.linalg.norm(rr[i + 1] - rr[i]) - 0.8) ** 2 for i in range(5))
    for q in range(6):
        assert np.allclose(F[q, :3], -fd_grad(lambda x: Ub(x, q), r[q]), rtol=1e-5, atol=1e-7)
    e.driveForce = 1.7
    Fd, _ = orc.filament_forces(p, e, zero, pos, zero, 1)
    drive = Fd[:, :3] - F[:, :3]
    for q in range(5):
        u = (r[q + 1] - r[q]) / np.linalg.norm(r[q + 1] - r[q])
        assert np.allclose(drive[q], float(np.float32(1.7)) * u, rtol=1e-9)
    assert np.allclose(drive[5], 0.0)


# --------------------------------------------------------------- integrator
def test_integrator_free_particle_and_kick():
    p = sim_params(grid=(8, 8, 8), cell=1.0, W=1, n=1)
    pos = np.array([[0.1, -0.2, 0.3, 0.5]], np.float32)
    vel = np.array([[1.0, 2.0, -1.0, 1.0]], np.float32)
    zero = np.zeros((1, 4), np.float32)
    po, vo, fo, foo = orc.integrate_filaments(p, pos, vel, zero, zero, 0.01)
    assert np.allclose(po[0, :3], pos[0, :3] + 0.01 * vel[0, :3], rtol=1e-12)
    assert np.allclose(vo[0, :3], vel[0, :3]) and po[0, 3] == 0.5 and vo[0, 3] == 1.0


def test_velocity_verlet_energy_and_reversibility():
    """Harmonic dimer driven through the oracle's integrator: energy stays within
    O(dt^2) and the trajectory retraces itself after a velocity flip."""
    p = sim_params(grid=(8, 8, 8), cell=1.0, W=1, n=2)
    k, r0, dt = 10.0, 1.0, 0.01
    pos = np.array([[-0.6, 0, 0, 0.5], [0.6, 0.1, 0, 0.5]], np.float64)
    vel = np.array([[0, 0.3, 0, 1], [0, -0.3, 0, 1]], np.float64)

    def force(r):
        d = r[1, :3] - r[0, :3]
        dist = np.linalg.norm(d)
        f = np.zeros((2, 4))
        f[0, :3] = k * (dist - r0) * d / dist
        f[1, :3] = -f[0, :3]
        return f

    def energy(r, v):
        return 0.5 * k * (np.linalg.norm(r[1, :3] - r[0, :3]) - r0) ** 2 + 0.5 * np.sum(v[:, :3] ** 2)

    f_old = force(pos)
    f = f_old.copy()
    E0 = energy(pos, vel)
    traj = []
    for _ in range(500):
        po, vo, _, foo = orc.integrate_filaments(p, pos.astype(np.float32), vel.astype(np.float32),
                                                 f.astype(np.float32), f_old.astype(np.float32), dt)
        pos, vel, f_old = po, vo, foo
        f = force(pos)
        traj.append(energy(pos, vel + 0.5 * (f + f_old) * dt * np.array([1, 1, 1, 0])))
    assert max(abs(e - E0) for e in traj) < 5e-3 * E0
    # reversibility: complete the velocity, flip it, run back
    vel[:, :3] += 0.5 * (f[:, :3] + f_old[:, :3]) * dt
    vel[:, :3] *= -1
    # the next call kicks by (f + f_old) dt / 2 = f dt before drifting
    f_old = f.copy()
    vel[:, :3] -= f[:, :3] * dt
    start = pos.copy()
    for _ in range(500):
        po, vo, _, foo = orc.integrate_filaments(p, pos.astype(np.float32), vel.astype(np.float32),
                                                 f.astype(np.float32), f_old.astype(np.float32), dt)
        pos, vel, f_old = po, vo, foo
        f = force(pos)
    assert np.abs(pos[:, :3] - np.array([[-0.6, 0, 0], [0.6, 0.1, 0]])).max() < 1e-4
    assert np.abs(start).max() > 0


def test_boundaries():
    p = sim_params(grid=(8, 8, 8), cell=1.0, W=1, n=1, boundaryX=abi.PERIODIC, boundaryY=abi.WALL,
                   boundaryZ=abi.PERIODIC, boundaryDamping=-0.5)
    pos = np.array([[3.99, 3.95, 0.0, 0.5]], np.float32)
    vel = np.array([[2.0, 2.0, 0.0, 1.0]], np.float32)
    z = np.zeros((1, 4), np.float32)
    po, vo, _, _ = orc.integrate_filaments(p, pos, vel, z, z, 0.1)
    assert abs(po[0, 0] - (3.99 + 0.2 - 8.0)) < 1e-6          # periodic wrap
    assert po[0, 1] == 4.0 - 0.5 and vo[0, 1] == -1.0          # damped wall: clamp, v *= coef
    p.boundaryY = abi.WALL_NO_SLIP
    po, vo, _, _ = orc.integrate_filaments(p, pos, vel, z, z, 0.1)
    # the whole displacement is undone after x was already wrapped (axes are
    # processed x, y, z: particles_kernel_impl.cuh:99-134), so x ends at
    # 3.99 + 0.2 - 8 - 0.2; y is pinned at the wall
    assert po[0, 1] == 3.5 and abs(po[0, 0] - (3.99 - 8.0)) < 1e-5
    assert vo[0, 1] == -2.0 and vo[0, 0] == 0.0                # normal reflected, tangential removed


# -------------------------------------------------------------------- SRD
def srd_case(rng, Ns=4000):
    p = sim_params(grid=(16, 16, 1), cell=1.0, W=10, n=5)
    s = solvent_params(p, (8, 8, 1), (2.0, 2.0, 1.0), alpha=math.pi / 2, kbT=0.7, offset=(0.3, 1.1, 0.0))
    N = 50
    P = np.zeros((N + Ns, 4), np.float32)
    P[:, 0] = p.origin.x + rng.random(N + Ns) * 16
    P[:, 1] = p.origin.y + rng.random(N + Ns) * 16
    V = np.zeros((N + Ns, 4), np.float32)
    V[:, :2] = rng.normal(0, 1, (N + Ns, 2))
    V[:, 3] = rng.choice([1.0, 2.0], N + Ns)
    h, i = orc.calc_hash(p, s, P, srd=True)
    hs, is_ = orc.sort_pairs(h, i)
    cs, ce, sp, sv, _ = orc.reorder(hs, is_, P, V, None, s.numCells)
    return p, s, P, V, h, hs, is_, cs, ce, sv


def test_srd_com_is_mass_weighted_mean():
    rng = np.random.default_rng(1)
    p, s, P, V, h, hs, is_, cs, ce, sv = srd_case(rng)
    com = orc.srd_com(sv, cs, ce, s.numCells)
    m = V[:, 3].astype(float)
    M = np.bincount(h, m, s.numCells)
    for k in range(2):
        mom = np.bincount(h, m * V[:, k], s.numCells)
        assert np.allclose(com[:, k], np.where(M > 0, mom / np.maximum(M, 1e-30), 0.0), rtol=1e-12)


def test_srd_rotation_conserves_cell_momentum_and_energy():
    rng = np.random.default_rng(2)
    p, s, P, V, h, hs, is_, cs, ce, sv = srd_case(rng)
    com = orc.srd_com(sv, cs, ce, s.numCells, prec="f32")
    uni = orc.random_uniform(5, 0, s.numCells)
    e = ext_params(srdCollideFilaments=1)
    Vn = orc.collide_solvent(p, s, e, P, V, com, uni, s.numCells)
    m = V[:, 3].astype(float)
    for k in range(2):
        assert np.allclose(np.bincount(h, m * Vn[:, k], s.numCells), np.bincount(h, m * V[:, k], s.numCells),
                           atol=1e-5)
    u = com[h, :2].astype(float)
    assert np.allclose(np.linalg.norm(Vn[:, :2] - u, axis=1), np.linalg.norm(V[:, :2] - u, axis=1), rtol=1e-6)
    # alpha = 90 deg: v - u turns by exactly +-90 degrees, sign from the cell's draw
    d0, d1 = V[:, :2] - u, Vn[:, :2] - u
    sgn = np.where(uni[h] > 0.5, 1.0, -1.0)
    assert np.allclose(d1[:, 0], -sgn * d0[:, 1], atol=1e-6) and np.allclose(d1[:, 1], sgn * d0[:, 0], atol=1e-6)
    # reference mode: beads do not rotate
    Vr = orc.collide_solvent(p, s, ext_params(), P, V, com, uni, s.numCells)
    assert np.array_equal(Vr[: p.numParticles], V[: p.numParticles].astype(np.float64))


def test_isokinetic_postcondition():
    rng = np.random.default_rng(3)
    p, s, P, V, h, hs, is_, cs, ce, sv = srd_case(rng)
    vo, so = orc.srd_isokinetic(s, V, sv, cs, ce, is_, s.numCells)
    KE = np.bincount(h, 0.5 * V[:, 3] * np.sum(vo[:, :3] ** 2, axis=1), s.numCells)
    occupied = np.bincount(h, None, s.numCells) > 0
    assert np.allclose(KE[occupied], float(np.float32(0.7)), rtol=1e-10)
    assert np.array_equal(so, vo[is_])


# --------------------------------------------------------------- Langevin
def test_langevin_closed_form_and_fdt_equipartition():
    rng = np.random.default_rng(4)
    n = 20000
    p = sim_params(W=n, n=1, gamma=2.0)
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 1, (n, 3))
    force = np.zeros((n, 4), np.float32)
    noise = rng.normal(0, 1, (n, 4))
    F = orc.langevin(p, force, vel, noise)
    assert np.allclose(F[:, :3], -2.0 * vel[:, :3] + noise[:, :3], rtol=1e-12)
    # free beads under VV + Langevin with sigma = sqrt(2 gamma kT / dt) relax to <v^2> = kT per component
    kT, dt, gamma = 0.4, 0.01, 2.0
    p.gamma = gamma
    sd = math.sqrt(2 * gamma * kT / dt)
    pos = np.zeros((n, 4), np.float32)
    vel[:, :3] = 0.0
    f = np.zeros((n, 4), np.float32)
    fo = f.copy()
    off = 0
    acc = []
    for step in range(600):
        pos, vel, _, fo = [a.astype(np.float32) for a in orc.integrate_filaments(p, pos, vel, f, fo, dt)]
        pos[:, :3] = 0.0
        z = orc.random_gaussian(7, off, 4 * n, 0.0, sd).reshape(n, 4)
        off += n
        f = orc.langevin(p, np.zeros((n, 4), np.float32), vel, z).astype(np.float32)
        if step >= 300:
            acc.append(np.mean(vel[:, :3].astype(float) ** 2))
    # the VV discretisation biases <v^2> by O(gamma dt); 2 gamma dt = 4 %
    assert abs(np.mean(acc) / kT - 1.0) < 0.05


# ---------------------------------------------------------------- reversal
def test_reversal_reverses_selected_worms():
    rng = np.random.default_rng(6)
    W, n = 50, 7
    p = sim_params(W=W, n=n, reverseProbability=0.4)
    arrs = [rng.random((W * n, 4)).astype(np.float32) for _ in range(4)]
    uni = rng.random(W).astype(np.float32)
    out = orc.reverse_filaments(p, uni, W, *arrs)
    for a, b in zip(arrs, out):
        for w in range(W):
            seg_in, seg_out = a[w * n:(w + 1) * n], b[w * n:(w + 1) * n]
            assert np.array_equal(seg_out, seg_in[::-1] if uni[w] <= 0.4 else seg_in)
    twice = orc.reverse_filaments(p, uni, W, *out)
    assert all(np.array_equal(x, y) for x, y in zip(twice, arrs))


# ---- order parameters (orc.observables) --------------------------------------
def _straight_worms(p, angles, centres, bond=0.8):
    """Straight worms along the given in-plane angles, centred at `centres`."""
    n = p.filamentSize
    pos = np.zeros((len(angles) * n, 4), np.float32)
    for w, (th, c) in enumerate(zip(angles, centres)):
        u = np.array([np.cos(th), np.sin(th), 0.0])
        for q in range(n):
            pos[w * n + q, :3] = c + (q - (n - 1) / 2) * bond * u
    pos[:, :3] = nwtest.wrap_into_box(p, pos[:, :3])
    return pos


def test_observables_aligned_and_antialigned_worms():
    p = sim_params(grid=(40, 40, 1), cell=1.0, W=8, n=6)
    rng = np.random.default_rng(3)
    centres = np.c_[rng.random((8, 2)) * 40 - 20, np.zeros(8)]
    vel = np.zeros((48, 4), np.float32)
    pos = _straight_worms(p, [0.3] * 8, centres)
    o = orc.observables(p, pos, vel, 8)
    assert o["nematicS"] == pytest.approx(1.0, abs=1e-6) and o["polarP"] == pytest.approx(1.0, abs=1e-6)
    assert o["bondS"] == pytest.approx(1.0, abs=1e-6)
    # half the worms reversed: nematic order stays, polar order vanishes
    pos = _straight_worms(p, [0.3, 0.3 + np.pi] * 4, centres)
    o = orc.observables(p, pos, vel, 8)  # FP32 positions: tangents to ~1e-7
    assert o["nematicS"] == pytest.approx(1.0, abs=1e-6) and o["polarP"] == pytest.approx(0.0, abs=1e-6)


def test_observables_isotropic_plane_and_periodic_worms():
    """M straight worms at angles k pi / M: <cos 2 theta> = <sin 2 theta> = 0, so
    the 3D Q tensor of in-plane tangents has eigenvalues 1/4, 1/4, -1/2; a
    worm across the periodic x face keeps its tangent under the minimum image."""
    M = 12
    p = sim_params(grid=(30, 30, 1), cell=1.0, W=M, n=5)
    centres = np.zeros((M, 3))
    centres[:, 0] = np.linspace(-12, 12, M)
    pos = _straight_worms(p, np.arange(M) * np.pi / M, centres)
    o = orc.observables(p, pos, np.zeros((M * 5, 4), np.float32), M)
    assert o["nematicS"] == pytest.approx(0.25, abs=1e-6)
    q = sim_params(grid=(30, 30, 1), cell=1.0, W=1, n=5)
    pos = _straight_worms(q, [0.0], np.array([[14.9, 0.0, 0.0]]))
    assert pos[0, 0] > 13.0 and pos[-1, 0] < -13.0  # the worm wraps
    vel = np.zeros((5, 4), np.float32)
    vel[:, 0] = 2.0
    o = orc.observables(q, pos, vel, 1)
    assert o["nematicS"] == pytest.approx(1.0, abs=1e-6) and o["polarP"] == pytest.approx(1.0, abs=1e-6)
    assert o["polarSpeed"] == pytest.approx(2.0, abs=1e-6) and o["meanV2"] == pytest.approx(4.0)


def test_library_sym_eig3_matches_numpy():
    """nwSymEig3Max (host code in libnwamd.so, no device) against LAPACK."""
    L = abi.load()
    rng = np.random.default_rng(11)
    for k in range(200):
        A = rng.normal(size=(3, 3))
        A = A + A.T
        if k % 4 == 0:
            A[0, 1] = A[1, 0] = A[0, 2] = A[2, 0] = A[1, 2] = A[2, 1] = 0.0
        a6 = (ctypes.c_double * 6)(A[0, 0], A[0, 1], A[0, 2], A[1, 1], A[1, 2], A[2, 2])
        assert L.nwSymEig3Max(a6) == pytest.approx(np.linalg.eigvalsh(A)[-1], abs=1e-12)

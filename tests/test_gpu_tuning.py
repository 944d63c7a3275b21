"""Launch shapes and the autotuner (nwSetTuning, nwAutotuneFused): every
tuned launch shape of the fused steps gives the default's results bit for
bit -- positions, velocities, forces, tangents, the cell order, the sorted
records, SRD cell momenta and signs -- and the tuner picks, applies and
caches a choice without touching the caller's state."""
import ctypes
import os

import numpy as np
import pytest

from nwamd import abi
from nwamd.abi import DeviceArray, NWTuning
from nwtest import ext_params, lattice_system_2d, random_forces, solvent_params

pytestmark = pytest.mark.gpu

W, NB, G = 3000, 20, 260
NS_PER_CELL = 5


@pytest.fixture
def lib(nwlib):
    yield nwlib
    nwlib.nwSetTuning(ctypes.byref(NWTuning()))  # back to the built-in shapes


def system(lib):
    rng = np.random.default_rng(31)
    p, pos, vel = lattice_system_2d(W, NB, G, rng)
    e = ext_params(pairPotential=abi.NW_PAIR_WCA, bondedMode=abi.NW_BONDED_CHAIN, driveForce=1.0,
                   stencil=abi.NW_STENCIL_2D, srdCollideFilaments=1, srdThermostat=abi.NW_SRD_THERMO_NONE)
    L = p.boxSize.x
    Gs = int(L)
    Ns = NS_PER_CELL * Gs * Gs
    s = solvent_params(p, (Gs, Gs, 1), (L / Gs, L / Gs, p.boxSize.z), alpha=np.pi / 2, kbT=0.1,
                       offset=(0.17, -0.31, 0.0), ns=Ns)
    sp = np.zeros((Ns, 4), np.float32)
    sp[:, :2] = rng.random((Ns, 2)) * L + p.origin.x
    sp[:, 2] = pos[0, 2]
    sv = np.zeros((Ns, 4), np.float32)
    sv[:, :2] = rng.normal(0, 0.3, (Ns, 2))
    sv[:, 3] = 1.0
    N = W * NB
    f, fo = random_forces(rng, N, 0.5), random_forces(rng, N, 0.5)
    f[:, 3] = fo[:, 3] = 0
    lib.setParameters(ctypes.byref(p))
    lib.setSolventParameters(ctypes.byref(s))
    lib.nwSetExtParameters(ctypes.byref(e))
    return p, s, np.concatenate([pos, sp]), np.concatenate([vel, sv]), f, fo


def step(lib, p, s, P, V, f, fo, tuning):
    """one fused MD step and one fused SRD collision under `tuning`: every output as raw bits"""
    lib.nwSetTuning(ctypes.byref(tuning))
    assert lib.nwGetLastError() == 0
    N, NT = p.numParticles, len(P)
    d = {k: DeviceArray.from_numpy(lib, a) for k, a in dict(P=P, V=V, f=f, fo=fo, t=np.zeros_like(f)).items()}
    h, i = DeviceArray(lib, 4 * NT), DeviceArray(lib, 4 * NT)
    sp, sv, st = (DeviceArray(lib, 16 * N) for _ in range(3))
    com = DeviceArray(lib, 16 * s.numCells)
    lib.nwRngSetSeed(5, 50)
    lib.nwFusedMdStep(d["P"].p, d["V"].p, d["f"].p, d["fo"].p, d["t"].p, h.p, i.p, sp.p, sv.p, st.p, 1e-3, 2.0, 0.0,
                      0, W)
    out = {"i_md": i.download(np.uint32, (N,)), "h_md": h.download(np.uint32, (N,)),
           "sp": sp.download(np.uint32, (N, 4))}
    lib.nwFusedSrdStep(d["P"].p, d["V"].p, com.p, h.p, i.p, 1e-2, N, NT - N, 0)
    lib.threadSync()
    for k in ("P", "V", "f", "fo", "t"):
        out[k] = d[k].download(np.uint32, (len(P) if k in "PV" else N, 4))
    out["com"] = com.download(np.uint32, (s.numCells, 4))
    out["i_srd"] = i.download(np.uint32, (NT,))
    out["h_srd"] = h.download(np.uint32, (NT,))
    return out


def candidates():
    yield "default", NWTuning()
    for fb in (64, 128, 512, 1024):
        yield f"frontBlock={fb}", NWTuning(frontBlock=fb)
    for items in (1024, 2048, 4096, 8192):
        for mem in (256, 512, 2048):
            yield f"md {items}/{mem}", NWTuning(mdTileItems=items, mdMembers=mem)
    for pb in (128, 256, 512):
        for stage in (0, 2 * pb):
            yield f"pairBlock={pb} pairStage={stage}", NWTuning(pairBlock=pb, pairStage=stage)
    for items in (4096, 8192, 16384):
        for mem in (768, 2048):
            yield f"srd {items}/{mem}", NWTuning(srdTileItems=items, srdMembers=mem)


def test_every_tuning_gives_the_default_results_bit_for_bit(lib):
    p, s, P, V, f, fo = system(lib)
    ref = step(lib, p, s, P, V, f, fo, NWTuning())
    assert not np.array_equal(ref["P"], P.view(np.uint32))
    plans = set()
    for name, t in candidates():
        got = step(lib, p, s, P, V, f, fo, t)
        for k in ref:
            assert np.array_equal(got[k], ref[k]), (name, k)
        plan = (ctypes.c_uint * 4)()
        lib.nwFusedPlanInfo(p.numParticles, p.numCells, NB, plan)
        plans.add(tuple(plan))
    assert len(plans) >= 6  # the tile and bucket choices took effect


def test_invalid_tuning_is_refused(lib):
    lib.nwSetTuning(ctypes.byref(NWTuning(pairBlock=512)))
    lib.nwSetTuning(ctypes.byref(NWTuning(pairBlock=100)))
    assert lib.nwGetLastError() == -22
    t = NWTuning()
    lib.nwGetTuning(ctypes.byref(t))
    assert t.pairBlock == 512  # unchanged


def test_autotuner_picks_applies_and_caches(lib, tmp_path):
    p, s, P, V, f, fo = system(lib)
    N = p.numParticles
    d = {k: DeviceArray.from_numpy(lib, a) for k, a in dict(P=P, V=V, f=f, fo=fo, t=np.zeros_like(f)).items()}
    lib.nwRngSetSeed(9, 99)
    cache = str(tmp_path / "tune" / "tuning.txt").encode()
    out = NWTuning()
    args = (d["P"].p, d["V"].p, d["f"].p, d["fo"].p, d["t"].p, W, len(P) - N, 1e-3, 2.0, 0.0, 0, 1e-2, 0, 3)
    assert lib.nwAutotuneFused(*args, cache, 1, ctypes.byref(out)) == 1
    cur = NWTuning()
    lib.nwGetTuning(ctypes.byref(cur))
    assert cur.as_dict() == out.as_dict()
    lines = open(cache.decode()).read().splitlines()
    assert len(lines) == 1 and lines[0].split("\t")[1] == f"{N} {p.numCells} {NB} {len(P) - N} {s.numCells}"
    # the caller's state and the generator are untouched
    assert np.array_equal(d["P"].download(np.float32, P.shape), P)
    assert np.array_equal(d["f"].download(np.float32, f.shape), f)
    seed, off = ctypes.c_uint64(), ctypes.c_uint64()
    lib.nwRngGetState(ctypes.byref(seed), ctypes.byref(off))
    assert (seed.value, off.value) == (9, 99)
    # from the cache: no timing, the same choice
    lib.nwSetTuning(ctypes.byref(NWTuning()))
    again = NWTuning()
    assert lib.nwAutotuneFused(*args, cache, 1, ctypes.byref(again)) == 2
    assert again.as_dict() == out.as_dict()
    print("tuned:", out.as_dict(), flush=True)

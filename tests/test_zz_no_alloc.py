"""nwFusedMdStepNoAlloc (the allocation-free variant of nwFusedMdStep: the
cell sort's scratch in a fixed device array of the library) against
nwFusedMdStep through the C-ABI: every output bit for bit, on two calls in a
row -- the second with a smaller system, so anything the first call left in
the shared array would show -- with and without the caller's cell array;
a system beyond the array's capacity is refused before any launch."""
import ctypes

import numpy as np
import pytest

from nwamd import abi
from nwamd.abi import DeviceArray
from nwtest import ext_params, lattice_system_2d as lattice_system, random_forces

pytestmark = pytest.mark.gpu

OUT = ("pos", "vel", "f", "fo", "tan", "i", "sp")


def run(lib, entry, p, e, pos, vel, f, fo, with_hash):
    N = len(pos)
    d = {k: DeviceArray.from_numpy(lib, a) for k, a in dict(pos=pos, vel=vel, f=f, fo=fo,
                                                               tan=np.zeros_like(pos)).items()}
    d["h"] = DeviceArray(lib, 4 * N)
    d["i"] = DeviceArray(lib, 4 * N)
    d["sp"] = DeviceArray(lib, 16 * N)
    lib.setParameters(ctypes.byref(p))
    lib.nwSetExtParameters(ctypes.byref(e))
    lib.nwRngSetSeed(99, 4321)
    getattr(lib, entry)(d["pos"].p, d["vel"].p, d["f"].p, d["fo"].p, d["tan"].p, d["h"].p if with_hash else None,
                        d["i"].p, d["sp"].p, None, None, 1e-3, 2.0, 100.0, 1, p.numFilaments)
    lib.threadSync()
    out = {k: d[k].download(np.uint32, (N, 4) if k not in ("i",) else (N,)) for k in OUT}
    if with_hash:
        out["h"] = d["h"].download(np.uint32, (N,))
    seed, off = ctypes.c_uint64(), ctypes.c_uint64()
    lib.nwRngGetState(ctypes.byref(seed), ctypes.byref(off))
    out["rng"] = off.value
    return out


@pytest.mark.parametrize("with_hash", [True, False])
def test_no_alloc_step_equals_fused_step(nwlib, with_hash):
    lib = nwlib
    rng = np.random.default_rng(7)
    e = ext_params(pairPotential=abi.NW_PAIR_WCA, bondedMode=abi.NW_BONDED_CHAIN, driveForce=1.0,
                   stencil=abi.NW_STENCIL_2D)
    # the larger system first (1.2M beads: ~21 MB of the 32 MiB array), then
    # a smaller one over the same array
    for W, n, G in ((60000, 20, 1100), (1500, 13, 160)):
        p, pos, vel = lattice_system(W, n, G, rng)
        N = W * n
        f, fo = random_forces(rng, N, 0.5), random_forces(rng, N, 0.5)
        f[:, 3] = fo[:, 3] = 0
        a = run(lib, "nwFusedMdStep", p, e, pos, vel, f, fo, with_hash)
        b = run(lib, "nwFusedMdStepNoAlloc", p, e, pos, vel, f, fo, with_hash)
        for k in a:
            assert np.array_equal(a[k], b[k]), (W, n, k)
        assert not np.array_equal(a["pos"], pos.view(np.uint32))  # the step moved the worms


def test_no_alloc_step_refuses_systems_beyond_the_array(nwlib):
    """2.4M beads over 1550^2 cells need ~40 MB of scratch: refused with the
    library's error, before anything is launched (the arrays are full size,
    so nothing could run out of bounds either way)."""
    lib = nwlib
    rng = np.random.default_rng(8)
    W, n, G = 120000, 20, 1550
    p, pos, vel = lattice_system(W, n, G, rng)
    N = W * n
    e = ext_params(pairPotential=abi.NW_PAIR_WCA, bondedMode=abi.NW_BONDED_CHAIN, driveForce=1.0,
                   stencil=abi.NW_STENCIL_2D)
    zero = np.zeros((N, 4), np.float32)
    d = [DeviceArray.from_numpy(lib, a) for a in (pos, vel, zero, zero, zero)]
    i, sp = DeviceArray(lib, 4 * N), DeviceArray(lib, 16 * N)
    lib.setParameters(ctypes.byref(p))
    lib.nwSetExtParameters(ctypes.byref(e))
    assert lib.nwGetLastError() == 0
    lib.nwFusedMdStepNoAlloc(d[0].p, d[1].p, d[2].p, d[3].p, d[4].p, None, i.p, sp.p, None, None, 1e-3, 2.0, 100.0,
                             1, W)
    assert lib.nwGetLastError() == -22
    assert b"scratch" in lib.nwGetLastErrorString()
    lib.threadSync()
    assert np.array_equal(d[0].download(np.float32, (N, 4)), pos)  # nothing ran

"""Order parameters on the device (nwObservables) against the oracle's
restatement, and the long-run statistics of the GPU step against the CPU
oracle's step on the same initial state and noise."""
import ctypes

import numpy as np
import pytest

import orc
from nwamd import abi, system
from nwamd.abi import DeviceArray
from nwtest import random_worms, sim_params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("W,n,grid", [(300, 10, (40, 40, 1)), (37, 64, (30, 30, 1)), (50, 3, (20, 20, 4)),
                                      (10, 1, (10, 10, 1)), (91, 20, (60, 60, 1))])
def test_observables_match_oracle(nwlib, W, n, grid):
    rng = np.random.default_rng([W, n, grid[2]])
    p = sim_params(grid=grid, cell=1.0, W=W, n=n)
    pos, vel, _ = random_worms(p, rng, twod=grid[2] == 1)
    lib = nwlib
    lib.setParameters(ctypes.byref(p))
    dP, dV = DeviceArray.from_numpy(lib, pos), DeviceArray.from_numpy(lib, vel)
    o = abi.NWObservables()
    lib.nwObservables(dP.p, dV.p, W, ctypes.byref(o))
    ref = orc.observables(p, pos, vel, W)
    # the same FP32 differences, FP64 sums in another order
    np.testing.assert_allclose(np.array(o.sums), ref["sums"], rtol=1e-10, atol=1e-9)
    for k in ("nematicS", "polarP", "bondS", "polarSpeed", "meanV2"):
        assert getattr(o, k) == pytest.approx(ref[k], rel=1e-9, abs=1e-9), k
    o2 = abi.NWObservables()
    lib.nwObservables(dP.p, dV.p, W, ctypes.byref(o2))
    assert bytes(o2) == bytes(o)  # fixed reduction order


def test_long_run_order_statistics_match_oracle(nwlib):
    """400 worms x 10 beads (WCA + bonds + bending + drive + Langevin), 2000
    steps on the GPU and on the CPU oracle from one initial state with the
    same Philox streams: over the second half, the mean nematic order S and
    mean speed along the tangent agree within 5 standard errors (sampled every
    100 steps) plus 0.02."""
    cfg = system.config_dry_4k(seed=2024)
    s = system.System(cfg)
    system.apply_worm_physics(s, kappa=5.0, gamma=2.0, kT=0.1)
    s.reset(system.NW_INIT_RANDOM, 0.1)
    p = s.params
    seed, offset = s.rng_state()
    dt, chunk, chunks = 1e-3, 100, 20
    cs = orc.CpuSystem(p, s.solvent, cfg.ext, s.get(system.POSITION), s.get(system.VELOCITY),
                       s.get(system.TANGENT), s.get(system.FORCE), s.get(system.FORCE_OLD), num_solvent=0,
                       srd=False, srd_gap=5, dt=dt, seed=seed, offset=offset, step=s.step_count,
                       langevin_fdt=True)
    W = p.numFilaments
    g_S, g_v, c_S, c_v = [], [], [], []
    for k in range(chunks):
        s.update(dt, chunk)
        cs.step(chunk)
        if k >= chunks // 2:
            og = s.observables()
            oc = orc.observables(p, cs.get("pos"), cs.get("vel"), W)
            g_S.append(og["nematicS"])
            g_v.append(og["polarSpeed"])
            c_S.append(oc["nematicS"])
            c_v.append(oc["polarSpeed"])
    s.close()
    for a, b in ((g_S, c_S), (g_v, c_v)):
        a, b = np.array(a), np.array(b)
        se = np.sqrt(a.var(ddof=1) / len(a) + b.var(ddof=1) / len(b))
        assert abs(a.mean() - b.mean()) <= 5 * se + 0.02, (a.mean(), b.mean(), se)
    # the drive f_a = 1 against gamma = 2 sets the speed along the tangent near f_a / gamma
    assert 0.2 < np.mean(g_v) < 0.8

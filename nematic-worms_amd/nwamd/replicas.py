"""Replica sets: many independent 2D worm systems of uneven size stepped
together by the batched fused step (include/nw_system.h, NWReplicaSet)."""
import ctypes
import math

import numpy as np

from . import abi
from .system import load

NW_REPLICAS_GRAPH, NW_REPLICAS_LAUNCHES, NW_REPLICAS_ONE_PER_LAUNCH = 0, 1, 2


class NWReplicaSpec(ctypes.Structure):
    _fields_ = [("numFilaments", ctypes.c_uint32), ("filamentSize", ctypes.c_uint32), ("boxSize", ctypes.c_float),
                ("driveForce", ctypes.c_float), ("bendingK", ctypes.c_float), ("activity", ctypes.c_float),
                ("seed", ctypes.c_uint64), ("reserved", ctypes.c_uint32 * 4)]


class NWReplicaPhysics(ctypes.Structure):
    _fields_ = [("bondK", ctypes.c_float), ("bondL", ctypes.c_float), ("wcaEpsilon", ctypes.c_float),
                ("wcaSigma", ctypes.c_float), ("gamma", ctypes.c_float), ("kT", ctypes.c_float),
                ("cellMin", ctypes.c_float), ("langevinFdt", ctypes.c_int), ("reserved", ctypes.c_uint32 * 8)]


def worm_physics(gamma=2.0, kT=0.1, fdt=True):
    """Bonds k = 57, r0 = 0.8; WCA eps = sigma = 1; Langevin gamma, kT."""
    p = NWReplicaPhysics()
    p.bondK, p.bondL, p.wcaEpsilon, p.wcaSigma = 57.0, 0.8, 1.0, 1.0
    p.gamma, p.kT = gamma, kT
    p.cellMin = 2.0 ** (1.0 / 6.0)
    p.langevinFdt = 1 if fdt else 0
    return p


def box_for(beads, area_fraction=0.3, sigma=1.0):
    """Side of the square box holding `beads` discs of diameter sigma at the area fraction."""
    return math.sqrt(beads * math.pi * sigma * sigma / (4.0 * area_fraction))


def replica_specs_config2(R=256, seed=20240601):
    """Independent replicas: worms U[500, 8000], beads per worm U{10..30},
    box at area fraction 0.3, drive f_a log-uniform in [0.1, 10], kappa
    uniform in [1, 20], one seed each."""
    rng = np.random.default_rng(seed)
    specs = (NWReplicaSpec * R)()
    for r in range(R):
        W = int(rng.integers(500, 8001))
        n = int(rng.integers(10, 31))
        s = specs[r]
        s.numFilaments, s.filamentSize = W, n
        s.boxSize = box_for(W * n)
        s.driveForce = float(np.exp(rng.uniform(np.log(0.1), np.log(10.0))))
        s.bendingK = float(rng.uniform(1.0, 20.0))
        s.activity = 0.0
        s.seed = int(rng.integers(1, 2 ** 62))
    return specs


class ReplicaSet:
    def __init__(self, specs, phys):
        self.lib = load()
        self.specs = specs
        self.h = self.lib.nwReplicaSetCreate(ctypes.cast(specs, ctypes.c_void_p), len(specs), ctypes.byref(phys))
        if not self.h:
            raise RuntimeError("nwReplicaSetCreate failed")

    @property
    def R(self):
        return self.lib.nwReplicaSetNumReplicas(self.h)

    def n(self, r=-1):
        return self.lib.nwReplicaSetNumParticles(self.h, r)

    def reset(self, mode, kT):
        self.lib.nwReplicaSetReset(self.h, mode, kT)

    def update(self, dt, steps, form=NW_REPLICAS_GRAPH):
        self.lib.nwReplicaSetUpdate(self.h, dt, steps, form)

    def sync(self):
        self.lib.nwReplicaSetSync(self.h)

    def get(self, which, r):
        out = np.empty((self.n(r), 4), np.float32)
        self.lib.nwReplicaSetGetArray(self.h, which, r, out.ctypes.data_as(ctypes.c_void_p))
        return out

    def set(self, which, r, arr):
        a = np.ascontiguousarray(arr, np.float32)
        assert a.shape == (self.n(r), 4)
        self.lib.nwReplicaSetSetArray(self.h, which, r, a.ctypes.data_as(ctypes.c_void_p))

    def params(self, r):
        p, e = abi.SimParams(), abi.NWExtParams()
        self.lib.nwReplicaSetParams(self.h, r, ctypes.byref(p), ctypes.byref(e))
        return p, e

    def rng_state(self, r):
        s, o = ctypes.c_uint64(), ctypes.c_uint64()
        self.lib.nwReplicaSetRngState(self.h, r, ctypes.byref(s), ctypes.byref(o))
        return s.value, o.value

    def observables(self, r):
        o = abi.NWObservables()
        self.lib.nwReplicaSetObservables(self.h, r, ctypes.byref(o))
        return dict(nematicS=o.nematicS, polarP=o.polarP, bondS=o.bondS, polarSpeed=o.polarSpeed, meanV2=o.meanV2)

    def close(self):
        if self.h:
            self.lib.nwReplicaSetDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

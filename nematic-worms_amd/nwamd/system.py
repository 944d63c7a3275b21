"""ctypes binding of libnwhost.so (include/nw_system.h): the ParticleSystem
host driver, plus the benchmark workloads built from it."""
import ctypes
import math
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnwhost.so")

NW_INIT_REFERENCE, NW_INIT_RANDOM, NW_INIT_LATTICE_RANDOM = 0, 1, 2
POSITION, VELOCITY, TANGENT, FORCE, FORCE_OLD = 0, 1, 2, 3, 4


class NWSystemConfig(ctypes.Structure):
    _fields_ = [
        ("numFilaments", ctypes.c_uint32), ("filamentSize", ctypes.c_uint32), ("gridSize", abi.UInt3),
        ("cellSize", ctypes.c_float), ("boxSize", abi.Float3), ("srd", ctypes.c_uint32),
        ("srdGridSize", abi.UInt3), ("srdCellSize", ctypes.c_float), ("srdDensity", ctypes.c_float),
        ("srdGap", ctypes.c_uint32), ("srdAlpha", ctypes.c_float), ("srdShiftCentred", ctypes.c_uint32),
        ("langevinFdt", ctypes.c_uint32), ("seed", ctypes.c_uint64), ("ext", abi.NWExtParams),
        ("perCallSequence", ctypes.c_uint32), ("solventSortEvery", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 6),
    ]


_lib = None


def load():
    global _lib
    if _lib is None:
        abi.load()  # libnwamd.so first (RTLD_GLOBAL)
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libnwhost.so not built at {LIB_PATH}")
        L = ctypes.CDLL(LIB_PATH)
        P, u32, u64, f = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float
        protos = {
            "nwSystemCreate": (P, [ctypes.POINTER(NWSystemConfig)]),
            "nwSystemDestroy": (None, [P]),
            "nwSystemParams": (ctypes.POINTER(abi.SimParams), [P]),
            "nwSystemSolventParams": (ctypes.POINTER(abi.SolventParams), [P]),
            "nwSystemReset": (None, [P, ctypes.c_int, f]),
            "nwSystemUpdate": (None, [P, f, u32]),
            "nwSystemSync": (None, [P]),
            "nwSystemNumParticles": (u32, [P]),
            "nwSystemNumSolvent": (u32, [P]),
            "nwSystemNumCells": (u32, [P]),
            "nwSystemStep": (u64, [P]),
            "nwSystemGetArray": (None, [P, ctypes.c_int, P]),
            "nwSystemSetArray": (None, [P, ctypes.c_int, P]),
            "nwSystemDeviceArray": (P, [P, ctypes.c_int]),
            "nwSystemRngState": (None, [P, ctypes.POINTER(u64), ctypes.POINTER(u64)]),
            "nwSystemSetRngState": (None, [P, u64]),
            "nwSystemWriteOutputs": (None, [P, ctypes.c_char_p, ctypes.c_int, f]),
            "nwSystemObservables": (None, [P, ctypes.POINTER(abi.NWObservables)]),
            "nwSystemAutotune": (ctypes.c_int, [P, f, ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p]),
            "nwReplicaSetCreate": (P, [ctypes.c_void_p, u32, ctypes.c_void_p]),
            "nwReplicaSetDestroy": (None, [P]),
            "nwReplicaSetReset": (None, [P, ctypes.c_int, f]),
            "nwReplicaSetUpdate": (None, [P, f, u32, ctypes.c_int]),
            "nwReplicaSetSync": (None, [P]),
            "nwReplicaSetNumReplicas": (u32, [P]),
            "nwReplicaSetNumParticles": (u64, [P, ctypes.c_int]),
            "nwReplicaSetGetArray": (None, [P, ctypes.c_int, ctypes.c_int, P]),
            "nwReplicaSetSetArray": (None, [P, ctypes.c_int, ctypes.c_int, P]),
            "nwReplicaSetParams": (None, [P, ctypes.c_int, ctypes.POINTER(abi.SimParams),
                                          ctypes.POINTER(abi.NWExtParams)]),
            "nwReplicaSetRngState": (None, [P, ctypes.c_int, ctypes.POINTER(u64), ctypes.POINTER(u64)]),
            "nwReplicaSetObservables": (None, [P, ctypes.c_int, ctypes.POINTER(abi.NWObservables)]),
        }
        for name, (res, args) in protos.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class System:
    """One ParticleSystem (simulation/particleSystem.cpp) on the device."""

    def __init__(self, cfg):
        self.lib = load()
        self.cfg = cfg
        self.h = self.lib.nwSystemCreate(ctypes.byref(cfg))
        if not self.h:
            raise RuntimeError("nwSystemCreate failed")

    @property
    def params(self):
        return self.lib.nwSystemParams(self.h).contents

    @property
    def solvent(self):
        return self.lib.nwSystemSolventParams(self.h).contents

    @property
    def n(self):
        return self.lib.nwSystemNumParticles(self.h)

    @property
    def ns(self):
        return self.lib.nwSystemNumSolvent(self.h)

    @property
    def num_cells(self):
        return self.lib.nwSystemNumCells(self.h)

    @property
    def step_count(self):
        return self.lib.nwSystemStep(self.h)

    def reset(self, mode=NW_INIT_REFERENCE, kT=0.0):
        self.lib.nwSystemReset(self.h, mode, kT)

    def update(self, dt, steps=1):
        self.lib.nwSystemUpdate(self.h, dt, steps)

    def sync(self):
        self.lib.nwSystemSync(self.h)

    def _count(self, which):
        return self.n + self.ns if which in (POSITION, VELOCITY) else self.n

    def get(self, which):
        out = np.empty((self._count(which), 4), np.float32)
        self.lib.nwSystemGetArray(self.h, which, out.ctypes.data_as(ctypes.c_void_p))
        return out

    def set(self, which, arr):
        a = np.ascontiguousarray(arr, np.float32)
        assert a.shape == (self._count(which), 4)
        self.lib.nwSystemSetArray(self.h, which, a.ctypes.data_as(ctypes.c_void_p))

    def rng_state(self):
        s, o = ctypes.c_uint64(), ctypes.c_uint64()
        self.lib.nwSystemRngState(self.h, ctypes.byref(s), ctypes.byref(o))
        return s.value, o.value

    def device(self, which):
        """Device pointer of one of the system's arrays (POSITION ... FORCE_OLD)."""
        return self.lib.nwSystemDeviceArray(self.h, which)

    def observables(self):
        """nwObservables of the current configuration as a dict."""
        o = abi.NWObservables()
        self.lib.nwSystemObservables(self.h, ctypes.byref(o))
        return dict(nematicS=o.nematicS, polarP=o.polarP, bondS=o.bondS, polarSpeed=o.polarSpeed,
                    meanV2=o.meanV2, beads=o.beads, bonds=o.bonds, sums=list(o.sums))

    def autotune(self, dt, cache_path=None, use_cache=True):
        """The library's autotuner on a copy of this system's state: returns
        (status, the chosen NWTuning as a dict); status 2 = from the cache,
        1 = timed, 0 = not tuned."""
        t = abi.NWTuning()
        path = None if cache_path is None else cache_path.encode()
        rc = self.lib.nwSystemAutotune(self.h, dt, path, 1 if use_cache else 0, ctypes.byref(t))
        return rc, t.as_dict()

    def write_outputs(self, path, iteration, dt):
        self.lib.nwSystemWriteOutputs(self.h, path.encode(), iteration, dt)

    def close(self):
        if self.h:
            self.lib.nwSystemDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def wca_ext(drive=1.0, stencil=abi.NW_STENCIL_2D, srd_beads=1):
    e = abi.NWExtParams()
    e.pairPotential = abi.NW_PAIR_WCA
    e.bondedMode = abi.NW_BONDED_CHAIN
    e.stencil = stencil
    e.srdCollideFilaments = srd_beads
    e.srdThermostat = abi.NW_SRD_THERMO_NONE
    e.wcaEpsilon = 1.0
    e.wcaSigma = 1.0
    e.wcaCutoff = 2.0 ** (1.0 / 6.0)
    e.driveForce = drive
    return e


def apply_worm_physics(sys_, kappa=5.0, gamma=2.0, kT=0.1):
    """Force parameters shared by the worm workloads: harmonic bonds k = 57,
    r0 = 0.8; bending kappa; Langevin gamma, kT; no extensile pair force."""
    p = sys_.params
    p.bondSpringK = 57.0
    p.bondSpringL = 0.8
    p.bondBendingK = kappa
    p.activity = 0.0
    p.reverseProbability = 0.0
    p.gamma = gamma
    p.kbT = kT
    sys_.solvent.kbT = kT


def config_dry_4k(seed=1337, cells=64):
    """400 worms x 10 beads, 2D periodic L = 100, WCA + bonds + bending + drive,
    Langevin (FDT noise), dt = 1e-3; pow2 grid 64^2 of cell 1.5625 >= r_c."""
    c = NWSystemConfig()
    c.numFilaments, c.filamentSize = 400, 10
    c.gridSize = abi.UInt3(cells, cells, 1)
    c.boxSize = abi.Float3(100.0, 100.0, 100.0 / cells)
    c.langevinFdt = 1
    c.seed = seed
    c.ext = wca_ext(stencil=abi.NW_STENCIL_2D, srd_beads=0)
    return c


def config_srd_800k(seed=1337, md_cells=890, solvent_sort_every=0):
    """40,000 worms x 20 beads in 2D L = 1000 with 5,000,000 SRD particles
    (5 per unit cell, 1000^2 cells of size 1), collision every 10 steps,
    rotation 90 deg with random sign per cell, grid shift in [-0.5, 0.5),
    beads in the collision with mass 1; MD grid md_cells^2 (cell 1000/md_cells
    >= r_c = 2^(1/6)); the solvent kept in cell order on the device
    (re-sorted every solvent_sort_every collisions) when that is > 0."""
    assert 1000.0 / md_cells >= 2.0 ** (1.0 / 6.0)
    c = NWSystemConfig()
    c.numFilaments, c.filamentSize = 40_000, 20
    c.gridSize = abi.UInt3(md_cells, md_cells, 1)
    c.boxSize = abi.Float3(1000.0, 1000.0, 1000.0 / md_cells)
    c.srd = 1
    c.srdGridSize = abi.UInt3(1000, 1000, 1)
    c.srdCellSize = 1.0
    c.srdDensity = 5.0
    c.srdGap = 10
    c.srdAlpha = math.pi / 2
    c.srdShiftCentred = 1
    c.seed = seed
    c.ext = wca_ext(stencil=abi.NW_STENCIL_2D, srd_beads=1)
    c.solventSortEvery = solvent_sort_every
    return c

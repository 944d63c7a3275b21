"""ctypes bindings of libnwamd.so (include/nw_abi.h).

The C-ABI is the product boundary; this module only declares it for Python
callers (tests, bench, the PyTorch extension).  Loading fails loudly when the
library is missing: there is no Python fallback for any kernel.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnwamd.so")

c_uint = ctypes.c_uint
c_float = ctypes.c_float
c_int = ctypes.c_int
c_void_p = ctypes.c_void_p
c_size_t = ctypes.c_size_t
c_u64 = ctypes.c_uint64
c_bool = ctypes.c_bool

PERIODIC, WALL, WALL_NO_SLIP = 0, 1, 2
NW_PAIR_REF_DEM, NW_PAIR_WCA = 0, 1
NW_BONDED_REF, NW_BONDED_CHAIN = 0, 1
NW_STENCIL_REF27, NW_STENCIL_2D = 0, 1
NW_SRD_THERMO_ISOKINETIC, NW_SRD_THERMO_NONE = 0, 1
NW_ERRORS_EXIT, NW_ERRORS_RETURN = 0, 1


class Float3(ctypes.Structure):
    _fields_ = [("x", c_float), ("y", c_float), ("z", c_float)]


class UInt3(ctypes.Structure):
    _fields_ = [("x", c_uint), ("y", c_uint), ("z", c_uint)]


class SimParams(ctypes.Structure):
    """simulation/particles_kernel.cuh:26-60 (144 bytes)."""
    _fields_ = [
        ("gravity", Float3), ("particleRadius", c_float), ("particleMass", c_float),
        ("gridSize", UInt3), ("numCells", c_uint),
        ("origin", Float3), ("boxSize", Float3), ("cellSize", Float3),
        ("boundaryX", c_int), ("boundaryY", c_int), ("boundaryZ", c_int),
        ("numParticles", c_uint), ("numFilaments", c_uint), ("filamentSize", c_uint),
        ("bondSpringK", c_float), ("bondSpringL", c_float), ("bondBendingK", c_float),
        ("activity", c_float), ("reverseProbability", c_float), ("kbT", c_float), ("gamma", c_float),
        ("spring", c_float), ("damping", c_float), ("shear", c_float), ("attraction", c_float),
        ("boundaryDamping", c_float),
    ]


class SolventParams(ctypes.Structure):
    """simulation/particles_kernel.cuh:62-73 (60 bytes)."""
    _fields_ = [
        ("enabled", c_bool), ("gridSize", UInt3), ("numCells", c_uint), ("randOffset", Float3),
        ("cellSize", Float3), ("numParticles", c_uint), ("alpha", c_float), ("kbT", c_float),
        ("particleMass", c_float),
    ]


class NWExtParams(ctypes.Structure):
    _fields_ = [
        ("pairPotential", c_uint), ("bondedMode", c_uint), ("stencil", c_uint),
        ("srdCollideFilaments", c_uint), ("srdThermostat", c_uint),
        ("wcaEpsilon", c_float), ("wcaSigma", c_float), ("wcaCutoff", c_float), ("driveForce", c_float),
        ("collideKernel", c_uint), ("reserved", c_uint * 6),
    ]


NW_COLLIDE_AUTO, NW_COLLIDE_GLOBAL, NW_COLLIDE_LDS = 0, 1, 2


class NWTuning(ctypes.Structure):
    """include/nw_abi.h NWTuning: launch shapes of the fused kernels (0: built-in)."""
    _fields_ = [("mdTileItems", c_uint), ("mdMembers", c_uint), ("srdTileItems", c_uint), ("srdMembers", c_uint),
                ("pairBlock", c_uint), ("frontBlock", c_uint), ("pairStage", c_uint), ("reserved", c_uint * 9)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


assert ctypes.sizeof(SimParams) == 144
assert ctypes.sizeof(SolventParams) == 60
assert ctypes.sizeof(NWExtParams) == 64

NW_OBS_SUMS = 19


class NWReplicaParams(ctypes.Structure):
    _fields_ = [("params", SimParams), ("ext", NWExtParams), ("seed", ctypes.c_uint64), ("offset", ctypes.c_uint64),
                ("langevinGamma", ctypes.c_float), ("langevinKbT", ctypes.c_float), ("langevin", ctypes.c_int),
                ("reserved", ctypes.c_uint32 * 5)]


class NWReplicaSolvent(ctypes.Structure):
    _fields_ = [("solvent", SolventParams), ("numSolvent", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 3)]


class NWObservables(ctypes.Structure):
    _fields_ = [("nematicS", ctypes.c_double), ("polarP", ctypes.c_double), ("bondS", ctypes.c_double),
                ("polarSpeed", ctypes.c_double), ("meanV2", ctypes.c_double), ("beads", ctypes.c_double),
                ("bonds", ctypes.c_double), ("sums", ctypes.c_double * NW_OBS_SUMS)]

P = c_void_p
# name -> (restype, argtypes); the reference surface first, then additions.
PROTOTYPES = {
    "cudaInit": (None, [c_int, ctypes.POINTER(ctypes.c_char_p)]),
    "randomizeUniform": (None, [P, c_uint]),
    "randomizeGaussian": (None, [P, c_uint, c_float, c_float]),
    "allocateArray": (None, [ctypes.POINTER(c_void_p), c_size_t]),
    "freeArray": (None, [P]),
    "threadSync": (None, []),
    "copyArrayFromDevice": (None, [P, P, P, c_int]),
    "copyArrayToDevice": (None, [P, P, c_int, c_int]),
    "setParameters": (None, [ctypes.POINTER(SimParams)]),
    "setSolventParameters": (None, [ctypes.POINTER(SolventParams)]),
    "integrateFilaments": (None, [P, P, P, P, c_float, c_uint]),
    "integrateSolvent": (None, [P, P, c_float, c_uint, c_uint]),
    "calcHash": (None, [P, P, P, c_int, c_bool]),
    "reorderDataAndFindCellStart": (None, [P, P, P, P, P, P, P, P, P, P, c_uint, c_uint]),
    "collideFilaments": (None, [P, P, P, P, P, P, P, c_uint, c_uint]),
    "collideSolvent": (None, [P, P, P, P, c_uint, c_uint]),
    "filamentForces": (None, [P, P, P, c_float]),
    "sortParticles": (None, [P, P, c_uint]),
    "reverseFilaments": (None, [P, P, P, P, P, c_uint]),
    "langevinThermostat": (None, [P, P, P, c_float, c_float, c_uint]),
    "solventCellCentorOfMomentum": (None, [P, P, P, P, P, c_uint]),
    "solventIsokineticThermostat": (None, [P, P, P, P, P, c_uint]),
    "iDivUp": (c_uint, [c_uint, c_uint]),
    "computeGridSize": (None, [c_uint, c_uint, ctypes.POINTER(c_uint), ctypes.POINTER(c_uint)]),
    "integrateSystem": (None, [P, P, P, P, c_float, c_uint]),
    "collide": (None, [P, P, P, P, P, P, P, c_uint, c_uint]),
    "nwSetErrorMode": (None, [c_int]),
    "nwGetLastError": (c_int, []),
    "nwGetLastErrorString": (ctypes.c_char_p, []),
    "nwDeviceAvailable": (c_int, []),
    "nwRngSetSeed": (None, [c_u64, c_u64]),
    "nwRngGetState": (None, [ctypes.POINTER(c_u64), ctypes.POINTER(c_u64)]),
    "nwPhiloxBlock": (None, [c_u64, c_u64, ctypes.POINTER(ctypes.c_uint32)]),
    "nwSetExtParameters": (None, [ctypes.POINTER(NWExtParams)]),
    "nwGetExtParameters": (None, [ctypes.POINTER(NWExtParams)]),
    "nwGetStream": (c_void_p, []),
    "nwSetStream": (None, [c_void_p]),
    "nwSortParticlesBounded": (None, [P, P, c_uint, c_uint]),
    "nwRadixSortPairs": (None, [P, P, c_uint, c_uint]),
    "nwFusedMdStep": (None, [P, P, P, P, P, P, P, P, P, P, c_float, c_float, c_float, c_int, c_uint]),
    "nwFusedMdStepSwapForces": (None, [P, P, P, P, P, P, P, P, P, P, c_float, c_float, c_float, c_int, c_uint]),
    "nwFusedMdStepNoAlloc": (None, [P, P, P, P, P, P, P, P, P, P, c_float, c_float, c_float, c_int, c_uint]),
    "nwFilamentTangents": (None, [P, P, c_uint]),
    "nwFusedSrdStep": (None, [P, P, P, P, P, c_float, c_uint, c_uint, c_int]),
    "nwFusedPlanFits": (c_int, [c_uint, c_uint, c_uint]),
    "nwFusedPlanInfo": (c_int, [c_uint, c_uint, c_uint, P]),
    "nwSetTuning": (None, [ctypes.c_void_p]),
    "nwGetTuning": (None, [ctypes.c_void_p]),
    "nwAutotuneFused": (c_int, [P, P, P, P, P, c_uint, c_uint, c_float, c_float, c_float, c_int, c_float, c_int,
                                c_uint, ctypes.c_char_p, c_int, ctypes.c_void_p]),
    "nwObservables": (None, [ctypes.c_void_p, ctypes.c_void_p, c_uint, ctypes.POINTER(NWObservables)]),
    "nwBatchSetReplicas": (None, [ctypes.POINTER(NWReplicaParams), c_uint]),
    "nwBatchLayout": (None, [ctypes.c_void_p, ctypes.c_void_p]),
    "nwBatchMdStep": (None, [ctypes.c_void_p] * 9 + [c_float, c_uint, c_int]),
    "nwBatchRngState": (None, [c_uint, ctypes.POINTER(ctypes.c_uint64)]),
    "nwAdjointForward": (None, [ctypes.c_void_p] * 5 + [c_uint, c_float, c_float, c_float, c_int, c_uint, c_uint]),
    "nwAdjointBackward": (None, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "nwAdjointForwardSrd": (None, [ctypes.c_void_p] * 5 + [c_uint, c_uint, c_float, c_uint,
                                                         ctypes.POINTER(c_float), c_uint, c_uint]),
    "nwBatchSetSolvent": (None, [ctypes.POINTER(NWReplicaSolvent), c_uint]),
    "nwBatchSolventLayout": (None, [ctypes.c_void_p, ctypes.c_void_p]),
    "nwBatchSrdStep": (None, [ctypes.c_void_p] * 5 + [c_float, ctypes.POINTER(c_float), c_uint]),
    "nwBatchAdjointForward": (None, [ctypes.c_void_p] * 5 + [c_float, c_uint, ctypes.POINTER(c_float), c_uint,
                                                           c_uint]),
    "nwBatchAdjointBackward": (None, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "nwBatchPolarSpeedGrad": (None, [ctypes.c_void_p] * 4 + [ctypes.POINTER(ctypes.c_double)]),
    "nwBondOrderGrad": (ctypes.c_double, [ctypes.c_void_p, c_uint, ctypes.c_void_p]),
    "nwSortSolvent": (None, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, c_uint, c_uint]),
    "nwPolarSpeedGrad": (ctypes.c_double, [ctypes.c_void_p, ctypes.c_void_p, c_uint, ctypes.c_void_p,
                                           ctypes.c_void_p]),
    "nwSymEig3Max": (ctypes.c_double, [ctypes.POINTER(ctypes.c_double)]),
    "nwGetDeviceFlags": (c_uint, [c_int]),
    "nwTimerEnable": (None, [ctypes.c_char_p, c_int]),
    "nwTimerRead": (None, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(c_u64)]),
}

_lib = None


def load(path=None):
    """Loads libnwamd.so and declares every prototype.  Raises if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libnwamd.so not built at {p}: run __graft_entry__.build()")
    lib = ctypes.CDLL(p, mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class DeviceArray:
    """A device allocation made through allocateArray, with numpy transfer helpers."""

    def __init__(self, lib, nbytes):
        self.lib = lib
        self.nbytes = int(nbytes)
        self.ptr = c_void_p()
        if self.nbytes > 0:
            lib.allocateArray(ctypes.byref(self.ptr), c_size_t(self.nbytes))

    @classmethod
    def from_numpy(cls, lib, arr):
        a = np.ascontiguousarray(arr)
        d = cls(lib, a.nbytes)
        d.upload(a)
        return d

    def upload(self, arr, offset=0):
        a = np.ascontiguousarray(arr)
        if a.nbytes:
            self.lib.copyArrayToDevice(self.ptr, a.ctypes.data_as(c_void_p), int(offset), int(a.nbytes))

    def download(self, dtype, shape, offset_bytes=0):
        out = np.empty(shape, dtype=dtype)
        if out.nbytes:
            src = c_void_p((self.ptr.value or 0) + offset_bytes)
            self.lib.copyArrayFromDevice(out.ctypes.data_as(c_void_p), src, None, int(out.nbytes))
        return out

    @property
    def p(self):
        return self.ptr

    def free(self):
        if self.ptr.value:
            self.lib.freeArray(self.ptr)
            self.ptr = c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def timer_read(lib, name):
    ms = ctypes.c_double()
    n = c_u64()
    lib.nwTimerRead(name.encode(), ctypes.byref(ms), ctypes.byref(n))
    return ms.value, n.value

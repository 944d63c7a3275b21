"""ctypes binding of oracle/liborc.so -- TEST INFRASTRUCTURE ONLY.

The oracle restates the reference's kernels on the CPU (see nw_oracle.h).  It
is the checker for libnwamd.so and the CPU baseline; only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "nematic-worms_amd"))
from nwamd.abi import SimParams, SolventParams, NWExtParams  # noqa: E402  (struct layouts only)

P = ctypes.c_void_p
u32 = ctypes.c_uint32
f32 = ctypes.c_float
u64 = ctypes.c_uint64


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        sp, ss, se = ctypes.POINTER(SimParams), ctypes.POINTER(SolventParams), ctypes.POINTER(NWExtParams)
        protos = {
            "orc_calc_hash": [sp, ss, P, u32, ctypes.c_int, P, P],
            "orc_sort_pairs": [P, P, u32],
            "orc_reorder": [P, P, P, P, P, P, P, P, P, P, u32, u32],
            "orc_reverse_filaments": [sp, P, u32, P, P, P, P],
            "orc_philox4x32_10": [P, P, P],
            "orc_random_uniform": [u64, u64, u32, P],
            "orc_random_gaussian_f64": [u64, u64, u32, ctypes.c_double, ctypes.c_double, P],
            "orc_srd_shift": [u64, u64, ctypes.c_int, P, P],
            "orc_system_step_f32": [P, u32],
        }
        for sfx in ("f64", "f32"):
            protos.update({
                f"orc_integrate_filaments_{sfx}": [sp, P, P, P, P, f32, u32, P, P, P, P],
                f"orc_integrate_solvent_{sfx}": [sp, P, P, f32, u32, u32, P, P],
                f"orc_collide_filaments_{sfx}": [sp, se, P, P, P, P, P, P, u32, P, P],
                f"orc_filament_forces_{sfx}": [sp, se, P, P, u32, P, P],
                f"orc_langevin_{sfx}": [sp, P, P, P, u32, P],
                f"orc_srd_com_{sfx}": [P, P, P, u32, P],
                f"orc_srd_isokinetic_{sfx}": [ss, P, P, P, P, P, u32, u32, P, P],
                f"orc_collide_solvent_{sfx}": [sp, ss, se, P, P, P, P, u32, u32, P],
            })
        for name, args in protos.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int if name == "orc_system_step_f32" else None
        _lib = L
    return _lib


def ptr(a):
    return None if a is None else a.ctypes.data_as(P)


def _rt(prec):
    return np.float64 if prec == "f64" else np.float32


# ---- thin numpy wrappers ----------------------------------------------------
def calc_hash(p, s, pos, srd=False):
    n = pos.shape[0]
    h = np.empty(n, np.uint32)
    idx = np.empty(n, np.uint32)
    lib().orc_calc_hash(ctypes.byref(p), ctypes.byref(s) if s is not None else None,
                        ptr(np.ascontiguousarray(pos, np.float32)), n, int(srd), ptr(h), ptr(idx))
    return h, idx


def sort_pairs(keys, vals):
    k = np.ascontiguousarray(keys, np.uint32).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    lib().orc_sort_pairs(ptr(k), ptr(v), k.shape[0])
    return k, v


def reorder(hash_, index, pos, vel, tan, num_cells, cell_end_init=None):
    n = hash_.shape[0]
    cs = np.empty(num_cells, np.uint32)
    ce = (np.zeros(num_cells, np.uint32) if cell_end_init is None else cell_end_init.copy())
    sp = np.empty((n, 4), np.float32)
    sv = np.empty((n, 4), np.float32) if vel is not None else None
    st = np.empty((n, 4), np.float32) if tan is not None else None
    lib().orc_reorder(ptr(cs), ptr(ce), ptr(sp), ptr(sv), ptr(st), ptr(hash_), ptr(index), ptr(pos),
                      ptr(vel), ptr(tan), n, num_cells)
    return cs, ce, sp, sv, st


def integrate_filaments(p, pos, vel, f, fold, dt, prec="f64"):
    n = pos.shape[0]
    R = _rt(prec)
    outs = [np.empty((n, 4), R) for _ in range(4)]
    getattr(lib(), f"orc_integrate_filaments_{prec}")(
        ctypes.byref(p), ptr(pos), ptr(vel), ptr(f), ptr(fold), dt, n, *[ptr(o) for o in outs])
    return outs


def integrate_solvent(p, pos, vel, dt, start, ns, prec="f64"):
    R = _rt(prec)
    po = pos.astype(R)
    vo = vel.astype(R)
    getattr(lib(), f"orc_integrate_solvent_{prec}")(ctypes.byref(p), ptr(pos), ptr(vel), dt, start, ns,
                                                    ptr(po), ptr(vo))
    return po, vo


def collide_filaments(p, e, spos, svel, stan, sidx, cs, ce, force, prec="f64"):
    n = spos.shape[0]
    out = np.empty((n, 4), _rt(prec))
    getattr(lib(), f"orc_collide_filaments_{prec}")(
        ctypes.byref(p), ctypes.byref(e), ptr(spos), ptr(svel), ptr(stan), ptr(sidx), ptr(cs), ptr(ce), n,
        ptr(force), ptr(out))
    return out


def filament_forces(p, e, force, pos, tangent_init, W, prec="f64"):
    R = _rt(prec)
    out = np.empty(force.shape, R)
    tan = tangent_init.astype(R)
    getattr(lib(), f"orc_filament_forces_{prec}")(ctypes.byref(p), ctypes.byref(e), ptr(force), ptr(pos), W,
                                                  ptr(out), ptr(tan))
    return out, tan


def langevin(p, force, vel, noise, prec="f64"):
    n = force.shape[0]
    R = _rt(prec)
    out = np.empty((n, 4), R)
    nz = np.ascontiguousarray(noise, R)
    getattr(lib(), f"orc_langevin_{prec}")(ctypes.byref(p), ptr(force), ptr(vel), ptr(nz), n, ptr(out))
    return out


def srd_com(svel, cs, ce, num_cells, prec="f64"):
    out = np.empty((num_cells, 4), _rt(prec))
    getattr(lib(), f"orc_srd_com_{prec}")(ptr(svel), ptr(cs), ptr(ce), num_cells, ptr(out))
    return out


def srd_isokinetic(s, vel, svel, cs, ce, index, num_cells, prec="f64"):
    R = _rt(prec)
    n = vel.shape[0]
    vo = np.empty((n, 4), R)
    so = np.empty((n, 4), R)
    getattr(lib(), f"orc_srd_isokinetic_{prec}")(ctypes.byref(s), ptr(vel), ptr(svel), ptr(cs), ptr(ce),
                                                 ptr(index), num_cells, n, ptr(vo), ptr(so))
    return vo, so


def collide_solvent(p, s, e, pos, vel, com, uniform, num_cells, prec="f64"):
    n = pos.shape[0]
    out = np.empty((n, 4), _rt(prec))
    getattr(lib(), f"orc_collide_solvent_{prec}")(
        ctypes.byref(p), ctypes.byref(s), ctypes.byref(e) if e is not None else None, ptr(pos), ptr(vel),
        ptr(np.ascontiguousarray(com, np.float32)), ptr(np.ascontiguousarray(uniform, np.float32)), num_cells, n,
        ptr(out))
    return out


def reverse_filaments(p, uniform, W, force, fold, vel, pos):
    arrs = [a.copy() for a in (force, fold, vel, pos)]
    lib().orc_reverse_filaments(ctypes.byref(p), ptr(np.ascontiguousarray(uniform, np.float32)), W,
                                *[ptr(a) for a in arrs])
    return arrs


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.empty(4, np.uint32)
    lib().orc_philox4x32_10(ptr(c), ptr(k), ptr(out))
    return out


def random_uniform(seed, offset, n):
    out = np.empty(n, np.float32)
    lib().orc_random_uniform(seed, offset, n, ptr(out))
    return out


def random_gaussian(seed, offset, n, mean=0.0, std=1.0):
    out = np.empty(n, np.float64)
    lib().orc_random_gaussian_f64(seed, offset, n, mean, std, ptr(out))
    return out


def srd_shift(seed, k, centred, cell):
    out = np.empty(3, np.float32)
    c = np.ascontiguousarray(cell, np.float32)
    lib().orc_srd_shift(seed, k, int(centred), ptr(c), ptr(out))
    return out


# ---- order parameters ------------------------------------------------------
def _mi_axis32(b, a, L, periodic):
    """mi_axis in FP32 (nw_oracle_impl.h), vectorised."""
    b, a = np.asarray(b, np.float32), np.asarray(a, np.float32)
    d = (b - a).astype(np.float32)
    if periodic:
        h = np.float32(L) / np.float32(2.0)
        up = ((b - h).astype(np.float32) - (a + h).astype(np.float32)).astype(np.float32)
        dn = ((b + h).astype(np.float32) - (a - h).astype(np.float32)).astype(np.float32)
        d = np.where(d > h, up, np.where(d < -h, dn, d))
    return d


def _min_image32(p, a, b):
    per = (p.boundaryX == 0, p.boundaryY == 0, p.boundaryZ == 0)
    L = (p.boxSize.x, p.boxSize.y, p.boxSize.z)
    return np.stack([_mi_axis32(b[:, k], a[:, k], L[k], per[k]) for k in range(3)], axis=1)


OBS_SUMS = 19


def observables(p, pos, vel, W):
    """The sums of nwObservables (same tangent definitions: the reference's
    analysis, analysis/afa/utils.py:35-49 and analyze_correlations.py:13-38,
    with the minimum image) in FP64 from the FP32 minimum-image differences,
    and S, P from numpy's symmetric eigensolver."""
    n = p.filamentSize
    r = np.asarray(pos, np.float32)[:W * n, :3].reshape(W, n, 3)
    v = np.asarray(vel, np.float32)[:W * n, :3].reshape(W, n, 3).astype(np.float64)
    out = np.zeros(OBS_SUMS)
    res = dict(sums=out, nematicS=0.0, polarP=0.0, bondS=0.0, polarSpeed=0.0, meanV2=0.0)
    if n < 2 or W == 0:
        return res
    lo = np.concatenate([r[:, :1], r[:, :-1]], axis=1)    # r[p-1], r[0] at the head
    hi = np.concatenate([r[:, 1:], r[:, -1:]], axis=1)    # r[p+1], r[n-1] at the tail
    d = _min_image32(p, lo.reshape(-1, 3), hi.reshape(-1, 3)).astype(np.float64)
    t = d / np.sqrt((d * d).sum(1))[:, None]
    e = _min_image32(p, r[:, :-1].reshape(-1, 3), r[:, 1:].reshape(-1, 3)).astype(np.float64)
    b = e / np.sqrt((e * e).sum(1))[:, None]
    iu = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    out[0:6] = [(t[:, i] * t[:, j]).sum() for i, j in iu]
    out[6:9] = t.sum(0)
    out[9] = len(t)
    out[10:16] = [(b[:, i] * b[:, j]).sum() for i, j in iu]
    out[16] = len(b)
    vv = v.reshape(-1, 3)
    out[17] = (vv * t).sum()
    out[18] = (vv * vv).sum()

    def smax(s6, cnt):
        Q = np.empty((3, 3))
        for k, (i, j) in enumerate(iu):
            Q[i, j] = Q[j, i] = 1.5 * s6[k] / cnt
        Q -= 0.5 * np.eye(3)
        return float(np.linalg.eigvalsh(Q)[-1])

    res.update(nematicS=smax(out[0:6], out[9]), polarP=float(np.linalg.norm(out[6:9] / out[9])),
               bondS=smax(out[10:16], out[16]), polarSpeed=out[17] / out[9], meanV2=out[18] / out[9])
    return res


# ---- whole-step CPU system (the CPU baseline) --------------------------------
class OrcSystemC(ctypes.Structure):
    _fields_ = [
        ("p", SimParams), ("s", SolventParams), ("e", NWExtParams),
        ("numSolvent", u32), ("srdGap", u32), ("srdEnabled", u32), ("langevinFdt", u32), ("srdShiftCentred", u32),
        ("dt", f32), ("seed", u64), ("offset", u64), ("step", u64),
        ("pos", P), ("vel", P), ("force", P), ("forceOld", P), ("tangent", P),
        ("sPos", P), ("sVel", P), ("sTan", P), ("com", P), ("rnd", P),
        ("hash", P), ("index", P), ("cellStart", P), ("cellEnd", P),
        ("numCellsAlloc", u32),
    ]


class CpuSystem:
    """FP32 OpenMP restatement of ParticleSystem::update (simulation/particleSystem.cpp:249-413)."""

    def __init__(self, p, s, e, pos, vel, tangent, force, force_old, *, num_solvent, srd, srd_gap, dt,
                 seed, offset, step=0, langevin_fdt=False, shift_centred=False):
        N = p.numParticles
        NT = N + (num_solvent if srd else 0)
        Cs = s.numCells if srd else 0
        self.arrays = dict(
            pos=np.ascontiguousarray(pos, np.float32).copy(), vel=np.ascontiguousarray(vel, np.float32).copy(),
            force=np.ascontiguousarray(force, np.float32).copy(),
            forceOld=np.ascontiguousarray(force_old, np.float32).copy(),
            tangent=np.ascontiguousarray(tangent, np.float32).copy(),
            sPos=np.empty((NT, 4), np.float32), sVel=np.empty((NT, 4), np.float32), sTan=np.empty((N, 4), np.float32),
            com=np.empty((max(Cs, 1), 4), np.float32),
            rnd=np.empty(max(4 * N, p.numFilaments, Cs, 4), np.float32),
            hash=np.empty(NT, np.uint32), index=np.empty(NT, np.uint32))
        ncells = max(p.numCells, Cs)
        self.arrays["cellStart"] = np.empty(ncells, np.uint32)
        self.arrays["cellEnd"] = np.zeros(ncells, np.uint32)
        c = OrcSystemC()
        c.p, c.s, c.e = p, s, e
        c.numSolvent, c.srdGap, c.srdEnabled = num_solvent, srd_gap, int(srd)
        c.langevinFdt, c.srdShiftCentred = int(langevin_fdt), int(shift_centred)
        c.dt, c.seed, c.offset, c.step = dt, seed, offset, step
        for k, a in self.arrays.items():
            setattr(c, k, a.ctypes.data)
        c.numCellsAlloc = ncells
        self.c = c

    def step(self, k=1):
        return lib().orc_system_step_f32(ctypes.byref(self.c), k)

    def get(self, name):
        return self.arrays[name]

"""The C-ABI libraries load (no GPU needed) and export every function that
include/*.h declares; the parameter structs have the reference's layout."""
import ctypes
import os
import re
import subprocess

import pytest

from nwamd import abi, replicas, system

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def declared_functions(header):
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set()
    for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}]*)\)\s*;", text):
        name = m.group(1)
        if name not in ("if", "while", "for", "sizeof", "return"):
            names.add(name)
    return names


@pytest.mark.parametrize("header,lib", [("nw_abi.h", "libnwamd.so"), ("nw_system.h", "libnwhost.so")])
def test_library_exports_every_declared_symbol(header, lib):
    path = os.path.join(ROOT, "nematic-worms_amd", "nwamd", lib)
    if lib == "libnwhost.so":
        abi.load()
    L = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    names = declared_functions(header)
    assert len(names) > (25 if header == "nw_abi.h" else 10)
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, f"{lib} lacks {missing}"
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert names <= exported


def test_reference_surface_is_complete():
    """The 22 entry points of simulation/particleSystem.cuh:12-112 (plus the
    helpers inside the extern "C" block of particleSystem_cuda.cu:105-115)."""
    ref = {"cudaInit", "randomizeUniform", "randomizeGaussian", "allocateArray", "freeArray", "threadSync",
           "copyArrayFromDevice", "copyArrayToDevice", "setParameters", "setSolventParameters",
           "integrateFilaments", "integrateSolvent", "calcHash", "reorderDataAndFindCellStart",
           "collideFilaments", "collideSolvent", "filamentForces", "sortParticles", "reverseFilaments",
           "langevinThermostat", "solventCellCentorOfMomentum", "solventIsokineticThermostat",
           "iDivUp", "computeGridSize", "integrateSystem", "collide"}
    assert ref <= declared_functions("nw_abi.h")
    L = abi.load()
    assert set(abi.PROTOTYPES) >= ref
    assert L.iDivUp(10, 4) == 3 and L.iDivUp(8, 4) == 2
    nb, nt = ctypes.c_uint(), ctypes.c_uint()
    L.computeGridSize(1000, 256, ctypes.byref(nb), ctypes.byref(nt))
    assert (nb.value, nt.value) == (4, 256)
    L.computeGridSize(100, 256, ctypes.byref(nb), ctypes.byref(nt))
    assert (nb.value, nt.value) == (1, 100)


def test_struct_layout_matches_c(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(r"""
#include <stdio.h>
#include <stddef.h>
#include "nw_system.h"
#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m));
int main(void) {
  printf("SimParams %zu\nSolventParams %zu\nNWExtParams %zu\nNWSystemConfig %zu\n", sizeof(SimParams),
         sizeof(SolventParams), sizeof(NWExtParams), sizeof(NWSystemConfig));
  F(SimParams, gridSize) F(SimParams, numCells) F(SimParams, origin) F(SimParams, boundaryX)
  F(SimParams, numParticles) F(SimParams, bondSpringK) F(SimParams, activity) F(SimParams, spring)
  F(SimParams, boundaryDamping) F(SolventParams, gridSize) F(SolventParams, randOffset) F(SolventParams, alpha)
  F(SolventParams, particleMass) F(NWSystemConfig, seed) F(NWSystemConfig, ext)
  printf("NWObservables %zu\nNWReplicaParams %zu\nNWReplicaSpec %zu\nNWReplicaPhysics %zu\n", sizeof(NWObservables),
         sizeof(NWReplicaParams), sizeof(NWReplicaSpec), sizeof(NWReplicaPhysics));
  F(NWObservables, sums) F(NWReplicaParams, ext) F(NWReplicaParams, seed) F(NWReplicaParams, langevin)
  F(NWReplicaSpec, seed) F(NWReplicaPhysics, langevinFdt)
  return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", f"-I{INC}", str(src), "-o", str(exe)], check=True)
    out = dict(ln.rsplit(" ", 1) for ln in subprocess.run([str(exe)], capture_output=True, text=True,
                                                           check=True).stdout.strip().splitlines())
    assert out["SimParams"] == "144" and out["SolventParams"] == "60"
    assert int(out["NWExtParams"]) == ctypes.sizeof(abi.NWExtParams)
    assert int(out["NWSystemConfig"]) == ctypes.sizeof(system.NWSystemConfig)
    classes = {"SimParams": abi.SimParams, "SolventParams": abi.SolventParams, "NWSystemConfig": system.NWSystemConfig,
               "NWObservables": abi.NWObservables, "NWReplicaParams": abi.NWReplicaParams,
               "NWReplicaSpec": replicas.NWReplicaSpec, "NWReplicaPhysics": replicas.NWReplicaPhysics}
    for t in ("NWObservables", "NWReplicaParams", "NWReplicaSpec", "NWReplicaPhysics"):
        assert int(out[t]) == ctypes.sizeof(classes[t]), t
    for key, val in out.items():
        if "." in key:
            t, m = key.split(".")
            cls = classes[t]
            assert getattr(cls, m).offset == int(val), key


def test_no_device_errors_are_reported_not_fatal():
    L = abi.load()
    L.nwSetErrorMode(abi.NW_ERRORS_RETURN)
    if L.nwDeviceAvailable():
        pytest.skip("a device is present")
    ptr = ctypes.c_void_p()
    L.allocateArray(ctypes.byref(ptr), 64)
    assert L.nwGetLastError() != 0
    assert b"device" in L.nwGetLastErrorString()


def test_fused_plan_limits():
    """nwFusedPlanFits does no device work: the benchmark's systems fit, a
    grid beyond the level-1 histogram and worms over 64 beads do not."""
    L = abi.load()
    assert L.nwFusedPlanFits(800000, 890 * 890, 20) == 1
    assert L.nwFusedPlanFits(5800000, 1000 * 1000, 0) == 1
    assert L.nwFusedPlanFits(0, 16, 10) == 1
    assert L.nwFusedPlanFits(1000, 16, 65) == 0
    assert L.nwFusedPlanFits(1000, 0, 10) == 0
    assert L.nwFusedPlanFits(1000, 32768 * 2048 + 1, 10) == 0
    assert L.nwFusedPlanFits(21_760_000, 45_000_000, 20) == 1  # 256 replicas: a tile pass of its own
    assert L.nwFusedPlanFits(2 ** 30, 2 ** 24, 0) == 0

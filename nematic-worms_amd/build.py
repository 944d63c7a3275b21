"""Builds the native parts in-tree with hipcc for gfx950.

Outputs (next to the Python bindings in nwamd/):
  libnwamd.so   the C-ABI (include/nw_abi.h) and its HIP kernels
  libnwhost.so  the C++ host driver above the C-ABI (ParticleSystem, xyzv I/O)
  nw            the command-line driver (reference flags, simulation/main.cpp)

Run: python nematic-worms_amd/build.py  (or __graft_entry__.build()).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "nwamd")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

# Device code: no FMA contraction unless a kernel opts in, IEEE-correct fp32
# division/sqrt (bit-exact cell hashing), packed-fp32 VALU forms disabled.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]

LIB_SOURCES = ["nw_kernels.hip", "nw_sort.hip", "nw_fused.hip", "nw_bucket.hip", "nw_observe.hip", "nw_adjoint.hip",
               "nw_abi.hip", "nw_tune.hip"]
HOST_SOURCES = ["host/particle_system.cpp", "host/worm_init.cpp", "host/replica_set.cpp", "host/system_capi.cpp"]
CLI_SOURCES = ["host/main.cpp"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _compile_objects(sources, objdir, extra):
    objs = []
    for src in sources:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        _run([HIPCC] + HIP_FLAGS + extra + ["-c", path, "-o", obj])
        objs.append(obj)
    return objs


def build(verbose=True):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(OUT, exist_ok=True)
    lib = os.path.join(OUT, "libnwamd.so")
    objs = _compile_objects(LIB_SOURCES, objdir, [])
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", lib] + objs)

    # the host driver is plain C++ above the C-ABI: no HIP in it
    host_flags = [CXX, "-std=c++17", "-O2", "-fPIC", "-Wall", f"-I{INCLUDE}", "-I" + os.path.join(CSRC, "host")]
    hobjs = []
    for src in HOST_SOURCES:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        _run(host_flags + ["-c", os.path.join(CSRC, src), "-o", obj])
        hobjs.append(obj)
    hostlib = os.path.join(OUT, "libnwhost.so")
    _run([CXX, "-shared", "-o", hostlib] + hobjs + [f"-L{OUT}", "-lnwamd", "-Wl,-rpath,$ORIGIN"])
    cobj = os.path.join(objdir, "main.cpp.o")
    _run(host_flags + ["-c", os.path.join(CSRC, CLI_SOURCES[0]), "-o", cobj])
    _run([CXX, "-o", os.path.join(OUT, "nw"), cobj, f"-L{OUT}", "-lnwhost", "-lnwamd", "-Wl,-rpath,$ORIGIN"])
    return lib


if __name__ == "__main__":
    build()
    sys.exit(0)

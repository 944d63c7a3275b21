"""A host compiled against the reference header (simulation/particleSystem.cuh:18,
`void allocateArray(void **devPtr, int size)`; copy sizes and offsets int as
well, :23-24) links against libnwamd.so and round-trips data through the
device.  The C caller is built with gcc at test time against prototypes
written out as the reference declares them."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "nematic-worms_amd", "nwamd")

SRC = r"""
#include <stdio.h>
#include <stdlib.h>
/* the reference header's prototypes (int sizes), not include/nw_abi.h */
void cudaInit(int argc, char **argv);
void allocateArray(void **devPtr, int size);
void freeArray(void *devPtr);
void threadSync(void);
void copyArrayFromDevice(void *host, const void *device, void *vbo, int size);
void copyArrayToDevice(void *device, const void *host, int offset, int size);

int main(int argc, char **argv) {
    cudaInit(argc, argv);
    /* sizes formed by int arithmetic, as particleSystem.cpp forms them */
    int n = argc > 1 ? atoi(argv[1]) : 12345;
    int bytes = n * 4 * (int)sizeof(float);
    float *h = (float *)malloc((size_t)bytes), *back = (float *)malloc((size_t)bytes);
    for (int i = 0; i < 4 * n; ++i) h[i] = (float)i * 0.5f - 3.0f;
    void *d = NULL;
    allocateArray(&d, bytes);
    if (!d) { printf("allocation failed\n"); return 2; }
    copyArrayToDevice(d, h, 0, bytes / 2);
    copyArrayToDevice(d, (const char *)h + bytes / 2, bytes / 2, bytes - bytes / 2);
    threadSync();
    copyArrayFromDevice(back, d, NULL, bytes);
    for (int i = 0; i < 4 * n; ++i)
        if (back[i] != h[i]) { printf("mismatch at %d\n", i); return 3; }
    freeArray(d);
    printf("ok %d\n", bytes);
    return 0;
}
"""


@pytest.mark.parametrize("n", [1, 12345, 1 << 20])
def test_reference_int_prototypes_round_trip(tmp_path, n):
    src = tmp_path / "int_caller.c"
    src.write_text(SRC)
    exe = tmp_path / "int_caller"
    subprocess.run(["gcc", "-O2", str(src), "-o", str(exe), f"-L{LIBDIR}", "-lnwamd", f"-Wl,-rpath,{LIBDIR}"],
                   check=True)
    r = subprocess.run([str(exe), str(n)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == f"ok {n * 16}"

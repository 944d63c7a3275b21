"""Writes tests/golden/philox_kat.json: Philox4x32-10 known-answer vectors
computed with PyTorch's host implementation (ATen/core/PhiloxRNGEngine.h,
an independent implementation of Salmon et al., SC'11).  Counter words
(c0, c1, c2, c3) map to torch's (offset = c0 | c1 << 32, subsequence =
c2 | c3 << 32); key = (lo32(seed), hi32(seed)).  Run: python tests/golden/gen_philox_kat.py
"""
import json
import os
import subprocess
import tempfile

import torch.utils.cpp_extension as ce

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = [
    # (counter[4], key[2])
    ([0, 0, 0, 0], [0, 0]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF, 0xFFFFFFFF]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]),
    ([1, 0, 0, 0], [1337, 0]),
    ([12345, 0, 0, 0], [1337, 0]),
    ([7, 3, 0, 0], [0xDEADBEEF, 0x01234567]),
    ([0xFFFFFFFF, 0, 0, 0], [42, 42]),
]

SRC = r"""
#include <ATen/core/PhiloxRNGEngine.h>
#include <cstdio>
#include <cstdint>
int main() {
  unsigned long long c0, c1, c2, c3, k0, k1;
  while (scanf("%llu %llu %llu %llu %llu %llu", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
    uint64_t seed = k0 | (k1 << 32), sub = c2 | (c3 << 32), off = c0 | (c1 << 32);
    at::Philox4_32 e(seed, sub, off);
    uint32_t a = e(), b = e(), c = e(), d = e();
    printf("%u %u %u %u\n", a, b, c, d);
  }
  return 0;
}
"""


def main():
    inc = ce.include_paths()
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "kat.cpp")
        exe = os.path.join(td, "kat")
        open(src, "w").write(SRC)
        subprocess.run(["g++", "-std=c++17", "-O1", src, "-o", exe] + [f"-I{p}" for p in inc], check=True)
        inp = "".join(" ".join(str(x) for x in c + k) + "\n" for c, k in CASES)
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    vectors = []
    for (c, k), line in zip(CASES, out):
        vectors.append(dict(counter=c, key=k, out=[int(x) for x in line.split()]))
    with open(os.path.join(HERE, "philox_kat.json"), "w") as f:
        json.dump(dict(source="torch ATen/core/PhiloxRNGEngine.h (host)", vectors=vectors), f, indent=1)
    for v in vectors:
        print([hex(x) for x in v["out"]])


if __name__ == "__main__":
    main()

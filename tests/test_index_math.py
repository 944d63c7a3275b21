"""The integer division the kernels use for bead -> worm and cell -> row /
layer (div_exact in nematic-worms_amd/csrc/nw_device.hpp): q = mulhi(x, m)
with m = ceil(2^32 / d), minus one when q d > x.  The claim is that this
equals floor(x / d) for every d >= 2 and x + d < 2^32; checked here against
Python's integer division, exhaustively for small x and at the boundaries
(multiples of d and their neighbours) up to 2^31."""
import numpy as np
import pytest


def magic(d):
    return 0 if d < 2 else ((1 << 32) + d - 1) // d


def div_exact(x, d):
    m = np.uint64(magic(d))
    x = x.astype(np.uint64)
    if m == 0:
        return x
    q = (x * m) >> np.uint64(32)
    return q - ((q * np.uint64(d)) > x).astype(np.uint64)


DIVISORS = [1, 2, 3, 5, 7, 10, 16, 20, 64, 89, 97, 120, 890, 1000, 4096, 7919, 65537,
            792100, 1 << 20, (1 << 20) + 7, 123456789, (1 << 30) + 3]


@pytest.mark.parametrize("d", DIVISORS)
def test_div_exact_small_x(d):
    x = np.arange(1 << 18, dtype=np.uint64)
    np.testing.assert_array_equal(div_exact(x, d), x // np.uint64(d))


@pytest.mark.parametrize("d", DIVISORS)
def test_div_exact_boundaries_to_2_31(d):
    top = (1 << 31) - 1
    k = np.unique(np.concatenate([np.arange(0, 4096), np.linspace(0, top // d, 20000).astype(np.int64)]))
    base = k.astype(np.uint64) * np.uint64(d)
    x = np.concatenate([base, base + 1, base + np.uint64(d - 1), base - 1])
    x = x[(x <= top) & (x + np.uint64(d) < np.uint64(1 << 32))]
    np.testing.assert_array_equal(div_exact(x, d), x // np.uint64(d))


def test_mulhi_alone_is_not_exact_for_the_benchmark_layer():
    """Why the correction: for the 890 x 890 layer the bare mulhi overshoots
    for cell ids far below the grid size."""
    d = 890 * 890
    x = np.arange(d * 3, dtype=np.uint64)
    q = (x * np.uint64(magic(d))) >> np.uint64(32)
    assert np.any(q != x // np.uint64(d))
    np.testing.assert_array_equal(div_exact(x, d), x // np.uint64(d))

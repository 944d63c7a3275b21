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


# ---------------------------------------------------------------------------
# floor_div_rc (nw_device.hpp): floor(d / c) of the correctly rounded quotient
# from t = d * rc, rc the hardware reciprocal of c (within 1 ulp).  The fast
# path is taken when t is farther than fma(|t|, 2^-19, 2^-100) from the
# nearest integer; the claim is that floor(t) then equals floor(RN(d / c)).
# Checked with rc = RN(1/c) and its float neighbours (the 1-ulp reciprocal),
# on random operands and on operands at and next to multiples of c (cell faces).
# ---------------------------------------------------------------------------
def _fast_path_floor_ok(d, c):
    d = d.astype(np.float32)
    c = np.float32(c)
    q = np.floor(d / c)  # float32 division: correctly rounded
    r0 = np.float32(1.0) / c
    for rc in (np.nextafter(r0, np.float32(0)), r0, np.nextafter(r0, np.float32(np.inf))):
        t = d * rc  # float32 product: correctly rounded
        dist = np.abs(t - np.rint(t))
        bound = (np.abs(t).astype(np.float64) * 2.0 ** -19 + 2.0 ** -100).astype(np.float32)
        fast = dist > bound
        assert np.array_equal(np.floor(t)[fast], q[fast]), (c, rc)
        # the fast path is the common one away from the faces
        yield fast.mean()


@pytest.mark.parametrize("c", [1.0, 1.1235955, 1.1236, 0.7, 1.5625, 3.0, 1e-3, 1234.5])
def test_floor_div_fast_path_random(c):
    rng = np.random.default_rng(7)
    d = np.concatenate([rng.uniform(-2000 * c, 2000 * c, 200_000), rng.uniform(-4 * c, 4 * c, 50_000)])
    fracs = list(_fast_path_floor_ok(d, c))
    assert min(fracs) > 0.99


@pytest.mark.parametrize("c", [1.1235955, 1.0, 0.7, 1.5625])
def test_floor_div_fast_path_at_cell_faces(c):
    k = np.arange(-1000, 1001, dtype=np.float32)
    g = (k * np.float32(c)).astype(np.float32)
    d = [g]
    for _ in range(3):
        d.append(np.nextafter(d[-1], np.float32(np.inf)))
    e = [g]
    for _ in range(3):
        e.append(np.nextafter(e[-1], np.float32(-np.inf)))
    list(_fast_path_floor_ok(np.concatenate(d + e[1:]), c))

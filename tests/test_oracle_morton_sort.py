"""Pins the oracle's Morton keys (barnes_hut.py:27-46) and stable order (:247)."""
import numpy as np
import pytest


def test_part1by2_kats(orc):
    L = orc.lib()
    assert L.orc_part1by2(0) == 0
    assert L.orc_part1by2(0x1FFFFF) == 0x1249249249249249
    assert int(orc.part1by2(np.array([0x1FFFFF]))[0]) == 0x1249249249249249
    for b in range(21):
        assert L.orc_part1by2(1 << b) == 1 << (3 * b)
    # bits above 21 are dropped
    assert L.orc_part1by2(1 << 21) == 0


def test_part1by2_matches_bitloop(orc):
    rng = np.random.default_rng(0)
    xs = rng.integers(0, 1 << 21, 2000, dtype=np.uint64)
    got = orc.part1by2(xs)
    for x, g in zip(xs[:300], got[:300]):
        assert int(g) == orc.part1by2_bitloop(int(x))
        assert orc.lib().orc_part1by2(int(x)) == int(g)
    # bijective on 21 bits: distinct inputs give distinct spreads
    assert len(set(got.tolist())) == len(set(xs.tolist()))


def test_unit_cube_corner_kat(orc):
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]])
    for keys in (orc.morton_codes(pos), orc.morton_c(pos)):
        assert int(keys[0]) == 0
        # scale = 2097151 / 1.0000001 quantises the max corner to 2097150
        assert int(keys[1]) == 0x7FFFFFFFFFFFFFF8


def test_axis_kats(orc):
    # a point on each axis at the far end of a unit span
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)
    k = orc.morton_c(pos)
    q = 2097150
    assert int(k[1]) == orc.part1by2_bitloop(q)
    assert int(k[2]) == orc.part1by2_bitloop(q) << 1
    assert int(k[3]) == orc.part1by2_bitloop(q) << 2


@pytest.mark.parametrize("bits", [21, 10, 1])
def test_c_matches_numpy(orc, bits):
    rng = np.random.default_rng(bits)
    pos = rng.normal(size=(5000, 3)) * np.array([3.0, 1.0, 0.2]) + 7.5
    pos[:50] = pos[50:100]  # duplicate positions
    a = orc.morton_codes(pos, bits)
    b = orc.morton_c(pos, bits)
    assert a.dtype == np.uint64 and np.array_equal(a, b)
    assert int(a.max()) < (1 << (3 * bits))


def test_degenerate_span(orc):
    pos = np.full((7, 3), 2.5)
    assert np.array_equal(orc.morton_c(pos), np.zeros(7, np.uint64))
    assert np.array_equal(orc.morton_codes(pos), np.zeros(7, np.uint64))


def test_order_preserving_along_axis(orc):
    x = np.sort(np.random.default_rng(3).uniform(-5, 5, 1000))
    pos = np.column_stack((x, np.zeros_like(x), np.zeros_like(x)))
    k = orc.morton_c(pos).astype(np.float64)
    assert np.all(np.diff(k) >= 0)


def test_stable_argsort_matches_numpy(orc):
    rng = np.random.default_rng(1)
    keys = rng.integers(0, 50, 20000).astype(np.uint64) << np.uint64(40)
    keys |= rng.integers(0, 3, 20000).astype(np.uint64)
    order = orc.stable_argsort_c(keys)
    assert np.array_equal(order, np.argsort(keys, kind="stable").astype(np.int32))
    big = rng.integers(0, 1 << 63, 5000, dtype=np.uint64)
    assert np.array_equal(orc.stable_argsort_c(big), np.argsort(big, kind="stable"))

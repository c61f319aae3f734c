"""Pins the oracle's Karras build (barnes_hut.py:50-121), aggregation
(:124-177) and walk (:180-217) against independent checks."""
import numpy as np
import pytest


def test_clz_identities(orc):
    rng = np.random.default_rng(0)
    vals = [1, 2, 3, (1 << 62), (1 << 62) | 1, (1 << 63) - 1] + [
        int(v) for v in rng.integers(1, 1 << 63, 500, dtype=np.int64)]
    for v in vals:
        # counted from bit 62: 62 - floor(log2 v)
        assert orc.clz64_ref(v) == 63 - v.bit_length()
    for v in [1, 2, 3, (1 << 30), (1 << 31) - 1] + list(rng.integers(1, 1 << 31, 500)):
        v = int(v)
        assert orc.clz_i32_ref(v) == 31 - v.bit_length()


def _keys(n, seed, dup=False, bits=21):
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(n, 3))
    if dup:
        pos[n // 2:] = pos[: n - n // 2]
    keys = orc_morton(pos, bits)
    return pos, keys


def orc_morton(pos, bits):
    from oracle import np_oracle
    return np_oracle.morton_c(pos, bits)


@pytest.mark.parametrize("n,seed,dup,bits", [
    (2, 0, False, 21), (3, 1, False, 21), (8, 2, True, 21), (64, 3, False, 21),
    (257, 4, True, 21), (1000, 5, False, 10), (513, 6, True, 10), (300, 7, False, 1)])
def test_build_matches_brute_force(orc, n, seed, dup, bits):
    pos, keys = _keys(n, seed, dup, bits)
    order = orc.stable_argsort_c(keys)
    code = keys[order].astype(np.int64)
    left, right, parent = orc.build_radix(code)
    bl, br, bp = orc.brute_force_radix_tree(code)
    assert np.array_equal(left, bl)
    assert np.array_equal(right, br)
    assert np.array_equal(parent, bp)


def test_build_all_equal_keys(orc):
    code = np.zeros(37, dtype=np.int64)
    left, right, parent = orc.build_radix(code)
    bl, br, bp = orc.brute_force_radix_tree(code)
    assert np.array_equal(left, bl) and np.array_equal(right, br) and np.array_equal(parent, bp)


def test_tree_invariants(orc):
    pos, keys = _keys(2000, 11, dup=True)
    order = orc.stable_argsort_c(keys)
    code = keys[order].astype(np.int64)
    n = len(code)
    left, right, parent = orc.build_radix(code)
    kids = np.concatenate((left, right))
    # every node except the root appears exactly once as a child
    assert np.array_equal(np.sort(kids), np.delete(np.arange(2 * n - 1), n))
    assert parent[n] == -1 and np.all(parent[np.arange(2 * n - 1) != n] >= n)
    lo, hi = orc.leaf_ranges(left, right, n)
    assert lo[n] == 0 and hi[n] == n - 1
    # children partition the parent's range
    assert np.all(lo[left] == lo[n:]) and np.all(hi[right] == hi[n:])
    assert np.all(hi[left] + 1 == lo[right])


def test_aggregate_closed_form(orc):
    rng = np.random.default_rng(5)
    n = 700
    pos = rng.normal(size=(n, 3)) * 3
    mass = rng.uniform(0.1, 2.0, n)
    mass[:10] = 0.0
    t = orc.bh_tree(pos, mass)
    lo, hi = orc.leaf_ranges(t["left"], t["right"], n)
    o = t["order"]
    for v in range(n, 2 * n - 1, 7):
        idx = o[lo[v]:hi[v] + 1]
        m = mass[idx].sum()
        assert t["nmass"][v] == pytest.approx(m, rel=1e-13)
        com = (mass[idx, None] * pos[idx]).sum(0) / m
        np.testing.assert_allclose(t["com"][v], com, rtol=1e-12, atol=1e-12)
        assert np.array_equal(t["nmin"][v], pos[idx].min(0))
        assert np.array_equal(t["nmax"][v], pos[idx].max(0))
    # leaves are the sorted particles
    assert np.array_equal(t["com"][:n], pos[o])
    assert np.array_equal(t["nmass"][:n], mass[o])


def test_traverse_theta_zero_is_direct(orc):
    rng = np.random.default_rng(8)
    n = 400
    pos = rng.normal(size=(n, 3))
    mass = rng.uniform(0.5, 1.5, n)
    acc, visits, ovf = orc.bh_acc(pos, mass, G=1.0, theta=0.0, eps=0.05)
    assert ovf == 0
    assert np.all(visits == 2 * n - 1)  # every node examined once
    ref = orc.direct_acc(pos, mass, 1.0, 0.05 ** 2)
    np.testing.assert_allclose(acc, ref, rtol=1e-11, atol=1e-13)


def test_traverse_theta_error_levels(orc):
    from rgsim.ic import make_plummer
    s = make_plummer(3000, total_mass=1.0, scale_radius=1.0, seed=42, G=1.0)
    ref = orc.direct_acc(s.pos, s.mass, 1.0, 0.01 ** 2)
    mag = np.linalg.norm(ref, axis=1)
    errs = {}
    for th in (0.2, 0.5, 0.8):
        acc, _, ovf = orc.bh_acc(s.pos, s.mass, 1.0, th, 0.01)
        assert ovf == 0
        errs[th] = np.mean(np.linalg.norm(acc - ref, axis=1) / mag)
    assert errs[0.2] < errs[0.5] < errs[0.8]
    assert errs[0.5] < 0.02  # reference reports ~0.5-0.6% at theta 0.5


def test_traverse_stack_overflow_is_detected(orc):
    from rgsim.ic import make_plummer
    s = make_plummer(2000, total_mass=1.0, seed=1, G=1.0)
    _, _, ovf = orc.bh_acc(s.pos, s.mass, 1.0, 0.5, 0.01, max_stack=4)
    assert ovf > 0

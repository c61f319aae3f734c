"""The FP32 walk records sit in right-first DFS preorder (the order the stack
walk of core/solvers/barnes_hut.py:180-217 pops nodes in).  The library gets
each slot from a scan instead of a climb to the root:
    pre(leaf p) = 2 (n-1-p) + p - F(p),  F(p) = #internal nodes with hi < p,
    pre(v)      = pre(leaf lo(v)) - 2 (hi(v) - lo(v))  for an internal node.
These tests check both formulas against an explicit right-first DFS over the
oracle's radix trees (barnes_hut.py:84-121), including duplicate keys, 10-bit
keys and the smallest trees."""
import numpy as np
import pytest

from oracle import np_oracle as orc


def dfs_preorder(left, right, n):
    pre = np.empty(2 * n - 1, dtype=np.int64)
    stack, c = [n], 0
    while stack:
        v = stack.pop()
        pre[v] = c
        c += 1
        if v >= n:  # right child popped first
            stack.append(int(left[v - n]))
            stack.append(int(right[v - n]))
    return pre


def scan_preorder(left, right, n):
    lo, hi = orc.leaf_ranges(left, right, n)
    ends = np.bincount(hi[n:], minlength=n)           # internal nodes ending at each leaf
    F = np.concatenate([[0], np.cumsum(ends)[:-1]])   # ... ending before it
    p = np.arange(n)
    pre = np.empty(2 * n - 1, dtype=np.int64)
    pre[:n] = 2 * (n - 1 - p) + p - F
    pre[n:] = pre[lo[n:]] - 2 * (hi[n:] - lo[n:])
    return pre


def inputs():
    from rgsim.ic import make_disk_bulge, make_plummer
    rng = np.random.default_rng(5)
    d = make_disk_bulge(n_disk=2500, n_bulge=500, seed=2)
    s = make_plummer(2000, total_mass=1.0, seed=7, G=1.0)
    dup = rng.normal(size=(600, 3))
    dup[300:] = dup[:300]
    return [
        ("diskbulge21", d.pos, 21), ("diskbulge10", d.pos, 10),
        ("plummer", s.pos, 21), ("quantised", np.round(s.pos * 4) / 4, 21),
        ("duplicates", dup, 21), ("n2", np.array([[0.0, 0, 0], [1.0, 1, 1]]), 21),
        ("n3", rng.normal(size=(3, 3)), 21),
    ]


@pytest.mark.parametrize("name,pos,bits", inputs(), ids=[c[0] for c in inputs()])
def test_scan_preorder_equals_dfs(name, pos, bits):
    t = orc.bh_tree(pos, np.ones(len(pos)), bits)
    n = len(pos)
    assert np.array_equal(scan_preorder(t["left"], t["right"], n),
                          dfs_preorder(t["left"], t["right"], n))


def test_scan_preorder_catches_a_wrong_offset():
    """A plausible slip (counting the internal nodes ending at or before p)
    must fail the comparison."""
    pos = np.random.default_rng(1).normal(size=(200, 3))
    t = orc.bh_tree(pos, np.ones(200), 21)
    left, right, n = t["left"], t["right"], 200
    lo, hi = orc.leaf_ranges(left, right, n)
    F_incl = np.cumsum(np.bincount(hi[n:], minlength=n))
    p = np.arange(n)
    wrong = 2 * (n - 1 - p) + p - F_incl
    assert not np.array_equal(wrong, dfs_preorder(left, right, n)[:n])

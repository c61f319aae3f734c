"""The FP32 chunked walk steps a wave of lanes through the right-first
preorder together (the per-lane stack walk of core/solvers/barnes_hut.py:180-217
run as one union walk).  Its next slot is not a minimum over the lanes: it is
k + 1 when an active lane opened slot k, else skip(k), since a lane that is not
at k waits at the skip of a node it accepted earlier, whose subtree holds k.
These tests check that rule against the brute-force minimum of every lane's
next slot, on the oracle's radix trees, for arbitrary accept patterns (leaves
always accepted), and that a plausible wrong rule fails the same check."""
import numpy as np
import pytest

from oracle import np_oracle as orc


def preorder_and_skip(left, right, n):
    """Node at each right-first preorder slot, and each slot's skip slot."""
    lo, hi = orc.leaf_ranges(left, right, n)
    m = 2 * n - 1
    node_at = np.empty(m, dtype=np.int64)
    stack, c = [n if n > 1 else 0], 0
    while stack:
        v = stack.pop()
        node_at[c] = v
        c += 1
        if v >= n:
            stack.append(int(left[v - n]))
            stack.append(int(right[v - n]))
    size = 2 * (hi[node_at] - lo[node_at] + 1) - 1  # nodes in the subtree
    skip = np.arange(m) + size
    return node_at, skip


def union_walk(node_at, skip, n, rng, lanes, p_accept, rule):
    """Steps the union walk; returns the number of steps, asserting at each
    that `rule` gives the smallest next slot over the lanes."""
    m = len(node_at)
    nxt = np.zeros(lanes, dtype=np.int64)
    k, steps = 0, 0
    while k < m:
        act = nxt <= k
        leaf = node_at[k] < n
        acc = act & (leaf | (rng.random(lanes) < p_accept))
        opened = act & ~acc
        nxt = np.where(acc, skip[k], nxt)  # an opening lane stays <= k + 1
        true_next = int(np.min(np.where(opened, k + 1, nxt)))
        kn = rule(k, opened, acc, skip)
        assert kn == true_next, (k, kn, true_next)
        k = kn
        steps += 1
    assert np.all(nxt >= m)  # every lane has finished the tree
    return steps


def rule_walk(k, opened, acc, skip):
    return k + 1 if opened.any() else int(skip[k])


def trees():
    from rgsim.ic import make_disk_bulge, make_plummer
    rng = np.random.default_rng(11)
    d = make_disk_bulge(n_disk=1500, n_bulge=300, seed=4)
    s = make_plummer(1200, total_mass=1.0, seed=8, G=1.0)
    dup = rng.normal(size=(400, 3))
    dup[200:] = dup[:200]
    return [("diskbulge", d.pos, 21), ("diskbulge10", d.pos, 10), ("plummer", s.pos, 21),
            ("duplicates", dup, 21), ("n2", np.array([[0.0, 0, 0], [1.0, 1, 1]]), 21),
            ("n3", rng.normal(size=(3, 3)), 21)]


@pytest.mark.parametrize("name,pos,bits", trees(), ids=[t[0] for t in trees()])
@pytest.mark.parametrize("p_accept", [0.0, 0.3, 0.9])
def test_next_slot_rule_is_the_lanes_minimum(name, pos, bits, p_accept):
    t = orc.bh_tree(pos, np.ones(len(pos)), bits)
    n = len(pos)
    node_at, skip = preorder_and_skip(t["left"], t["right"], n)
    rng = np.random.default_rng(3)
    steps = union_walk(node_at, skip, n, rng, 64, p_accept, rule_walk)
    if p_accept == 0.0:  # nothing accepted but leaves: every node is examined
        assert steps == 2 * n - 1


def test_a_wrong_next_slot_rule_fails():
    """Taking skip(k) as soon as any lane accepts (instead of only when none
    opened) must be caught."""
    pos = np.random.default_rng(2).normal(size=(300, 3))
    t = orc.bh_tree(pos, np.ones(300), 21)
    node_at, skip = preorder_and_skip(t["left"], t["right"], 300)

    def wrong(k, opened, acc, skip):
        return int(skip[k]) if acc.any() else k + 1

    with pytest.raises(AssertionError):
        union_walk(node_at, skip, 300, np.random.default_rng(3), 64, 0.5, wrong)

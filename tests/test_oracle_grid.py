"""Pins of the oracle's restatement of NeighborGrid.rebuild / rebuild_all
(core/solvers/neighbor_grid.py:52-136): invariants of a binning that a
wrong index, axis order or clip would break, the grid-shape rules, small
hand-computed grids, and brute-force completeness of the 3x3x3 stencil."""
import itertools

import numpy as np
import pytest


def gas_case(n, seed, star_frac=0.3, clustered=True):
    rng = np.random.default_rng(seed)
    pos = rng.normal(scale=3.0, size=(n, 3))
    if clustered:
        pos[: n // 4] *= 0.05
    h = rng.uniform(0.05, 0.3, n)
    is_star = (rng.random(n) < star_frac).astype(np.int32)
    return pos, h, is_star


def check_invariants(g, pos, h, is_star, min_radius):
    n = len(h)
    gas = np.arange(n) if is_star is None else np.nonzero(is_star == 0)[0]
    ngas, (nx, ny, nz) = g["ngas"], g["dims"]
    assert ngas == len(gas)
    gs = g["gsort"]
    # every gas particle exactly once, zero padding after
    assert sorted(gs[:ngas].tolist()) == sorted(gas.tolist())
    assert not gs[ngas:].any()
    # the grid: cell size rules and budget
    assert g["cell"] >= 2 * h[gas].max() - 1e-15 and g["cell"] >= min_radius
    assert nx * ny * nz == g["n_cells"] <= g["max_cells"] == max(64, 4 * n)
    cs = g["cell_start"]
    assert cs[0] == 0 and cs[g["n_cells"]] == ngas and not cs[g["n_cells"] + 1:].any()
    assert np.all(np.diff(cs[: g["n_cells"] + 1]) >= 0)
    lo = np.array(g["lo"])
    assert np.array_equal(lo, pos[gas].min(axis=0))
    # each particle sits in the cell whose range holds it (the top cell of an
    # axis also takes the points past it), and within a cell in index order
    for c in range(g["n_cells"]):
        members = gs[cs[c]:cs[c + 1]]
        assert np.all(np.diff(members) > 0)
        i, rem = divmod(c, ny * nz)
        j, k = divmod(rem, nz)
        for a, (q, d) in enumerate(((i, nx), (j, ny), (k, nz))):
            lower = lo[a] + q * g["cell"]
            x = pos[members, a]
            assert np.all(x >= lower - 1e-12 * max(1.0, abs(lower)))
            if q < d - 1:
                assert np.all(x < lower + g["cell"] * (1 + 1e-12))


@pytest.mark.parametrize("n,seed,star", [(1, 0, 0.0), (2, 1, 0.0), (300, 2, 0.3), (2000, 3, 0.5),
                                         (2000, 4, 0.0)])
def test_grid_invariants(orc, n, seed, star):
    pos, h, is_star = gas_case(n, seed, star)
    is_star[0] = 0  # at least one gas particle
    for st in (is_star, None):
        g = orc.grid_rebuild(pos, h, st, min_radius=0.0)
        check_invariants(g, pos, h, st, 0.0)


def test_grid_budget_growth(orc):
    """Cells of 2 h_max over a wide box exceed 4n: the cell grows by 1.3 per
    step until the grid fits, and the last smaller cell did not fit."""
    rng = np.random.default_rng(5)
    pos = rng.uniform(-50, 50, size=(500, 3))
    h = np.full(500, 0.01)
    g = orc.grid_rebuild(pos, h, None)
    assert g["n_cells"] <= g["max_cells"] == 2000
    k = round(np.log(g["cell"] / 0.02) / np.log(1.3))
    assert k > 0 and g["cell"] == pytest.approx(0.02 * 1.3 ** k, rel=1e-12)
    smaller = g["cell"] / 1.3
    span = pos.max(axis=0) - pos.min(axis=0)
    dims = np.maximum(np.floor(span / smaller).astype(np.int64) + 1, 1)
    assert dims.prod() > g["max_cells"]
    check_invariants(g, pos, h, None, 0.0)


def test_grid_min_radius_and_degenerate(orc):
    pos = np.array([[1.0, 2.0, 3.0]] * 5)  # zero span: one cell
    g = orc.grid_rebuild(pos, np.full(5, 0.1), None, min_radius=0.7)
    assert g["cell"] == 0.7 and g["dims"] == (1, 1, 1)
    assert g["cell_start"][:2].tolist() == [0, 5] and g["gsort"].tolist() == [0, 1, 2, 3, 4]
    # h = 0 and no min radius: the cell falls back to 1.0
    g = orc.grid_rebuild(pos, np.zeros(5), None)
    assert g["cell"] == 1.0
    assert orc.grid_rebuild(pos, np.full(5, 0.1), np.ones(5, np.int32)) is None


def test_grid_hand_computed(orc):
    """Four gas particles on the x axis with cell 1: dims (3, 1, 1), cells
    0, 0, 1, 2; the star (index 2) is left out."""
    pos = np.array([[0.0, 0, 0], [0.5, 0, 0], [9.0, 9, 9], [1.5, 0, 0], [2.2, 0, 0]])
    h = np.array([0.5, 0.25, 5.0, 0.25, 0.25])
    is_star = np.array([0, 0, 1, 0, 0], np.int32)
    g = orc.grid_rebuild(pos, h, is_star)
    assert g["cell"] == 1.0 and g["dims"] == (3, 1, 1)
    assert g["gsort"].tolist() == [0, 1, 3, 4, 0]
    assert g["cell_start"][:4].tolist() == [0, 2, 3, 4]


def test_grid_stencil_is_complete(orc):
    """Every gas pair within the cell size of each other lies in the same or
    an adjacent cell: the 3x3x3 search the SPH kernels run finds them all."""
    pos, h, is_star = gas_case(600, 7, 0.2)
    g = orc.grid_rebuild(pos, h, is_star)
    cell_of = {}
    nx, ny, nz = g["dims"]
    cs, gs = g["cell_start"], g["gsort"]
    for c in range(g["n_cells"]):
        for p in gs[cs[c]:cs[c + 1]]:
            i, rem = divmod(c, ny * nz)
            cell_of[int(p)] = (i, *divmod(rem, nz))
    gas = np.nonzero(is_star == 0)[0]
    d = np.linalg.norm(pos[gas, None, :] - pos[None, gas, :], axis=-1)
    close = np.argwhere(d <= g["cell"])
    assert len(close) > len(gas)  # non-trivial
    for a, b in close:
        ca, cb = cell_of[int(gas[a])], cell_of[int(gas[b])]
        assert all(abs(x - y) <= 1 for x, y in zip(ca, cb))

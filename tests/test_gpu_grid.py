"""GPU parity of the SPH neighbour grid (core/solvers/neighbor_grid.py:52-136)
against the oracle's numpy restatement: gsort, cell_start, grid shape,
origin and 1/cell bit-identical, on gas/star mixes, all-gas, duplicates,
degenerate spans, the cell-budget growth loop and a min radius; no gas
leaves the fields as they were; non-finite input is rejected."""
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).cuda()


def case(name):
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    if name.startswith("mix"):
        n = int(name[3:])
        pos = rng.normal(scale=3.0, size=(n, 3))
        pos[: n // 4] *= 0.05
        h = rng.uniform(0.05, 0.3, n)
        st = (rng.random(n) < 0.3).astype(np.int32)
        st[0] = 0
        return pos, h, st, 0.0
    if name == "galaxy120k":
        from rgsim.ic import make_disk_bulge
        d = make_disk_bulge(100000, 20000, seed=1)
        h = 0.1 * np.cbrt(rng.uniform(0.2, 1.0, d.n))
        st = (rng.random(d.n) < 0.5).astype(np.int32)
        return d.pos, h, st, 0.0
    if name == "dups":
        pos = rng.normal(size=(1000, 3))
        pos[500:] = pos[:500]
        return pos, np.full(1000, 0.2), np.zeros(1000, np.int32), 0.0
    if name == "same_point":
        return np.tile([[1.0, -2.0, 3.0]], (7, 1)), np.full(7, 0.1), np.zeros(7, np.int32), 0.0
    if name == "wide_box":  # 2 h_max cells over a wide box: the budget loop runs
        return (rng.uniform(-50, 50, size=(500, 3)), np.full(500, 0.01),
                np.zeros(500, np.int32), 0.0)
    if name == "min_radius":
        return rng.normal(size=(800, 3)), rng.uniform(0.0, 0.05, 800), np.zeros(800, np.int32), 0.9
    if name == "one":
        return np.array([[0.5, 0.5, 0.5]]), np.array([0.1]), np.zeros(1, np.int32), 0.0
    raise KeyError(name)


# mix600000 is above the key count where the sort switches to its larger tile
NAMES = ["mix2", "mix300", "mix5000", "galaxy120k", "mix600000", "dups", "same_point",
         "wide_box", "min_radius", "one"]


@pytest.mark.parametrize("all_gas", [False, True])
@pytest.mark.parametrize("name", NAMES)
def test_grid_bitexact_vs_oracle(orc, name, all_gas):
    from rgsim.solvers.neighbor_grid import NeighborGrid
    print(f"grid {name} all_gas={all_gas}", flush=True)
    pos, h, st, rmin = case(name)
    n = len(h)
    g = NeighborGrid(n, min_radius=rmin)
    if all_gas:
        ngas = g.rebuild_all(dev(pos), dev(h))
        ref = orc.grid_rebuild(pos, h, None, rmin)
    else:
        ngas = g.rebuild(dev(pos), dev(h), dev(st, torch.int32))
        ref = orc.grid_rebuild(pos, h, st, rmin)
    assert ngas == ref["ngas"] == g.ngas
    assert (g.nx, g.ny, g.nz) == ref["dims"]
    assert g.lo == ref["lo"] and g.inv_cell == ref["inv_cell"] and g.max_cells == ref["max_cells"]
    assert np.array_equal(g.gsort.cpu().numpy()[:n], ref["gsort"])
    assert np.array_equal(g.cell_start.cpu().numpy(), ref["cell_start"])


def test_grid_no_gas_leaves_fields(orc):
    from rgsim.solvers.neighbor_grid import NeighborGrid
    pos, h, st, _ = case("mix300")
    g = NeighborGrid(300)
    g.rebuild(dev(pos), dev(h), dev(st, torch.int32))
    before = (g.gsort.clone(), g.cell_start.clone(), g.nx, g.ny, g.nz, g.lo, g.inv_cell)
    assert g.rebuild(dev(pos), dev(h), dev(np.ones(300, np.int32), torch.int32)) == 0
    assert g.ngas == 0
    assert torch.equal(g.gsort, before[0]) and torch.equal(g.cell_start, before[1])
    assert (g.nx, g.ny, g.nz, g.lo, g.inv_cell) == before[2:]


def test_grid_reuse_and_errors(orc):
    from rgsim.solvers.neighbor_grid import NeighborGrid
    pos, h, st, _ = case("mix5000")
    g = NeighborGrid(5000)
    g.rebuild(dev(pos), dev(h), dev(st, torch.int32))
    pos2 = pos * 0.5 + 1.0
    g.rebuild(dev(pos2), dev(h), dev(st, torch.int32))  # second call, new geometry
    ref = orc.grid_rebuild(pos2, h, st)
    assert np.array_equal(g.gsort.cpu().numpy(), ref["gsort"])
    assert np.array_equal(g.cell_start.cpu().numpy(), ref["cell_start"])
    bad = pos.copy()
    bad[np.nonzero(st == 0)[0][3], 1] = np.nan
    with pytest.raises(RuntimeError, match="non-finite"):
        g.rebuild(dev(bad), dev(h), dev(st, torch.int32))
    bad = pos.copy()
    bad[np.nonzero(st == 1)[0][0], 0] = np.inf  # a star's coordinate is not binned
    g.rebuild(dev(bad), dev(h), dev(st, torch.int32))
    assert g.ngas == ref["ngas"]
    empty = NeighborGrid(0)
    assert empty.rebuild_all(torch.zeros((0, 3), dtype=torch.float64, device="cuda"),
                             torch.zeros(0, dtype=torch.float64, device="cuda")) == 0

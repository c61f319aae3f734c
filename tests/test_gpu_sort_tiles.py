"""Radix-sort tile shapes (256 threads with 4, 8 or 16 keys per thread, or
1024 threads with 4; with and without tile tickets) give the stable order
of the Morton keys, as np.argsort(kind="stable") of the oracle's keys
(barnes_hut.py:38-46, 247), at the sizes where each shape is the default
and forced at a small size."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


@pytest.fixture
def sort_params():
    from rgsim import _lib
    yield _lib.set_param
    _lib.set_param("sort_items", 0)
    _lib.set_param("sort_ticket", 0)
    _lib.set_param("sort_block", 0)


@pytest.mark.parametrize("n,items,ticket,block", [
    (1_300_001, 0, 0, 0),   # default 1024 x 4, more tiles than CUs: ticketed
    (800_000, 0, 0, 0),     # default 1024 x 4, every tile resident
    (600_000, 0, 1, 0),     # default 1024 x 4, tickets forced
    (120_001, 0, 0, 0),     # default 256 x 4
    # 256-thread tiles at the sizes where they were the defaults
    (1_300_001, 0, 0, 256), (800_000, 0, 0, 256), (600_000, 0, 1, 256),
    (120_001, 16, 0, 0), (120_001, 8, 1, 0), (5000, 16, 1, 0),
    # 1024-thread tiles forced small: resident, ticketed, ragged
    (120_001, 0, 0, 1024), (5000, 0, 1, 1024), (4097, 0, 0, 1024)])
def test_order_matches_stable_argsort(orc, sort_params, n, items, ticket, block):
    from rgsim.solvers.barnes_hut import BarnesHut
    print(f"sort tiles n={n} items={items} ticket={ticket} block={block}", flush=True)
    rng = np.random.default_rng(n)
    pos = rng.normal(size=(n, 3))
    pos[n // 2:n // 2 + n // 10] = pos[:n // 10]  # a tenth duplicated: equal keys
    sort_params("sort_items", items)
    sort_params("sort_ticket", ticket)
    sort_params("sort_block", block)
    bh = BarnesHut(n, theta=0.5, softening=0.05, G=1.0, precision="fp32")
    p = dev(pos)
    acc = torch.empty_like(p)
    bh.compute(p, dev(np.ones(n)), acc)
    keys, order = (t.cpu().numpy() for t in bh.export_tree()[:2])
    k_ref = orc.morton_c(pos)
    o_ref = np.argsort(k_ref, kind="stable").astype(np.int32)
    assert np.array_equal(order, o_ref)
    assert np.array_equal(keys.astype(np.uint64), k_ref[o_ref])


def test_sort_items_range():
    from rgsim import _lib
    with pytest.raises(RuntimeError, match="sort_items"):
        _lib.set_param("sort_items", 2)

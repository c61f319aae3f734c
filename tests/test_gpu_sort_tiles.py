"""Radix-sort tile shapes (256 threads with 4, 8 or 16 keys per thread, or
1024 threads with 4; with and without tile tickets) give the stable order
of the Morton keys, as np.argsort(kind="stable") of the oracle's keys
(barnes_hut.py:38-46, 247), at the sizes where each shape is the default
and forced at a small size; and so does the sample sort that a handle uses
from its second compute on, with buckets in LDS and with a bucket too large
for LDS (sorted in global memory)."""
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
    _lib.set_param("sort_sample", -1)


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


def _check_order(orc, bh, pos, bits):
    keys, order = (t.cpu().numpy() for t in bh.export_tree()[:2])
    k_ref = orc.morton_c(pos, bits)
    o_ref = np.argsort(k_ref, kind="stable").astype(np.int32)
    assert np.array_equal(order, o_ref)
    assert np.array_equal(keys.astype(np.uint64), k_ref[o_ref])
    return keys.astype(np.uint64), order


def _bucket_sizes(prev_keys, prev_order, keys, nbk):
    """Bucket sizes of the sample sort: splitters (key, index) at every
    stride-th position of the previous sorted keys and order; particle i
    goes to the number of splitters <= (keys[i], i)."""
    n = len(keys)
    stride = -(-n // nbk)
    sk = prev_keys[stride * np.arange(1, nbk)]
    si = prev_order[stride * np.arange(1, nbk)].astype(np.int64)
    o = np.lexsort((si, sk))
    sk, si = sk[o], si[o]
    b = np.searchsorted(sk, keys, side="right")
    lo = np.searchsorted(sk, keys, side="left")
    tie = np.nonzero(b > lo)[0]  # equal to splitter keys: compare indices
    for l0 in np.unique(lo[tie]):
        j = tie[lo[tie] == l0]
        run = si[l0:b[j[0]]]  # ascending within equal keys
        b[j] = l0 + np.searchsorted(run, j, side="right")
    return np.bincount(b, minlength=nbk)


@pytest.mark.parametrize("n,bits,force", [
    (120_001, 21, 0), (120_001, 10, 0), (1_000_003, 21, 0), (1_000_003, 10, 0),
    (9000, 21, 0), (5000, 21, 1), (1000, 21, 1), (3, 21, 1)])
def test_sample_sort_matches_stable_argsort(orc, sort_params, n, bits, force):
    """First compute: the LSD radix sort.  Second, after a small move: the
    sample sort with the first result's splitters, every bucket in LDS.
    Third, with 99% of the particles collapsed onto a small cluster inside
    the box: the cluster's buckets (above 8,192 keys) are sorted in global
    memory.  The spill count is checked against the bucket sizes computed
    here from the exported previous result.  A tenth of the particles
    duplicate others (equal keys, ordered by index); 30-bit keys put many
    more on equal keys."""
    from rgsim.solvers.barnes_hut import BarnesHut
    print(f"sample sort n={n} bits={bits} force={force}", flush=True)
    if force:
        sort_params("sort_sample", 1)
    rng = np.random.default_rng(n + bits)
    pos = rng.normal(size=(n, 3))
    pos[n // 2:n // 2 + n // 10] = pos[:n // 10]
    mass = dev(np.ones(n))
    bh = BarnesHut(n, theta=0.5, softening=0.05, G=1.0, precision="fp32", morton_bits=bits)
    acc = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    bh.compute(dev(pos), mass, acc)
    k1, o1 = _check_order(orc, bh, pos, bits)
    st = bh.stats()
    assert st["sort_passes"] == (8 if bits == 21 else 4) and st["sort_buckets"] == 0

    nbk = min(256, max(1, n // 1024))
    sample = force == 1 or n >= 4096
    pos2 = pos + 1e-3 * rng.normal(size=(n, 3))
    bh.compute(dev(pos2), mass, acc)
    k2, o2 = _check_order(orc, bh, pos2, bits)
    st = bh.stats()
    assert st["sort_passes"] == (1 if sample else (8 if bits == 21 else 4))
    assert st["sort_buckets"] == (nbk if sample else 0)
    if sample:
        sizes = _bucket_sizes(k1, o1, orc.morton_c(pos2, bits), nbk)
        assert st["sort_spilled"] == np.count_nonzero(sizes > 8192) == 0

    pos3 = pos2.copy()
    cl = np.arange(n) % 100 != 0
    pos3[cl] = np.array([0.1, 0.2, 0.3]) + 1e-4 * rng.normal(size=(np.count_nonzero(cl), 3))
    bh.compute(dev(pos3), mass, acc)
    _check_order(orc, bh, pos3, bits)
    st = bh.stats()
    if sample:
        sizes = _bucket_sizes(k2, o2, orc.morton_c(pos3, bits), nbk)
        assert st["sort_spilled"] == np.count_nonzero(sizes > 8192)
        if n >= 9000:
            assert st["sort_spilled"] >= 1  # the global-memory path ran
    assert torch.all(torch.isfinite(acc))


def test_sort_sample_range():
    from rgsim import _lib
    with pytest.raises(RuntimeError, match="sort_sample"):
        _lib.set_param("sort_sample", 2)

"""rg_direct_acc_noalloc (the direct sum with its slice partials in a static
device array, for callers that cannot allocate) against rg_direct_acc, the
entry point it is a variant of: bit for bit in every precision, on two calls
in a row, the second with a smaller input; a size whose split needs more
than the static array is rejected before any launch."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def call(name, pos, mass, acc, n, prec):
    from rgsim import _lib
    L = _lib.load()
    return getattr(L, name)(pos.data_ptr(), mass.data_ptr(), acc.data_ptr(), n, 1.0,
                            0.01 ** 2, _lib.precision_code(prec), _lib.stream_ptr(pos))


@pytest.mark.parametrize("prec", ["fp32c", "fp32", "fp64"])
def test_noalloc_matches_direct_acc(prec):
    from rgsim import _lib
    from rgsim.ic import make_plummer
    for n in (70001, 3001):  # the second call is smaller: stale partials must not leak
        print(f"noalloc {prec} n={n}", flush=True)
        s = make_plummer(n, total_mass=1.0, seed=n, G=1.0)
        pos = torch.from_numpy(s.pos).cuda()
        mass = torch.from_numpy(s.mass).cuda()
        ref = torch.full_like(pos, np.nan)
        got = torch.full_like(pos, np.nan)
        _lib.check(call("rg_direct_acc", pos, mass, ref, n, prec), "rg_direct_acc")
        _lib.check(call("rg_direct_acc_noalloc", pos, mass, got, n, prec),
                   "rg_direct_acc_noalloc")
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (prec, n, (got - ref).abs().max().item())
        assert torch.all(torch.isfinite(got))


def test_noalloc_rejects_sizes_beyond_its_scratch():
    from rgsim import _lib
    n = 200000  # FP32C default split: 64 slices x 200,000 x 24 B = 307 MB > 128 MiB
    pos = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    mass = torch.ones(n, dtype=torch.float64, device="cuda")
    acc = torch.full_like(pos, 5.0)
    st = call("rg_direct_acc_noalloc", pos, mass, acc, n, "fp32c")
    assert st == _lib.RG_ERR_CAPACITY
    assert b"static" in _lib.load().rg_last_error()
    torch.cuda.synchronize()
    assert torch.all(acc == 5.0)
    # FP64 sums are never split: no scratch, any n
    s = torch.randn((n, 3), dtype=torch.float64, device="cuda")
    _lib.check(call("rg_direct_acc_noalloc", s, mass, acc, n, "fp64"), "rg_direct_acc_noalloc")
    ref = torch.empty_like(acc)
    _lib.check(call("rg_direct_acc", s, mass, ref, n, "fp64"), "rg_direct_acc")
    torch.cuda.synchronize()
    assert torch.equal(acc, ref)

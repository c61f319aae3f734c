"""The symmetric FP32/FP32C direct sum (each pair evaluated once, its term
added to both particles; Newton's third law), against the oracle's FP64 sum
of gravity.py:13-26 at sizes where the oracle finishes in seconds, with the
path forced on ("direct_sym" 1; by default it runs from n = 131,072, which
the precision-form test covers at N=262,144).  Ragged i-blocks and j-tiles,
the epilogues (add, analytic add, closing kick), hi/lo positions and
run-to-run determinism are covered."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


@pytest.fixture(params=[1, 3], ids=["lds_fp64", "rotating_fp32"])
def sym_on(request):
    """The symmetric path forced on, with the j-side sums in FP64 LDS
    accumulators (1) or rotating through the lanes' registers (3)."""
    from rgsim import _lib
    _lib.set_param("direct_sym", request.param)
    yield request.param
    _lib.set_param("direct_sym", 0)


def plummer(n, seed):
    from rgsim.ic import make_plummer
    return make_plummer(n, total_mass=1.0, scale_radius=1.0, seed=seed, G=1.0)


def rel_rows(a, r):
    return np.linalg.norm(a - r, axis=1) / np.linalg.norm(r, axis=1)


@pytest.mark.parametrize("precision,tol_max,tol_rms", [("fp32c", 3e-6, 3e-7), ("fp32", 2e-3, 1e-4)])
@pytest.mark.parametrize("n", [2, 700, 1025, 3000, 5003, 20000])
def test_sym_vs_oracle(orc, sym_on, n, precision, tol_max, tol_rms):
    from rgsim.solvers import gravity
    print(f"sym {precision} n={n}", flush=True)
    s = plummer(n, seed=n)
    ref = orc.direct_acc(s.pos, s.mass, 1.0, 0.01 ** 2)
    p, m = dev(s.pos), dev(s.mass)
    acc = torch.full_like(p, np.nan)
    gravity.compute_acc_direct(p, m, acc, n, 1.0, 0.01 ** 2, precision)
    rel = rel_rows(acc.cpu().numpy(), ref)
    print(f"  max {rel.max():.2e} rms {np.sqrt(np.mean(rel ** 2)):.2e}", flush=True)
    assert rel.max() < tol_max and np.sqrt(np.mean(rel ** 2)) < tol_rms
    again = torch.empty_like(p)
    gravity.compute_acc_direct(p, m, again, n, 1.0, 0.01 ** 2, precision)
    assert torch.equal(acc, again)  # fixed summation order


def test_sym_epilogues_and_hilo(orc, sym_on):
    """The reduce pass's epilogues use the same sums as the store: add is
    acc0 + a, the fused kick is the store then the kick, bit for bit; hi/lo
    positions with zero offsets equal plain positions."""
    from rgsim.integrators import leapfrog
    from rgsim.solvers import gravity
    n = 5003
    s = plummer(n, seed=4)
    p, m = dev(s.pos), dev(s.mass)
    a = torch.empty_like(p)
    gravity.compute_acc_direct(p, m, a, n, 1.0, 1e-4, "fp32c")
    acc0 = torch.randn_like(p)
    b = acc0.clone()
    gravity.add_acc_direct(p, m, b, n, 1.0, 1e-4, "fp32c")
    assert torch.equal(b, acc0 + a)
    v = torch.randn_like(p)
    v2 = v.clone()
    c = torch.empty_like(p)
    gravity.compute_acc_direct_kick(p, m, c, v, n, 1.0, 1e-4, 0.5e-3, "fp32c")
    leapfrog.kick(v2, a, n, 0.5e-3)
    assert torch.equal(c, a) and torch.equal(v, v2)
    lo = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    h = torch.empty_like(p)
    gravity.compute_acc_direct_hilo(p, lo, m, h, n, 1.0, 1e-4)
    assert torch.equal(h, a)


def test_sym_default_range():
    """Off below 131,072 particles by default; an explicit split selects the
    j-split kernel (the batched kernel's per-system equality relies on it)."""
    from rgsim import _lib
    with pytest.raises(RuntimeError):
        _lib.set_param("direct_sym", 4)

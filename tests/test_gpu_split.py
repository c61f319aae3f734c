"""The j-split of the FP32/FP32C direct sums and of the adjoint VJP (blocks
over (i-block, j-slice), slice sums added in slice order in FP64).

Pinned against the oracle's FP64 restatement of gravity.py:13-26 at every
split count (ragged slices included), through each epilogue that follows
the slice reduction (store, add, analytic add, kick), against the closed
form of the VJP, and for bitwise repeatability at a fixed split."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def plummer(n, seed=0):
    from rgsim.ic import make_plummer
    return make_plummer(n, total_mass=1.0, scale_radius=1.0, seed=seed, G=1.0)


@pytest.fixture
def split():
    from rgsim import _lib

    def set_split(v):
        _lib.set_param("direct_split", v)
    yield set_split
    _lib.set_param("direct_split", 0)


def rel_rows(got, ref):
    return np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)


@pytest.mark.parametrize("n", [257, 1000, 4096, 5003])
@pytest.mark.parametrize("ns", [1, 2, 3, 7, 0])
@pytest.mark.parametrize("precision,tol_max,tol_rms", [("fp32c", 3e-6, 3e-7),
                                                        ("fp32", 2e-3, 1e-4)])
def test_split_sum_vs_oracle(orc, split, n, ns, precision, tol_max, tol_rms):
    from rgsim.solvers import gravity
    print(f"split {precision} n={n} ns={ns}", flush=True)
    s = plummer(n, seed=n)
    ref = orc.direct_acc(s.pos, s.mass, 1.0, 0.01 ** 2)
    split(ns)
    p, m = dev(s.pos), dev(s.mass)
    acc = torch.empty_like(p)
    gravity.compute_acc_direct(p, m, acc, n, 1.0, 0.01 ** 2, precision)
    got = acc.cpu().numpy()
    rel = rel_rows(got, ref)
    assert rel.max() < tol_max and np.sqrt(np.mean(rel ** 2)) < tol_rms, (rel.max(), ns)
    again = torch.empty_like(p)
    gravity.compute_acc_direct(p, m, again, n, 1.0, 0.01 ** 2, precision)
    assert torch.equal(acc, again)


def test_split_epilogues(orc, split):
    """The add, analytic-add and kick epilogues run after the slice
    reduction on the same FP64 sum the store form writes."""
    from rgsim.solvers import gravity
    n, G, eps2, dth = 3001, 43009.1725, 0.05 ** 2, 5e-5
    print(f"split epilogues n={n}", flush=True)
    s = plummer(n, seed=4)
    rng = np.random.default_rng(4)
    acc0 = rng.normal(size=(n, 3))
    split(5)
    p, m = dev(s.pos), dev(s.mass)
    stored = torch.empty_like(p)
    gravity.compute_acc_direct(p, m, stored, n, G, eps2, "fp32c")
    a = stored.cpu().numpy()
    ref = orc.direct_acc(s.pos, s.mass, G, eps2)
    assert rel_rows(a, ref).max() < 3e-6

    added = dev(acc0)
    gravity.add_acc_direct(p, m, added, n, G, eps2, "fp32c")
    assert np.array_equal(added.cpu().numpy(), acc0 + a)

    v = dev(s.vel)
    kicked = torch.empty_like(p)
    gravity.compute_acc_direct_kick(p, m, kicked, v, n, G, eps2, dth, "fp32c")
    assert np.array_equal(kicked.cpu().numpy(), a)
    assert np.array_equal(v.cpu().numpy(), orc.kick(s.vel, a, dth))

    pot = gravity.potential_struct(5.0, 3.0, 0.3, 12.0, 8.0, 0, 0.1)
    ext = dev(acc0)
    gravity.add_acc_direct_ext(p, m, ext, n, G, eps2, pot, "fp32c")
    split(1)
    ext1 = dev(acc0)
    gravity.add_acc_direct_ext(p, m, ext1, n, G, eps2, pot, "fp32c")
    e, e1 = ext.cpu().numpy(), ext1.cpu().numpy()
    scale = np.abs(e1 - acc0).max()
    assert np.abs(e - e1).max() <= 2e-6 * scale


def vjp_closed_form(x, m, lam, G, eps2):
    """g_x[k] = G sum_j T_kj (m_k lam_j - m_j lam_k); g_m[k] = -G sum_j lam_j.d_kj/s^3."""
    d = x[None, :, :] - x[:, None, :]
    s2 = (d * d).sum(-1) + eps2
    inv3 = s2 ** -1.5
    inv5 = s2 ** -2.5
    v = m[:, None, None] * lam[None, :, :] - m[None, :, None] * lam[:, None, :]
    dv = (d * v).sum(-1)
    gx = G * ((v * inv3[..., None]).sum(1) - 3 * ((dv * inv5)[..., None] * d).sum(1))
    gm = -G * ((lam[None, :, :] * d).sum(-1) * inv3).sum(1)
    return gx, gm


@pytest.mark.parametrize("ns", [1, 3, 0])
@pytest.mark.parametrize("precision,rtol", [("fp64", 1e-11), ("fp32c", 2e-5)])
def test_split_vjp_vs_closed_form(split, ns, precision, rtol):
    from rgsim.adjoint import direct_vjp
    n = 2500
    print(f"split vjp {precision} n={n} ns={ns}", flush=True)
    s = plummer(n, seed=6)
    rng = np.random.default_rng(6)
    mass = s.mass * rng.uniform(0.5, 1.5, n)
    lam = rng.normal(size=(n, 3))
    gx_ref, gm_ref = vjp_closed_form(torch.tensor(s.pos), torch.tensor(mass),
                                     torch.tensor(lam), 1.0, 1e-4)
    split(ns)
    args = (dev(s.pos), dev(mass), dev(lam), 1.0, 1e-4, precision)
    gx, gm = direct_vjp(*args)
    assert (gx.cpu() - gx_ref).abs().max() <= rtol * gx_ref.abs().max()
    assert (gm.cpu() - gm_ref).abs().max() <= rtol * gm_ref.abs().max()
    gx2, gm2 = direct_vjp(*args)
    assert torch.equal(gx, gx2) and torch.equal(gm, gm2)



def test_split_on_two_streams(split):
    """Split sums issued on two streams at once each use their own scratch:
    both equal the same sums issued one after another on one stream."""
    from rgsim.solvers import gravity
    print("split two streams", flush=True)
    systems = [plummer(n, seed=20 + n) for n in (6000, 7001)]
    split(0)
    bufs = []
    for s in systems:
        p, m = dev(s.pos), dev(s.mass)
        bufs.append((p, m, torch.empty_like(p), torch.empty_like(p)))
    for p, m, ref, _ in bufs:
        gravity.compute_acc_direct(p, m, ref, len(m), 1.0, 1e-4, "fp32c")
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for _ in range(3):
        for (p, m, _, out), st in zip(bufs, streams):
            with torch.cuda.stream(st):
                gravity.compute_acc_direct(p, m, out, len(m), 1.0, 1e-4, "fp32c")
        torch.cuda.synchronize()
        for p, m, ref, out in bufs:
            assert torch.equal(out, ref)


def test_direct_autotuner_keeps_results_in_tolerance(orc):
    """tune_direct times i-particles per lane x j-slices and leaves the
    fastest set; the sum at that setting stays within the FP32C tolerance of
    the oracle.  FP64 is tuned over i-particles per lane only and stays
    bit-exact."""
    from rgsim import _lib
    from rgsim.autotune import tune_direct
    from rgsim.solvers import gravity
    n = 3000
    print(f"tune_direct n={n}", flush=True)
    s = plummer(n, seed=31)
    p, m = dev(s.pos), dev(s.mass)
    ref = orc.direct_acc(s.pos, s.mass, 1.0, 1e-4)
    try:
        best = tune_direct(p, m, 1.0, 1e-4, "fp32c", ipts=(1, 2), splits=(0, 1, 4), reps=1)
        assert len(best["trials"]) == 6 and best["split"] in (0, 1, 4)
        acc = torch.empty_like(p)
        gravity.compute_acc_direct(p, m, acc, n, 1.0, 1e-4, "fp32c")
        assert rel_rows(acc.cpu().numpy(), ref).max() < 3e-6
        best = tune_direct(p, m, 1.0, 1e-4, "fp64", ipts=(1, 2), reps=1)
        assert len(best["trials"]) == 2 and best["split"] == 0
        gravity.compute_acc_direct(p, m, acc, n, 1.0, 1e-4, "fp64")
        assert np.array_equal(acc.cpu().numpy(), ref)
    finally:
        _lib.set_param("direct_ipt", 0)
        _lib.set_param("direct_split", 0)


def term_magnitude_sums(pos, mass, eps2, idx):
    """sum_j |m_j d_ij| / (|d_ij|^2 + eps2)^(3/2) (G = 1) for the particles idx."""
    out = np.empty(len(idx))
    for k, i in enumerate(idx):
        d = pos - pos[i]
        q = (d * d).sum(1)
        r2 = q + eps2
        out[k] = (mass * np.sqrt(q) / (r2 * np.sqrt(r2))).sum()
    return out


@pytest.mark.parametrize("n,precision,seed,tol_max,tol_rms,tol_cond", [
    (65536, "fp32c", 3, 1e-6, 1e-7, 2e-7), (65536, "fp32c", 0, 1e-5, 1e-7, 2e-7),
    (70001, "fp32c", 3, 1e-6, 1e-7, 2e-7), (131073, "fp32", 3, 2e-3, 1e-4, 5e-5)])
def test_default_shape_large_n(orc, n, precision, seed, tol_max, tol_rms, tol_cond):
    """The default launch shape at large n (FP32C four i-particles per lane
    and 64 slices from n = 65,536; FP32 two per lane from 131,072), ragged
    sizes included, against the oracle.  The maximum relative error is set
    by particles whose forces nearly cancel, so it depends on the
    realization (seed 3: within the 1e-6 the precision form is held to;
    seed 0 puts it at 4e-6 on a particle with |a| 1/40 of the median).  The
    error over the sum of the term magnitudes, on a 128-particle sample plus
    the worst particle, does not: FP32C stays below 2e-7 in both."""
    from rgsim.solvers import gravity
    print(f"default shape {precision} n={n} seed={seed}", flush=True)
    s = plummer(n, seed=seed)
    ref = orc.direct_acc(s.pos, s.mass, 1.0, 0.005 ** 2)
    p, m = dev(s.pos), dev(s.mass)
    acc = torch.empty_like(p)
    gravity.compute_acc_direct(p, m, acc, n, 1.0, 0.005 ** 2, precision)
    got = acc.cpu().numpy()
    rel = rel_rows(got, ref)
    idx = np.concatenate([np.random.default_rng(5).choice(n, 128, replace=False),
                          [int(np.argmax(rel))]])
    cond = np.linalg.norm(got[idx] - ref[idx], axis=1) / term_magnitude_sums(
        s.pos, s.mass, 0.005 ** 2, idx)
    print(f"  max {rel.max():.2e} rms {np.sqrt(np.mean(rel ** 2)):.2e} "
          f"cond {cond.max():.2e}", flush=True)
    assert rel.max() < tol_max and np.sqrt(np.mean(rel ** 2)) < tol_rms
    assert cond.max() < tol_cond

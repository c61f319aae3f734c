"""GPU parity of the direct sum, KDK and energies against the oracle
(core/solvers/gravity.py, core/integrators/leapfrog.py)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def plummer(n, seed=0, G=1.0, M=1.0, a=1.0):
    from rgsim.ic import make_plummer
    return make_plummer(n, total_mass=M, scale_radius=a, seed=seed, G=G)


def gpu_acc(pos, mass, G, eps2, precision):
    from rgsim.solvers import gravity
    p, m = dev(pos), dev(mass)
    acc = torch.empty_like(p)
    gravity.compute_acc_direct(p, m, acc, len(mass), G, eps2, precision)
    return acc.cpu().numpy()


@pytest.mark.parametrize("n", [1, 2, 3, 255, 256, 257, 1000, 4096])
def test_fp64_bitexact_vs_oracle(orc, n):
    print(f"direct fp64 n={n}", flush=True)
    s = plummer(n, seed=n)
    G = 43009.1725
    ref = orc.direct_acc(s.pos, s.mass, G, 0.05 ** 2)
    got = gpu_acc(s.pos, s.mass, G, 0.05 ** 2, "fp64")
    assert np.array_equal(got, ref), np.abs(got - ref).max()


@pytest.mark.parametrize("ipt", [1, 2, 4])
@pytest.mark.parametrize("precision,tol_max,tol_rms", [("fp32c", 3e-6, 3e-7), ("fp32", 2e-3, 1e-4)])
def test_fp32_tolerance(orc, precision, tol_max, tol_rms, ipt):
    from rgsim import _lib
    print(f"direct {precision} ipt={ipt}", flush=True)
    _lib.set_param("direct_ipt", ipt)
    try:
        s = plummer(4096, seed=7)
        ref = orc.direct_acc(s.pos, s.mass, 1.0, 0.01 ** 2)
        got = gpu_acc(s.pos, s.mass, 1.0, 0.01 ** 2, precision)
    finally:
        _lib.set_param("direct_ipt", 0)
    rel = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert rel.max() < tol_max and np.sqrt(np.mean(rel ** 2)) < tol_rms, (rel.max(), rel.mean())


def test_fp32c_far_from_origin(orc):
    # offset cluster: hi/lo splitting must keep the separation accurate
    s = plummer(2048, seed=3)
    pos = s.pos * 0.01 + np.array([1000.0, -2000.0, 500.0])
    ref = orc.direct_acc(pos, s.mass, 1.0, 1e-6)
    got = gpu_acc(pos, s.mass, 1.0, 1e-6, "fp32c")
    rel = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert rel.max() < 1e-5, rel.max()


def test_fused_kick_bitexact(orc):
    from rgsim.solvers import gravity
    s = plummer(1500, seed=5)
    p, m, v = dev(s.pos), dev(s.mass), dev(s.vel)
    acc = torch.empty_like(p)
    gravity.compute_acc_direct_kick(p, m, acc, v, 1500, 1.0, 1e-4, 0.5e-3, "fp64")
    ref_acc = orc.direct_acc(s.pos, s.mass, 1.0, 1e-4)
    assert np.array_equal(acc.cpu().numpy(), ref_acc)
    assert np.array_equal(v.cpu().numpy(), orc.kick(s.vel, ref_acc, 0.5e-3))


def test_kick_drift_bitexact(orc):
    from rgsim.integrators import leapfrog
    rng = np.random.default_rng(0)
    x, v, a = (rng.normal(size=(777, 3)) for _ in range(3))
    X, V, A = dev(x), dev(v), dev(a)
    leapfrog.kick_drift(X, V, A, 777, 0.25, 0.5)
    v1 = orc.kick(v, a, 0.25)
    x1 = orc.drift(x, v1, 0.5)
    assert np.array_equal(V.cpu().numpy(), v1) and np.array_equal(X.cpu().numpy(), x1)
    leapfrog.kick(V, A, 777, 0.1)
    leapfrog.drift(X, V, 777, 0.2)
    v2 = orc.kick(v1, a, 0.1)
    assert np.array_equal(V.cpu().numpy(), v2)
    assert np.array_equal(X.cpu().numpy(), orc.drift(x1, v2, 0.2))


@pytest.mark.parametrize("n", [1, 2, 300, 2049])
def test_energy_vs_oracle(orc, n):
    from rgsim.solvers import gravity
    s = plummer(n, seed=11 + n, G=43009.1725, M=10.0)
    G, eps2 = 43009.1725, 0.05 ** 2
    ke, pe, mom = gravity.energy_terms(dev(s.pos), dev(s.vel), dev(s.mass), n, G, eps2)
    ke_ref = orc.kinetic_energy(s.vel, s.mass)
    pe_ref, pe_i = orc.potential_energy(s.pos, s.mass, G, eps2, per_particle=True)
    assert ke == pytest.approx(ke_ref, rel=1e-13)
    if n > 1:
        assert pe == pytest.approx(pe_ref, rel=1e-13)
    else:
        assert pe == 0.0
    np.testing.assert_allclose(mom, (s.mass[:, None] * s.vel).sum(0), rtol=1e-10, atol=1e-9)
    # fixed-order reduction: bitwise reproducible
    again = gravity.energy_terms(dev(s.pos), dev(s.vel), dev(s.mass), n, G, eps2)
    assert again == (ke, pe, mom)


def test_errors_are_reported():
    from rgsim.solvers import gravity
    s = plummer(64)
    p, m = dev(s.pos), dev(s.mass)
    with pytest.raises(RuntimeError, match="eps2"):
        gravity.compute_acc_direct(p, m, torch.empty_like(p), 64, 1.0, 0.0)
    with pytest.raises(RuntimeError, match="alias"):
        gravity.compute_acc_direct(p, m, p, 64, 1.0, 0.01)
    with pytest.raises(TypeError):
        gravity.compute_acc_direct(p.float(), m, torch.empty_like(p), 64, 1.0, 0.01)
    with pytest.raises(ValueError):
        gravity.compute_acc_direct(p, m, torch.empty_like(p), 65, 1.0, 0.01)

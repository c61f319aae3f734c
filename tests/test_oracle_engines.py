"""Pins of the oracle's restatement of the gas-disk and living-galaxy
engines' gravity kernels (core/gas_disk_engine.py:20-89,
core/living_galaxy_engine.py:101-150).

The analytic accelerations are checked against exact gradients of the
potentials they come from, differentiated symbolically (sympy) and evaluated
at 50 digits (mpmath):
  Miyamoto-Nagai  Phi = -G M / sqrt(R^2 + (a + sqrt(z^2 + b^2))^2)
  Plummer         Phi = -G M / sqrt(r^2 + a^2)
  Hernquist       Phi = -G M / (r + a)                     (r > 1e-6)
  point mass      Phi = -G M / (r + e) + G M e / (2 (r + e)^2),  e = 1e-6,
                  whose gradient is G M x / (r + e)^3 (the softened form the
                  engines use).
A dropped term, a wrong sign or a lost factor (the disk's (a+S)/S on z)
fails these.
"""
import numpy as np
import pytest

sympy = pytest.importorskip("sympy")
mpmath = pytest.importorskip("mpmath")

G_KPC = 43009.1725
DEFAULT_POT = dict(M_d=5.0, a=3.0, b=0.3, M_h=12.0, a_h=8.0)                # living_galaxy_engine.py:567
HERNQUIST_SMBH = dict(M_d=0.0, a=1.0, b=1.0, M_h=40.0, a_h=15.0, is_hernquist=1,
                      smbh_mass=0.1)                                        # scenarios.py:331
DISK_HERNQUIST_SMBH = dict(M_d=5.0, a=3.0, b=0.3, M_h=40.0, a_h=15.0, is_hernquist=1,
                           smbh_mass=0.1)
NO_POTENTIAL = dict(M_d=0.0, a=1.0, b=1.0, M_h=0.0, a_h=1.0)                # scenarios.py:227


def _exact_acc_fn(pot, G):
    x, y, z = sympy.symbols("x y z", real=True)
    r = sympy.sqrt(x * x + y * y + z * z)
    Gs = sympy.Float(G, 50)
    md, a, b = (sympy.Float(pot[k], 50) for k in ("M_d", "a", "b"))
    mh, ah = sympy.Float(pot["M_h"], 50), sympy.Float(pot["a_h"], 50)
    phi = -Gs * md / sympy.sqrt(x * x + y * y + (a + sympy.sqrt(z * z + b * b)) ** 2)
    if pot.get("is_hernquist", 0) == 1:
        phi += -Gs * mh / (r + ah)
    else:
        phi += -Gs * mh / sympy.sqrt(r * r + ah * ah)
    ms = pot.get("smbh_mass", 0.0)
    if ms > 0:
        e = sympy.Float(1e-6, 50)
        msf = sympy.Float(ms, 50)
        phi += -Gs * msf / (r + e) + Gs * msf * e / (2 * (r + e) ** 2)
    grads = [sympy.lambdify((x, y, z), -sympy.diff(phi, v), "mpmath") for v in (x, y, z)]

    def f(p):
        with mpmath.workdps(50):
            return np.array([[float(g(*(mpmath.mpf(float(c)) for c in q))) for g in grads]
                             for q in p])
    return f


def _points(rng, k=24):
    # disk-like points (kpc): radii 0.05..40, heights up to a few kpc, plus
    # points on the plane and on the axis
    R = np.exp(rng.uniform(np.log(0.05), np.log(40.0), k))
    ph = rng.uniform(0, 2 * np.pi, k)
    z = rng.normal(0, 1.5, k)
    p = np.stack([R * np.cos(ph), R * np.sin(ph), z], 1)
    p[0, 2] = 0.0
    p[1, :2] = 0.0
    return p


@pytest.mark.parametrize("pot", [DEFAULT_POT, HERNQUIST_SMBH, DISK_HERNQUIST_SMBH],
                         ids=["mn_plummer", "hernquist_smbh", "mn_hernquist_smbh"])
def test_analytic_is_exact_gradient(orc, pot):
    p = _points(np.random.default_rng(3))
    got = orc.analytic_add(p, np.zeros_like(p), G_KPC, pot)
    want = _exact_acc_fn(pot, G_KPC)(p)
    err = np.linalg.norm(got - want, axis=1) / np.linalg.norm(want, axis=1)
    assert err.max() < 1e-13, err.max()


def test_mn_with_a_zero_is_plummer_sphere(orc):
    """A Miyamoto-Nagai disk with a = 0 is a Plummer sphere of scale b."""
    p = _points(np.random.default_rng(4))
    disk = orc.analytic_add(p, np.zeros_like(p), 1.0,
                            dict(M_d=2.0, a=0.0, b=1.7, M_h=0.0, a_h=1.0))
    halo = orc.analytic_add(p, np.zeros_like(p), 1.0,
                            dict(M_d=0.0, a=1.0, b=1.0, M_h=2.0, a_h=1.7))
    np.testing.assert_allclose(disk, halo, rtol=1e-14, atol=0)


def test_mn_in_plane_circular_speed(orc):
    """In the plane, R |a_R| = G M R^2 / (R^2 + (a+b)^2)^{3/2} (the speed
    core/ic/gas_disk.py:54 sets up)."""
    R = np.linspace(0.1, 30, 40)
    p = np.stack([R, np.zeros_like(R), np.zeros_like(R)], 1)
    pot = dict(M_d=5.0, a=3.0, b=0.3, M_h=0.0, a_h=8.0)
    a = orc.analytic_add(p, np.zeros_like(p), G_KPC, pot)
    vc2 = G_KPC * 5.0 * R ** 2 / (R ** 2 + 3.3 ** 2) ** 1.5
    np.testing.assert_allclose(-a[:, 0] * R, vc2, rtol=1e-14)
    assert np.all(a[:, 1] == 0) and np.all(a[:, 2] == 0)


def test_massless_potential_leaves_acc_unchanged(orc):
    rng = np.random.default_rng(5)
    p = _points(rng)
    acc = rng.normal(size=p.shape)
    assert np.array_equal(orc.analytic_add(p, acc, G_KPC, NO_POTENTIAL), acc)


def test_self_direct_adds_the_direct_sum(orc):
    rng = np.random.default_rng(6)
    p = rng.normal(size=(300, 3))
    m = rng.uniform(0.5, 1.5, 300)
    acc = rng.normal(size=p.shape)
    got = orc.grav_self_direct(p, m, acc, 1.3, 1e-3)
    assert np.array_equal(got, acc + orc.direct_acc(p, m, 1.3, 1e-3))


def test_grav_ext_is_self_gravity_plus_analytic(orc):
    """The fused form sums ((N^2 + disk) + halo) + point mass; the split form
    adds (disk + halo) + point mass to the N^2 result: same terms, different
    association (a few ulp)."""
    rng = np.random.default_rng(7)
    p = 5 * rng.normal(size=(257, 3))
    m = rng.uniform(1e-5, 2e-5, 257)
    acc0 = np.zeros_like(p)
    for pot in (DEFAULT_POT, DISK_HERNQUIST_SMBH):
        fused = orc.grav_ext_add(p, m, acc0, G_KPC, 0.05 ** 2, pot)
        split = orc.analytic_add(p, orc.grav_self_direct(p, m, acc0, G_KPC, 0.05 ** 2),
                                 G_KPC, pot)
        scale = np.linalg.norm(fused, axis=1)
        assert (np.linalg.norm(fused - split, axis=1) <= 1e-15 * scale).all()

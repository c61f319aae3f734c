"""Pins of the seeded input generators (every benchmark and parity input
rests on them) against the distributions they restate, by Kolmogorov-Smirnov
tests on closed-form CDFs and moment checks; CPU only.

- make_plummer (core/ic/plummer.py:32-58): radii follow the Plummer mass
  profile M(<r)/M = r^3 / (r^2 + a^2)^{3/2}; directions are isotropic;
  the speed fraction q = |v| / v_esc(r), v_esc = sqrt(2 G M / a) (1 + r^2 /
  a^2)^{-1/4}, has density proportional to q^2 (1 - q^2)^{7/2}, so q^2 is
  Beta(3/2, 9/2) distributed.
- make_disk_bulge (core/ic/disk.py:94-111, core/ic/realistic_galaxy.py:
  45-53): disk radii follow the exponential-disk mass profile 1 - (1 + x)
  e^{-x}, x = R / R_d; heights the sech^2 CDF (1 + tanh(z / z_0)) / 2
  restricted to u in [0.001, 0.999]; azimuths are uniform; bulge radii the
  Plummer profile truncated at x_1 < 0.999; the disk dispersions are sigma_R
  = 0.1 v_c, sigma_phi = sigma_R / 2, sigma_z = 0.35 sigma_R about the
  circular speed (core/ic/disk.py:56-70).

Each KS pin is shown to have power: the same test rejects a sample drawn
with a plausible mistake (a wrong exponent, a dropped factor).
"""
import numpy as np
from scipy import stats

P_MIN = 1e-3


def ks_p(sample, cdf):
    return stats.kstest(sample, cdf).pvalue


def plummer_cdf(a, trunc=1.0):
    def f(r):
        r = np.asarray(r, dtype=np.float64)
        return (r / a) ** 3 / (1.0 + (r / a) ** 2) ** 1.5 / trunc
    return f


def test_plummer_radii_directions_speeds():
    from rgsim.ic import make_plummer
    n, a, M, G = 20000, 1.3, 2.0, 1.0
    s = make_plummer(n, total_mass=M, scale_radius=a, seed=11, G=G)
    r = np.linalg.norm(s.pos, axis=1)
    assert ks_p(r, plummer_cdf(a)) > P_MIN
    # isotropy: cos(theta) uniform on [-1, 1], phi uniform on [-pi, pi]
    assert ks_p(s.pos[:, 2] / r, stats.uniform(-1.0, 2.0).cdf) > P_MIN
    phi = np.arctan2(s.pos[:, 1], s.pos[:, 0])
    assert ks_p(phi, stats.uniform(-np.pi, 2 * np.pi).cdf) > P_MIN
    v = np.linalg.norm(s.vel, axis=1)
    v_esc = np.sqrt(2.0 * G * M / a) * (1.0 + (r / a) ** 2) ** -0.25
    q = v / v_esc
    assert q.max() < 1.0
    assert ks_p(q ** 2, stats.beta(1.5, 4.5).cdf) > P_MIN
    vz = s.vel[:, 2] / v
    assert ks_p(vz, stats.uniform(-1.0, 2.0).cdf) > P_MIN
    assert np.allclose(s.mass, M / n) and np.isclose(s.mass.sum(), M)


def test_plummer_pins_have_power():
    """Samples with plausible mistakes fail the same tests."""
    rng = np.random.default_rng(5)
    a, n = 1.3, 20000
    x = rng.random(n)
    wrong_r = a / np.sqrt(x ** (-1.0 / 3.0) - 1.0)        # exponent -1/3 for -2/3
    assert ks_p(wrong_r, plummer_cdf(a)) < P_MIN
    # q drawn by rejection on q^2 (1 - q^2)^{5/2} (exponent 5/2 for 7/2)
    q = []
    while len(q) < n:
        x4, x5 = rng.random(n), rng.random(n)
        q.extend(x4[0.1 * x5 < x4 ** 2 * (1 - x4 ** 2) ** 2.5])
    assert ks_p(np.array(q[:n]) ** 2, stats.beta(1.5, 4.5).cdf) < P_MIN
    # direction cosines of theta drawn uniform on [0, pi] (not area-uniform)
    cos_wrong = np.cos(rng.uniform(0.0, np.pi, n))
    assert ks_p(cos_wrong, stats.uniform(-1.0, 2.0).cdf) < P_MIN


def _disk_cdfs(Rd, z0):
    def f_R(R):
        x = np.asarray(R) / Rd
        return 1.0 - (1.0 + x) * np.exp(-x)

    def f_z(z):
        u = 0.5 * (1.0 + np.tanh(np.asarray(z) / z0))
        return np.clip((u - 0.001) / 0.998, 0.0, 1.0)
    return f_R, f_z


def test_disk_bulge_positions():
    from rgsim.ic import make_disk_bulge
    nd, nb, Rd, z0, ab = 50000, 10000, 3.0, 0.3, 0.5
    s = make_disk_bulge(n_disk=nd, n_bulge=nb, disk_scale=Rd, disk_height=z0,
                        bulge_scale=ab, seed=2)
    disk, bulge = s.pos[:nd], s.pos[nd:]
    f_R, f_z = _disk_cdfs(Rd, z0)
    R = np.linalg.norm(disk[:, :2], axis=1)
    assert ks_p(R, f_R) > P_MIN
    assert ks_p(disk[:, 2], f_z) > P_MIN
    u = 0.5 * (1.0 + np.tanh(disk[:, 2] / z0))
    assert u.min() >= 0.001 - 1e-12 and u.max() <= 0.999 + 1e-12
    phi = np.arctan2(disk[:, 1], disk[:, 0])
    assert ks_p(phi, stats.uniform(-np.pi, 2 * np.pi).cdf) > P_MIN
    rb = np.linalg.norm(bulge, axis=1)
    assert ks_p(rb, plummer_cdf(ab, trunc=0.999)) > P_MIN
    assert rb.max() <= ab / np.sqrt(0.999 ** (-2.0 / 3.0) - 1.0) * (1 + 1e-12)
    assert np.isclose(s.mass[:nd].sum(), 5.0) and np.isclose(s.mass[nd:].sum(), 1.0)


def test_disk_pins_have_power():
    rng = np.random.default_rng(8)
    Rd, z0, n = 3.0, 0.3, 50000
    f_R, f_z = _disk_cdfs(Rd, z0)
    # radii from the surface-density profile e^{-x} instead of the mass CDF
    assert ks_p(rng.exponential(Rd, n), f_R) < P_MIN
    # sech^2 draw without the factor 2 (artanh(u - 0.5) for artanh(2u - 1))
    uz = rng.uniform(0.001, 0.999, n)
    assert ks_p(z0 * np.arctanh(uz - 0.5), f_z) < P_MIN
    # an exponential vertical profile (core/ic/disk.py:103) is not sech^2
    assert ks_p(rng.exponential(z0, n) * rng.choice((-1.0, 1.0), n), f_z) < P_MIN
    # the truncation moves the CDF by 0.1% only, below a KS test's reach at
    # this size; it is pinned by the largest radius, which an untruncated
    # sample of the same size passes
    x = rng.random(10000)
    r_trunc = 0.5 / np.sqrt(0.999 ** (-2.0 / 3.0) - 1.0)
    assert (0.5 / np.sqrt(x ** (-2.0 / 3.0) - 1.0)).max() > r_trunc
    # bulge radii with the scale radius squared in (a^2 for a) fail
    assert ks_p(0.25 / np.sqrt(x ** (-2.0 / 3.0) - 1.0), plummer_cdf(0.5, trunc=0.999)) < P_MIN


def test_disk_velocity_dispersions():
    """Disk velocities about the circular speed: sigma_R = 0.1 v_c, sigma_phi
    = 0.05 v_c, sigma_z = 0.035 v_c, means zero (core/ic/disk.py:56-70); the
    bulge speeds are |N(0, 0.3 v_c)| in isotropic directions (:72-76); the
    total momentum is zero (:79).  v_c is the circular speed of the enclosed
    particle mass with softening 0.05 R_d (:45-52), which bounds v_phi."""
    from rgsim.ic import make_disk_bulge
    nd, nb, Rd = 50000, 10000, 3.0
    G = 1.0
    s = make_disk_bulge(n_disk=nd, n_bulge=nb, disk_scale=Rd, seed=4, G=G)
    pos, vel, mass = s.pos, s.vel, s.mass
    r = np.linalg.norm(pos, axis=1)
    o = np.argsort(r)
    m_enc = np.empty_like(mass)
    m_enc[o] = np.cumsum(mass[o])
    v_c = np.sqrt(G * m_enc / np.sqrt(r ** 2 + (0.05 * Rd) ** 2))
    d, vd, vc = pos[:nd], vel[:nd], v_c[:nd]
    R = np.linalg.norm(d[:, :2], axis=1)
    vR = (d[:, 0] * vd[:, 0] + d[:, 1] * vd[:, 1]) / R
    vphi = (d[:, 0] * vd[:, 1] - d[:, 1] * vd[:, 0]) / R
    # the net-momentum removal shifts every velocity by one small vector
    # (|shift| / v_c ~ 1e-3 here), well inside these bounds
    for name, x, sig in (("R", vR / vc, 0.1), ("phi", vphi / vc - 1.0, 0.05),
                         ("z", vd[:, 2] / vc, 0.035)):
        assert abs(np.std(x) / sig - 1.0) < 0.03, (name, np.std(x))
        assert abs(np.mean(x)) < 0.05 * sig + 3e-3, (name, np.mean(x))
    # standardised radial deviations are normal (sub-sample keeps the tiny
    # momentum shift below the test's resolution)
    assert ks_p((vR / (0.1 * vc))[:5000], stats.norm.cdf) > P_MIN
    vb = np.linalg.norm(vel[nd:], axis=1) / (0.3 * v_c[nd:])
    assert ks_p(vb[:5000], stats.halfnorm.cdf) > P_MIN
    assert np.linalg.norm((mass[:, None] * vel).sum(axis=0)) < 1e-9 * np.abs(mass[:, None] * vel).sum()

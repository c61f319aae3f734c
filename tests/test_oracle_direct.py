"""Pins the oracle's direct sum, energies and KDK (gravity.py, leapfrog.py,
engine.py:108-124) with closed forms and invariants."""
import numpy as np
import pytest


def test_two_body_closed_form(orc):
    for sep, eps2, m1, m2 in [(1.0, 0.01, 1.0, 2.0), (0.3, 1e-4, 3.0, 0.5), (5.0, 0.0025, 1, 1)]:
        pos = np.array([[0.0, 0, 0], [sep, 0, 0]])
        mass = np.array([m1, m2])
        G = 43009.1725
        acc = orc.direct_acc(pos, mass, G, eps2)
        np.testing.assert_allclose(acc, orc.two_body_acc(m1, m2, sep, G, eps2), rtol=1e-14)
        pe = orc.potential_energy(pos, mass, G, eps2)
        assert pe == pytest.approx(orc.two_body_pe(m1, m2, sep, G, eps2), rel=1e-14)


def test_self_term_is_zero_single_particle(orc):
    acc = orc.direct_acc(np.array([[1.0, 2.0, 3.0]]), np.array([5.0]), 1.0, 0.01)
    assert np.all(acc == 0.0)


def test_invariants(orc):
    rng = np.random.default_rng(2)
    n = 500
    pos = rng.normal(size=(n, 3))
    mass = rng.uniform(0.1, 1, n)
    acc = orc.direct_acc(pos, mass, 1.0, 0.01)
    # momentum: sum m a = 0 (pairwise antisymmetry)
    f = (mass[:, None] * acc).sum(0)
    assert np.abs(f).max() < 1e-12 * np.abs(mass[:, None] * acc).sum()
    # translation invariance
    acc2 = orc.direct_acc(pos + 3.0, mass, 1.0, 0.01)
    np.testing.assert_allclose(acc2, acc, rtol=1e-9, atol=1e-10)
    # an independently ordered numpy evaluation agrees to rounding
    np.testing.assert_allclose(orc.direct_acc_np(pos, mass, 1.0, 0.01), acc, rtol=1e-11,
                               atol=1e-12)


def test_symmetric_cube_centre_force_zero(orc):
    c = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], float)
    pos = np.vstack(([0.0, 0, 0], c))
    acc = orc.direct_acc(pos, np.ones(9), 1.0, 0.01)
    assert np.abs(acc[0]).max() < 1e-15


def test_kinetic_energy(orc):
    rng = np.random.default_rng(4)
    vel = rng.normal(size=(100, 3))
    mass = rng.uniform(0.1, 1, 100)
    assert orc.kinetic_energy(vel, mass) == pytest.approx(
        0.5 * np.einsum("i,ij,ij->", mass, vel, vel), rel=1e-14)


def test_kdk_two_body_circular_orbit(orc):
    # equal masses m on a circle of separation a: v = sqrt(G m / (2 a)) with
    # softening folded into the force law; check radius stays ~constant.
    G, m, a, eps2 = 1.0, 1.0, 1.0, 1e-8
    f = a / (a * a + eps2) ** 1.5  # |a1| = G m f
    v = np.sqrt(G * m * f * a / 2)
    pos = np.array([[-a / 2, 0, 0], [a / 2, 0, 0]])
    vel = np.array([[0, -v, 0], [0, v, 0]])
    mass = np.array([m, m])
    acc = orc.direct_acc(pos, mass, G, eps2)
    e0 = orc.kinetic_energy(vel, mass) + orc.potential_energy(pos, mass, G, eps2)
    for _ in range(2000):
        pos, vel, acc = orc.kdk_step(pos, vel, acc, mass, 1e-3, G, eps2)
    sep = np.linalg.norm(pos[1] - pos[0])
    assert sep == pytest.approx(a, rel=1e-5)
    e1 = orc.kinetic_energy(vel, mass) + orc.potential_energy(pos, mass, G, eps2)
    assert abs((e1 - e0) / e0) < 1e-8


def test_kdk_time_reversible(orc):
    from rgsim.ic import make_plummer
    s = make_plummer(200, total_mass=1.0, seed=3, G=1.0)
    pos, vel, mass = s.pos, s.vel, s.mass
    acc = orc.direct_acc(pos, mass, 1.0, 1e-4)
    p, v, a = pos, vel, acc
    for _ in range(50):
        p, v, a = orc.kdk_step(p, v, a, mass, 1e-3, 1.0, 1e-4)
    for _ in range(50):
        p, v, a = orc.kdk_step(p, v, a, mass, -1e-3, 1.0, 1e-4)
    np.testing.assert_allclose(p, pos, atol=1e-10)
    np.testing.assert_allclose(v, vel, atol=1e-9)


def test_verify_phase1_thresholds_cpu(orc):
    """verify_phase1.py:30-51 at reduced length: Plummer N=3000, M=10, a=1,
    seed 42, eps 0.05, dt 3e-5; drift and virial thresholds."""
    from rgsim.ic import make_plummer
    from rgsim.units import G
    s = make_plummer(3000, total_mass=10.0, scale_radius=1.0, seed=42)
    eps2 = 0.05 ** 2
    pos, vel, mass = s.pos, s.vel, s.mass
    acc = orc.direct_acc(pos, mass, G, eps2)
    e0 = orc.kinetic_energy(vel, mass) + orc.potential_energy(pos, mass, G, eps2)
    for _ in range(20):
        pos, vel, acc = orc.kdk_step(pos, vel, acc, mass, 3e-5, G, eps2)
    ke = orc.kinetic_energy(vel, mass)
    pe = orc.potential_energy(pos, mass, G, eps2)
    assert abs((ke + pe - e0) / e0) < 1e-3
    assert 0.8 < -2 * ke / pe < 1.2

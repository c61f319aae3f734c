"""GPU parity of the self-gravity entry points the gas-disk and living-galaxy
engines call (core/gas_disk_engine.py:20-89, core/living_galaxy_engine.py:
101-150, and their tree path :1016-1018 / :218-219), against the oracle.

Bars: the accumulate forms of the direct sum and the field add are bit-exact
in FP64 (same j loop, same adds).  The analytic potential goes through
pow(x, -1.5), which neither libm rounds correctly: a one-ulp difference there
can flip the rounding of each later add, so results that include it are held
to 8 ulp (8 * 2^-53) of the summed magnitudes of the terms, per particle --
the forward error bound of the four-term sum.  The tree add is bitwise
compute-then-add in every precision."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

G_KPC = 43009.1725
POTS = {
    "mn_plummer": dict(M_d=5.0, a=3.0, b=0.3, M_h=12.0, a_h=8.0),            # :567 default
    "hernquist_smbh": dict(M_d=0.0, a=1.0, b=1.0, M_h=40.0, a_h=15.0, is_hernquist=1,
                           smbh_mass=0.1),                                   # scenarios.py:331
    "mn_hernquist_smbh": dict(M_d=5.0, a=3.0, b=0.3, M_h=40.0, a_h=15.0, is_hernquist=1,
                              smbh_mass=0.1),
    "none": dict(M_d=0.0, a=1.0, b=1.0, M_h=0.0, a_h=1.0),                  # scenarios.py:227
}


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def gas_like(n, seed):
    """Disk + bulge positions in kpc with gas-particle masses (1e10 Msun units)."""
    from rgsim.ic import make_disk_bulge
    rng = np.random.default_rng(seed)
    if n < 12:  # too few for both components
        return (rng.normal(scale=3.0, size=(n, 3)), np.full(n, 6.0 / max(n, 1)),
                rng.normal(scale=1e3, size=(n, 3)))
    nb = n // 6
    d = make_disk_bulge(n_disk=n - nb, n_bulge=nb, seed=seed)
    return d.pos, d.mass, rng.normal(scale=1e3, size=(n, 3))


ULP8 = 8 * 2.0 ** -53


def check_close(got, want, *terms):
    """|got - want| <= 8 ulp of sum |term| per particle; reports the worst ratio."""
    scale = sum(np.linalg.norm(t, axis=1) for t in terms)
    ratio = np.linalg.norm(got - want, axis=1) / (ULP8 * np.maximum(scale, np.finfo(float).tiny))
    print(f"  worst error / (8 ulp of the terms) = {ratio.max():.3f}")
    assert ratio.max() <= 1.0


@pytest.mark.parametrize("n", [1, 2, 257, 3000])
def test_grav_self_direct_bitexact(orc, n):
    from rgsim.living_galaxy_engine import grav_self_direct
    print(f"grav_self_direct n={n}", flush=True)
    pos, mass, acc0 = gas_like(n, 11)
    acc = dev(acc0)
    grav_self_direct(dev(pos), dev(mass), acc, n, 0.05 ** 2)
    want = orc.grav_self_direct(pos, mass, acc0, G_KPC, 0.05 ** 2)
    assert np.array_equal(acc.cpu().numpy(), want)


@pytest.mark.parametrize("precision,rel", [("fp32c", 1e-6), ("fp32", 5e-3)])
def test_grav_self_direct_reduced_precision(orc, precision, rel):
    """acc0 + the reduced-precision sum, held to that sum's own error bar
    (relative to the self-gravity magnitude)."""
    from rgsim.solvers import gravity
    pos, mass, acc0 = gas_like(3000, 12)
    acc = dev(acc0)
    gravity.add_acc_direct(dev(pos), dev(mass), acc, 3000, G_KPC, 0.05 ** 2, precision)
    selfg = orc.direct_acc(pos, mass, G_KPC, 0.05 ** 2)
    err = np.linalg.norm(acc.cpu().numpy() - (acc0 + selfg), axis=1)
    assert (err <= rel * np.linalg.norm(selfg, axis=1) + 1e-12 * np.abs(acc0).max()).all()


@pytest.mark.parametrize("pot", list(POTS))
@pytest.mark.parametrize("n", [1, 257, 3000])
def test_add_grav_ext_vs_oracle(orc, pot, n):
    from rgsim.gas_disk_engine import add_grav_ext
    print(f"add_grav_ext {pot} n={n}", flush=True)
    pos, mass, acc0 = gas_like(n, 13)
    p = POTS[pot]
    acc = dev(acc0)
    add_grav_ext(dev(pos), dev(mass), acc, n, 0.05 ** 2, p["M_d"], p["a"], p["b"], p["M_h"],
                 p["a_h"], p.get("is_hernquist", 0), p.get("smbh_mass", 0.0))
    want = orc.grav_ext_add(pos, mass, acc0, G_KPC, 0.05 ** 2, p)
    got = acc.cpu().numpy()
    zero = np.zeros_like(acc0)
    check_close(got, want, acc0, orc.direct_acc(pos, mass, G_KPC, 0.05 ** 2),
                orc.analytic_add(pos, zero, G_KPC, p))
    if pot == "none":  # massless potential: the self-gravity sum alone, bitwise
        assert np.array_equal(got, orc.grav_self_direct(pos, mass, acc0, G_KPC, 0.05 ** 2))


@pytest.mark.parametrize("pot", list(POTS))
def test_analytic_vs_oracle(orc, pot):
    from rgsim.gas_disk_engine import add_analytic
    from rgsim.living_galaxy_engine import grav_analytic
    pos, _, acc0 = gas_like(5000, 14)
    pos[0] = 0.0          # the centre: the halo's and point mass's safe radii
    pos[1, :2] = 0.0      # on the axis
    pos[2, 2] = 0.0       # in the plane
    p = POTS[pot]
    want = orc.analytic_add(pos, acc0, G_KPC, p)
    ext = orc.analytic_add(pos, np.zeros_like(acc0), G_KPC, p)
    args = (p["M_d"], p["a"], p["b"], p["M_h"], p["a_h"], p.get("is_hernquist", 0),
            p.get("smbh_mass", 0.0))
    for fn in (add_analytic, grav_analytic):
        acc = dev(acc0)
        fn(dev(pos), acc, 5000, *args)
        got = acc.cpu().numpy()
        check_close(got, want, acc0, ext)
    if pot == "none":
        assert np.array_equal(got, acc0)


def test_add_field_bitexact():
    from rgsim.gas_disk_engine import add_field
    from rgsim.living_galaxy_engine import add_acc
    rng = np.random.default_rng(15)
    a0, src = rng.normal(size=(1001, 3)), rng.normal(size=(1001, 3))
    for fn in (add_field, add_acc):
        acc = dev(a0)
        fn(acc, dev(src), 1001)
        assert np.array_equal(acc.cpu().numpy(), a0 + src)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("n", [256, 6000])
def test_bh_compute_add_is_compute_then_add(precision, n):
    """living_galaxy_engine.py:1017-1018: compute into acc_g, then acc += acc_g."""
    from rgsim.solvers.barnes_hut import BarnesHut
    print(f"bh compute_add {precision} n={n}", flush=True)
    pos, mass, acc0 = gas_like(n, 16)
    bh = BarnesHut(n, theta=0.6, softening=0.05, precision=precision)
    p, m = dev(pos), dev(mass)
    acc_g = torch.empty_like(p)
    bh.compute(p, m, acc_g)
    acc = dev(acc0)
    bh.compute_add(p, m, acc)
    assert np.array_equal(acc.cpu().numpy(), acc0 + acc_g.cpu().numpy())
    st = bh.stats()
    assert st["n"] == n and st["visits_total"] > 0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("n", [256, 6000])
def test_bh_step_is_kick_drift_then_compute_kick(precision, n):
    """rg_bh_step (the Engine's tree step, core/engine.py:108-116) equals
    leapfrog.kick_drift followed by compute_kick bit for bit over several
    steps: positions, velocities, forces and the walk statistics, including
    the non-finite count of the drifted positions."""
    from rgsim.integrators import leapfrog
    from rgsim.solvers.barnes_hut import BarnesHut
    print(f"bh step {precision} n={n}", flush=True)
    pos, mass, vel = gas_like(n, 18)
    vel *= 1e-3
    m = dev(mass)
    runs = []
    for fused in (False, True):
        bh = BarnesHut(n, theta=0.6, softening=0.05, precision=precision)
        p, v = dev(pos), dev(vel)
        a = torch.empty_like(p)
        bh.compute(p, m, a)
        stats = []
        for _ in range(3):
            if fused:
                bh.step(p, v, m, a, 5e-5, 1e-4)
            else:
                leapfrog.kick_drift(p, v, a, n, 5e-5, 1e-4)
                bh.compute_kick(p, m, a, v, 5e-5)
            st = bh.stats()
            stats.append((st["visits_total"], st["visits_max"], st["nonfinite"]))
        runs.append((p.cpu().numpy(), v.cpu().numpy(), a.cpu().numpy(), stats))
    for x, y in zip(runs[0][:3], runs[1][:3]):
        assert np.array_equal(x, y)
    assert runs[0][3] == runs[1][3]
    # a non-finite velocity drifts into a non-finite position: both paths count it
    counts = []
    for fused in (False, True):
        bh = BarnesHut(n, theta=0.6, softening=0.05, precision=precision)
        p, v = dev(pos), dev(vel)
        a = torch.zeros_like(p)
        v[3, 1] = float("inf")
        if fused:
            bh.step(p, v, m, a, 5e-5, 1e-4)
        else:
            leapfrog.kick_drift(p, v, a, n, 5e-5, 1e-4)
            bh.compute_kick(p, m, a, v, 5e-5)
        counts.append(bh.stats()["nonfinite"])
    assert counts[0] == counts[1] == 1


def test_engine_entry_points_reject_bad_arguments():
    from rgsim import _lib
    from rgsim.solvers import gravity
    pos, mass, _ = gas_like(64, 17)
    p, m = dev(pos), dev(mass)
    acc = torch.zeros_like(p)
    L = _lib.load()
    st = L.rg_direct_ext_acc_add(p.data_ptr(), m.data_ptr(), acc.data_ptr(), 64, 1.0, 1e-4,
                                 None, 0, _lib.stream_ptr(p))
    assert st == _lib.RG_ERR_ARG and b"potential" in L.rg_last_error()
    with pytest.raises(RuntimeError, match="overlap"):
        gravity.add_field(acc, acc, 64)
    with pytest.raises(RuntimeError, match="alias"):
        gravity.add_acc_analytic(p, p, 64, 1.0, gravity.potential_struct(1, 1, 1, 1, 1, 0, 0))
    with pytest.raises(RuntimeError, match="eps2"):
        gravity.add_acc_direct(p, m, acc, 64, 1.0, 0.0)
    gravity.add_acc_direct(p[:0], m[:0], acc[:0], 0, 1.0, 1e-4)  # n = 0 is a no-op


@pytest.mark.parametrize("flavor", ["living_galaxy", "gas_disk"])
@pytest.mark.parametrize("mode", ["direct", "bh"])
def test_engine_gravity_half(orc, flavor, mode):
    """The engines' whole gravity half (living_galaxy_engine.py:727-738 and
    :1015-1026; gas_disk_engine.py:169-173 and :216-230) through
    LivingGalaxyGravity / GasDiskGravity, as the particle count changes
    (3000 -> 2000 -> 200): the direct path matches the oracle's call sequence
    in each engine's add order, the tree path is bitwise BarnesHut.compute
    into acc_g + add + the analytic add, a count change rebuilds the tree,
    and below 256 particles the direct sum runs."""
    from rgsim.gas_disk_engine import GasDiskGravity
    from rgsim.living_galaxy_engine import LivingGalaxyGravity
    from rgsim.solvers import gravity
    from rgsim.solvers.barnes_hut import BarnesHut
    p = POTS["mn_plummer"]
    cls = LivingGalaxyGravity if flavor == "living_galaxy" else GasDiskGravity
    sg = cls(0.05, gravity_mode=mode, theta=0.5, pot=p)
    eps2 = 0.05 ** 2
    for k, n in enumerate((3000, 2000, 200)):
        print(f"engine gravity {flavor} {mode} n={n}", flush=True)
        pos, mass, acc0 = gas_like(n, 20 + k)
        acc = dev(acc0)
        sg.add_forces(dev(pos), dev(mass), acc, n)
        got = acc.cpu().numpy()
        tree = mode == "bh" and n >= 256
        assert (sg.tree is not None) == tree and sg.n == n
        zero = np.zeros_like(acc0)
        if tree:
            acc_g = torch.empty((n, 3), dtype=torch.float64, device="cuda")
            BarnesHut(n, theta=0.5, softening=0.05).compute(dev(pos), dev(mass), acc_g)
            want = dev(acc0)
            gravity.add_field(want, acc_g, n)
            gravity.add_acc_analytic(dev(pos), want, n, G_KPC, sg.pot)
            assert np.array_equal(got, want.cpu().numpy())
        elif flavor == "gas_disk":
            check_close(got, orc.grav_ext_add(pos, mass, acc0, G_KPC, eps2, p), acc0,
                        orc.direct_acc(pos, mass, G_KPC, eps2), orc.analytic_add(pos, zero, G_KPC, p))
        else:
            selfg = orc.grav_self_direct(pos, mass, acc0, G_KPC, eps2)
            check_close(got, orc.analytic_add(pos, selfg, G_KPC, p), acc0,
                        orc.direct_acc(pos, mass, G_KPC, eps2), orc.analytic_add(pos, zero, G_KPC, p))
    assert sg.rebuilds == (2 if mode == "bh" else 0)

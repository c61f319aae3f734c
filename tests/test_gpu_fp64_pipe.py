"""The pipelined FP64 direct sum (several waves compute one 64-particle
block's pair terms, three waves add them in j order): the same doubles as
the oracle's restatement of gravity.py:13-26 for every pipeline width and
ragged block, through the store, add (living_galaxy_engine.py:101-113),
analytic-add (gas_disk_engine.py:20-51) and kick (leapfrog.py:11-15)
epilogues."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

G_KPC = 43009.1725


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def plummer(n, seed=0, M=10.0):
    from rgsim.ic import make_plummer
    return make_plummer(n, total_mass=M, scale_radius=1.0, seed=seed, G=G_KPC)


@pytest.fixture
def pipe():
    from rgsim import _lib

    def set_pipe(v):
        _lib.set_param("fp64_pipe", v)
    yield set_pipe
    _lib.set_param("fp64_pipe", 0)


@pytest.mark.parametrize("waves", [0, 1, 4, 8, 16])
@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 1000, 3001])
def test_pipe_bitexact_vs_oracle(orc, pipe, waves, n):
    from rgsim.solvers import gravity
    print(f"fp64 pipe waves={waves} n={n}", flush=True)
    s = plummer(n, seed=40 + n)
    ref = orc.direct_acc(s.pos, s.mass, G_KPC, 0.05 ** 2)
    pipe(waves)
    p, m = dev(s.pos), dev(s.mass)
    acc = torch.empty_like(p)
    gravity.compute_acc_direct(p, m, acc, n, G_KPC, 0.05 ** 2, "fp64")
    got = acc.cpu().numpy()
    assert np.array_equal(got, ref), np.abs(got - ref).max()


@pytest.mark.parametrize("n", [65536, 81921, 120000, 131072])
def test_default_choice_bitexact_large_n(orc, n):
    """At these sizes the library picks the 8-wave pipeline (fewer than two
    one-wave waves per SIMD, or an uneven spread: 65,536, 81,921) or the
    one-wave kernel (120,000, 131,072) by SIMD occupancy; either gives the
    oracle's doubles."""
    from rgsim.solvers import gravity
    print(f"fp64 default choice n={n}", flush=True)
    s = plummer(n, seed=3)
    ref = orc.direct_acc(s.pos, s.mass, G_KPC, 0.05 ** 2)
    p, m = dev(s.pos), dev(s.mass)
    acc = torch.empty_like(p)
    gravity.compute_acc_direct(p, m, acc, n, G_KPC, 0.05 ** 2, "fp64")
    assert np.array_equal(acc.cpu().numpy(), ref)


@pytest.mark.parametrize("waves", [4, 8, 16])
def test_pipe_epilogues_bitexact(orc, pipe, waves):
    from rgsim.solvers import gravity
    n, dth = 2113, 5e-5
    print(f"fp64 pipe epilogues waves={waves}", flush=True)
    s = plummer(n, seed=9)
    rng = np.random.default_rng(9)
    acc0 = rng.normal(size=(n, 3)) * 1e3
    pipe(waves)
    p, m = dev(s.pos), dev(s.mass)
    eps2 = 0.05 ** 2

    acc = dev(acc0)
    gravity.add_acc_direct(p, m, acc, n, G_KPC, eps2, "fp64")
    assert np.array_equal(acc.cpu().numpy(), orc.grav_self_direct(s.pos, s.mass, acc0, G_KPC, eps2))

    # Hernquist halo + point mass, no disk: no pow() in the analytic terms,
    # so the device and host doubles agree bit for bit
    pot = {"M_d": 0.0, "a": 1.0, "b": 1.0, "M_h": 40.0, "a_h": 15.0, "is_hernquist": 1,
           "smbh_mass": 0.1}
    acc = dev(acc0)
    gravity.add_acc_direct_ext(p, m, acc, n, G_KPC, eps2,
                               gravity.potential_struct(**{k: pot[k] for k in
                                                           ("M_d", "a", "b", "M_h", "a_h",
                                                            "is_hernquist", "smbh_mass")}),
                               "fp64")
    assert np.array_equal(acc.cpu().numpy(), orc.grav_ext_add(s.pos, s.mass, acc0, G_KPC, eps2, pot))

    v = dev(s.vel)
    acc = torch.empty_like(p)
    gravity.compute_acc_direct_kick(p, m, acc, v, n, G_KPC, eps2, dth, "fp64")
    ref = orc.direct_acc(s.pos, s.mass, G_KPC, eps2)
    assert np.array_equal(acc.cpu().numpy(), ref)
    assert np.array_equal(v.cpu().numpy(), orc.kick(s.vel, ref, dth))


def test_pipe_param_range():
    from rgsim import _lib
    with pytest.raises(RuntimeError, match="fp64_pipe"):
        _lib.set_param("fp64_pipe", 2)

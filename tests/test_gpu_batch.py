"""Batching form on the GPU: a batched force is bitwise the per-system
direct kernel run unsplit (and, in FP64, the oracle's gravity.py:13-26
restatement);
batched KDK steps, hipGraph replays and every launch shape agree bitwise;
per-system FP64 energies match the oracle."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture
def unsplit():
    """The per-system direct sums with one j-slice: the batched kernel keeps
    that summation order."""
    from rgsim import _lib
    _lib.set_param("direct_split", 1)
    yield
    _lib.set_param("direct_split", 0)


def small_systems(sizes=(1, 2, 37, 256, 300, 1000, 129, 2500), G=1.0):
    from rgsim.ic import make_plummer
    out = []
    for k, n in enumerate(sizes):
        a = 0.5 + 0.25 * k
        s = make_plummer(n, total_mass=1.0, scale_radius=a, seed=100 + k, G=G)
        out.append({"pos": s.pos, "vel": s.vel, "mass": s.mass, "G": G * (1 + 0.1 * k),
                    "eps": 0.02 * a, "dt": 1e-3 * a ** 1.5})
    return out


@pytest.mark.parametrize("precision", ["fp32c", "fp64"])
@pytest.mark.parametrize("tile", [64, 128, 256, 512, 1024])
def test_batch_acc_equals_per_system_direct(orc, unsplit, precision, tile):
    from rgsim.batch import Ensemble
    from rgsim.solvers import gravity
    print(f"batch acc {precision} tile={tile}", flush=True)
    systems = small_systems()
    ens = Ensemble(systems, precision=precision)
    ens.set_shape(tile, 7)
    ens.compute_acc()
    got = ens.acc.cpu().numpy()
    for k, s in enumerate(systems):
        a, b = ens.offsets[k], ens.offsets[k + 1]
        p = torch.from_numpy(s["pos"]).cuda()
        m = torch.from_numpy(s["mass"]).cuda()
        acc = torch.empty_like(p)
        gravity.compute_acc_direct(p, m, acc, len(s["mass"]), s["G"], s["eps"] ** 2, precision)
        assert np.array_equal(got[a:b], acc.cpu().numpy()), k
        if precision == "fp64":
            assert np.array_equal(got[a:b], orc.direct_acc(s["pos"], s["mass"], s["G"],
                                                           s["eps"] ** 2)), k


@pytest.mark.parametrize("precision", ["fp32c", "fp64"])
def test_batch_steps_equal_per_system_steps(unsplit, precision):
    from rgsim.batch import Ensemble
    systems = small_systems()
    e1 = Ensemble(systems, precision=precision)
    e2 = Ensemble(systems, precision=precision)
    e1.step(7)
    e2.per_system_step(7)
    assert torch.equal(e1.pos, e2.pos) and torch.equal(e1.vel, e2.vel)
    assert torch.equal(e1.acc, e2.acc)


def test_graph_replays_match_eager():
    from rgsim.batch import Ensemble
    systems = small_systems()
    eg = Ensemble(systems)
    ref = Ensemble(systems)
    eg.capture(5)
    for r in range(3):
        eg.replay()
        ref.step(5)
        torch.cuda.synchronize()
        # eager work on the device between replays
        _ = (eg.pos * 2).sum().item()
        assert torch.equal(eg.pos, ref.pos), r
        assert torch.equal(eg.vel, ref.vel), r


def test_batch_energy_vs_oracle(orc):
    from rgsim.batch import Ensemble
    systems = small_systems()
    ens = Ensemble(systems, precision="fp64")
    e = ens.energies()
    for k, s in enumerate(systems):
        ke = orc.kinetic_energy(s["vel"], s["mass"])
        pe = orc.potential_energy(s["pos"], s["mass"], s["G"], s["eps"] ** 2)
        assert e[k, 0] == pytest.approx(ke, rel=1e-13)
        assert e[k, 1] == pytest.approx(pe, rel=1e-12, abs=1e-300)
        np.testing.assert_allclose(e[k, 2:], (s["mass"][:, None] * s["vel"]).sum(0),
                                   rtol=1e-10, atol=1e-12)


def test_ensemble_energy_drift_small():
    from rgsim.batch import Ensemble, plummer_ensemble
    systems = plummer_ensemble(nsys=16, n_min=256, n_max=2048, master_seed=7)
    ens = Ensemble(systems, precision="fp32c")
    e0 = ens.energies()
    ens.capture(20)
    for _ in range(5):
        ens.replay()
    e1 = ens.energies()
    tot0, tot1 = e0[:, 0] + e0[:, 1], e1[:, 0] + e1[:, 1]
    drift = np.abs((tot1 - tot0) / tot0)
    print(f"max per-system drift after 100 steps: {drift.max():.2e}")
    assert drift.max() < 1e-4


def test_autotuner_picks_a_valid_shape():
    from rgsim.autotune import tune_batch
    from rgsim.batch import Ensemble
    ens = Ensemble(small_systems())
    best = tune_batch(ens, grids_per_cu=(1, 2), reps=1)
    assert best["tile"] in (64, 128, 256, 512, 1024) and best["ms"] > 0
    assert len(best["trials"]) == 10
    ens.compute_acc()
    assert torch.isfinite(ens.acc).all()


def test_ensemble_full_config_drift():
    """The batching form at full size: 512 Plummer systems, N log-uniform in
    [256, 16384] (master seed 2026), 100 FP32C KDK steps in one hipGraph;
    every system's FP64 energy drift stays small and matches the drift of
    the same steps in FP64 to within 1e-6 (the FP32C error is below the
    integration error)."""
    from rgsim.batch import Ensemble, plummer_ensemble
    print("ensemble 512 systems, 100 graph steps", flush=True)
    systems = plummer_ensemble(nsys=512, n_min=256, n_max=16384, master_seed=2026)
    out = {}
    for prec in ("fp32c", "fp64"):
        ens = Ensemble(systems, precision=prec)
        e0 = ens.energies()
        ens.capture(100)
        ens.replay()
        e1 = ens.energies()
        t0, t1 = e0[:, 0] + e0[:, 1], e1[:, 0] + e1[:, 1]
        out[prec] = np.abs((t1 - t0) / t0)
        assert ens.ntot > 1_500_000
    print(f"  drift max fp32c {out['fp32c'].max():.2e} fp64 {out['fp64'].max():.2e}", flush=True)
    assert out["fp32c"].max() < 1e-5
    assert np.abs(out["fp32c"] - out["fp64"]).max() < 1e-6

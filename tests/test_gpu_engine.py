"""The drop-in Engine (core/engine.py:32-124) on the GPU: the reference's
acceptance scripts restated (verify_phase1.py:30-64, verify_bh_galaxy.py:17-55)
and bit-exact KDK against the oracle."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_verify_phase1_plummer():
    """verify_phase1.py:30-51: N=3000, M=10, a=1, seed 42, eps 0.05,
    dt 3e-5, 1000 steps; |dE/E| < 1e-3 and 0.8 < -2KE/PE < 1.2."""
    from rgsim.engine import Engine
    from rgsim.ic import make_plummer
    state = make_plummer(n=3000, total_mass=10.0, scale_radius=1.0, seed=42)
    eng = Engine(softening=0.05)
    eng.load_state(state)
    assert eng._bh is None
    ke0, pe0, e0 = eng.energies()
    for _ in range(1000):
        eng.step(3e-5)
    ke1, pe1, e1 = eng.energies()
    drift = abs((e1 - e0) / e0)
    print(f"phase1 drift={drift:.2e} virial={-2 * ke1 / pe1:.3f}")
    assert drift < 1e-3
    assert 0.8 < -2.0 * ke1 / pe1 < 1.2
    out = eng.to_state()
    assert out.step == 1000 and out.seed == 42 and np.array_equal(out.ids, state.ids)


def test_engine_kdk_bitexact_vs_oracle(orc):
    from rgsim.engine import Engine
    from rgsim.ic import make_plummer
    from rgsim.units import G
    s = make_plummer(n=600, total_mass=10.0, seed=5)
    eng = Engine(softening=0.05)
    eng.load_state(s)
    pos, vel, mass = s.pos, s.vel, s.mass
    acc = orc.direct_acc(pos, mass, G, 0.05 ** 2)
    for _ in range(10):
        eng.step(1e-4)
        pos, vel, acc = orc.kdk_step(pos, vel, acc, mass, 1e-4, G, 0.05 ** 2)
    out = eng.to_state()
    assert np.array_equal(out.pos, pos) and np.array_equal(out.vel, vel)
    assert eng.step_count == 10 and eng.time == pytest.approx(1e-3)


@pytest.mark.parametrize("precision,tol", [("fp64", 0.0), ("fp32c", 1e-7), ("fp32", 1e-3)])
def test_plummer4096_kdk_100_steps(orc, precision, tol):
    """The small plain-method workload: Plummer N=4096, M=1, G=1, a=1,
    seed 42, eps 0.01, KDK dt 1e-3 for 100 steps (engine.py:108-116 over
    gravity.py:13-26).  FP64 is bit-exact against the oracle's trajectory;
    FP32C and FP32 stay within tol of it (relative to the position scale);
    the FP64 energy drift stays under verify_phase1.py:50's 1e-3."""
    from rgsim.engine import Engine
    from rgsim.ic import make_plummer
    print(f"plummer4096 kdk {precision}", flush=True)
    s = make_plummer(n=4096, total_mass=1.0, scale_radius=1.0, seed=42, G=1.0)
    eng = Engine(softening=0.01, precision=precision, G=1.0)
    eng.load_state(s)
    _, _, e0 = eng.energies()
    for _ in range(100):
        eng.step(1e-3)
    _, _, e1 = eng.energies()
    out = eng.to_state()
    pos, vel, mass = s.pos, s.vel, s.mass
    acc = orc.direct_acc(pos, mass, 1.0, 1e-4)
    for _ in range(100):
        pos, vel, acc = orc.kdk_step(pos, vel, acc, mass, 1e-3, 1.0, 1e-4)
    drift = abs((e1 - e0) / e0)
    print(f"plummer4096 {precision}: |dE/E| {drift:.2e}, "
          f"max |dx| {np.abs(out.pos - pos).max():.2e}")
    if precision == "fp64":
        assert np.array_equal(out.pos, pos) and np.array_equal(out.vel, vel)
    else:
        assert np.abs(out.pos - pos).max() <= tol * np.abs(pos).max()
    assert drift < 1e-3


def test_verify_bh_galaxy_restated(orc):
    """verify_bh_galaxy.py:17-55 with the disk+bulge generator in place of
    the SPH galaxy scenario: mean relative force error < 5% at theta 0.6,
    KE finite and positive after 20 steps, _bh None in direct mode."""
    from rgsim.engine import Engine
    from rgsim.ic import make_disk_bulge
    s = make_disk_bulge(n_disk=6000, n_bulge=2000, seed=1)
    eng_d = Engine(softening=0.05, gravity_mode="direct")
    eng_b = Engine(softening=0.05, gravity_mode="bh", theta=0.6)
    eng_d.load_state(s)
    eng_b.load_state(s)
    assert eng_d._bh is None and eng_b._bh is not None
    a_d = eng_d.acc.cpu().numpy()
    a_b = eng_b.acc.cpu().numpy()
    rel = np.linalg.norm(a_b - a_d, axis=1) / np.maximum(np.linalg.norm(a_d, axis=1), 1e-30)
    print(f"bh galaxy mean rel err {rel.mean():.3%}")
    assert rel.mean() < 0.05
    for _ in range(20):
        eng_d.step(1e-4)
        eng_b.step(1e-4)
    ke_b, _, _ = eng_b.energies()
    assert np.isfinite(ke_b) and ke_b > 0


def test_bh_small_n_falls_back_to_direct():
    from rgsim.engine import Engine
    from rgsim.ic import make_plummer
    eng = Engine(softening=0.05, gravity_mode="bh")
    eng.load_state(make_plummer(255, seed=0))
    assert eng._bh is None
    eng.load_state(make_plummer(256, seed=0))
    assert eng._bh is not None and eng._bh.n == 256
    eng.load_state(make_plummer(300, seed=0))  # N change rebuilds the handle
    assert eng._bh.n == 300


def test_engine_bh_nonfinite_check():
    from rgsim.engine import Engine
    from rgsim.ic import make_plummer
    s = make_plummer(500, seed=0)
    s.pos[3, 0] = np.inf
    eng = Engine(softening=0.05, gravity_mode="bh", check_finite=True)
    with pytest.raises(FloatingPointError):
        eng.load_state(s)


def test_resume_is_bitexact(tmp_path):
    """Kill-and-resume: a run stopped at step 20, saved and reloaded, then
    continued to 40 equals an uninterrupted 40-step run bit for bit
    (direct FP64; verify_phase2.py:70-80's resume check)."""
    from rgsim.engine import Engine
    from rgsim.ic import make_plummer
    from rgsim.state import State
    s = make_plummer(800, seed=9)
    a = Engine(softening=0.05)
    a.load_state(s)
    for _ in range(40):
        a.step(2e-5)
    b = Engine(softening=0.05)
    b.load_state(s)
    for _ in range(20):
        b.step(2e-5)
    path = str(tmp_path / "snap.npz")
    b.to_state().save_npz(path)
    c = Engine(softening=0.05)
    c.load_state(State.load_npz(path))
    for _ in range(20):
        c.step(2e-5)
    assert np.array_equal(a.to_state().pos, c.to_state().pos)
    assert np.array_equal(a.to_state().vel, c.to_state().vel)
    assert c.step_count == 40


@pytest.mark.parametrize("mode,precision", [("bh", "fp32"), ("bh", "fp64"), ("direct", "fp32c")])
def test_resume_from_snapshot_keeps_solver(tmp_path, mode, precision):
    """A snapshot records the engine's settings; Engine.from_state resumes
    with the same solver and precision, and the continued run equals the
    uninterrupted one bit for bit (the snapshot-and-resume row of the tree
    and reduced-precision paths)."""
    from rgsim.engine import Engine
    from rgsim.ic import make_disk_bulge
    from rgsim.state import State
    print(f"resume {mode} {precision}", flush=True)
    d = make_disk_bulge(n_disk=2500, n_bulge=500, seed=5)
    kw = dict(softening=0.05, gravity_mode=mode, theta=0.5, precision=precision)
    a = Engine(**kw)
    a.load_state(d)
    for _ in range(12):
        a.step(1e-4)
    b = Engine(**kw)
    b.load_state(d)
    for _ in range(6):
        b.step(1e-4)
    path = str(tmp_path / "snap.npz")
    b.to_state().save_npz(path)
    c = Engine.from_state(State.load_npz(path))
    assert (c.gravity_mode, c.precision, c.theta) == (mode, precision, 0.5)
    assert (c._bh is not None) == (mode == "bh")
    for _ in range(6):
        c.step(1e-4)
    sa, sc = a.to_state(), c.to_state()
    assert np.array_equal(sa.pos, sc.pos) and np.array_equal(sa.vel, sc.vel)
    assert c.step_count == 12 and sc.meta["engine_precision"] == precision
    if precision == "fp32c":  # the hi/lo offsets travel with the snapshot
        assert sc.pos_lo is not None and np.array_equal(sa.pos_lo, sc.pos_lo)
        assert np.any(sc.pos_lo != 0)

"""Host-side checks of the engine factory (core/simulation.py:20-33 with the
tree solver wired in) and of the snapshot containers; no GPU needed."""
import numpy as np
import pytest


def test_make_engine_maps_solver_names():
    from rgsim.simulation import make_engine
    e = make_engine("direct", 0.15)
    assert e.gravity_mode == "direct" and e.softening == 0.15 and e._bh is None
    for name in ("tree", "bh"):
        e = make_engine(name, 0.05, {"theta": 0.5}, precision="fp32")
        assert e.gravity_mode == "bh" and e.theta == 0.5 and e.precision == "fp32"
    assert make_engine("tree", 0.05).theta == 0.6
    with pytest.raises(NotImplementedError):
        make_engine("test_particle", 0.0)
    with pytest.raises(ValueError):
        make_engine("sph", 0.1)


def test_engine_settings_round_trip_through_npz(tmp_path):
    """The engine_* meta entries survive save_npz/load_npz, and from_state's
    arguments come from them (checked without loading onto a GPU)."""
    from rgsim.engine import Engine
    from rgsim.state import State
    with pytest.raises(ValueError):
        Engine(gravity_mode="bh", precision="fp32c")  # no compensated tree walk
    e = Engine(softening=0.07, gravity_mode="bh", theta=0.45, precision="fp32",
               morton_bits=10)
    rng = np.random.default_rng(3)
    s = State(pos=rng.normal(size=(9, 3)), vel=rng.normal(size=(9, 3)), mass=np.ones(9),
              ptype=np.zeros(9), ids=np.arange(9), time=1.5, step=7, seed=4,
              age=rng.random(9), meta={"level": "medium", **e.settings()},
              pos_lo=rng.normal(size=(9, 3)) * 1e-17)
    assert s.pos_lo.dtype == np.float32
    path = str(tmp_path / "s.npz")
    s.save_npz(path)
    t = State.load_npz(path)
    assert t.meta == s.meta and t.step == 7 and t.time == 1.5 and t.seed == 4
    for name in ("pos", "vel", "mass", "ptype", "ids", "age", "pos_lo"):
        assert np.array_equal(getattr(t, name), getattr(s, name))
    assert t.temperature is None
    with pytest.raises(ValueError):
        State(pos=np.zeros((2, 3)), vel=np.zeros((2, 3)), mass=np.ones(2), ptype=np.zeros(2),
              ids=np.arange(2), pos_lo=np.zeros((3, 3)))


def test_hdf5_snapshot_needs_h5py(tmp_path):
    """save/load keep the reference's HDF5 layout and import h5py on use."""
    from rgsim.state import State
    s = State(pos=np.zeros((2, 3)), vel=np.zeros((2, 3)), mass=np.ones(2), ptype=np.zeros(2),
              ids=np.arange(2), meta={"engine_precision": "fp32c"})
    try:
        import h5py  # noqa: F401
    except ModuleNotFoundError:
        with pytest.raises(ModuleNotFoundError):
            s.save(str(tmp_path / "s.h5"))
        return
    s.save(str(tmp_path / "s.h5"))
    t = State.load(str(tmp_path / "s.h5"))
    assert t.meta == s.meta and np.array_equal(t.mass, s.mass)


def test_engine_gravity_rules_without_gpu():
    """The engines' tree rule (living_galaxy_engine.py:734, gas_disk_engine.py:
    169): a tree iff gravity_mode == "bh" and n >= 256; bad settings raise
    before any device work."""
    from rgsim.gas_disk_engine import GasDiskGravity
    from rgsim.living_galaxy_engine import LivingGalaxyGravity
    from rgsim.solvers.self_gravity import SelfGravity, uses_tree
    assert uses_tree("bh", 256) and not uses_tree("bh", 255) and not uses_tree("direct", 10 ** 6)
    assert LivingGalaxyGravity(0.05).flavor == "living_galaxy"
    assert GasDiskGravity(0.05, "bh", 0.4, {"M_d": 5.0}).pot.M_d == 5.0
    with pytest.raises(ValueError):
        SelfGravity(0.05, flavor="sph")
    with pytest.raises(ValueError):
        SelfGravity(0.05, gravity_mode="tree")

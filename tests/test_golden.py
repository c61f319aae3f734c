"""Golden fixtures (tests/golden/golden_n*.npz, written by make_golden.py
from the oracle): the oracle keeps reproducing them bit for bit, and the GPU
product matches them (FP64 direct and KDK, Morton keys, order and topology
bit-exact; aggregation and BH walk to rounding)."""
import glob
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = sorted(glob.glob(os.path.join(HERE, "golden_n*.npz")))


def load(path):
    with np.load(path, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def test_fixtures_present():
    assert len(FILES) == 6


@pytest.mark.parametrize("path", FILES, ids=os.path.basename)
def test_oracle_reproduces_fixture(orc, path):
    g = load(path)
    G, eps = float(g["G"]), float(g["eps"])
    assert np.array_equal(orc.direct_acc(g["pos"], g["mass"], G, eps ** 2), g["acc_direct"])
    assert orc.kinetic_energy(g["vel"], g["mass"]) == g["ke"]
    assert orc.potential_energy(g["pos"], g["mass"], G, eps ** 2) == g["pe"]
    for bits in (21, 10):
        t = orc.bh_tree(g["pos"], g["mass"], bits)
        for k in ("keys", "order", "left", "right", "parent"):
            assert np.array_equal(t[k], g[f"{k}{bits}"]), (k, bits)
    # numpy twin of morton_codes agrees with the C restatement on the fixture
    assert np.array_equal(orc.morton_codes(g["pos"]), g["keys21"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=os.path.basename)
def test_gpu_matches_fixture(path):
    import torch

    from rgsim.engine import Engine
    from rgsim.solvers import gravity
    from rgsim.solvers.barnes_hut import BarnesHut
    from rgsim.state import State
    g = load(path)
    n = len(g["mass"])
    G, eps, theta = float(g["G"]), float(g["eps"]), float(g["theta"])
    pos = torch.from_numpy(g["pos"]).cuda()
    mass = torch.from_numpy(g["mass"]).cuda()
    acc = torch.empty_like(pos)
    gravity.compute_acc_direct(pos, mass, acc, n, G, eps ** 2)
    assert np.array_equal(acc.cpu().numpy(), g["acc_direct"])
    eng = Engine(softening=eps, G=G)
    eng.load_state(State(pos=g["pos"], vel=g["vel"], mass=g["mass"],
                         ptype=np.zeros(n, np.int32), ids=np.arange(n)))
    eng.step(1e-4)
    st = eng.to_state()
    assert np.array_equal(st.pos, g["kdk_pos"]) and np.array_equal(st.vel, g["kdk_vel"])
    if n < 2:
        return
    for bits in (21, 10):
        bh = BarnesHut(n, theta=theta, softening=eps, G=G, morton_bits=bits)
        bh.compute(pos, mass, acc)
        keys, order, left, right, parent = (t.cpu().numpy() for t in bh.export_tree())
        assert np.array_equal(keys.astype(np.uint64), g[f"keys{bits}"][g[f"order{bits}"]])
        assert np.array_equal(order, g[f"order{bits}"])
        assert np.array_equal(left, g[f"left{bits}"]) and np.array_equal(right, g[f"right{bits}"])
        assert np.array_equal(parent, g[f"parent{bits}"])
    bh = BarnesHut(n, theta=theta, softening=eps, G=G)
    bh.compute(pos, mass, acc)
    com, nmass, nmin, nmax = (t.cpu().numpy() for t in bh.export_nodes())
    assert np.array_equal(nmin, g["nmin"]) and np.array_equal(nmax, g["nmax"])
    np.testing.assert_allclose(nmass, g["nmass"], rtol=1e-13)
    np.testing.assert_allclose(com, g["com"], rtol=1e-12, atol=1e-12)
    rel = np.linalg.norm(acc.cpu().numpy() - g["acc_bh"], axis=1) / np.linalg.norm(g["acc_bh"], axis=1)
    assert rel.max() < 1e-9
    assert bh.stats()["visits_total"] == g["visits_bh"].sum()

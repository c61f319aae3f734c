"""The bench's measurement helpers (CPU): the sum of term magnitudes behind
the condition-normalised force error against a pairwise loop, and the VALU
and scalar issue figures read from the committed rocprofv3 summaries."""
import importlib.util
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_term_magnitude_sums_pairwise(bench):
    rng = np.random.default_rng(2)
    pos = rng.normal(size=(70, 3))
    mass = rng.uniform(0.5, 1.5, 70)
    idx = np.array([0, 5, 69, 33])
    got = bench.term_magnitude_sums(pos, mass, 2.5, 1e-3, idx, rows=3)
    for k, i in enumerate(idx):
        ref = 0.0
        for j in range(70):
            d = pos[j] - pos[i]
            r2 = d @ d + 1e-3
            ref += 2.5 * mass[j] * np.sqrt(d @ d) / r2 ** 1.5
        assert got[k] == pytest.approx(ref, rel=1e-13)


def test_cond_err_zero_and_scale(bench):
    ref = np.ones((4, 3))
    a = ref.copy()
    a[2, 0] += 3.0
    e = bench.cond_err(a, ref, np.array([1, 2]), np.array([10.0, 30.0]))
    assert e["max"] == pytest.approx(0.1) and e["sample"] == 2


def test_issue_figures_from_profiles(bench):
    # FP32C direct sum at N=262,144 (symmetric: each pair evaluated once):
    # about 13 VALU instructions per interaction
    d = bench.direct_issue(262144, "fp32c", 19.0)
    assert d is not None and 12 < d["valu_instr_per_interaction"] < 14
    assert d["kernel"] == "k_direct_sym<2, false, false>"
    assert d["peak"] == pytest.approx(1228.8)
    assert d["achieved"] == pytest.approx(d["valu_instr_per_launch"] / 19e-3 / 1e9)
    f = bench.direct_issue(262144, "fp64", 104.0)
    assert f["peak"] == pytest.approx(614.4) and 45 < f["valu_instr_per_interaction"] < 52
    assert bench.direct_issue(1000, "fp32c", 1.0) is None
    b = bench.kernel_issue("k_batch_force_512systems_sq.json", "k_batch_force<2, 256, 1>",
                           6.0, 15509199223)
    assert b is not None and 19 < b["valu_instr_per_interaction"] < 22
    assert bench.kernel_issue("k_batch_force_512systems_sq.json", "no_such_kernel", 1.0, 1) is None


def test_walk_salu_issue_from_profile(bench):
    """The walk's scalar-unit issue: SQ_INSTS_SALU per launch from the
    committed summary over the launch time, against one scalar instruction
    per cycle per CU (256 CUs at 2.4 GHz)."""
    import json
    with open(os.path.join(ROOT, "profiles", "k_walk32_chunked4_true_true_diskbulge120k_sq.json")) as f:
        rec = json.load(f)
    s = bench.salu_issue(rec, 0.3)
    assert s["peak"] == pytest.approx(614.4)
    assert s["achieved"] == pytest.approx(rec["per_launch"]["SQ_INSTS_SALU"] / 0.3e-3 / 1e9)
    assert 0.3 < s["frac"] < 0.5
    assert bench.salu_issue({"per_launch": {}}, 1.0) is None


def test_phase_rooflines_sort_bytes(bench):
    """The sort phase's algorithmic bytes follow the path that ran: 248 B per
    particle for the 8-pass LSD sort (bbox read, keys, 8 passes of keys and
    indices read and written), 152 for 4 passes, 102 for the sample sort."""
    ms = {"sort": 1.0, "build": 1.0, "aggregate": 1.0}
    n = 1000
    for passes, per in ((8, 248), (4, 152), (1, 102)):
        r = bench.phase_rooflines(n, ms, passes)
        assert r["sort"]["algorithmic_bytes"] == n * per
        assert abs(r["sort"]["GB_per_s"] - n * per / 1e-3 / 1e9) < 1e-9
    assert bench.phase_rooflines(n, ms)["sort"]["algorithmic_bytes"] == n * 248

"""Benchmark of the gravity hot path on one MI355X.

Workload (the Barnes-Hut drop-in configuration): exponential sech^2 disk
N=100,000 (R_d 3 kpc, z_0 0.3 kpc, 5e10 Msun) in a live Plummer bulge
N=20,000 (a 0.5 kpc, 1e10 Msun), seed 1, G = 43009.1725, eps 0.05 kpc,
theta 0.5, 63-bit Morton keys, FP32 walk over FP64 aggregation, KDK dt 1e-4.
A step is one full KDK step: kick+drift, Morton keys, radix sort, Karras
build, aggregation, walk, kick.  value = N * steps / time in Mparticle-steps/s.

Also reported: the direct N^2 sum on a Plummer sphere N=262,144 (a=1, G=M=1,
eps 0.005) in FP64, FP32 and FP32-compensated, in G interactions/s with the
force error of each against FP64; and the CPU baseline (the oracle's C/OpenMP
restatement of the reference algorithm on the host cores).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "real-astronomy-simulation_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = ("direct Ginteractions/s & BH Mparticle-steps/s vs plain FP64 GPU, "
          "at stated force err")
HBM_PEAK_GBS = 8000.0
# wave64 VALU instruction issue, G/s: an FP32 instruction takes 2 cycles on
# each of the 1,024 SIMDs at 2.4 GHz (157.3 TFLOP/s FP32 is one wave64 FMA per
# SIMD every 2 cycles), an FP64 one 4 cycles (78.6 TFLOP/s)
VALU_PEAK_GINST = 1024 * 2.4 / 2
VALU64_PEAK_GINST = 1024 * 2.4 / 4
# scalar instruction issue, G/s: each CU's scalar unit issues about one
# instruction per cycle for all the waves of its 4 SIMDs (a probe of
# independent s_add chains at full occupancy issued 581 G/s over the GPU)
SALU_PEAK_GINST = 256 * 2.4
# direct-sum kernel each precision launches at n = 262,144 (default shape)
DIRECT_KERNELS = {"fp64": "k_direct<0, 1, false>", "fp32": "k_direct_sym<1, false, false>",
                  "fp32c": "k_direct_sym<2, false, false>"}
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
FLOP_PER_INTERACTION = 20
WALK_RECORD_BYTES = 32          # FP32 node record fetched per visit
# FP32 walk variant (rg_bh_stats walk_mode) -> the kernel it launches
WALK_KERNELS = {0: "k_walk32", 1: "k_walk32_pair", 2: "k_walk32_wave", 3: "k_walk32_wave2",
                4: "k_walk32_wave_split<2>", 5: "k_walk32_wave_split<4>",
                6: "k_walk32_chunked<{w}>", 7: "k_walk32_chunked<{w}, true, true>",
                8: "k_walk32_chunked2<{w}>", -1: "k_walk64"}
WALK_PARTICLE_BYTES = 16 + 4 + 24  # own leaf record + order index + FP64 acceleration
HILO_KD_BYTES = (24 + 12 + 24 + 24) + (24 + 12 + 24)  # hi/lo kick + drift, read + written
WALK_NODE_BYTES = 16 + 8           # FP32 record pair (com, G m) + (threshold, skip), per node
WALK_IO_BYTES = 4 + 24             # order index read + FP64 acceleration written, per particle


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def rel_err(a, ref):
    import numpy as np
    r = np.linalg.norm(a - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-300)
    return {"max": float(r.max()), "rms": float(np.sqrt(np.mean(r * r))),
            "mean": float(r.mean())}


def term_magnitude_sums(pos, mass, G, eps2, idx, rows=32):
    """S_i = sum_j |G m_j d_ij| / (|d_ij|^2 + eps2)^(3/2) for the particles idx,
    in FP64 numpy: the sum of the magnitudes of the terms whose vector sum is
    a_i.  |a_i - a_ref_i| / S_i is the force error normalised by the
    summation's own condition, which stays meaningful where the terms cancel
    (near the centre of a Plummer sphere |a_i| -> 0 while S_i does not)."""
    import numpy as np
    out = np.empty(len(idx))
    for k in range(0, len(idx), rows):
        i = idx[k:k + rows]
        d = pos[None, :, :] - pos[i, None, :]
        q = np.einsum("rnk,rnk->rn", d, d)
        r2 = q + eps2
        out[k:k + rows] = (G * mass[None, :] * np.sqrt(q) / (r2 * np.sqrt(r2))).sum(1)
    return out


def cond_err(a, ref, idx, sums):
    import numpy as np
    e = np.linalg.norm(a[idx] - ref[idx], axis=1) / sums
    return {"max": float(e.max()), "rms": float(np.sqrt(np.mean(e * e))), "sample": len(idx)}


def bench_bh(steps, warmup):
    import numpy as np
    import torch

    from rgsim.engine import Engine
    from rgsim.ic import make_disk_bulge
    from rgsim.solvers import gravity

    state = make_disk_bulge(n_disk=100000, n_bulge=20000, seed=1)
    n = state.n
    eng = Engine(softening=0.05, gravity_mode="bh", theta=0.5, precision="fp32")
    eng.load_state(state)
    # force error of the drop-in against the FP64 direct sum on the IC
    ref = torch.empty_like(eng.pos)
    gravity.compute_acc_direct(eng.pos, eng.mass, ref, n, eng.G, 0.05 ** 2, "fp64")
    err = rel_err(eng.acc.cpu().numpy(), ref.cpu().numpy())
    st = eng._bh.stats()
    dt = 1e-4
    for _ in range(warmup):
        eng.step(dt)
    torch.cuda.synchronize()
    # the timed steps run without phase events (each event adds a gap)
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.step(dt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    # phase times and the walk's launch duration from HIP events on the
    # compute stream, over further steps (back to back: no synchronisation
    # between them); visits of the last of those steps
    phase_steps = min(steps, 20)
    eng._bh.set_timing(True)
    for _ in range(phase_steps):
        eng.step(dt)
    tm = eng._bh.timing()
    eng._bh.set_timing(False)
    st_last = eng._bh.stats()
    visits = st_last["visits_total"]
    walk_ms = tm["walk_ms"] / max(tm["count"], 1)
    walk_bytes = visits * WALK_RECORD_BYTES + n * WALK_PARTICLE_BYTES
    achieved = walk_bytes / (walk_ms * 1e-3) / 1e9
    kernel = WALK_KERNELS[st_last["walk_mode"]].format(w=st_last["walk_waves"])
    traffic = None
    fname = re.sub(r"[^A-Za-z0-9_]", "", kernel.replace(", ", "_"))
    pmc = os.path.join(ROOT, "profiles", f"{fname}_diskbulge120k_pmc.json")
    if os.path.exists(pmc):  # rocprofv3 FETCH_SIZE / WRITE_SIZE of this kernel, copy-scaled
        with open(pmc) as f:
            rec = json.load(f)
        if rec.get("kernel") == kernel and rec.get("n") == n:
            traffic = rec.get("hbm_bytes_per_launch")
    # the walk against vector-instruction issue: its VALU wave instructions
    # per launch (rocprofv3 SQ_INSTS_VALU, committed summary) over the
    # measured launch time, against the FP32 issue peak
    roofline_valu = None
    sq = os.path.join(ROOT, "profiles", f"{fname}_diskbulge120k_sq.json")
    if os.path.exists(sq):
        with open(sq) as f:
            rec = json.load(f)
        if rec.get("kernel") == kernel and rec.get("n") == n:
            rate = rec["valu_instr_per_launch"] / (walk_ms * 1e-3) / 1e9
            roofline_valu = {"bound": "valu", "achieved": rate, "peak": VALU_PEAK_GINST,
                             "unit": "G wave-instructions/s", "frac": rate / VALU_PEAK_GINST,
                             "kernel": kernel,
                             "valu_instr_per_launch": rec["valu_instr_per_launch"],
                             "salu_issue": salu_issue(rec, walk_ms)}
    # The walk against HBM: the bytes it must move at least (every FP32 node
    # record once, 24 B, and per particle its order index and FP64
    # acceleration, 28 B) and the bytes it did move past the L2 (rocprofv3
    # FETCH_SIZE + WRITE_SIZE, copy-scaled).  The per-visit record figure
    # (effective_GBs) counts LDS, register-window and cache hits too, so it
    # is a reuse rate, not an HBM fraction.
    compulsory = (2 * n - 1) * WALK_NODE_BYTES + n * WALK_IO_BYTES
    hbm = {"compulsory_bytes_per_launch": compulsory,
           "compulsory_GBs": compulsory / (walk_ms * 1e-3) / 1e9,
           "traffic_bytes_per_launch": traffic,
           "traffic_GBs": traffic / (walk_ms * 1e-3) / 1e9 if traffic else None,
           "traffic_frac_of_hbm_peak": (traffic / (walk_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
                                        if traffic else None),
           "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    if roofline_valu is not None:
        # the bound the counters show: VALU issue (the walk is far from HBM
        # peak at every size; its waves wait on dependent LDS and scalar
        # results most of their lifetime)
        roofline = dict(roofline_valu)
        roofline.update({"traffic": traffic, "avg_launch_ms": walk_ms,
                         "effective_GBs": achieved,
                         "algorithmic_record_bytes_per_launch": walk_bytes})
    else:
        roofline = {"bound": "hbm", "achieved": hbm["compulsory_GBs"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": hbm["compulsory_GBs"] / HBM_PEAK_GBS,
                    "traffic": traffic, "kernel": kernel, "avg_launch_ms": walk_ms,
                    "effective_GBs": achieved}
    return {
        "n": n, "wall_s": wall, "ms_per_step": 1e3 * wall / steps,
        "mps": n * steps / wall / 1e6,
        "force_err_vs_fp64_direct": err,
        "visits_per_particle": visits / n, "visits_max": st_last["visits_max"],
        "visits_ic": st["visits_total"],
        "phase_ms": {"sort": tm["sort_ms"] / max(tm["count"], 1),
                     "build": tm["build_ms"] / max(tm["count"], 1),
                     "aggregate": tm["aggregate_ms"] / max(tm["count"], 1),
                     "walk": walk_ms},
        "phase_hbm": phase_rooflines(n, {"sort": tm["sort_ms"] / max(tm["count"], 1),
                                         "build": tm["build_ms"] / max(tm["count"], 1),
                                         "aggregate": tm["aggregate_ms"] / max(tm["count"], 1)},
                                     st_last["sort_passes"]),
        "sort": {"global_passes": st_last["sort_passes"], "buckets": st_last["sort_buckets"],
                 "buckets_in_global_memory": st_last["sort_spilled"]},
        "roofline": roofline,
        "walk_hbm": hbm,
        "energy": None,
    }


def bench_bh_fp64(state, dt=1e-4, steps=20):
    """The BH drop-in with the reference's FP64 arithmetic end to end: FP64
    aggregation and the FP64 stackless walk (k_walk64, one lane per particle,
    the reference's criterion and pop order): the plain FP64 GPU treecode the
    FP32 walk is compared with."""
    import torch

    from rgsim.engine import Engine

    n = state.n
    e64 = Engine(softening=0.05, gravity_mode="bh", theta=0.5, precision="fp64")
    e64.load_state(state)
    for _ in range(3):
        e64.step(dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        e64.step(dt)
    torch.cuda.synchronize()
    w = time.perf_counter() - t0
    return {"precision": "fp64", "steps": steps, "ms_per_step": 1e3 * w / steps,
            "mps": n * steps / w / 1e6}


def bench_bh30(state, dt=1e-4, steps=50):
    """The BH workload with 30-bit keys (10 bits per axis, the option the
    workload text names; 4 sort passes instead of 8), 50 steps.  Its own
    section, so that a profile of the bh section holds only 63-bit walks."""
    import torch

    from rgsim.engine import Engine

    n = state.n
    e30 = Engine(softening=0.05, gravity_mode="bh", theta=0.5, precision="fp32",
                 morton_bits=10)
    e30.load_state(state)
    for _ in range(5):
        e30.step(dt)
    torch.cuda.synchronize()
    t30 = time.perf_counter()
    for _ in range(steps):
        e30.step(dt)
    torch.cuda.synchronize()
    w30 = time.perf_counter() - t30
    return {"morton_bits": 10, "steps": steps, "ms_per_step": 1e3 * w30 / steps,
            "mps": n * steps / w30 / 1e6}


def salu_issue(rec, ms):
    """Scalar-unit issue rate of a kernel: its scalar instructions per launch
    (rocprofv3 SQ_INSTS_SALU in the committed summary `rec`) over the measured
    launch time `ms`, against the CUs' scalar issue peak.  None when the
    summary has no SQ_INSTS_SALU."""
    instr = rec.get("per_launch", {}).get("SQ_INSTS_SALU")
    if not instr:
        return None
    rate = instr / (ms * 1e-3) / 1e9
    return {"achieved": rate, "peak": SALU_PEAK_GINST, "unit": "G instructions/s",
            "frac": rate / SALU_PEAK_GINST, "salu_instr_per_launch": instr}


def phase_rooflines(n, phase_ms, passes=8):
    """HBM rooflines of the tree-build phases from their algorithmic bytes.
    Keys + sort per particle: 24 (bbox read) + 24 + 8 (Morton) + passes x 24
    for the LSD radix sort; for the sample sort (passes == 1) 24 (bbox) +
    24 + 8 + 1 (keys, bucket id) + 8 + 1 + 12 (bucket pass) + 24 (each bucket
    read and written once).  Build 24 B per internal node + 4 B preorder slot
    per node; aggregation 2 x 80 B per node."""
    m = 2 * n - 1
    sort_b = 24 + 24 + 8 + 1 + 8 + 1 + 12 + 24 if passes == 1 else 24 + 24 + 8 + passes * 24
    nbytes = {"sort": n * sort_b,
              "build": 24 * (n - 1) + 4 * m,
              "aggregate": 160 * m}
    out = {}
    for k, b in nbytes.items():
        gbs = b / (phase_ms[k] * 1e-3) / 1e9
        out[k] = {"algorithmic_bytes": b, "GB_per_s": gbs, "frac_of_hbm_peak": gbs / HBM_PEAK_GBS}
    return out


def kernel_issue(profile, kernel, ms, interactions, peak=None):
    """VALU issue rate of one launch of `kernel`: its wave64 VALU instructions
    per launch from the committed rocprofv3 SQ_INSTS_VALU summary `profile`
    over the launch time `ms` measured here, against the FP32 issue peak
    (or `peak`)."""
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        k = json.load(f).get("kernels", {}).get(kernel)
    if k is None:
        return None
    instr = k["valu_instr_per_launch"]
    rate = instr / (ms * 1e-3) / 1e9
    peak = VALU_PEAK_GINST if peak is None else peak
    return {"kernel": kernel, "valu_instr_per_launch": instr,
            "valu_instr_per_interaction": instr * 64 / interactions,
            "achieved": rate, "peak": peak, "unit": "G wave-instructions/s",
            "frac": rate / peak}


def direct_issue(n, prec, ms):
    """VALU issue rate of one direct-sum evaluation: the kernel's wave64 VALU
    instructions per launch (rocprofv3 SQ_INSTS_VALU, committed summary) over
    the measured evaluation time, against the issue peak of its precision
    (every instruction counted at that precision's rate)."""
    path = os.path.join(ROOT, "profiles", "k_direct_plummer262144_sq.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    k = rec.get("kernels", {}).get(DIRECT_KERNELS[prec])
    if rec.get("n") != n or k is None:
        return None
    instr = k["valu_instr_per_launch"]
    rate = instr / (ms * 1e-3) / 1e9
    peak = VALU64_PEAK_GINST if prec == "fp64" else VALU_PEAK_GINST
    return {"kernel": DIRECT_KERNELS[prec], "valu_instr_per_launch": instr,
            "valu_instr_per_interaction": instr * 64 / (n * n),
            "achieved": rate, "peak": peak, "unit": "G wave-instructions/s",
            "frac": rate / peak}


def bench_direct(n=262144, evals=3, kdk_steps=50):
    import numpy as np
    import torch

    from rgsim.engine import Engine
    from rgsim.ic import make_plummer
    from rgsim.solvers import gravity

    s = make_plummer(n, total_mass=1.0, scale_radius=1.0, seed=0, G=1.0)
    pos = torch.from_numpy(s.pos).cuda()
    mass = torch.from_numpy(s.mass).cuda()
    eps2 = 0.005 ** 2
    out = {"n": n}
    ref = None
    for prec in ("fp64", "fp32", "fp32c"):
        acc = torch.empty_like(pos)
        gravity.compute_acc_direct(pos, mass, acc, n, 1.0, eps2, prec)  # warm-up
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(evals):
            gravity.compute_acc_direct(pos, mass, acc, n, 1.0, eps2, prec)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / evals
        a = acc.cpu().numpy()
        if prec == "fp64":
            ref = a
            # condition-normalised errors on a fixed sample of 256 particles
            cidx = np.random.default_rng(11).choice(n, size=min(n, 256), replace=False)
            csums = term_magnitude_sums(s.pos, s.mass, 1.0, eps2, cidx)
        gint = n * n / (ms * 1e-3) / 1e9
        peak = FP64_PEAK_TFLOPS if prec == "fp64" else FP32_PEAK_TFLOPS
        tflops = gint * FLOP_PER_INTERACTION / 1e3
        entry = {"ms_per_eval": ms, "Gint_per_s": gint, "TFLOP_s_20flop": tflops,
                 "frac_of_peak": tflops / peak}
        issue = direct_issue(n, prec, ms)
        if issue is not None:
            entry["valu_issue"] = issue
        if prec != "fp64":
            entry["force_err_vs_fp64"] = rel_err(a, ref)
            entry["force_err_vs_fp64"]["cond_normalised"] = cond_err(a, ref, cidx, csums)
        out[prec] = entry
        log(f"direct {prec}: {ms:.2f} ms/eval, {gint:.1f} Gint/s")
    out["speedup_fp32c_vs_fp64"] = out["fp64"]["ms_per_eval"] / out["fp32c"]["ms_per_eval"]
    out["speedup_fp32_vs_fp64"] = out["fp64"]["ms_per_eval"] / out["fp32"]["ms_per_eval"]
    # the FP32/FP32C sums above run split over j (64 slices at this n, added
    # in order in FP64, four i-particles per lane for FP32C); the same FP32C
    # sum with one slice and one i-particle per lane, for comparison
    from rgsim import _lib
    _lib.set_param("direct_split", 1)
    _lib.set_param("direct_ipt", 1)
    try:
        acc = torch.empty_like(pos)
        gravity.compute_acc_direct(pos, mass, acc, n, 1.0, eps2, "fp32c")
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(evals):
            gravity.compute_acc_direct(pos, mass, acc, n, 1.0, eps2, "fp32c")
        e1.record()
        torch.cuda.synchronize()
    finally:
        _lib.set_param("direct_split", 0)
        _lib.set_param("direct_ipt", 0)
    ms1 = e0.elapsed_time(e1) / evals
    out["fp32c_one_slice_one_per_lane"] = {"ms_per_eval": ms1, "Gint_per_s": n * n / (ms1 * 1e-3) / 1e9,
                            "force_err_vs_fp64": rel_err(acc.cpu().numpy(), ref)}
    # precision form: 50 KDK steps dt 5e-4, FP64 energy and momentum
    eng = Engine(softening=0.005, precision="fp32c", G=1.0)
    eng.load_state(s)
    ke0, pe0, e_0 = eng.energies()
    p0 = np.array(eng.momentum())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(kdk_steps):
        eng.step(5e-4)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ke1, pe1, e_1 = eng.energies()
    p1 = np.array(eng.momentum())
    out["fp32c_kdk"] = {"steps": kdk_steps, "dt": 5e-4, "ms_per_step": 1e3 * wall / kdk_steps,
                        "positions": "FP64 hi + FP32 lo (compensated drift)",
                        "energy_drift": abs((e_1 - e_0) / e_0), "virial": -2 * ke1 / pe1,
                        "momentum_change": float(np.linalg.norm(p1 - p0))}
    log(f"direct fp32c KDK: {out['fp32c_kdk']['ms_per_step']:.2f} ms/step, "
        f"|dE/E| {out['fp32c_kdk']['energy_drift']:.2e}")
    out["hilo_drift"] = bench_hilo_drift(eng)
    return out


def bench_hilo_drift(eng, reps=50):
    """The opening kick + compensated drift of hi/lo positions
    (k_kick_drift_hilo): reads x_hi 24 + x_lo 12 + v 24 + a 24 B and writes
    x_hi 24 + x_lo 12 + v 24 B per particle, 144 B; at the precision form's
    n and at 2^22 particles, where the pass is long enough to stream at HBM
    rate.  dt = 0 leaves the positions unchanged."""
    import torch

    from rgsim.integrators import leapfrog

    def timed(pos, lo, vel, acc, n):
        leapfrog.kick_drift_hilo(pos, lo, vel, acc, n, 0.0, 0.0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            leapfrog.kick_drift_hilo(pos, lo, vel, acc, n, 0.0, 0.0)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        gbs = HILO_KD_BYTES * n / (ms * 1e-3) / 1e9
        return {"n": n, "ms": ms, "GB_per_s": gbs, "frac_of_hbm_peak": gbs / HBM_PEAK_GBS}

    out = {"bytes_per_particle": HILO_KD_BYTES,
           "at_precision_form_n": timed(eng.pos, eng.pos_lo, eng.vel, eng.acc, eng.n)}
    big = 1 << 22
    pos = torch.zeros((big, 3), dtype=torch.float64, device="cuda")
    lo = torch.zeros((big, 3), dtype=torch.float32, device="cuda")
    vel, acc = torch.zeros_like(pos), torch.zeros_like(pos)
    out["at_4M"] = timed(pos, lo, vel, acc, big)
    log(f"hi/lo kick+drift: {out['at_4M']['GB_per_s']:.0f} GB/s at 2^22 "
        f"({out['at_4M']['frac_of_hbm_peak']:.2f} of HBM peak)")
    return out


def bench_adjoint(n=16384, steps=400, ckpt=20):
    """Differentiation form: forward KDK step vs adjoint (reverse) step."""
    import torch

    from rgsim.adjoint import direct_vjp, kdk_trajectory
    from rgsim.ic import make_plummer
    from rgsim.solvers import gravity

    s = make_plummer(n, total_mass=1.0, scale_radius=1.0, seed=7, G=1.0)
    x = torch.from_numpy(s.pos).cuda()
    m = torch.from_numpy(s.mass).cuda()
    lam = torch.randn_like(x)
    eps2 = 0.01 ** 2
    acc = torch.empty_like(x)
    reps = 50

    def timed(fn):
        # warm up for at least 50 ms of calls, so that the timed calls run at
        # the clocks a sustained run sees
        t_end = time.perf_counter() + 0.05
        while time.perf_counter() < t_end:
            fn()
            torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    out = {"n": n, "steps": steps, "checkpoint_every": ckpt, "precision": "fp32c"}
    out["force_ms"] = timed(lambda: gravity.compute_acc_direct(x, m, acc, n, 1.0, eps2, "fp32c"))
    out["vjp_ms"] = timed(lambda: direct_vjp(x, m, lam, 1.0, eps2, "fp32c"))
    issue = kernel_issue("k_vjp32c_plummer16384_sq.json", "k_vjp32c", out["vjp_ms"], n * n) \
        if n == 16384 else None
    if issue is not None:
        out["vjp_valu_issue"] = issue
    out["vjp_over_force"] = out["vjp_ms"] / out["force_ms"]
    x0 = torch.from_numpy(s.pos).cuda().requires_grad_(True)
    v0 = torch.from_numpy(s.vel).cuda().requires_grad_(True)
    mm = torch.from_numpy(s.mass).cuda().requires_grad_(True)
    target = x0.detach() + 1e-3 * torch.randn_like(x0)
    # an untimed 2-segment trajectory forward and backward first, so that the
    # timed one starts from the same warm state for both directions
    xw, _ = kdk_trajectory(x0, v0, mm, 2 * ckpt, 1e-3, 1.0, 0.01, ckpt, "fp32c")
    ((xw - target) ** 2).sum(1).mean().backward()
    x0.grad = v0.grad = mm.grad = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xN, _ = kdk_trajectory(x0, v0, mm, steps, 1e-3, 1.0, 0.01, ckpt, "fp32c")
    loss = ((xN - target) ** 2).sum(1).mean()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    out["forward_ms_per_step"] = 1e3 * (t1 - t0) / steps
    out["backward_ms_per_step"] = 1e3 * (t2 - t1) / steps
    out["backward_over_forward"] = (t2 - t1) / (t1 - t0)
    out["backward_includes"] = ("segment recompute (19 forces per 20-step segment: the force at a "
                                "segment's end is not needed) + 1 VJP/step + update")
    log(f"adjoint: force {out['force_ms']:.3f} ms, vjp {out['vjp_ms']:.3f} ms "
        f"({out['vjp_over_force']:.2f}x); trajectory bwd/fwd {out['backward_over_forward']:.2f}x")
    return out


def bench_batch(steps=100):
    """Batching form: 512 Plummer systems, N log-uniform in [256, 16384]."""
    import numpy as np
    import torch

    from rgsim.autotune import tune_batch
    from rgsim.batch import Ensemble, plummer_ensemble

    systems = plummer_ensemble(nsys=512, n_min=256, n_max=16384, master_seed=2026)
    ens = Ensemble(systems, precision="fp32c")
    tune = tune_batch(ens, reps=2)
    e0 = ens.energies()
    ens.capture(steps)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    ens.replay()
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    e1 = ens.energies()
    inter = ens.interactions_per_step()
    # one-launch-per-system baseline, same kernels
    ps_steps = 5
    ens2 = Ensemble(systems, precision="fp32c")
    ens2.per_system_step(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens2.per_system_step(ps_steps)
    torch.cuda.synchronize()
    ps_ms = 1e3 * (time.perf_counter() - t0) / ps_steps
    # FP64 batched run of the same steps for the drift comparison
    ens64 = Ensemble(systems, precision="fp64")
    f0 = ens64.energies()
    ens64.capture(steps)
    ens64.replay()
    f1 = ens64.energies()
    d32 = np.abs((e1[:, 0] + e1[:, 1] - e0[:, 0] - e0[:, 1]) / (e0[:, 0] + e0[:, 1]))
    d64 = np.abs((f1[:, 0] + f1[:, 1] - f0[:, 0] - f0[:, 1]) / (f0[:, 0] + f0[:, 1]))
    # the tuned tile -> the kernel instance the graph replays
    tile = tune.get("tile")
    kname = {64: "k_batch_force<2, 64, 1>", 128: "k_batch_force<2, 128, 1>",
             256: "k_batch_force<2, 256, 1>", 512: "k_batch_force<2, 256, 2>",
             1024: "k_batch_force<2, 256, 4>"}.get(tile)
    issue = kernel_issue("k_batch_force_512systems_sq.json", kname, ms, inter) \
        if kname and ens.nsys == 512 else None
    out = {"nsys": ens.nsys, "ntot": ens.ntot, "interactions_per_step": inter,
           "ms_per_step_graph": ms, "Gint_per_s_graph": inter / (ms * 1e-3) / 1e9,
           "ms_per_step_per_system_launches": ps_ms,
           "Gint_per_s_per_system_launches": inter / (ps_ms * 1e-3) / 1e9,
           "speedup_vs_per_system": ps_ms / ms,
           "autotune": {k: v for k, v in tune.items() if k != "trials"},
           "energy_drift_fp32c": {"max": float(d32.max()), "median": float(np.median(d32))},
           "energy_drift_fp64": {"max": float(d64.max()), "median": float(np.median(d64))}}
    if issue is not None:
        out["valu_issue"] = issue
    log(f"batch: {out['Gint_per_s_graph']:.0f} Gint/s graph vs "
        f"{out['Gint_per_s_per_system_launches']:.0f} per-system; drift fp32c max {d32.max():.1e}")
    return out


def bench_bh_1m(steps=5):
    """BH at N=1,000,000: disk:bulge 5:1, theta 0.5, FP32 walk."""
    import torch

    from rgsim.engine import Engine
    from rgsim.ic import make_disk_bulge
    from rgsim.solvers import gravity

    s = make_disk_bulge(n_disk=833334, n_bulge=166666, seed=1)
    n = s.n
    eng = Engine(softening=0.05, gravity_mode="bh", theta=0.5, precision="fp32")
    eng.load_state(s)
    ref = torch.empty_like(eng.pos)
    t0 = time.perf_counter()
    gravity.compute_acc_direct(eng.pos, eng.mass, ref, n, eng.G, 0.05 ** 2, "fp64")
    torch.cuda.synchronize()
    ref_s = time.perf_counter() - t0
    err = rel_err(eng.acc.cpu().numpy(), ref.cpu().numpy())
    del ref
    eng.step(1e-4)
    torch.cuda.synchronize()
    eng._bh.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.step(1e-4)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    tm = eng._bh.timing()
    eng._bh.set_timing(False)
    st = eng._bh.stats()
    c = max(tm["count"], 1)
    walk_ms = tm["walk_ms"] / c
    walk_bytes = st["visits_total"] * WALK_RECORD_BYTES + n * WALK_PARTICLE_BYTES
    phases = {"sort": tm["sort_ms"] / c, "build": tm["build_ms"] / c,
              "aggregate": tm["aggregate_ms"] / c}
    # the fused opening kick + drift (leapfrog.py:11-22): 120 B per particle
    # (x and v read and written, a read), HBM-bound at this size
    from rgsim.integrators import leapfrog
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    leapfrog.kick_drift(eng.pos, eng.vel, eng.acc, n, 0.0, 0.0)
    e0.record()
    for _ in range(20):
        leapfrog.kick_drift(eng.pos, eng.vel, eng.acc, n, 0.0, 0.0)
    e1.record()
    torch.cuda.synchronize()
    kd_ms = e0.elapsed_time(e1) / 20
    kd_gbs = 120 * n / (kd_ms * 1e-3) / 1e9
    out = {"n": n, "ms_per_step": 1e3 * wall / steps, "mps": n * steps / wall / 1e6,
           "phase_ms": dict(phases, walk=walk_ms),
           "phase_hbm": phase_rooflines(n, phases, st["sort_passes"]),
           "sort": {"global_passes": st["sort_passes"], "buckets": st["sort_buckets"],
                    "buckets_in_global_memory": st["sort_spilled"]},
           "kick_drift": {"ms": kd_ms, "GB_per_s": kd_gbs, "frac_of_hbm_peak": kd_gbs / HBM_PEAK_GBS},
           "visits_per_particle": st["visits_total"] / n,
           "walk_effective_GBs": walk_bytes / (walk_ms * 1e-3) / 1e9,
           "walk_compulsory_GBs": ((2 * n - 1) * WALK_NODE_BYTES + n * WALK_IO_BYTES)
                                  / (walk_ms * 1e-3) / 1e9,
           "force_err_vs_direct": err,
           "force_err_reference": f"FP64 direct sum (the reference's arithmetic), {ref_s:.2f} s"}
    # bytes past the L2 per walk launch (rocprofv3 FETCH_SIZE + WRITE_SIZE,
    # copy-scaled, committed summary) over the measured walk time
    pmc = os.path.join(ROOT, "profiles", "k_walk32_chunked4_true_true_diskbulge1M_pmc.json")
    if os.path.exists(pmc) and st["walk_mode"] == 7 and st["walk_waves"] == 4:
        with open(pmc) as f:
            rec = json.load(f)
        if rec.get("n") == n:
            gbs = rec["hbm_bytes_per_launch"] / (walk_ms * 1e-3) / 1e9
            out["walk_traffic"] = {"bytes_per_launch": rec["hbm_bytes_per_launch"],
                                   "GB_per_s": gbs, "frac_of_hbm_peak": gbs / HBM_PEAK_GBS}
    frac = out.get("walk_traffic", {}).get("frac_of_hbm_peak")
    log(f"bh 1M: {out['ms_per_step']:.2f} ms/step, walk {walk_ms:.2f} ms"
        + (f" (measured traffic {frac:.2f} of HBM peak)" if frac is not None else "")
        + f", force error mean {err['mean']:.2e} against FP64 direct")
    return out


def bench_engines(n_gas=65536, n_tree=120000, reps=5):
    """Self-gravity of the gas-disk and living-galaxy engines: the fused
    N^2 + analytic potential force (add_grav_ext) on a gas disk in both
    precisions, the analytic-only and field-add passes, and the tree path
    with the add fused into the walk against compute + separate add."""
    import torch

    from rgsim.ic import make_disk_bulge
    from rgsim.solvers import gravity
    from rgsim.solvers.barnes_hut import BarnesHut
    from rgsim.units import G

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    pot = gravity.potential_struct(5.0, 3.0, 0.3, 12.0, 8.0, 0, 0.0)  # the engines' default
    s = make_disk_bulge(n_disk=n_gas * 7 // 8, n_bulge=n_gas // 8, seed=3)
    n = s.n
    pos = torch.from_numpy(s.pos).cuda()
    mass = torch.from_numpy(s.mass).cuda()
    acc = torch.zeros_like(pos)
    out = {"gas_disk": {"n": n, "potential": "Miyamoto-Nagai M_d 5 a 3 b 0.3 + Plummer halo "
                                             "M_h 12 a_h 8"}}
    for prec in ("fp64", "fp32c"):
        ms = timed(lambda: gravity.add_acc_direct_ext(pos, mass, acc, n, G, 0.05 ** 2, pot, prec))
        gint = n * n / (ms * 1e-3) / 1e9
        peak = FP64_PEAK_TFLOPS if prec == "fp64" else FP32_PEAK_TFLOPS
        out["gas_disk"][f"add_grav_ext_{prec}"] = {
            "ms": ms, "Gint_per_s": gint,
            "frac_of_peak": gint * FLOP_PER_INTERACTION / 1e3 / peak}
        log(f"engines: add_grav_ext {prec} N={n}: {ms:.3f} ms, {gint:.0f} Gint/s")
    ms = timed(lambda: gravity.add_acc_analytic(pos, acc, n, G, pot))
    out["gas_disk"]["add_analytic"] = {"ms": ms, "GB_per_s": 72 * n / (ms * 1e-3) / 1e9}
    src = torch.ones_like(pos)
    ms = timed(lambda: gravity.add_field(acc, src, n))
    out["gas_disk"]["add_field"] = {"ms": ms, "GB_per_s": 72 * n / (ms * 1e-3) / 1e9}

    t = make_disk_bulge(n_disk=n_tree * 5 // 6, n_bulge=n_tree // 6, seed=1)
    pt = torch.from_numpy(t.pos).cuda()
    mt = torch.from_numpy(t.mass).cuda()
    at = torch.zeros_like(pt)
    acc_g = torch.empty_like(pt)
    bh = BarnesHut(t.n, theta=0.5, softening=0.05, precision="fp32")

    def split():
        bh.compute(pt, mt, acc_g)
        gravity.add_field(at, acc_g, t.n)

    out["living_galaxy_tree"] = {
        "n": t.n, "theta": 0.5,
        "compute_then_add_ms": timed(split),
        "compute_add_ms": timed(lambda: bh.compute_add(pt, mt, at))}
    log(f"engines: tree N={t.n}: compute+add {out['living_galaxy_tree']['compute_then_add_ms']:.3f}"
        f" ms, fused {out['living_galaxy_tree']['compute_add_ms']:.3f} ms")
    return out


def bench_grid(reps=20):
    """SPH neighbour-grid rebuild (core/solvers/neighbor_grid.py:52-92) on the
    disk + bulge N=120000 with half the particles gas: the device rebuild
    (which returns the gas count to the host, as the reference does) against
    the reference's numpy binning restated in the oracle, on the host."""
    import numpy as np
    import torch

    from oracle import np_oracle as orc
    from rgsim.ic import make_disk_bulge
    from rgsim.solvers.neighbor_grid import NeighborGrid

    s = make_disk_bulge(n_disk=100000, n_bulge=20000, seed=1)
    rng = np.random.default_rng(3)
    h = 0.1 * np.cbrt(rng.uniform(0.2, 1.0, s.n))
    st = (rng.random(s.n) < 0.5).astype(np.int32)
    pos = torch.from_numpy(s.pos).cuda()
    ht = torch.from_numpy(h).cuda()
    stt = torch.from_numpy(st).cuda()
    g = NeighborGrid(s.n)
    g.rebuild(pos, ht, stt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.rebuild(pos, ht, stt)
    wall = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(3):
        orc.grid_rebuild(s.pos, h, st)
    host = (time.perf_counter() - t0) / 3
    # algorithmic bytes: pos, h, is_star read twice (bounds, keys); keys + index
    # through the sort passes (read and write, 12 B each); gsort and cell_start
    passes = 3
    nbytes = s.n * (2 * (24 + 8 + 4) + passes * 2 * 12 + 4) + 4 * (g.max_cells + 1)
    out = {"n": s.n, "ngas": g.ngas, "cells": g.nx * g.ny * g.nz, "ms_per_rebuild": 1e3 * wall,
           "host_numpy_ms": 1e3 * host, "speedup_vs_host": host / wall,
           "algorithmic_bytes": nbytes, "GB_per_s": nbytes / wall / 1e9}
    log(f"grid: rebuild {out['ms_per_rebuild']:.3f} ms (host numpy {out['host_numpy_ms']:.1f} ms)")
    return out


def copy_calibration(mib=1024, reps=3):
    """A plain device copy of known size, for scaling FETCH_SIZE/WRITE_SIZE."""
    import torch
    a = torch.empty(mib * 1024 * 1024 // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": a.numel() * 4,
                      "reps": reps}))


def host_cpu_model():
    """The host CPU's model name with its family/model/stepping numbers
    (/proc/cpuinfo), or None."""
    fields = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    break  # first processor only
                k, _, v = line.partition(":")
                fields[k.strip()] = v.strip()
    except OSError:
        return None
    name = fields.get("model name")
    if name is None:
        return None
    ids = [f"{k} {fields[k]}" for k in ("cpu family", "model", "stepping") if k in fields]
    return f"{name} ({', '.join(ids)})" if ids else name


def host_cpus():
    """(cores this process may run on, cores the host has)."""
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return usable, os.cpu_count() or usable


def host_threads():
    """Host threads for the CPU baselines: every core this process may run
    on (its affinity mask), unless OMP_NUM_THREADS is set lower -- the GPU
    box sets its per-GPU CPU share there."""
    usable, _ = host_cpus()
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return min(int(env), usable)
    return usable


def host_cpu_info(threads):
    usable, nproc = host_cpus()
    env = os.environ.get("OMP_NUM_THREADS", "")
    return {"cpu_model": host_cpu_model(), "nproc": nproc, "affinity_cores": usable,
            "threads_used": threads,
            "threads_from": ("OMP_NUM_THREADS (the per-GPU CPU share)"
                             if env.isdigit() and int(env) > 0 and int(env) < usable
                             else "affinity mask")}


def cpu_baseline_direct(n=32768):
    """The oracle's C/OpenMP FP64 direct sum (gravity.py:13-26 restated,
    outer loop parallel) on a Plummer sample of n particles: one evaluation
    on the host threads and one on a single thread."""
    from oracle import np_oracle as orc
    from rgsim.ic import make_plummer

    orc.build()
    threads = host_threads()
    s = make_plummer(n, total_mass=1.0, scale_radius=1.0, seed=0, G=1.0)
    out = {}
    for label, t in (("all", threads), ("single_thread", 1)):
        orc.set_threads(t)
        sub = s.pos if t > 1 else s.pos[: n // 8]
        msub = s.mass if t > 1 else s.mass[: n // 8]
        t0 = time.perf_counter()
        orc.direct_acc(sub, msub, 1.0, 0.005 ** 2)
        w = time.perf_counter() - t0
        out[label] = {"value": len(msub) ** 2 / w / 1e9, "unit": "Ginteractions/s",
                      "sample": f"one FP64 evaluation at N={len(msub)}, {w:.2f} s on {t} thread(s)"}
    orc.set_threads(threads)
    return {"value": out["all"]["value"], "unit": "Ginteractions/s", "cores": threads,
            **host_cpu_info(threads), "kind": "port", "sample": out["all"]["sample"] + " (the oracle's C/OpenMP "
            "restatement of gravity.py:13-26)", "single_thread": out["single_thread"]}


def cpu_baseline(steps=10):
    """The oracle's C/OpenMP restatement of the reference BH step on the host
    cores: Morton + stable sort + build + aggregation + stack walk + KDK.
    `steps` KDK steps on all threads (about 1 s of wall time, 15-20 s of CPU
    work on 16 threads), plus one step on a single thread."""
    from oracle import np_oracle as orc
    from rgsim.ic import make_disk_bulge
    from rgsim.units import G

    orc.build()
    threads = host_threads()
    s = make_disk_bulge(n_disk=100000, n_bulge=20000, seed=1)

    def run(nsteps, nthreads):
        orc.set_threads(nthreads)
        pos, vel, mass = s.pos.copy(), s.vel.copy(), s.mass
        acc, _, _ = orc.bh_acc(pos, mass, G, 0.5, 0.05)
        dt = 1e-4
        t0 = time.perf_counter()
        for _ in range(nsteps):
            vel = orc.kick(vel, acc, 0.5 * dt)
            pos = orc.drift(pos, vel, dt)
            acc, _, _ = orc.bh_acc(pos, mass, G, 0.5, 0.05)
            vel = orc.kick(vel, acc, 0.5 * dt)
        return time.perf_counter() - t0

    wall = run(steps, threads)
    cores = orc.get_threads()
    wall1 = run(1, 1)
    orc.set_threads(threads)
    return {"value": s.n * steps / wall / 1e6, "unit": "Mparticle-steps/s",
            "cores": cores, **host_cpu_info(cores), "kind": "port",
            "sample": f"{steps} KDK steps of the FP64 BH oracle (C/OpenMP restatement of "
                      f"barnes_hut.py + leapfrog.py) on the same N={s.n} disk+bulge, "
                      f"{wall:.1f} s on {cores} threads",
            "single_thread": {"value": s.n / wall1 / 1e6, "unit": "Mparticle-steps/s",
                              "sample": f"1 KDK step on 1 thread, {wall1:.1f} s"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-direct", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="only the BH drop-in workload (no direct/adjoint/batch/1M sections)")
    ap.add_argument("--sections", default="bh,bh30,bhfp64,direct,adjoint,batch,bh1m,engines,grid",
                    help="comma-separated subset of bh,bh30,bhfp64,direct,adjoint,batch,bh1m,"
                         "engines,grid")
    ap.add_argument("--copy-calibration", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(a run covers one GPU's share of the work)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    if args.copy_calibration:
        copy_calibration()
        return

    sections = set(args.sections.split(","))
    if args.no_extra:
        sections &= {"bh", "bh30", "bhfp64"}
    if args.no_direct:
        sections.discard("direct")
    bh = direct = adjoint = batch = bh1m = engines = grid = None
    if "bh" in sections:
        log("bench: BH disk+bulge N=120000 ...")
        bh = bench_bh(args.steps, args.warmup)
        log(f"bench: BH {bh['mps']:.2f} Mparticle-steps/s, {bh['ms_per_step']:.3f} ms/step")
        if "bh30" in sections or "bhfp64" in sections:
            from rgsim.ic import make_disk_bulge
            dbs = make_disk_bulge(n_disk=100000, n_bulge=20000, seed=1)
        if "bh30" in sections:
            bh["morton_30bit"] = bench_bh30(dbs)
            log(f"bench: BH 30-bit keys {bh['morton_30bit']['ms_per_step']:.3f} ms/step")
        if "bhfp64" in sections:
            f64 = bench_bh_fp64(dbs)
            f64["speedup_of_fp32_walk"] = f64["ms_per_step"] / bh["ms_per_step"]
            bh["plain_fp64_gpu"] = f64
            log(f"bench: BH FP64 walk {f64['ms_per_step']:.3f} ms/step "
                f"({f64['speedup_of_fp32_walk']:.1f}x slower than the FP32 walk)")
    if "direct" in sections:
        log("bench: direct Plummer N=262144 ...")
        direct = bench_direct()
        if not args.no_cpu_baseline:
            direct["cpu_baseline_fp64"] = cpu_baseline_direct()
            log(f"bench: direct CPU FP64 {direct['cpu_baseline_fp64']['value']:.2f} Gint/s on "
                f"{direct['cpu_baseline_fp64']['cores']} threads")
    if "adjoint" in sections:
        log("bench: adjoint Plummer N=16384 ...")
        adjoint = bench_adjoint()
    if "batch" in sections:
        log("bench: batch 512 systems ...")
        batch = bench_batch()
    if "bh1m" in sections:
        log("bench: BH N=1M ...")
        bh1m = bench_bh_1m()
    if "engines" in sections:
        log("bench: engine self-gravity ...")
        engines = bench_engines()
    if "grid" in sections:
        log("bench: SPH neighbour grid ...")
        grid = bench_grid()
    if bh is None:
        print(json.dumps({"sections": sorted(sections), "direct_plummer_262144": direct,
                          "adjoint_plummer_16384": adjoint, "batch_512_systems": batch,
                          "bh_disk_bulge_1M": bh1m, "engines": engines,
                          "neighbor_grid": grid}))
        return
    cpu = None
    if not args.no_cpu_baseline:
        log("bench: CPU baseline ...")
        cpu = cpu_baseline()
        log(f"bench: CPU {cpu['value']:.3f} Mparticle-steps/s on {cpu['cores']} threads")

    line = {
        "metric": METRIC,
        "value": bh["mps"],
        "unit": "Mparticle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": bh["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": (bh["plain_fp64_gpu"]["ms_per_step"] / bh["ms_per_step"]
                        if bh.get("plain_fp64_gpu") else None),
        "vs_baseline_basis": "the plain FP64 GPU treecode (the reference's arithmetic, "
                             "criterion and pop order: FP64 aggregation and walk) on the same "
                             "workload in the same run, as the metric names",
        "dtype": "f32",
        "data": "synthetic (seeded disk+bulge and Plummer generators)",
        "config": {"workload": "BH drop-in: sech2 exp disk N=100000 + Plummer bulge N=20000, "
                               "theta 0.5, eps 0.05 kpc, 63-bit Morton, KDK dt 1e-4",
                   "n": bh["n"], "theta": 0.5, "softening": 0.05, "morton_bits": 21,
                   "walk_precision": "fp32", "aggregation_precision": "fp64",
                   "parallelism": "single GPU"},
        "roofline": bh["roofline"],
        "cpu_baseline": cpu,
        "bh": {k: v for k, v in bh.items() if k != "roofline"},
        "direct_plummer_262144": direct,
        "adjoint_plummer_16384": adjoint,
        "batch_512_systems": batch,
        "bh_disk_bulge_1M": bh1m,
        "engines": engines,
        "neighbor_grid": grid,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()

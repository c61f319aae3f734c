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
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "real-astronomy-simulation_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = ("direct Ginteractions/s & BH Mparticle-steps/s vs plain FP64 GPU, "
          "at stated force err")
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
FLOP_PER_INTERACTION = 20
WALK_RECORD_BYTES = 32          # FP32 node record fetched per visit
WALK_PARTICLE_BYTES = 16 + 4 + 24  # own leaf record + order index + FP64 acceleration


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def rel_err(a, ref):
    import numpy as np
    r = np.linalg.norm(a - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-300)
    return {"max": float(r.max()), "rms": float(np.sqrt(np.mean(r * r))),
            "mean": float(r.mean())}


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
    eng._bh.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.step(dt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    tm = eng._bh.timing()
    eng._bh.set_timing(False)
    st_last = eng._bh.stats()
    wall = t1 - t0
    walk_ms = tm["walk_ms"] / max(tm["count"], 1)
    visits = st_last["visits_total"]
    walk_bytes = visits * WALK_RECORD_BYTES + n * WALK_PARTICLE_BYTES
    achieved = walk_bytes / (walk_ms * 1e-3) / 1e9
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
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_walk32", "algorithmic_bytes_per_launch": walk_bytes,
                     "avg_launch_ms": walk_ms},
        "energy": None,
    }


def bench_direct(n=262144, evals=3):
    import numpy as np
    import torch

    from rgsim.ic import make_plummer
    from rgsim.solvers import gravity

    s = make_plummer(n, total_mass=1.0, scale_radius=1.0, seed=0, G=1.0)
    pos = torch.from_numpy(s.pos).cuda()
    mass = torch.from_numpy(s.mass).cuda()
    eps2 = 0.005 ** 2
    out = {}
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
        gint = n * n / (ms * 1e-3) / 1e9
        peak = FP64_PEAK_TFLOPS if prec == "fp64" else FP32_PEAK_TFLOPS
        tflops = gint * FLOP_PER_INTERACTION / 1e3
        out[prec] = {"ms_per_eval": ms, "Gint_per_s": gint, "TFLOP_s_20flop": tflops,
                     "frac_of_peak": tflops / peak,
                     "force_err_vs_fp64": rel_err(a, ref) if prec != "fp64" else None}
        log(f"direct {prec}: {ms:.2f} ms/eval, {gint:.1f} Gint/s")
    out["speedup_fp32c_vs_fp64"] = out["fp64"]["ms_per_eval"] / out["fp32c"]["ms_per_eval"]
    out["speedup_fp32_vs_fp64"] = out["fp64"]["ms_per_eval"] / out["fp32"]["ms_per_eval"]
    out["n"] = n
    return out


def cpu_baseline(steps=2):
    """The oracle's C/OpenMP restatement of the reference BH step on the host
    cores: Morton + stable sort + build + aggregation + stack walk + KDK."""
    from oracle import np_oracle as orc
    from rgsim.ic import make_disk_bulge
    from rgsim.units import G

    orc.build()
    threads = os.cpu_count() or 1
    orc.set_threads(threads)
    s = make_disk_bulge(n_disk=100000, n_bulge=20000, seed=1)
    pos, vel, mass = s.pos.copy(), s.vel.copy(), s.mass
    acc, _, _ = orc.bh_acc(pos, mass, G, 0.5, 0.05)
    dt = 1e-4
    t0 = time.perf_counter()
    for _ in range(steps):
        vel = orc.kick(vel, acc, 0.5 * dt)
        pos = orc.drift(pos, vel, dt)
        acc, _, _ = orc.bh_acc(pos, mass, G, 0.5, 0.05)
        vel = orc.kick(vel, acc, 0.5 * dt)
    wall = time.perf_counter() - t0
    return {"value": s.n * steps / wall / 1e6, "unit": "Mparticle-steps/s",
            "cores": orc.get_threads(), "kind": "port",
            "sample": f"{steps} KDK steps of the FP64 BH oracle (C/OpenMP restatement of "
                      f"barnes_hut.py + leapfrog.py) on the same N={s.n} disk+bulge, "
                      f"{wall:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-direct", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)

    log("bench: BH disk+bulge N=120000 ...")
    bh = bench_bh(args.steps, args.warmup)
    log(f"bench: BH {bh['mps']:.2f} Mparticle-steps/s, {bh['ms_per_step']:.3f} ms/step")
    direct = None
    if not args.no_direct:
        log("bench: direct Plummer N=262144 ...")
        direct = bench_direct()
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
        "vs_baseline": None,
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
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()

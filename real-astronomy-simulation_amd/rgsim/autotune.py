"""Launch-shape autotuner: times candidate shapes on the caller's actual
workload with HIP events (torch.cuda.Event on the launch stream) and keeps
the fastest.  No reference counterpart (the reference runs Taichi's default
block sizes; it sets no block_dim).

* ``tune_batch(ens)``  tile (i-particles per work item: one per thread up
  to 256, then 2 or 4 per thread on 256 threads) x persistent grid size of
  the batched force kernel.
* ``tune_direct(pos, mass, ...)``  i-particles per lane x j-slices of the
  direct kernels (FP64 sums are never split, so only i-particles per lane
  are tried for them).
* ``tune_walk(bh, pos, mass, acc)``  threads per block of the BH walk.

Results can be cached in a JSON file keyed by device and workload.
"""
from __future__ import annotations

import json
import os

import torch

from . import _lib


def _time(fn, reps):
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _cache_load(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def _cache_store(path, cache):
    if path:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        with open(path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)


def _device_key():
    return torch.cuda.get_device_name(torch.cuda.current_device())


def tune_batch(ens, tiles=(64, 128, 256, 512, 1024), grids_per_cu=(1, 2, 4, 8), reps=2,
               cache_path=None):
    key = f"batch|{_device_key()}|{ens.nsys}|{ens.ntot}|{ens.interactions_per_step()}|{ens.precision}"
    cache = _cache_load(cache_path)
    if key in cache:
        best = cache[key]
        ens.set_shape(best["tile"], best["grid"])
        return best
    cus = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
    trials = []
    for tile in tiles:
        for g in grids_per_cu:
            ens.set_shape(tile, g * cus)
            ms = _time(ens.compute_acc, reps)
            trials.append({"tile": tile, "grid": g * cus, "ms": ms})
    best = min(trials, key=lambda t: t["ms"])
    ens.set_shape(best["tile"], best["grid"])
    best = dict(best, trials=trials)
    cache[key] = best
    _cache_store(cache_path, cache)
    return best


def tune_direct(pos, mass, G, eps2, precision, ipts=(1, 2, 4), splits=(0, 8, 16, 32, 64),
                reps=2):
    from .solvers import gravity
    n = pos.shape[0]
    acc = torch.empty_like(pos)
    if precision == "fp64":
        splits = (0,)
    trials = []
    for ipt in ipts:
        _lib.set_param("direct_ipt", ipt)
        for split in splits:
            _lib.set_param("direct_split", split)
            ms = _time(lambda: gravity.compute_acc_direct(pos, mass, acc, n, G, eps2, precision),
                       reps)
            trials.append({"ipt": ipt, "split": split, "ms": ms})
    best = min(trials, key=lambda t: t["ms"])
    _lib.set_param("direct_ipt", best["ipt"])
    _lib.set_param("direct_split", best["split"])
    return dict(best, trials=trials)


def tune_walk(bh, pos, mass, acc, blocks=(64, 128, 256, 512), reps=3):
    trials = []
    for b in blocks:
        _lib.set_param("walk_block", b)
        bh.set_timing(True)
        for _ in range(reps):
            bh.compute(pos, mass, acc)
        tm = bh.timing()
        bh.set_timing(False)
        trials.append({"walk_block": b, "walk_ms": tm["walk_ms"] / max(tm["count"], 1)})
    best = min(trials, key=lambda t: t["walk_ms"])
    _lib.set_param("walk_block", best["walk_block"])
    return dict(best, trials=trials)

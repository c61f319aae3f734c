"""LBVH Barnes-Hut treecode over the HIP C ABI, mirroring
core/solvers/barnes_hut.py:220-259.

``BarnesHut(n, theta, softening, max_stack)`` owns the device scratch for n
particles and ``compute(pos, mass, acc)`` overwrites acc.  Morton keys,
the stable sort, the Karras build, the aggregation and the walk all run on
the device; nothing is copied to the host.  The walk is stackless, so
``max_stack`` is accepted for signature parity and has no limit to hit
(the reference silently drops children when its 64-entry stack is full,
barnes_hut.py:212).

Extra keyword arguments: ``G`` (the reference uses core.units.G),
``precision`` ("fp64": the reference's FP64 criterion and arithmetic;
"fp32": FP32 node records; there is no compensated walk, so "fp32c" is
rejected), ``morton_bits`` (21 -> 63-bit keys as the reference; 10 ->
30-bit keys).

Failures are never silent: a compute whose radix-sort look-back wait gave
up (the ``sort_spin_limit`` bound) computes nothing after the sort, and the
next call on the handle -- or ``stats()``, which synchronises -- raises.
"""
from __future__ import annotations

import ctypes

import torch

from .. import _lib
from .._tensors import vec1, vec3
from ..units import G as G_GAL


def morton_codes(pos, bits: int = 21):
    """Device twin of morton_codes (barnes_hut.py:38-46): uint64 keys as an
    int64 tensor (keys are < 2^63)."""
    n = pos.shape[0]
    vec3(pos, n, "pos")
    _lib.require_gpu()
    keys = torch.empty(n, dtype=torch.int64, device=pos.device)
    _lib.check(_lib.load().rg_morton_codes(pos.data_ptr(), n, int(bits), keys.data_ptr(),
                                           _lib.stream_ptr(pos)), "rg_morton_codes")
    return keys


class BarnesHut:
    def __init__(self, n: int, theta: float = 0.6, softening: float = 0.1,
                 max_stack: int = 64, *, G: float = G_GAL, precision: str = "fp64",
                 morton_bits: int = 21):
        _lib.require_gpu()
        self.n = int(n)
        self.theta = float(theta)
        self.theta2 = self.theta ** 2
        self.softening = float(softening)
        self.eps2 = self.softening ** 2
        self.max_stack = int(max_stack)
        self.G = float(G)
        self.precision = precision
        self.morton_bits = int(morton_bits)
        self.device = torch.cuda.current_device()
        self._h = ctypes.c_void_p()
        _lib.check(_lib.load().rg_bh_create(self.n, self.theta, self.softening, self.G,
                                            self.morton_bits,
                                            _lib.precision_code(precision),
                                            ctypes.byref(self._h)), "rg_bh_create")

    def compute(self, pos, mass, acc) -> None:
        """Fill acc with the treecode acceleration for pos (barnes_hut.py:242)."""
        n = self.n
        vec3(pos, n, "pos"), vec1(mass, n, "mass"), vec3(acc, n, "acc")
        _lib.check(_lib.load().rg_bh_compute(self._h, pos.data_ptr(), mass.data_ptr(),
                                             acc.data_ptr(), _lib.stream_ptr(pos)),
                   "rg_bh_compute")

    def compute_kick(self, pos, mass, acc, vel, dt_half) -> None:
        """compute(pos, mass, acc) then vel += acc * dt_half (leapfrog.py:11-15),
        the kick fused into the walk: the closing half of a KDK step."""
        n = self.n
        vec3(pos, n, "pos"), vec1(mass, n, "mass"), vec3(acc, n, "acc"), vec3(vel, n, "vel")
        _lib.check(_lib.load().rg_bh_compute_kick(self._h, pos.data_ptr(), mass.data_ptr(),
                                                  acc.data_ptr(), vel.data_ptr(),
                                                  float(dt_half), _lib.stream_ptr(pos)),
                   "rg_bh_compute_kick")

    def step(self, pos, vel, mass, acc, dt_half, dt) -> None:
        """One KDK step of the tree path (core/engine.py:108-116): kick, drift,
        compute(pos, mass, acc), kick -- bit-identical to
        leapfrog.kick_drift followed by compute_kick, with the opening
        kick + drift pass also boxing the positions for the tree.  acc holds
        the previous force on entry and the new one after."""
        n = self.n
        vec3(pos, n, "pos"), vec3(vel, n, "vel"), vec1(mass, n, "mass"), vec3(acc, n, "acc")
        _lib.check(_lib.load().rg_bh_step(self._h, pos.data_ptr(), vel.data_ptr(),
                                          mass.data_ptr(), acc.data_ptr(), float(dt_half),
                                          float(dt), _lib.stream_ptr(pos)),
                   "rg_bh_step")

    def compute_add(self, pos, mass, acc) -> None:
        """acc += the treecode acceleration: compute(pos, mass, acc_g) then
        acc += acc_g, as living_galaxy_engine.py:1017-1018 and
        gas_disk_engine.py:218-219 do, with the add fused into the walk."""
        n = self.n
        vec3(pos, n, "pos"), vec1(mass, n, "mass"), vec3(acc, n, "acc")
        _lib.check(_lib.load().rg_bh_compute_add(self._h, pos.data_ptr(), mass.data_ptr(),
                                                 acc.data_ptr(), _lib.stream_ptr(pos)),
                   "rg_bh_compute_add")

    def export_tree(self):
        """(keys_sorted, order, left, right, parent) of the last compute, as
        device tensors in the reference's encoding."""
        n = self.n
        dev = torch.device("cuda", self.device)
        keys = torch.empty(n, dtype=torch.int64, device=dev)
        order = torch.empty(n, dtype=torch.int32, device=dev)
        left = torch.empty(n - 1, dtype=torch.int32, device=dev)
        right = torch.empty(n - 1, dtype=torch.int32, device=dev)
        parent = torch.empty(2 * n - 1, dtype=torch.int32, device=dev)
        _lib.check(_lib.load().rg_bh_export_tree(self._h, keys.data_ptr(), order.data_ptr(),
                                                 left.data_ptr(), right.data_ptr(),
                                                 parent.data_ptr(), _lib.stream_ptr(keys)),
                   "rg_bh_export_tree")
        return keys, order, left, right, parent

    def export_nodes(self):
        """(com, nmass, nmin, nmax) FP64 node data of the last compute."""
        m = 2 * self.n - 1
        dev = torch.device("cuda", self.device)
        com = torch.empty((m, 3), dtype=torch.float64, device=dev)
        nmass = torch.empty(m, dtype=torch.float64, device=dev)
        nmin = torch.empty((m, 3), dtype=torch.float64, device=dev)
        nmax = torch.empty((m, 3), dtype=torch.float64, device=dev)
        _lib.check(_lib.load().rg_bh_export_nodes(self._h, com.data_ptr(), nmass.data_ptr(),
                                                  nmin.data_ptr(), nmax.data_ptr(),
                                                  _lib.stream_ptr(com)),
                   "rg_bh_export_nodes")
        return com, nmass, nmin, nmax

    def stats(self) -> dict:
        """Walk statistics of the last compute (synchronises its stream);
        raises if that compute aborted or its tree was corrupt."""
        s = _lib.BHStats()
        _lib.check(_lib.load().rg_bh_stats(self._h, ctypes.byref(s)), "rg_bh_stats")
        return {"n": s.n, "visits_total": s.visits_total, "visits_max": s.visits_max,
                "nonfinite": s.nonfinite, "walk_mode": s.walk_mode, "walk_waves": s.walk_waves,
                "sort_passes": s.sort_passes, "sort_buckets": s.sort_buckets,
                "sort_spilled": s.sort_spilled}

    def set_timing(self, enable: bool) -> None:
        """Record HIP events around each phase of every compute."""
        _lib.check(_lib.load().rg_bh_set_timing(self._h, int(bool(enable))), "rg_bh_set_timing")

    def timing(self) -> dict:
        """Summed per-phase milliseconds since the last read (waits for them)."""
        ms = (ctypes.c_double * 4)()
        cnt = ctypes.c_int64()
        _lib.check(_lib.load().rg_bh_timing(self._h, ms, ctypes.byref(cnt)), "rg_bh_timing")
        return {"count": cnt.value, "sort_ms": ms[0], "build_ms": ms[1],
                "aggregate_ms": ms[2], "walk_ms": ms[3]}

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().rg_bh_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

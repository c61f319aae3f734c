"""Batching form: many independent systems of uneven N in one persistent
launch per step, optionally replayed from a hipGraph (the C ABI's
rg_batch_*).  There is no reference counterpart: the reference advances one
system per Engine (core/engine.py:32-116); each system here follows that
Engine's KDK step and direct force law with its own G, softening and dt.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .ic.plummer import make_plummer


def plummer_ensemble(nsys=512, n_min=256, n_max=16384, master_seed=2026, G=1.0):
    """Seeded ensemble: N_k = round(exp(U[ln n_min, ln n_max])), a_k ~ U[0.5, 2],
    eps_k = 0.02 a_k, M_k = 1, seed_k = k, dt_k = 1e-3 a_k^{3/2}."""
    rng = np.random.default_rng(master_seed)
    ns = np.round(np.exp(rng.uniform(np.log(n_min), np.log(n_max), nsys))).astype(int)
    a = rng.uniform(0.5, 2.0, nsys)
    systems = []
    for k in range(nsys):
        s = make_plummer(int(ns[k]), total_mass=1.0, scale_radius=float(a[k]), seed=k, G=G)
        systems.append({"pos": s.pos, "vel": s.vel, "mass": s.mass, "G": G,
                        "eps": 0.02 * a[k], "dt": 1e-3 * a[k] ** 1.5})
    return systems


class Ensemble:
    """Concatenated device state of many systems plus the batch handle."""

    def __init__(self, systems, precision="fp32c", device=None):
        _lib.require_gpu()
        dev = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        self.nsys = len(systems)
        self.sizes = np.array([len(s["mass"]) for s in systems], dtype=np.int64)
        self.offsets = np.concatenate(([0], np.cumsum(self.sizes))).astype(np.int64)
        self.G = np.array([s["G"] for s in systems], dtype=np.float64)
        self.eps2 = np.array([s["eps"] ** 2 for s in systems], dtype=np.float64)
        self.dt = np.array([s["dt"] for s in systems], dtype=np.float64)
        self.precision = precision
        cat = lambda key: torch.from_numpy(np.ascontiguousarray(
            np.concatenate([s[key] for s in systems]), dtype=np.float64)).to(dev)
        self.pos, self.vel, self.mass = cat("pos"), cat("vel"), cat("mass")
        self.acc = torch.empty_like(self.pos)
        self._h = ctypes.c_void_p()
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        _lib.check(_lib.load().rg_batch_create(self.nsys, P(self.offsets), P(self.G),
                                               P(self.eps2), P(self.dt),
                                               _lib.precision_code(precision),
                                               ctypes.byref(self._h)), "rg_batch_create")
        self.compute_acc()

    @property
    def ntot(self) -> int:
        return int(self.offsets[-1])

    def _s(self):
        return _lib.stream_ptr(self.pos)

    def set_shape(self, tile: int, grid: int = 0) -> None:
        _lib.check(_lib.load().rg_batch_set_shape(self._h, int(tile), int(grid)),
                   "rg_batch_set_shape")

    def compute_acc(self) -> None:
        _lib.check(_lib.load().rg_batch_acc(self._h, self.pos.data_ptr(), self.mass.data_ptr(),
                                            self.acc.data_ptr(), self._s()), "rg_batch_acc")

    def step(self, nsteps: int = 1) -> None:
        _lib.check(_lib.load().rg_batch_step(self._h, self.pos.data_ptr(), self.vel.data_ptr(),
                                             self.acc.data_ptr(), self.mass.data_ptr(),
                                             int(nsteps), self._s()), "rg_batch_step")

    def capture(self, nsteps: int) -> None:
        """Capture nsteps steps into a hipGraph (replayed by replay())."""
        torch.cuda.synchronize()
        _lib.check(_lib.load().rg_batch_graph_capture(self._h, self.pos.data_ptr(),
                                                      self.vel.data_ptr(), self.acc.data_ptr(),
                                                      self.mass.data_ptr(), int(nsteps)),
                   "rg_batch_graph_capture")

    def replay(self) -> None:
        _lib.check(_lib.load().rg_batch_graph_launch(self._h, self._s()), "rg_batch_graph_launch")

    def energies(self) -> np.ndarray:
        """(nsys, 5) FP64 {KE, PE, Px, Py, Pz} per system."""
        out = torch.empty((self.nsys, 5), dtype=torch.float64, device=self.pos.device)
        _lib.check(_lib.load().rg_batch_energy(self._h, self.pos.data_ptr(), self.vel.data_ptr(),
                                               self.mass.data_ptr(), out.data_ptr(), self._s()),
                   "rg_batch_energy")
        return out.cpu().numpy()

    def per_system_step(self, nsteps: int = 1) -> None:
        """The same KDK steps issued as separate launches per system (the
        one-Engine-per-system baseline)."""
        from .integrators import leapfrog
        from .solvers import gravity
        for _ in range(int(nsteps)):
            for k in range(self.nsys):
                a, b = int(self.offsets[k]), int(self.offsets[k + 1])
                x, v, acc, m = self.pos[a:b], self.vel[a:b], self.acc[a:b], self.mass[a:b]
                dt = float(self.dt[k])
                leapfrog.kick_drift(x, v, acc, b - a, 0.5 * dt, dt)
                gravity.compute_acc_direct_kick(x, m, acc, v, b - a, float(self.G[k]),
                                                float(self.eps2[k]), 0.5 * dt, self.precision)

    def interactions_per_step(self) -> int:
        return int((self.sizes.astype(np.float64) ** 2).sum())

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().rg_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

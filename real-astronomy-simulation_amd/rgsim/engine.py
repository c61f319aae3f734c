"""Drop-in ``Engine`` for core/engine.py:32-124.

Same constructor, methods and attributes as the reference's Engine
(``Engine(softening, gravity_mode, theta)``; ``load_state``, ``to_state``,
``step``, ``energies``; ``n``, ``time``, ``step_count``, ``seed``, ``meta``,
``softening``, ``theta``, ``gravity_mode``, ``_bh``).  Fields live in torch
float64 CUDA tensors and every kernel runs through the HIP C ABI:

* Barnes-Hut is used iff ``gravity_mode == "bh" and n >= 256``
  (core/engine.py:75); otherwise ``_bh is None`` and the direct sum runs.
* ``step(dt)`` is the reference's KDK order (core/engine.py:108-116): the
  opening kick and the drift are one fused pass (which, for BH, also boxes
  the positions for the tree), and the closing kick is fused into the force
  kernel (direct) or the tree walk's store (BH).  The fusions are
  bit-identical to the separate reference kernels.
* ``energies()`` is FP64 (KE, PE, KE + PE) like core/engine.py:119-124.

Extra keyword arguments: ``precision`` ("fp64" default = the reference's
arithmetic, "fp32", "fp32c"; the tree walk has no compensated form, so
``gravity_mode="bh"`` takes "fp64" or "fp32"), ``G`` (default
core.units.G), ``morton_bits``.

With ``precision="fp32c"`` the engine holds each position as an FP64 value
plus an FP32 offset (``pos`` and ``pos_lo``): the drift adds v dt into the
pair with a compensated sum, the force kernels read both parts, and
``to_state`` returns their sum rounded to FP64.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .integrators import leapfrog
from .solvers import gravity
from .solvers.barnes_hut import BarnesHut
from .state import State
from .units import G as G_GAL

BH_MIN_N = 256


def init_taichi(arch: str = "cuda") -> None:
    """Compatibility shim for core/engine.py:21-29: there is no runtime to
    initialise beyond selecting the GPU; ``arch`` other than a GPU is an
    error because this solver has no CPU path."""
    if arch not in ("cuda", "gpu", "hip", "rocm"):
        raise RuntimeError(f"arch={arch!r}: rgsim runs on the GPU only")
    _lib.require_gpu()


class Engine:
    def __init__(self, softening: float = 0.05, gravity_mode: str = "direct",
                 theta: float = 0.6, *, precision: str = "fp64", G: float = G_GAL,
                 morton_bits: int = 21, check_finite: bool = False):
        self.softening = float(softening)
        self.gravity_mode = gravity_mode
        self.theta = float(theta)
        self.precision = precision
        self.G = float(G)
        self.morton_bits = int(morton_bits)
        self.check_finite = bool(check_finite)
        _lib.precision_code(precision)
        if gravity_mode == "bh" and precision == "fp32c":
            raise ValueError("the tree walk has no compensated form: gravity_mode='bh' takes "
                             "precision 'fp64' or 'fp32'")
        self._bh = None
        self.n = 0
        self.time = 0.0
        self.step_count = 0
        self.seed = 0
        self.meta: dict = {}
        self.pos = self.vel = self.acc = self.mass = None
        self.pos_lo = None  # FP32 position offsets (precision "fp32c")
        self._ptype = self._ids = None
        self._age = self._temperature = self._metallicity = None

    # ------------------------------------------------------------- setup
    def load_state(self, state: State) -> None:
        """Upload a State and prime the accelerations (core/engine.py:53-81)."""
        _lib.require_gpu()
        self.n = state.n
        self.time = state.time
        self.step_count = state.step
        self.seed = state.seed
        self.meta = dict(state.meta)
        self._ptype = state.ptype.copy()
        self._ids = state.ids.copy()
        self._age = state.age
        self._temperature = state.temperature
        self._metallicity = state.metallicity
        dev = torch.device("cuda", torch.cuda.current_device())
        self.pos = torch.from_numpy(np.ascontiguousarray(state.pos, np.float64)).to(dev)
        self.vel = torch.from_numpy(np.ascontiguousarray(state.vel, np.float64)).to(dev)
        self.mass = torch.from_numpy(np.ascontiguousarray(state.mass, np.float64)).to(dev)
        self.acc = torch.empty_like(self.pos)
        self.pos_lo = None
        if self.precision == "fp32c":
            lo = state.pos_lo if state.pos_lo is not None else np.zeros(state.pos.shape,
                                                                        np.float32)
            self.pos_lo = torch.from_numpy(np.ascontiguousarray(lo, np.float32)).to(dev)
        if self.gravity_mode == "bh" and self.n >= BH_MIN_N:
            if self._bh is None or self._bh.n != self.n:
                self._bh = BarnesHut(self.n, theta=self.theta, softening=self.softening,
                                     G=self.G, precision=self.precision,
                                     morton_bits=self.morton_bits)
        else:
            self._bh = None
        self._compute_acc()

    def to_state(self) -> State:
        """Download the state (core/engine.py:83-98).  ``meta`` also records
        this engine's solver settings (``engine_*`` keys), so that
        ``Engine.from_state`` resumes with the same precision and tree."""
        meta = dict(self.meta)
        meta.update(self.settings())
        if self._bh is not None:
            self._bh.stats()  # raises if a compute since the last check aborted
        return State(pos=self.pos.cpu().numpy(), vel=self.vel.cpu().numpy(),
                     mass=self.mass.cpu().numpy(), ptype=self._ptype, ids=self._ids,
                     time=self.time, step=self.step_count, seed=self.seed, age=self._age,
                     temperature=self._temperature, metallicity=self._metallicity,
                     meta=meta,
                     pos_lo=self.pos_lo.cpu().numpy() if self.pos_lo is not None else None)

    def settings(self) -> dict:
        """The constructor arguments, as snapshot meta entries."""
        return {"engine_softening": self.softening, "engine_gravity_mode": self.gravity_mode,
                "engine_theta": self.theta, "engine_precision": self.precision,
                "engine_G": self.G, "engine_morton_bits": self.morton_bits}

    @classmethod
    def from_state(cls, state: State, **overrides) -> "Engine":
        """An engine built with the settings a snapshot recorded (keyword
        arguments override them), loaded with the snapshot.  A snapshot
        without them gets the constructor defaults."""
        m = state.meta
        kw = {"softening": float(m.get("engine_softening", 0.05)),
              "gravity_mode": str(m.get("engine_gravity_mode", "direct")),
              "theta": float(m.get("engine_theta", 0.6)),
              "precision": str(m.get("engine_precision", "fp64")),
              "G": float(m.get("engine_G", G_GAL)),
              "morton_bits": int(m.get("engine_morton_bits", 21))}
        kw.update(overrides)
        eng = cls(**kw)
        eng.load_state(state)
        return eng

    # ---------------------------------------------------------- stepping
    def _compute_acc(self) -> None:
        if self._bh is not None:
            self._bh.compute(self.pos, self.mass, self.acc)
            if self.check_finite:
                self._check_bh()
        elif self.pos_lo is not None:
            gravity.compute_acc_direct_hilo(self.pos, self.pos_lo, self.mass, self.acc, self.n,
                                            self.G, self.softening ** 2)
        else:
            gravity.compute_acc_direct(self.pos, self.mass, self.acc, self.n, self.G,
                                       self.softening ** 2, self.precision)

    def _check_bh(self) -> None:
        st = self._bh.stats()
        if st["nonfinite"]:
            raise FloatingPointError(f"{st['nonfinite']} non-finite coordinates in the input")

    def step(self, dt: float) -> None:
        """One KDK step (core/engine.py:108-116)."""
        half = 0.5 * dt
        if self._bh is not None:
            # kick + drift (boxing the positions for the tree), force, and the
            # closing half-kick in the walk's store
            self._bh.step(self.pos, self.vel, self.mass, self.acc, half, dt)
            if self.check_finite:
                self._check_bh()
        elif self.pos_lo is not None:
            # hi/lo positions: compensated drift, FP32C force on both parts
            leapfrog.kick_drift_hilo(self.pos, self.pos_lo, self.vel, self.acc, self.n, half, dt)
            gravity.compute_acc_direct_kick_hilo(self.pos, self.pos_lo, self.mass, self.acc,
                                                 self.vel, self.n, self.G, self.softening ** 2,
                                                 half)
        else:
            leapfrog.kick_drift(self.pos, self.vel, self.acc, self.n, half, dt)
            gravity.compute_acc_direct_kick(self.pos, self.mass, self.acc, self.vel, self.n,
                                            self.G, self.softening ** 2, half, self.precision)
        self.time += dt
        self.step_count += 1

    # ------------------------------------------------------- diagnostics
    def energies(self) -> tuple[float, float, float]:
        """(kinetic, potential, total), FP64 (core/engine.py:119-124)."""
        ke, pe, _ = gravity.energy_terms(self.pos, self.vel, self.mass, self.n, self.G,
                                         self.softening ** 2)
        if self._bh is not None:
            self._bh.stats()  # raises if a compute since the last check aborted
        return ke, pe, ke + pe

    def momentum(self):
        """Total momentum sum m v (FP64)."""
        return gravity.energy_terms(self.pos, self.vel, self.mass, self.n, self.G,
                                    self.softening ** 2, with_pe=False)[2]

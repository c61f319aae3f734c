"""Simulation state container, as core/state.py:25-59.

The dtype/shape normalisation in ``__post_init__`` is the drop-in's input
contract (core/state.py:53-59): positions and velocities (N,3) float64,
masses (N,) float64, ptype int32, ids int64, all C-contiguous.  HDF5
save/load (core/state.py:62-110) is checkpoint I/O and not part of the
gravity path; ``save_npz``/``load_npz`` keep a lossless copy instead.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

PTYPE_DM = 0
PTYPE_STAR = 1
PTYPE_GAS = 2


@dataclass
class State:
    pos: np.ndarray
    vel: np.ndarray
    mass: np.ndarray
    ptype: np.ndarray
    ids: np.ndarray
    time: float = 0.0
    step: int = 0
    seed: int = 0
    age: Optional[np.ndarray] = None
    temperature: Optional[np.ndarray] = None
    metallicity: Optional[np.ndarray] = None
    u: Optional[np.ndarray] = None
    rho: Optional[np.ndarray] = None
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return self.pos.shape[0]

    def __post_init__(self) -> None:
        self.pos = np.ascontiguousarray(self.pos, dtype=np.float64).reshape(-1, 3)
        self.vel = np.ascontiguousarray(self.vel, dtype=np.float64).reshape(-1, 3)
        self.mass = np.ascontiguousarray(self.mass, dtype=np.float64).reshape(-1)
        self.ptype = np.ascontiguousarray(self.ptype, dtype=np.int32).reshape(-1)
        self.ids = np.ascontiguousarray(self.ids, dtype=np.int64).reshape(-1)

    def save_npz(self, path: str) -> None:
        extra = {k: getattr(self, k) for k in ("age", "temperature", "metallicity", "u", "rho")
                 if getattr(self, k) is not None}
        np.savez(path, pos=self.pos, vel=self.vel, mass=self.mass, ptype=self.ptype,
                 ids=self.ids, time=self.time, step=self.step, seed=self.seed, **extra)

    @classmethod
    def load_npz(cls, path: str) -> "State":
        with np.load(path, allow_pickle=False) as f:
            opt = {k: f[k] for k in ("age", "temperature", "metallicity", "u", "rho") if k in f}
            return cls(pos=f["pos"], vel=f["vel"], mass=f["mass"], ptype=f["ptype"],
                       ids=f["ids"], time=float(f["time"]), step=int(f["step"]),
                       seed=int(f["seed"]), **opt)

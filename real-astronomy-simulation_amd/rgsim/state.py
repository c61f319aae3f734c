"""Simulation state container, as core/state.py:25-59.

The dtype/shape normalisation in ``__post_init__`` is the drop-in's input
contract (core/state.py:53-59): positions and velocities (N,3) float64,
masses (N,) float64, ptype int32, ids int64, all C-contiguous.

Snapshots:

* ``save``/``load`` write and read the reference's HDF5 layout
  (core/state.py:62-110): attrs time/step/seed/n/units and ``meta_<key>``,
  gzip datasets pos/vel/mass/ptype/ids plus any optional per-particle
  array.  They need h5py, imported on use as the reference does.
* ``save_npz``/``load_npz`` keep the same fields, meta included (as JSON),
  with numpy only.

Both are lossless for the float64 fields, so a run resumed from a snapshot
continues bit for bit; the Engine stores its solver settings (precision,
tree parameters) in ``meta`` so that ``Engine.from_state`` rebuilds the same
solver.

``pos_lo`` (optional, (N,3) float32; not a field of the reference's State)
holds the FP32 offsets of an FP32C engine's hi/lo positions, X = pos +
pos_lo; ``pos`` alone is X rounded to FP64.  Both snapshot formats keep it,
so an FP32C run resumes bit for bit too.
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

PTYPE_DM = 0
PTYPE_STAR = 1
PTYPE_GAS = 2

OPTIONAL = ("age", "temperature", "metallicity", "u", "rho")


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
    pos_lo: Optional[np.ndarray] = None

    @property
    def n(self) -> int:
        return self.pos.shape[0]

    def __post_init__(self) -> None:
        self.pos = np.ascontiguousarray(self.pos, dtype=np.float64).reshape(-1, 3)
        self.vel = np.ascontiguousarray(self.vel, dtype=np.float64).reshape(-1, 3)
        self.mass = np.ascontiguousarray(self.mass, dtype=np.float64).reshape(-1)
        self.ptype = np.ascontiguousarray(self.ptype, dtype=np.int32).reshape(-1)
        self.ids = np.ascontiguousarray(self.ids, dtype=np.int64).reshape(-1)
        if self.pos_lo is not None:
            self.pos_lo = np.ascontiguousarray(self.pos_lo, dtype=np.float32).reshape(-1, 3)
            if self.pos_lo.shape != self.pos.shape:
                raise ValueError(f"pos_lo must have pos's shape {self.pos.shape}")

    # ------------------------------------------------------------- HDF5
    def save(self, path: str) -> None:
        """The reference's HDF5 snapshot (core/state.py:62-87)."""
        import h5py

        with h5py.File(path, "w") as f:
            f.attrs["time"] = float(self.time)
            f.attrs["step"] = int(self.step)
            f.attrs["seed"] = int(self.seed)
            f.attrs["n"] = int(self.n)
            f.attrs["units"] = "kpc, km/s, 1e10 Msun"
            for key, val in self.meta.items():
                # attrs take scalars and strings; anything else is stored as text
                f.attrs[f"meta_{key}"] = val if np.isscalar(val) else str(val)
            for name in ("pos", "vel", "mass", "ptype", "ids"):
                f.create_dataset(name, data=getattr(self, name), compression="gzip")
            for name in OPTIONAL:
                arr = getattr(self, name)
                if arr is not None:
                    f.create_dataset(name, data=np.asarray(arr, dtype=np.float64),
                                     compression="gzip")
            if self.pos_lo is not None:
                f.create_dataset("pos_lo", data=self.pos_lo, compression="gzip")

    @classmethod
    def load(cls, path: str) -> "State":
        """Read a snapshot written by ``save`` (core/state.py:89-110)."""
        import h5py

        with h5py.File(path, "r") as f:
            meta = {k[5:]: _plain(f.attrs[k]) for k in f.attrs if k.startswith("meta_")}
            opt = {name: f[name][:] for name in OPTIONAL + ("pos_lo",) if name in f}
            return cls(pos=f["pos"][:], vel=f["vel"][:], mass=f["mass"][:],
                       ptype=f["ptype"][:], ids=f["ids"][:], time=float(f.attrs["time"]),
                       step=int(f.attrs["step"]), seed=int(f.attrs["seed"]), meta=meta, **opt)

    # -------------------------------------------------------------- npz
    def save_npz(self, path: str) -> None:
        """The same fields with numpy only; meta as a JSON string."""
        extra = {k: getattr(self, k) for k in OPTIONAL + ("pos_lo",)
                 if getattr(self, k) is not None}
        np.savez(path, pos=self.pos, vel=self.vel, mass=self.mass, ptype=self.ptype,
                 ids=self.ids, time=self.time, step=self.step, seed=self.seed,
                 meta=np.array(json.dumps(self.meta, default=str)), **extra)

    @classmethod
    def load_npz(cls, path: str) -> "State":
        with np.load(path, allow_pickle=False) as f:
            opt = {k: f[k] for k in OPTIONAL + ("pos_lo",) if k in f}
            meta = json.loads(str(f["meta"])) if "meta" in f else {}
            return cls(pos=f["pos"], vel=f["vel"], mass=f["mass"], ptype=f["ptype"],
                       ids=f["ids"], time=float(f["time"]), step=int(f["step"]),
                       seed=int(f["seed"]), meta=meta, **opt)


def _plain(v):
    """HDF5 attribute value -> Python scalar or str."""
    if isinstance(v, bytes):
        return v.decode()
    if isinstance(v, np.generic):
        return v.item()
    return v

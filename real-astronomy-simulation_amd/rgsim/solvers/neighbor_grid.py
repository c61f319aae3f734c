"""Drop-in ``NeighborGrid`` for core/solvers/neighbor_grid.py:36-136: the
uniform grid over the SPH gas particles, binned on the device.

Same constructor ``NeighborGrid(n, min_radius=0.0)``, methods ``rebuild(pos,
h, is_star) -> ngas`` and ``rebuild_all(pos, h) -> ngas``, and attributes
(``n``, ``min_radius``, ``max_cells``, ``gsort``, ``cell_start``, ``nx``,
``ny``, ``nz``, ``lo``, ``inv_cell``, ``ngas``).  ``pos`` (n,3) and ``h``
(n,) are float64 CUDA tensors, ``is_star`` (n,) int32; ``gsort`` and
``cell_start`` are int32 CUDA tensors with the reference's contents: gas
indices in cell order (stable, 0 past ngas) and the first position of each
cell (0 past the cell count).  Like the reference, a rebuild with no gas
returns 0 and leaves the fields and geometry as they were.
"""
from __future__ import annotations

import ctypes

import torch

from .. import _lib
from .._tensors import vec1, vec3


class NeighborGrid:
    """Uniform grid over the gas particles, rebuilt each force evaluation."""

    def __init__(self, n: int, min_radius: float = 0.0):
        _lib.require_gpu()
        self.n = int(n)
        self.min_radius = float(min_radius)
        self._h = ctypes.c_void_p()
        _lib.check(_lib.load().rg_grid_create(self.n, self.min_radius, ctypes.byref(self._h)),
                   "rg_grid_create")
        mc = ctypes.c_int64()
        _lib.check(_lib.load().rg_grid_max_cells(self._h, ctypes.byref(mc)), "rg_grid_max_cells")
        self.max_cells = int(mc.value)
        dev = torch.device("cuda", torch.cuda.current_device())
        self.gsort = torch.zeros(max(self.n, 1), dtype=torch.int32, device=dev)
        self.cell_start = torch.zeros(self.max_cells + 1, dtype=torch.int32, device=dev)
        self.nx = self.ny = self.nz = 1
        self.lo = (0.0, 0.0, 0.0)
        self.inv_cell = 1.0
        self.ngas = 0

    def _rebuild(self, pos, h, is_star) -> int:
        n = self.n
        vec3(pos, n, "pos"), vec1(h, n, "h")
        if is_star is not None:
            if not (is_star.is_cuda and is_star.dtype == torch.int32 and is_star.is_contiguous()
                    and is_star.numel() == n):
                raise ValueError("is_star must be a contiguous int32 CUDA tensor of n entries")
        info = _lib.GridInfo()
        _lib.check(_lib.load().rg_grid_rebuild(
            self._h, pos.data_ptr(), h.data_ptr(),
            is_star.data_ptr() if is_star is not None else None,
            self.gsort.data_ptr(), self.cell_start.data_ptr(), ctypes.byref(info),
            _lib.stream_ptr(pos)), "rg_grid_rebuild")
        self.ngas = int(info.ngas)
        if self.ngas == 0:
            return 0
        self.nx, self.ny, self.nz = int(info.nx), int(info.ny), int(info.nz)
        self.lo = (float(info.lo[0]), float(info.lo[1]), float(info.lo[2]))
        self.inv_cell = float(info.inv_cell)
        return self.ngas

    def rebuild(self, pos, h, is_star) -> int:
        """Bin the gas particles (is_star == 0); return the gas count."""
        return self._rebuild(pos, h, is_star)

    def rebuild_all(self, pos, h) -> int:
        """Bin all particles (pure-gas engines)."""
        return self._rebuild(pos, h, None)

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().rg_grid_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

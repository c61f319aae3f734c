"""The gravity half of the SPH engines' force pass, as one solver object.

Both SPH engines of the reference add self-gravity and an analytic galaxy
potential to the hydro accelerations in ``_forces`` and size their
Barnes-Hut tree to the particle count in ``_rebuild_solvers``:

* core/living_galaxy_engine.py:727-738 (``_rebuild_solvers``: a tree iff
  ``gravity_mode == "bh" and n >= 256``, rebuilt whenever n changes) and
  :1015-1026 (``_forces``: tree -> ``BarnesHut.compute`` into acc_g, then
  ``add_acc``; direct -> ``grav_self_direct``; then ``grav_analytic``).
* core/gas_disk_engine.py:169-173 and :216-230 (the same tree rule; tree ->
  compute into acc_g, ``add_field``, ``add_analytic``; direct -> the fused
  ``add_grav_ext``).

``SelfGravity`` keeps the tree across calls, rebuilds it on an n change and
adds in each engine's order, so the accelerations are the same doubles as
the reference's call sequence on the same kernels: the tree add is fused
into the walk's store (``BarnesHut.compute_add``, the same doubles as
compute + add), and the direct paths are ``rg_direct_acc_add`` +
``rg_analytic_acc_add`` (living galaxy) or ``rg_direct_ext_acc_add`` (gas
disk, one kernel).
"""
from __future__ import annotations

from ..units import G as G_GAL
from . import gravity
from .barnes_hut import BarnesHut

BH_MIN_N = 256  # living_galaxy_engine.py:734, gas_disk_engine.py:169

# the two engines' direct-path add orders
LIVING_GALAXY = "living_galaxy"  # (acc + self) + ((disk + halo) + point mass)
GAS_DISK = "gas_disk"            # acc + (((self + disk) + halo) + point mass), one pass


def uses_tree(gravity_mode: str, n: int) -> bool:
    """The engines' rule for building a tree (living_galaxy_engine.py:734)."""
    return gravity_mode == "bh" and int(n) >= BH_MIN_N


class SelfGravity:
    """Self-gravity plus analytic potential for an engine whose particle
    count may change between calls.

    ``pot`` holds the engines' potential entries (``M_d``, ``a``, ``b``,
    ``M_h``, ``a_h``, optional ``is_hernquist`` and ``smbh_mass``); None or
    zero masses leave only self-gravity.  ``flavor`` is ``LIVING_GALAXY`` or
    ``GAS_DISK`` and sets the direct path's order of adds."""

    def __init__(self, softening: float, gravity_mode: str = "direct", theta: float = 0.6,
                 pot: dict | None = None, *, flavor: str = LIVING_GALAXY, G: float = G_GAL,
                 precision: str = "fp64", tree_precision: str = "fp64",
                 morton_bits: int = 21):
        if flavor not in (LIVING_GALAXY, GAS_DISK):
            raise ValueError(f"flavor must be {LIVING_GALAXY!r} or {GAS_DISK!r}")
        if gravity_mode not in ("direct", "bh"):
            raise ValueError(f"gravity_mode must be 'direct' or 'bh', got {gravity_mode!r}")
        self.softening = float(softening)
        self.eps2 = self.softening ** 2
        self.gravity_mode = gravity_mode
        self.theta = float(theta)
        self.flavor = flavor
        self.G = float(G)
        self.precision = precision
        self.tree_precision = tree_precision
        self.morton_bits = int(morton_bits)
        p = dict(pot or {})
        self.pot = gravity.potential_struct(p.get("M_d", 0.0), p.get("a", 1.0),
                                            p.get("b", 1.0), p.get("M_h", 0.0),
                                            p.get("a_h", 1.0), int(p.get("is_hernquist", 0)),
                                            float(p.get("smbh_mass", 0.0)))
        self._bh = None
        self.n = 0
        self.rebuilds = 0

    def rebuild(self, n: int) -> None:
        """Size the solver to n particles (``_rebuild_solvers``'s gravity
        half): a tree iff the mode is "bh" and n >= 256, rebuilt when n
        changes, none otherwise."""
        n = int(n)
        if uses_tree(self.gravity_mode, n):
            if self._bh is None or self._bh.n != n:
                self._bh = BarnesHut(n, theta=self.theta, softening=self.softening, G=self.G,
                                     precision=self.tree_precision,
                                     morton_bits=self.morton_bits)
                self.rebuilds += 1
        else:
            self._bh = None
        self.n = n

    @property
    def tree(self):
        """The Barnes-Hut handle in use, or None on the direct path."""
        return self._bh

    def add_forces(self, pos, mass, acc, n: int | None = None) -> None:
        """acc += self-gravity + analytic potential (``_forces``'s gravity
        half); n defaults to the last rebuild's, and a different n rebuilds
        first (the engines' dynamic particle count)."""
        n = self.n if n is None else int(n)
        if n != self.n or (self._bh is None and uses_tree(self.gravity_mode, n)):
            self.rebuild(n)
        if n == 0:
            return
        if self._bh is not None:
            # compute into acc_g + add_acc / add_field, fused into the walk
            self._bh.compute_add(pos, mass, acc)
            gravity.add_acc_analytic(pos, acc, n, self.G, self.pot)
        elif self.flavor == GAS_DISK:
            gravity.add_acc_direct_ext(pos, mass, acc, n, self.G, self.eps2, self.pot,
                                       self.precision)
        else:
            gravity.add_acc_direct(pos, mass, acc, n, self.G, self.eps2, self.precision)
            gravity.add_acc_analytic(pos, acc, n, self.G, self.pot)

"""The gravity kernels of core/living_galaxy_engine.py, same names and
argument order, on the HIP library.  Only these: SPH, star formation and the
engine class itself stay with the reference.

    grav_self_direct(pos, mass, acc, n, eps2)           :101-113
    add_acc(acc, acc_g, n)                               :116-120
    grav_analytic(pos, acc, n, M_d, a, b, M_h, a_h,
                  is_hernquist, smbh_mass)               :123-150

The reference reads the module constant G (core/units.py:21); here it is the
keyword ``G`` with that default.  ``precision`` selects the self-gravity pair
arithmetic as in ``rgsim.solvers.gravity`` (FP64 by default, the reference's).
The tree path of ``_forces`` (:1016-1018, BarnesHut.compute into acc_g, then
add_acc) is ``BarnesHut.compute_add`` in one call.

``LivingGalaxyGravity`` is the engine's whole gravity half -- the tree rule
and n-change rebuild of ``_rebuild_solvers`` (:727-738) and the adds of
``_forces`` (:1015-1026) in their order -- for a host engine that keeps its
SPH and star formation and hands the force pass its arrays.
"""
from __future__ import annotations

from .solvers import gravity as _g
from .solvers.self_gravity import LIVING_GALAXY, SelfGravity
from .units import G as _G


class LivingGalaxyGravity(SelfGravity):
    """SelfGravity with the living-galaxy engine's order of adds: (acc +
    self-gravity) + ((disk + halo) + point mass)."""

    def __init__(self, softening, gravity_mode="direct", theta=0.6, pot=None, **kw):
        super().__init__(softening, gravity_mode, theta, pot, flavor=LIVING_GALAXY, **kw)


def grav_self_direct(pos, mass, acc, n, eps2, *, G=_G, precision="fp64"):
    """acc += direct-N^2 self-gravity over all particles."""
    _g.add_acc_direct(pos, mass, acc, n, G, eps2, precision)


def add_acc(acc, acc_g, n):
    """acc += acc_g."""
    _g.add_field(acc, acc_g, n)


def grav_analytic(pos, acc, n, M_d, a, b, M_h, a_h, is_hernquist, smbh_mass, *, G=_G):
    """acc += Miyamoto-Nagai disk + Plummer/Hernquist halo + point mass."""
    _g.add_acc_analytic(pos, acc, n, G,
                        _g.potential_struct(M_d, a, b, M_h, a_h, is_hernquist, smbh_mass))

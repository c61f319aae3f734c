"""The gravity kernels of core/gas_disk_engine.py, same names and argument
order, on the HIP library.  Only these: SPH, cooling and the engine class
itself stay with the reference.

    add_grav_ext(pos, mass, acc, n, eps2, M_d, a, b, M_h, a_h,
                 is_hernquist, smbh_mass)                :20-51
    add_analytic(pos, acc, n, M_d, a, b, M_h, a_h,
                 is_hernquist, smbh_mass)                :54-82
    add_field(acc, src, n)                               :84-89

The reference reads the module constant G (core/units.py:21); here it is the
keyword ``G`` with that default.  ``precision`` selects the self-gravity pair
arithmetic (FP64 by default, the reference's); the analytic terms are FP64.
The tree path of ``_forces`` (:218-219, BarnesHut.compute into acc_g, then
add_field) is ``BarnesHut.compute_add`` in one call.

``GasDiskGravity`` is the engine's whole gravity half -- the tree rule and
n-change rebuild (:169-173) and the adds of ``_forces`` (:216-230): the tree
then the analytic potential, or the fused N^2 + analytic kernel.
"""
from __future__ import annotations

from .solvers import gravity as _g
from .solvers.self_gravity import GAS_DISK, SelfGravity
from .units import G as _G


class GasDiskGravity(SelfGravity):
    """SelfGravity with the gas-disk engine's order of adds: acc + (((self +
    disk) + halo) + point mass) in one kernel on the direct path."""

    def __init__(self, softening, gravity_mode="direct", theta=0.6, pot=None, **kw):
        super().__init__(softening, gravity_mode, theta, pot, flavor=GAS_DISK, **kw)


def add_grav_ext(pos, mass, acc, n, eps2, M_d, a, b, M_h, a_h, is_hernquist, smbh_mass, *,
                 G=_G, precision="fp64"):
    """acc += self-gravity (softened N^2) + analytic disk + halo + point mass."""
    _g.add_acc_direct_ext(pos, mass, acc, n, G, eps2,
                          _g.potential_struct(M_d, a, b, M_h, a_h, is_hernquist, smbh_mass),
                          precision)


def add_analytic(pos, acc, n, M_d, a, b, M_h, a_h, is_hernquist, smbh_mass, *, G=_G):
    """acc += analytic disk + halo + point mass only (the tree path's partner)."""
    _g.add_acc_analytic(pos, acc, n, G,
                        _g.potential_struct(M_d, a, b, M_h, a_h, is_hernquist, smbh_mass))


def add_field(acc, src, n):
    """acc += src."""
    _g.add_field(acc, src, n)

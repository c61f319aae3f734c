"""Direct N^2 gravity and FP64 energy diagnostics, mirroring
core/solvers/gravity.py:13-57 over torch device buffers.

    compute_acc_direct(pos, mass, acc, n, grav, eps2)   gravity.py:13-26
    kinetic_energy(vel, mass, n)                         gravity.py:29-36
    potential_energy(pos, mass, n, grav, eps2)           gravity.py:39-57

and the accumulate forms the gas-disk and living-galaxy engines use
(core/gas_disk_engine.py:20-89, core/living_galaxy_engine.py:101-150):

    add_acc_direct        acc += self-gravity sum
    add_acc_direct_ext    acc += self-gravity sum + analytic potential, fused
    add_acc_analytic      acc += analytic potential
    add_field             acc += src

``pos``/``acc`` are (n,3) and ``mass`` (n,) contiguous float64 CUDA tensors.
``precision`` selects the pair arithmetic: "fp64" (the reference's),
"fp32", or "fp32c" (FP32 on hi+lo positions with compensated sums).
``compute_acc_direct_hilo`` / ``compute_acc_direct_kick_hilo`` take the
positions as FP64 ``pos`` plus FP32 offsets ``pos_lo`` (FP32C arithmetic).
"""
from __future__ import annotations

import ctypes

import torch

from .. import _lib
from .._tensors import vec1, vec3, vec3_f32


def compute_acc_direct(pos, mass, acc, n, grav, eps2, precision="fp64"):
    """Overwrites acc with G sum_j m_j d_ij / (|d_ij|^2 + eps2)^{3/2}."""
    n = int(n)
    vec3(pos, n, "pos"), vec1(mass, n, "mass"), vec3(acc, n, "acc")
    _lib.require_gpu()
    _lib.check(_lib.load().rg_direct_acc(pos.data_ptr(), mass.data_ptr(), acc.data_ptr(), n,
                                         float(grav), float(eps2),
                                         _lib.precision_code(precision), _lib.stream_ptr(pos)),
               "rg_direct_acc")


def compute_acc_direct_kick(pos, mass, acc, vel, n, grav, eps2, dt_half, precision="fp64"):
    """compute_acc_direct followed by v += acc * dt_half in the same kernel."""
    n = int(n)
    vec3(pos, n, "pos"), vec1(mass, n, "mass"), vec3(acc, n, "acc"), vec3(vel, n, "vel")
    _lib.require_gpu()
    _lib.check(_lib.load().rg_direct_acc_kick(pos.data_ptr(), mass.data_ptr(), acc.data_ptr(),
                                              vel.data_ptr(), n, float(grav), float(eps2),
                                              _lib.precision_code(precision), float(dt_half),
                                              _lib.stream_ptr(pos)),
               "rg_direct_acc_kick")


def compute_acc_direct_hilo(pos, pos_lo, mass, acc, n, grav, eps2):
    """compute_acc_direct in FP32C on hi/lo positions X = pos + pos_lo."""
    n = int(n)
    vec3(pos, n, "pos"), vec3_f32(pos_lo, n, "pos_lo"), vec1(mass, n, "mass")
    vec3(acc, n, "acc")
    _lib.require_gpu()
    _lib.check(_lib.load().rg_direct_acc_hilo(pos.data_ptr(), pos_lo.data_ptr(), mass.data_ptr(),
                                              acc.data_ptr(), n, float(grav), float(eps2),
                                              _lib.stream_ptr(pos)),
               "rg_direct_acc_hilo")


def compute_acc_direct_kick_hilo(pos, pos_lo, mass, acc, vel, n, grav, eps2, dt_half):
    """compute_acc_direct_hilo followed by v += acc * dt_half in the same kernel."""
    n = int(n)
    vec3(pos, n, "pos"), vec3_f32(pos_lo, n, "pos_lo"), vec1(mass, n, "mass")
    vec3(acc, n, "acc"), vec3(vel, n, "vel")
    _lib.require_gpu()
    _lib.check(_lib.load().rg_direct_acc_kick_hilo(pos.data_ptr(), pos_lo.data_ptr(),
                                                   mass.data_ptr(), acc.data_ptr(),
                                                   vel.data_ptr(), n, float(grav), float(eps2),
                                                   float(dt_half), _lib.stream_ptr(pos)),
               "rg_direct_acc_kick_hilo")


def potential_struct(M_d, a, b, M_h, a_h, is_hernquist, smbh_mass):
    """rg_potential_t from the engines' potential arguments
    (gas_disk_engine.py:21-23)."""
    return _lib.Potential(float(M_d), float(a), float(b), float(M_h), float(a_h),
                          float(smbh_mass), int(is_hernquist))


def add_acc_direct(pos, mass, acc, n, grav, eps2, precision="fp64"):
    """acc += G sum_j m_j d_ij / (|d_ij|^2 + eps2)^{3/2}
    (living_galaxy_engine.py:101-113)."""
    n = int(n)
    vec3(pos, n, "pos"), vec1(mass, n, "mass"), vec3(acc, n, "acc")
    _lib.require_gpu()
    _lib.check(_lib.load().rg_direct_acc_add(pos.data_ptr(), mass.data_ptr(), acc.data_ptr(),
                                             n, float(grav), float(eps2),
                                             _lib.precision_code(precision),
                                             _lib.stream_ptr(pos)),
               "rg_direct_acc_add")


def add_acc_direct_ext(pos, mass, acc, n, grav, eps2, pot, precision="fp64"):
    """acc += self-gravity sum + analytic potential (gas_disk_engine.py:20-51);
    ``pot`` is an rg_potential_t (``potential_struct``)."""
    n = int(n)
    vec3(pos, n, "pos"), vec1(mass, n, "mass"), vec3(acc, n, "acc")
    _lib.require_gpu()
    _lib.check(_lib.load().rg_direct_ext_acc_add(pos.data_ptr(), mass.data_ptr(),
                                                 acc.data_ptr(), n, float(grav), float(eps2),
                                                 ctypes.byref(pot),
                                                 _lib.precision_code(precision),
                                                 _lib.stream_ptr(pos)),
               "rg_direct_ext_acc_add")


def add_acc_analytic(pos, acc, n, grav, pot):
    """acc += analytic disk + halo + point-mass acceleration
    (living_galaxy_engine.py:123-150, gas_disk_engine.py:54-82)."""
    n = int(n)
    vec3(pos, n, "pos"), vec3(acc, n, "acc")
    _lib.require_gpu()
    _lib.check(_lib.load().rg_analytic_acc_add(pos.data_ptr(), acc.data_ptr(), n, float(grav),
                                               ctypes.byref(pot), _lib.stream_ptr(pos)),
               "rg_analytic_acc_add")


def add_field(acc, src, n):
    """acc += src (living_galaxy_engine.py:116-120, gas_disk_engine.py:84-89)."""
    n = int(n)
    vec3(acc, n, "acc"), vec3(src, n, "src")
    _lib.require_gpu()
    _lib.check(_lib.load().rg_add_field(acc.data_ptr(), src.data_ptr(), n,
                                        _lib.stream_ptr(acc)),
               "rg_add_field")


def energy_terms(pos, vel, mass, n, grav, eps2, with_pe=True):
    """(KE, PE, momentum[3]) in FP64, reduced in a fixed order.  With
    with_pe=False the N^2 potential pass is skipped and PE is 0."""
    n = int(n)
    vec3(vel, n, "vel"), vec1(mass, n, "mass")
    if with_pe:
        vec3(pos, n, "pos")
    _lib.require_gpu()
    out = torch.empty(5, dtype=torch.float64, device=vel.device)
    _lib.check(_lib.load().rg_energy(pos.data_ptr() if with_pe else None, vel.data_ptr(),
                                     mass.data_ptr(), n, float(grav), float(eps2),
                                     out.data_ptr(), _lib.stream_ptr(vel)),
               "rg_energy")
    o = out.cpu().tolist()
    return o[0], o[1], o[2:5]


def kinetic_energy(vel, mass, n):
    """KE = 1/2 sum m v.v (gravity.py:29-36)."""
    return energy_terms(None, vel, mass, n, 0.0, 1.0, with_pe=False)[0]


def potential_energy(pos, mass, n, grav, eps2):
    """PE = -1/2 G sum_i m_i sum_{j != i} m_j / sqrt(r^2 + eps2) (gravity.py:39-57)."""
    zeros = torch.zeros_like(pos)
    return energy_terms(pos, zeros, mass, n, grav, eps2)[1]

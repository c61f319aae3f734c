"""Leapfrog KDK kernels, mirroring core/integrators/leapfrog.py:11-22 over
torch device buffers ((n,3) contiguous float64 CUDA tensors).

``drift_hilo``/``kick_drift_hilo`` drift hi/lo positions X = pos + pos_lo
(pos FP64, pos_lo an FP32 offset below half an ulp of pos) with a
compensated sum: the part of each increment v dt below the ulp of pos is
kept in pos_lo instead of being rounded away.
"""
from __future__ import annotations

from .. import _lib
from .._tensors import vec3, vec3_f32


def kick(vel, acc, n, dt_half):
    """vel += acc * dt_half (leapfrog.py:11-15)."""
    n = int(n)
    vec3(vel, n, "vel"), vec3(acc, n, "acc")
    _lib.require_gpu()
    _lib.check(_lib.load().rg_kick(vel.data_ptr(), acc.data_ptr(), n, float(dt_half),
                                   _lib.stream_ptr(vel)), "rg_kick")


def drift(pos, vel, n, dt):
    """pos += vel * dt (leapfrog.py:18-22)."""
    n = int(n)
    vec3(pos, n, "pos"), vec3(vel, n, "vel")
    _lib.require_gpu()
    _lib.check(_lib.load().rg_drift(pos.data_ptr(), vel.data_ptr(), n, float(dt),
                                    _lib.stream_ptr(pos)), "rg_drift")


def kick_drift(pos, vel, acc, n, dt_half, dt):
    """kick(vel, acc, n, dt_half) then drift(pos, vel, n, dt), one pass."""
    n = int(n)
    vec3(pos, n, "pos"), vec3(vel, n, "vel"), vec3(acc, n, "acc")
    _lib.require_gpu()
    _lib.check(_lib.load().rg_kick_drift(pos.data_ptr(), vel.data_ptr(), acc.data_ptr(), n,
                                         float(dt_half), float(dt), _lib.stream_ptr(pos)),
               "rg_kick_drift")


def drift_hilo(pos, pos_lo, vel, n, dt):
    """pos + pos_lo += vel * dt, compensated (leapfrog.py:18-22 on hi/lo positions)."""
    n = int(n)
    vec3(pos, n, "pos"), vec3_f32(pos_lo, n, "pos_lo"), vec3(vel, n, "vel")
    _lib.require_gpu()
    _lib.check(_lib.load().rg_drift_hilo(pos.data_ptr(), pos_lo.data_ptr(), vel.data_ptr(), n,
                                         float(dt), _lib.stream_ptr(pos)), "rg_drift_hilo")


def kick_drift_hilo(pos, pos_lo, vel, acc, n, dt_half, dt):
    """kick(vel, acc, n, dt_half) then drift_hilo(pos, pos_lo, vel, n, dt), one pass."""
    n = int(n)
    vec3(pos, n, "pos"), vec3_f32(pos_lo, n, "pos_lo"), vec3(vel, n, "vel")
    vec3(acc, n, "acc")
    _lib.require_gpu()
    _lib.check(_lib.load().rg_kick_drift_hilo(pos.data_ptr(), pos_lo.data_ptr(), vel.data_ptr(),
                                              acc.data_ptr(), n, float(dt_half), float(dt),
                                              _lib.stream_ptr(pos)),
               "rg_kick_drift_hilo")

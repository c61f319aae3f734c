"""Leapfrog KDK kernels, mirroring core/integrators/leapfrog.py:11-22 over
torch device buffers ((n,3) contiguous float64 CUDA tensors)."""
from __future__ import annotations

from .. import _lib
from .._tensors import vec3


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

"""Plummer sphere sampler, restating core/ic/plummer.py:19-76 with G as an
argument (the reference reads core.units.G).

Radii from the inverse mass CDF r = a / sqrt(x^{-2/3} - 1); isotropic
directions; speeds q * v_esc with q drawn by rejection on
g(q) = q^2 (1 - q^2)^{7/2} (accept when 0.1 x5 < g(x4)); equal masses.
The numpy call sequence is the reference's, so for G = core.units.G the
sample stream for a seed is the reference's.
"""
from __future__ import annotations

import numpy as np

from ..state import PTYPE_DM, State
from ..units import G as G_GAL


def random_directions(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform unit vectors (core/ic/plummer.py:69-76)."""
    cos_t = rng.uniform(-1.0, 1.0, n)
    sin_t = np.sqrt(1.0 - cos_t ** 2)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    return np.column_stack((sin_t * np.cos(phi), sin_t * np.sin(phi), cos_t))


def make_plummer(n: int = 10000, total_mass: float = 10.0, scale_radius: float = 1.0,
                 seed: int = 0, G: float = G_GAL) -> State:
    rng = np.random.default_rng(seed)
    a = float(scale_radius)
    M = float(total_mass)
    x1 = rng.random(n)
    r = a / np.sqrt(x1 ** (-2.0 / 3.0) - 1.0)
    pos = random_directions(rng, n) * r[:, None]
    q = np.empty(n)
    remaining = np.arange(n)
    while remaining.size:
        x4 = rng.random(remaining.size)
        x5 = rng.random(remaining.size)
        ok = 0.1 * x5 < x4 ** 2 * (1.0 - x4 ** 2) ** 3.5
        q[remaining[ok]] = x4[ok]
        remaining = remaining[~ok]
    v_esc = np.sqrt(2.0 * G * M / a) * (1.0 + (r / a) ** 2) ** -0.25
    vel = random_directions(rng, n) * (q * v_esc)[:, None]
    return State(pos=pos, vel=vel, mass=np.full(n, M / n),
                 ptype=np.full(n, PTYPE_DM, dtype=np.int32), ids=np.arange(n, dtype=np.int64),
                 time=0.0, step=0, seed=seed,
                 meta={"ic": "plummer", "total_mass": M, "scale_radius": a, "n": n, "G": G})

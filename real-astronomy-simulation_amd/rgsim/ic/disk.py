"""Disk galaxy samplers.

``make_disk_galaxy`` restates core/ic/disk.py:24-111 (exponential disk in a
live Plummer halo) with G as an argument.  ``make_disk_bulge`` is the
Barnes-Hut drop-in workload: an exponential disk with a sech^2 vertical
profile (the draw of core/ic/realistic_galaxy.py:45-53) inside a live
Plummer bulge, with velocities built as in core/ic/disk.py:40-79 (circular
speed from the spherically enclosed particle mass with softening 0.05 R_d,
disk dispersion sigma_R = f v_c, sigma_phi = sigma_R / 2, sigma_z = 0.35
sigma_R; bulge isotropic |N(0, 0.3 v_c)|; net momentum removed).
"""
from __future__ import annotations

import numpy as np

from ..state import PTYPE_DM, PTYPE_STAR, State
from ..units import G as G_GAL
from .plummer import random_directions


def _exp_disk_R(rng, n, Rd, xmax):
    u = rng.random(n)
    x = np.linspace(0.0, xmax, 100000)
    cdf = 1.0 - (1.0 + x) * np.exp(-x)
    return np.interp(u, cdf, x) * Rd


def _plummer_positions(rng, n, a):
    x1 = rng.uniform(0.0, 0.999, n)
    r = a / np.sqrt(x1 ** (-2.0 / 3.0) - 1.0)
    return random_directions(rng, n) * r[:, None]


def _velocities(rng, pos, mass, n_disk, disk_scale, sigma_frac, sph_frac, G):
    r_sph = np.linalg.norm(pos, axis=1)
    order = np.argsort(r_sph)
    m_cum = np.empty_like(mass)
    m_cum[order] = np.cumsum(mass[order])
    eps = 0.05 * disk_scale
    v_circ = np.sqrt(G * m_cum / np.sqrt(r_sph ** 2 + eps ** 2))
    vel = np.zeros_like(pos)
    dp = pos[:n_disk]
    R = np.linalg.norm(dp[:, :2], axis=1) + 1e-6
    e_R = np.column_stack((dp[:, 0] / R, dp[:, 1] / R, np.zeros(n_disk)))
    e_phi = np.column_stack((-dp[:, 1] / R, dp[:, 0] / R, np.zeros(n_disk)))
    vc = v_circ[:n_disk]
    sR = sigma_frac * vc
    d_R = rng.normal(0.0, sR)
    d_phi = rng.normal(0.0, 0.5 * sR)
    d_z = rng.normal(0.0, 0.35 * sR)
    vel[:n_disk] = e_R * d_R[:, None] + e_phi * (vc + d_phi)[:, None]
    vel[:n_disk, 2] = d_z
    n_sph = pos.shape[0] - n_disk
    vel[n_disk:] = random_directions(rng, n_sph) * np.abs(
        rng.normal(0.0, sph_frac * v_circ[n_disk:, None]))
    vel -= np.average(vel, axis=0, weights=mass)
    return vel


def make_disk_galaxy(n_disk=20000, n_halo=20000, disk_mass=5.0, disk_scale=3.0,
                     disk_height=0.25, halo_mass=12.0, halo_scale=8.0, sigma_frac=0.1,
                     seed=0, G=G_GAL) -> State:
    rng = np.random.default_rng(seed)
    R = _exp_disk_R(rng, n_disk, disk_scale, 20.0)
    phi = rng.uniform(0.0, 2.0 * np.pi, n_disk)
    z = rng.exponential(disk_height, n_disk) * rng.choice((-1.0, 1.0), n_disk)
    disk = np.column_stack((R * np.cos(phi), R * np.sin(phi), z))
    halo = _plummer_positions(rng, n_halo, halo_scale)
    pos = np.vstack((disk, halo))
    mass = np.concatenate((np.full(n_disk, disk_mass / n_disk), np.full(n_halo, halo_mass / n_halo)))
    vel = _velocities(rng, pos, mass, n_disk, disk_scale, sigma_frac, 0.3, G)
    ptype = np.concatenate((np.full(n_disk, PTYPE_STAR, np.int32), np.full(n_halo, PTYPE_DM, np.int32)))
    return State(pos=pos, vel=vel, mass=mass, ptype=ptype,
                 ids=np.arange(n_disk + n_halo, dtype=np.int64), seed=seed,
                 meta={"ic": "disk_galaxy", "n_disk": n_disk, "n_halo": n_halo})


def make_disk_bulge(n_disk=100000, n_bulge=20000, disk_mass=5.0, disk_scale=3.0,
                    disk_height=0.3, bulge_mass=1.0, bulge_scale=0.5, sigma_frac=0.1,
                    seed=1, G=G_GAL) -> State:
    rng = np.random.default_rng(seed)
    R = _exp_disk_R(rng, n_disk, disk_scale, 30.0)
    phi = rng.uniform(0.0, 2.0 * np.pi, n_disk)
    u_z = rng.uniform(0.001, 0.999, n_disk)
    z = disk_height * np.arctanh(2.0 * u_z - 1.0)
    disk = np.column_stack((R * np.cos(phi), R * np.sin(phi), z))
    bulge = _plummer_positions(rng, n_bulge, bulge_scale)
    pos = np.vstack((disk, bulge))
    mass = np.concatenate((np.full(n_disk, disk_mass / n_disk),
                           np.full(n_bulge, bulge_mass / n_bulge)))
    vel = _velocities(rng, pos, mass, n_disk, disk_scale, sigma_frac, 0.3, G)
    ptype = np.concatenate((np.full(n_disk, PTYPE_STAR, np.int32),
                            np.full(n_bulge, PTYPE_DM, np.int32)))
    return State(pos=pos, vel=vel, mass=mass, ptype=ptype,
                 ids=np.arange(n_disk + n_bulge, dtype=np.int64), seed=seed,
                 meta={"ic": "disk_bulge", "n_disk": n_disk, "n_bulge": n_bulge, "G": G})

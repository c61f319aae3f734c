"""CPU oracle for the gravity hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only as the checker.  The product package never
imports it and has no CPU fallback.

Two layers:

* ``C`` -- ctypes binding of ``liboracle.so`` (``rg_oracle.c``), the FP64
  restatement of the reference kernels (file:line in that file's header).
* numpy twins and independent checkers used to pin the C restatement:

  ==========================  ==============================================
  function                    pinned by
  ==========================  ==============================================
  morton_codes / part1by2     KATs (0 -> 0, 0x1FFFFF -> 0x1249249249249249,
                              unit cube corners -> 0 and 0x7FFFFFFFFFFFFFF8),
                              single-bit spreads, brute-force bit loop,
                              C restatement bit-equality
  stable_argsort              numpy ``argsort(kind="stable")`` (a library
                              routine: the reference's own call)
  clz / delta                 ``int.bit_length`` identities, exhaustively on
                              edge values
  build_radix                 ``brute_force_radix_tree`` (independent top-down
                              highest-differing-bit split over (key, index))
                              plus structural invariants
  aggregate                   closed-form sums over each node's leaf range
  direct_acc                  two-body closed form, momentum invariant
                              sum m a = 0, translation invariance, a numpy
                              pairwise evaluation in a different order
  potential / kinetic energy  two-body closed form, Plummer virial ratio
  traverse                    theta = 0 reproduces the direct sum exactly in
                              set (every leaf visited), FP64 direct-sum error
                              at theta > 0
  grid_rebuild                the binning invariants (every gas particle once,
                              each in the cell that contains it, cells in
                              order, stable within a cell, cell_start the
                              cumulative counts), the grid-shape rules
                              (cell >= 2 h_max and >= min_radius, budget,
                              growth by 1.3), hand-computed small grids, and
                              brute-force completeness of the 3x3x3 stencil
                              for every pair within the search radius
  grav_self_direct,           the direct sum plus the incoming acc, bitwise
  grav_ext_add                (same j loop); the analytic part as below
  analytic_add                exact gradients of the Miyamoto-Nagai, Plummer,
                              Hernquist and point-mass potentials (sympy,
                              50-digit mpmath); Miyamoto-Nagai a = 0 equals a
                              Plummer sphere of scale b; in-plane circular
                              speed closed form; no-op for M_d = M_h = 0
                              away from the b = z = 0 singular line
  ==========================  ==============================================
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64 = ctypes.c_int64
_d = ctypes.c_double


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB_PATH) or (
            os.path.getmtime(_LIB_PATH) < os.path.getmtime(os.path.join(_HERE, "rg_oracle.c"))):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"],
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_direct_acc.argtypes = [_dp, _dp, _dp, _i64, _d, _d]
        L.orc_kinetic_energy.argtypes = [_dp, _dp, _i64]
        L.orc_kinetic_energy.restype = _d
        L.orc_potential_energy.argtypes = [_dp, _dp, _i64, _d, _d, _dp]
        L.orc_potential_energy.restype = _d
        L.orc_kick.argtypes = [_dp, _dp, _i64, _d]
        L.orc_drift.argtypes = [_dp, _dp, _i64, _d]
        L.orc_part1by2.argtypes = [ctypes.c_uint64]
        L.orc_part1by2.restype = ctypes.c_uint64
        L.orc_morton.argtypes = [_dp, _i64, ctypes.c_int, _u64p]
        L.orc_morton.restype = _i64
        L.orc_stable_argsort.argtypes = [_u64p, _i64, _i32p]
        L.orc_clz64.argtypes = [ctypes.c_int64]
        L.orc_clz64.restype = ctypes.c_int
        L.orc_clz_i32.argtypes = [ctypes.c_int32]
        L.orc_clz_i32.restype = ctypes.c_int
        L.orc_delta.argtypes = [_i64, _i64, _i64, _i64p]
        L.orc_delta.restype = ctypes.c_int
        L.orc_build_radix.argtypes = [_i64p, _i64, _i32p, _i32p, _i32p]
        L.orc_aggregate.argtypes = [_i32p, _dp, _dp, _i32p, _i64, _dp, _dp, _dp, _dp]
        L.orc_traverse.argtypes = [_dp, _i32p, _dp, _dp, _dp, _dp, _i32p, _i32p, _dp,
                                   _i64, _d, _d, _d, ctypes.c_int, _i64p, _i64p]
        L.orc_traverse.restype = _i64
        L.orc_bh_compute.argtypes = [_dp, _dp, _dp, _i64, _d, _d, _d, ctypes.c_int,
                                     ctypes.c_int, _i64p]
        L.orc_bh_compute.restype = _i64
        _pot = [_d, _d, _d, _d, _d, ctypes.c_int, _d]
        L.orc_grav_self_direct.argtypes = [_dp, _dp, _dp, _i64, _d, _d]
        L.orc_analytic_add.argtypes = [_dp, _dp, _i64, _d] + _pot
        L.orc_grav_ext_add.argtypes = [_dp, _dp, _dp, _i64, _d, _d] + _pot
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(shape) if shape is not None else a


def set_threads(t: int) -> None:
    lib().orc_set_threads(int(t))


def get_threads() -> int:
    return int(lib().orc_get_threads())


# ---------------------------------------------------------------- C oracle
def direct_acc(pos, mass, G, eps2):
    """gravity.py:13-26 restated in C (FP64, sequential j, all j incl. i)."""
    pos = _f64(pos, (-1, 3))
    mass = _f64(mass, (-1,))
    acc = np.empty_like(pos)
    lib().orc_direct_acc(_p(pos, _dp), _p(mass, _dp), _p(acc, _dp), pos.shape[0],
                         float(G), float(eps2))
    return acc


def _pot_args(pot):
    """(M_d, a, b, M_h, a_h, is_hernquist, smbh_mass) of a potential dict
    shaped like the engines' ``self.pot`` (gas_disk_engine.py:220-224)."""
    return (float(pot["M_d"]), float(pot["a"]), float(pot["b"]), float(pot["M_h"]),
            float(pot["a_h"]), int(pot.get("is_hernquist", 0)), float(pot.get("smbh_mass", 0.0)))


def grav_self_direct(pos, mass, acc, G, eps2):
    """living_galaxy_engine.py:101-113 restated in C: returns acc + N^2 sum."""
    pos = _f64(pos, (-1, 3))
    mass = _f64(mass, (-1,))
    out = _f64(acc, (-1, 3)).copy()
    lib().orc_grav_self_direct(_p(pos, _dp), _p(mass, _dp), _p(out, _dp), pos.shape[0],
                               float(G), float(eps2))
    return out


def analytic_add(pos, acc, G, pot):
    """living_galaxy_engine.py:123-150 / gas_disk_engine.py:54-82 restated in
    C: returns acc + (disk + halo + point mass)."""
    pos = _f64(pos, (-1, 3))
    out = _f64(acc, (-1, 3)).copy()
    lib().orc_analytic_add(_p(pos, _dp), _p(out, _dp), pos.shape[0], float(G), *_pot_args(pot))
    return out


def grav_ext_add(pos, mass, acc, G, eps2, pot):
    """gas_disk_engine.py:20-51 restated in C: returns acc + (N^2 sum + disk +
    halo + point mass), summed in that order."""
    pos = _f64(pos, (-1, 3))
    mass = _f64(mass, (-1,))
    out = _f64(acc, (-1, 3)).copy()
    lib().orc_grav_ext_add(_p(pos, _dp), _p(mass, _dp), _p(out, _dp), pos.shape[0], float(G),
                           float(eps2), *_pot_args(pot))
    return out


def kinetic_energy(vel, mass):
    vel = _f64(vel, (-1, 3))
    mass = _f64(mass, (-1,))
    return float(lib().orc_kinetic_energy(_p(vel, _dp), _p(mass, _dp), vel.shape[0]))


def potential_energy(pos, mass, G, eps2, per_particle=False):
    pos = _f64(pos, (-1, 3))
    mass = _f64(mass, (-1,))
    pe_i = np.empty(pos.shape[0])
    pe = float(lib().orc_potential_energy(_p(pos, _dp), _p(mass, _dp), pos.shape[0],
                                          float(G), float(eps2), _p(pe_i, _dp)))
    return (pe, pe_i) if per_particle else pe


def kick(vel, acc, dt_half):
    vel = _f64(vel, (-1, 3)).copy()
    acc = _f64(acc, (-1, 3))
    lib().orc_kick(_p(vel, _dp), _p(acc, _dp), vel.shape[0], float(dt_half))
    return vel


def drift(pos, vel, dt):
    pos = _f64(pos, (-1, 3)).copy()
    vel = _f64(vel, (-1, 3))
    lib().orc_drift(_p(pos, _dp), _p(vel, _dp), pos.shape[0], float(dt))
    return pos


def kdk_step(pos, vel, acc, mass, dt, G, eps2):
    """engine.py:108-116 with the direct solver: kick(dt/2), drift(dt),
    force, kick(dt/2).  Returns new (pos, vel, acc)."""
    half = 0.5 * dt
    vel = kick(vel, acc, half)
    pos = drift(pos, vel, dt)
    acc = direct_acc(pos, mass, G, eps2)
    vel = kick(vel, acc, half)
    return pos, vel, acc


def morton_c(pos, bits=21):
    pos = _f64(pos, (-1, 3))
    keys = np.empty(pos.shape[0], dtype=np.uint64)
    lib().orc_morton(_p(pos, _dp), pos.shape[0], int(bits), _p(keys, _u64p))
    return keys


def stable_argsort_c(keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    order = np.empty(keys.shape[0], dtype=np.int32)
    lib().orc_stable_argsort(_p(keys, _u64p), keys.shape[0], _p(order, _i32p))
    return order


def clz64_ref(x: int) -> int:
    return int(lib().orc_clz64(int(x)))


def clz_i32_ref(x: int) -> int:
    return int(lib().orc_clz_i32(int(x)))


def build_radix(code):
    """barnes_hut.py:84-121 restated in C.  Returns (left, right, parent)."""
    code = np.ascontiguousarray(code, dtype=np.int64)
    n = code.shape[0]
    left = np.empty(max(n - 1, 1), dtype=np.int32)
    right = np.empty(max(n - 1, 1), dtype=np.int32)
    parent = np.empty(max(2 * n - 1, 1), dtype=np.int32)
    lib().orc_build_radix(_p(code, _i64p), n, _p(left, _i32p), _p(right, _i32p),
                          _p(parent, _i32p))
    return left[:n - 1], right[:n - 1], parent[:2 * n - 1]


def aggregate(order, pos, mass, parent):
    """barnes_hut.py:124-177.  Returns (com, nmass, nmin, nmax) per node id."""
    pos = _f64(pos, (-1, 3))
    mass = _f64(mass, (-1,))
    order = np.ascontiguousarray(order, dtype=np.int32)
    parent = np.ascontiguousarray(parent, dtype=np.int32)
    n = pos.shape[0]
    m = 2 * n - 1
    com = np.empty((m, 3))
    nmass = np.empty(m)
    nmin = np.empty((m, 3))
    nmax = np.empty((m, 3))
    lib().orc_aggregate(_p(order, _i32p), _p(pos, _dp), _p(mass, _dp), _p(parent, _i32p), n,
                        _p(com, _dp), _p(nmass, _dp), _p(nmin, _dp), _p(nmax, _dp))
    return com, nmass, nmin, nmax


def traverse(pos, order, com, nmass, nmin, nmax, left, right, G, theta, eps,
             max_stack=64):
    """barnes_hut.py:180-217.  Returns (acc, visits_per_particle, overflow)."""
    pos = _f64(pos, (-1, 3))
    n = pos.shape[0]
    acc = np.empty_like(pos)
    visits = np.empty(n, dtype=np.int64)
    ovf = ctypes.c_int64(0)
    args = [np.ascontiguousarray(a, dtype=t) for a, t in
            ((order, np.int32), (com, np.float64), (nmass, np.float64), (nmin, np.float64),
             (nmax, np.float64), (left, np.int32), (right, np.int32))]
    lib().orc_traverse(_p(pos, _dp), _p(args[0], _i32p), _p(args[1], _dp), _p(args[2], _dp),
                       _p(args[3], _dp), _p(args[4], _dp), _p(args[5], _i32p),
                       _p(args[6], _i32p), _p(acc, _dp), n, float(G),
                       float(theta) ** 2, float(eps) ** 2, int(max_stack),
                       _p(visits, _i64p), ctypes.byref(ovf))
    return acc, visits, int(ovf.value)


def bh_tree(pos, mass, bits=21):
    """Keys, order, sorted code, topology and node data for one BH call
    (barnes_hut.py:245-256)."""
    keys = morton_c(pos, bits)
    order = stable_argsort_c(keys)
    code = keys[order].astype(np.int64)
    left, right, parent = build_radix(code)
    com, nmass, nmin, nmax = aggregate(order, pos, mass, parent)
    return dict(keys=keys, order=order, code=code, left=left, right=right, parent=parent,
                com=com, nmass=nmass, nmin=nmin, nmax=nmax)


def bh_acc(pos, mass, G, theta, eps, bits=21, max_stack=64):
    """BarnesHut.compute restated (barnes_hut.py:242-259).
    Returns (acc, visits_per_particle, overflow)."""
    t = bh_tree(pos, mass, bits)
    return traverse(pos, t["order"], t["com"], t["nmass"], t["nmin"], t["nmax"],
                    t["left"], t["right"], G, theta, eps, max_stack)


# ------------------------------------------------------ numpy twins/checkers
_MASKS = (0x1F00000000FFFF, 0x1F0000FF0000FF, 0x100F00F00F00F00F,
          0x10C30C30C30C30C3, 0x1249249249249249)


def part1by2(x):
    """barnes_hut.py:27-35 in numpy uint64 arithmetic."""
    x = np.asarray(x).astype(np.uint64) & np.uint64(0x1FFFFF)
    for sh, mk in zip((32, 16, 8, 4, 2), _MASKS):
        x = (x | (x << np.uint64(sh))) & np.uint64(mk)
    return x


def part1by2_bitloop(x: int) -> int:
    """Independent spread: bit b of x goes to bit 3b."""
    out = 0
    for b in range(21):
        if (x >> b) & 1:
            out |= 1 << (3 * b)
    return out


def morton_codes(pos, bits=21):
    """barnes_hut.py:38-46 in numpy (same array expressions)."""
    pos = np.asarray(pos, dtype=np.float64)
    lo = pos.min(axis=0)
    span = float((pos.max(axis=0) - lo).max())
    span = span if span > 0 else 1.0
    scale = (2 ** bits - 1) / (span * 1.0000001)
    q = np.clip((pos - lo) * scale, 0, 2 ** bits - 1).astype(np.uint64)
    return part1by2(q[:, 0]) | (part1by2(q[:, 1]) << np.uint64(1)) | (
        part1by2(q[:, 2]) << np.uint64(2))


def brute_force_radix_tree(code):
    """Independent top-down binary radix tree over the distinct augmented keys
    (code, index): a range [lo, hi] splits after the last index whose
    augmented key has a 0 at the highest bit where key[lo] and key[hi]
    differ.  Node ids follow the Karras layout (root = internal 0, the
    node split at g has children g and g+1, leaf if its range is a single
    index, else internal).  Returns (left, right, parent) in the reference's
    encoding (internal ids offset by n)."""
    code = [int(c) for c in np.asarray(code, dtype=np.int64)]
    n = len(code)
    aug = [(c << 31) | i for i, c in enumerate(code)]
    left = [-1] * (n - 1)
    right = [-1] * (n - 1)
    parent = [-1] * (2 * n - 1)
    stack = [(0, n - 1, 0)]
    while stack:
        lo, hi, node = stack.pop()
        x = aug[lo] ^ aug[hi]
        bit = x.bit_length() - 1
        g = lo
        while g + 1 <= hi and not (aug[g + 1] >> bit) & 1:
            g += 1
        lc = g if lo == g else g + n
        rc = g + 1 if hi == g + 1 else g + 1 + n
        left[node] = lc
        right[node] = rc
        parent[lc] = node + n
        parent[rc] = node + n
        if lc >= n:
            stack.append((lo, g, g))
        if rc >= n:
            stack.append((g + 1, hi, g + 1))
    return (np.array(left, dtype=np.int32), np.array(right, dtype=np.int32),
            np.array(parent, dtype=np.int32))


def leaf_ranges(left, right, n):
    """[lo, hi] leaf range of every node id (leaves: [i, i])."""
    lo = np.zeros(2 * n - 1, dtype=np.int64)
    hi = np.zeros(2 * n - 1, dtype=np.int64)
    lo[:n] = np.arange(n)
    hi[:n] = np.arange(n)
    # post-order via explicit stack from the root
    order = []
    stack = [n]
    while stack:
        v = stack.pop()
        order.append(v)
        if v >= n:
            stack.append(int(left[v - n]))
            stack.append(int(right[v - n]))
    for v in reversed(order):
        if v >= n:
            lo[v] = lo[int(left[v - n])]
            hi[v] = hi[int(right[v - n])]
    return lo, hi


def direct_acc_np(pos, mass, G, eps2):
    """Vectorised FP64 direct sum (different summation order from the C
    restatement; agreement is to rounding, not bitwise)."""
    pos = np.asarray(pos, dtype=np.float64)
    mass = np.asarray(mass, dtype=np.float64)
    d = pos[None, :, :] - pos[:, None, :]
    r2 = (d * d).sum(-1) + eps2
    inv3 = r2 ** -1.5
    return G * np.einsum("ij,ijk->ik", mass[None, :] * inv3, d)


def two_body_acc(m1, m2, sep, G, eps2):
    """Closed form for two bodies on the x axis at 0 and sep: magnitude
    G m / (sep^2 + eps^2)^{3/2} * sep, pointing at the other body."""
    f = sep / (sep * sep + eps2) ** 1.5
    return np.array([[G * m2 * f, 0, 0], [-G * m1 * f, 0, 0]])


def two_body_pe(m1, m2, sep, G, eps2):
    return -G * m1 * m2 / np.sqrt(sep * sep + eps2)


# ------------------------------------------------------- SPH neighbour grid
def grid_rebuild(pos, h, is_star=None, min_radius=0.0, n=None):
    """NeighborGrid.rebuild / rebuild_all (core/solvers/neighbor_grid.py:52-136)
    restated in numpy.  Returns None when there is no gas (the reference
    returns 0 and leaves its fields untouched), else a dict with gsort (n,
    zero-padded past ngas), cell_start (max_cells + 1, zero-padded past
    n_cells), ngas, dims, lo, cell, inv_cell, max_cells."""
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    h = np.ascontiguousarray(h, np.float64).reshape(-1)
    n = pos.shape[0] if n is None else n
    max_cells = max(64, 4 * n)                                    # :42
    if is_star is None:                                           # :96-98 (rebuild_all)
        gas = np.arange(n, dtype=np.int32)
    else:                                                         # :57 (rebuild)
        gas = np.nonzero(np.asarray(is_star) == 0)[0].astype(np.int32)
    ngas = int(gas.size)
    if ngas == 0:
        return None
    gpos = pos[gas]
    h_max = float(h[gas].max())                                   # :63
    cell = max(2.0 * h_max, float(min_radius))                    # :64
    cell = cell if cell > 1e-9 else 1.0                           # :65
    lo = gpos.min(axis=0)                                         # :67
    span = gpos.max(axis=0) - lo
    dims = np.maximum(np.floor(span / cell).astype(np.int64) + 1, 1)
    while int(dims.prod()) > max_cells:                           # :71-73
        cell *= 1.3
        dims = np.maximum(np.floor(span / cell).astype(np.int64) + 1, 1)
    n_cells = int(dims.prod())
    ijk = np.clip(((gpos - lo) / cell).astype(np.int64), 0, dims - 1)   # :76
    flat = (ijk[:, 0] * dims[1] + ijk[:, 1]) * dims[2] + ijk[:, 2]
    order = np.argsort(flat, kind="stable")
    sorted_global = gas[order]
    counts = np.bincount(flat, minlength=n_cells)
    starts = np.empty(n_cells + 1, np.int32)
    starts[0] = 0
    starts[1:] = np.cumsum(counts).astype(np.int32)
    gsort = np.zeros(n, np.int32)                                 # :82-84 (padded)
    gsort[:ngas] = sorted_global
    cs = np.zeros(max_cells + 1, np.int32)                        # :85-86
    cs[:n_cells + 1] = starts
    return dict(gsort=gsort, cell_start=cs, ngas=ngas, dims=tuple(int(d) for d in dims),
                lo=tuple(float(x) for x in lo), cell=cell, inv_cell=1.0 / cell,
                max_cells=max_cells, n_cells=n_cells, flat=flat, gas=gas)

"""Hi/lo positions of the precision form (no reference counterpart: the
reference's drift, core/integrators/leapfrog.py:18-22, is x += v dt in
FP64).  X = pos + pos_lo, pos FP64 and pos_lo an FP32 offset.

- The compensated drift is checked against exact rational arithmetic: after
  K steps X must equal x0 + K fl(v dt) to within the bound the method
  guarantees, about K (2^-52 |v dt| + 2^-24 ulp(X)); the plain FP64 drift,
  which rounds every step to half an ulp of X, misses it by orders of
  magnitude more (so the check has power).
- The FP32C force on hi/lo positions equals rg_direct_acc's FP32C force bit
  for bit when the offsets are zero, and follows the offsets (not pos
  alone) when they are not.
"""
from fractions import Fraction

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def exact(x):
    return Fraction(float(x))


def test_drift_hilo_against_exact_arithmetic():
    from rgsim.integrators import leapfrog
    print("hilo drift vs exact arithmetic", flush=True)
    rng = np.random.default_rng(3)
    n, steps, dt = 512, 2000, 1e-3
    scale = 10.0 ** rng.uniform(-3, 3, (n, 1))
    x0 = rng.uniform(-1, 1, (n, 3)) * scale
    v = rng.normal(size=(n, 3)) * scale * 10.0 ** rng.uniform(-6, 0, (n, 1))
    pos = torch.from_numpy(x0.copy()).cuda()
    lo = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    vel = torch.from_numpy(v).cuda()
    plain = pos.clone()
    for _ in range(steps):
        leapfrog.drift_hilo(pos, lo, vel, n, dt)
        leapfrog.drift(plain, vel, n, dt)
    xh, xl, xp = pos.cpu().numpy(), lo.cpu().numpy(), plain.cpu().numpy()
    # |pos_lo| stays below half an ulp of pos (so pos is X rounded to FP64)
    assert np.all(np.abs(xl.astype(np.float64)) <= 0.5 * np.spacing(np.abs(xh)))
    inc = v * dt  # fl(v dt), the increment both drifts add each step
    idx = rng.choice(3 * n, 96, replace=False)
    ratio = []
    for k in idx:
        i, a = divmod(int(k), 3)
        want = exact(x0[i, a]) + steps * exact(inc[i, a])
        got = exact(xh[i, a]) + exact(np.float64(xl[i, a]))
        err = abs(got - want)
        bound = steps * (abs(inc[i, a]) * 2.0 ** -52
                         + np.spacing(abs(float(want)) + abs(x0[i, a])) * 2.0 ** -24)
        assert err <= Fraction(bound), (i, a, float(err), bound)
        err_plain = abs(exact(xp[i, a]) - want)
        if err_plain > 0:
            ratio.append(float(err_plain / max(err, Fraction(1, 10 ** 300))))
    # the plain FP64 drift is far less accurate on the same increments
    assert np.median(ratio) > 1e3, np.median(ratio)


@pytest.mark.parametrize("n", [1000, 1001, 1])
def test_kick_drift_hilo_is_kick_then_drift_hilo(n):
    """The fused pass (two components per thread, an odd last component on
    its own) equals kick then drift_hilo bit for bit."""
    from rgsim.integrators import leapfrog
    rng = np.random.default_rng(4)
    p = torch.from_numpy(rng.normal(size=(n, 3))).cuda()
    lo = torch.from_numpy((rng.normal(size=(n, 3)) * 1e-17).astype(np.float32)).cuda()
    v = torch.from_numpy(rng.normal(size=(n, 3))).cuda()
    a = torch.from_numpy(rng.normal(size=(n, 3))).cuda()
    p2, lo2, v2 = p.clone(), lo.clone(), v.clone()
    leapfrog.kick_drift_hilo(p, lo, v, a, n, 0.5e-3, 1e-3)
    leapfrog.kick(v2, a, n, 0.5e-3)
    leapfrog.drift_hilo(p2, lo2, v2, n, 1e-3)
    assert torch.equal(p, p2) and torch.equal(lo, lo2) and torch.equal(v, v2)


def test_hilo_force_zero_offsets_equal_fp32c():
    from rgsim.ic import make_plummer
    from rgsim.solvers import gravity
    for n in (300, 5003, 70001):
        s = make_plummer(n, total_mass=1.0, seed=n, G=1.0)
        pos = torch.from_numpy(s.pos).cuda()
        mass = torch.from_numpy(s.mass).cuda()
        lo = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
        a1, a2 = torch.empty_like(pos), torch.empty_like(pos)
        gravity.compute_acc_direct(pos, mass, a1, n, 1.0, 1e-4, "fp32c")
        gravity.compute_acc_direct_hilo(pos, lo, mass, a2, n, 1.0, 1e-4)
        assert torch.equal(a1, a2), n


def test_hilo_force_follows_the_offsets(orc):
    """Positions p (FP64) handed over as pos = fl(p - delta) plus pos_lo =
    p - pos rounded to FP32, with |delta| about 1e-4 of the position scale:
    the hi/lo force matches the FP64 force at p within the FP32C tolerance,
    while pos alone is 1e-4 away and misses it."""
    from rgsim.ic import make_plummer
    from rgsim.solvers import gravity
    n, eps2 = 4096, 0.01 ** 2
    s = make_plummer(n, total_mass=1.0, seed=9, G=1.0)
    rng = np.random.default_rng(9)
    p = s.pos
    delta = rng.normal(size=p.shape) * 1e-4
    hi = p - delta
    lo = (p - hi).astype(np.float32)
    ref = orc.direct_acc(p, s.mass, 1.0, eps2)
    mass = torch.from_numpy(s.mass).cuda()
    got = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    gravity.compute_acc_direct_hilo(torch.from_numpy(hi).cuda(), torch.from_numpy(lo).cuda(),
                                    mass, got, n, 1.0, eps2)
    rel = np.linalg.norm(got.cpu().numpy() - ref, axis=1) / np.linalg.norm(ref, axis=1)
    without = torch.empty_like(got)
    gravity.compute_acc_direct(torch.from_numpy(hi).cuda(), mass, without, n, 1.0, eps2, "fp32c")
    rel0 = np.linalg.norm(without.cpu().numpy() - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print(f"  hi/lo max rel {rel.max():.2e}; pos alone {rel0.max():.2e}", flush=True)
    assert rel.max() <= 3e-6
    assert rel0.max() > 100 * rel.max()

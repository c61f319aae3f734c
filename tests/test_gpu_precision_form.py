"""The precision form at its full size: Plummer N=262,144 (a=1, G=M=1,
eps 0.005, seed 0).  FP32C pair sums (FP32 on hi/lo positions, compensated
partials) against the FP64 sum in the reference's order (gravity.py:13-26,
bit-exact with the oracle at every size the oracle tests run), held to the
1e-6 maximum relative force error the form is specified with; then 50 KDK
steps at dt 5e-4 (engine.py:108-116) with FP64-hi + FP32-lo positions in
the drift, against the FP64 engine's trajectory and energy drift.  The
plain-FP32 variant of the same configuration is held to its own stated
bound."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_fp32c_full_size_error_and_kdk():
    from rgsim.engine import Engine
    from rgsim.ic import make_plummer
    from rgsim.solvers import gravity
    n, eps2 = 262144, 0.005 ** 2
    print(f"precision form n={n}", flush=True)
    s = make_plummer(n, total_mass=1.0, scale_radius=1.0, seed=0, G=1.0)
    pos = torch.from_numpy(s.pos).cuda()
    mass = torch.from_numpy(s.mass).cuda()
    ref = torch.empty_like(pos)
    gravity.compute_acc_direct(pos, mass, ref, n, 1.0, eps2, "fp64")
    got = torch.empty_like(pos)
    gravity.compute_acc_direct(pos, mass, got, n, 1.0, eps2, "fp32c")
    r, g = ref.cpu().numpy(), got.cpu().numpy()
    rel = np.linalg.norm(g - r, axis=1) / np.linalg.norm(r, axis=1)
    print(f"  fp32c max {rel.max():.2e} rms {np.sqrt(np.mean(rel ** 2)):.2e}", flush=True)
    assert rel.max() <= 1e-6 and np.sqrt(np.mean(rel ** 2)) <= 1e-7

    # 50 KDK steps with hi/lo positions (compensated drift, FP32C forces on
    # both parts) beside the FP64 engine (the reference's arithmetic, whose
    # trajectory is the oracle's bit for bit at every size the oracle runs)
    eng = Engine(softening=0.005, precision="fp32c", G=1.0)
    eng.load_state(s)
    e64 = Engine(softening=0.005, precision="fp64", G=1.0)
    e64.load_state(s)
    assert eng.pos_lo is not None and e64.pos_lo is None
    _, _, e0 = eng.energies()
    p0 = np.array(eng.momentum())
    for _ in range(50):
        eng.step(5e-4)
        e64.step(5e-4)
    ke1, pe1, e1 = eng.energies()
    p1 = np.array(eng.momentum())
    _, _, f1 = e64.energies()
    drift = abs((e1 - e0) / e0)
    drift64 = abs((f1 - e0) / e0)
    out, ref64 = eng.to_state(), e64.to_state()
    x = out.pos + out.pos_lo.astype(np.float64)
    dx = np.abs(x - ref64.pos).max() / np.abs(ref64.pos).max()
    dv = np.abs(out.vel - ref64.vel).max() / np.abs(ref64.vel).max()
    print(f"  50 KDK steps: |dE/E| {drift:.2e} (FP64 engine {drift64:.2e}), virial "
          f"{-2 * ke1 / pe1:.4f}, max |dx| {dx:.2e}, max |dv| {dv:.2e} of the scale",
          flush=True)
    assert drift < 1e-8 and drift <= 1.05 * drift64 + 1e-12
    assert dx <= 1e-12 and dv <= 1e-8
    assert 0.8 < -2.0 * ke1 / pe1 < 1.2
    assert np.linalg.norm(p1 - p0) < 1e-12


def test_fp32_full_size_error():
    """Plain FP32 pair sums at the precision form's size (N=262,144, eps
    0.005, seed 0), against the FP64 sum: FP32 positions lose |x| / |d| of
    the separation's accuracy for near pairs, and each of the 32 j-slices is
    summed naively (slices added in FP64).  Stated bound: maximum relative
    error 5e-5 (measured 1.7e-5), RMS 2e-6."""
    from rgsim.ic import make_plummer
    from rgsim.solvers import gravity
    n, eps2 = 262144, 0.005 ** 2
    print(f"precision form fp32 n={n}", flush=True)
    s = make_plummer(n, total_mass=1.0, scale_radius=1.0, seed=0, G=1.0)
    pos = torch.from_numpy(s.pos).cuda()
    mass = torch.from_numpy(s.mass).cuda()
    ref = torch.empty_like(pos)
    gravity.compute_acc_direct(pos, mass, ref, n, 1.0, eps2, "fp64")
    got = torch.empty_like(pos)
    gravity.compute_acc_direct(pos, mass, got, n, 1.0, eps2, "fp32")
    r, g = ref.cpu().numpy(), got.cpu().numpy()
    rel = np.linalg.norm(g - r, axis=1) / np.linalg.norm(r, axis=1)
    rms = np.sqrt(np.mean(rel ** 2))
    print(f"  fp32 max {rel.max():.2e} rms {rms:.2e}", flush=True)
    assert rel.max() <= 5e-5 and rms <= 2e-6

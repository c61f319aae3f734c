"""Differentiation form: the VJP of the direct acceleration and the
checkpointed KDK reverse, pinned by torch autograd of a dense FP64
implementation of gravity.py:13-26 / engine.py:108-116 and by FP64 central
finite differences."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def dense_acc(x, m, G, eps2):
    """gravity.py:13-26 as dense torch ops (differentiable)."""
    d = x[None, :, :] - x[:, None, :]
    r2 = (d * d).sum(-1) + eps2
    inv3 = r2 ** -1.5
    return G * torch.einsum("ij,ijk->ik", m[None, :] * inv3, d)


def vjp_closed_form(x, m, lam, G, eps2):
    """g_x[k] = G sum_j T_kj (m_k lam_j - m_j lam_k); g_m[k] = -G sum_j lam_j.d_kj/s^3."""
    d = x[None, :, :] - x[:, None, :]          # d[k, j] = x_j - x_k
    s2 = (d * d).sum(-1) + eps2
    inv3 = s2 ** -1.5
    inv5 = s2 ** -2.5
    v = m[:, None, None] * lam[None, :, :] - m[None, :, None] * lam[:, None, :]
    dv = (d * v).sum(-1)
    gx = G * ((v * inv3[..., None]).sum(1) - 3 * ((dv * inv5)[..., None] * d).sum(1))
    gm = -G * ((lam[None, :, :] * d).sum(-1) * inv3).sum(1)
    return gx, gm


def plummer(n, seed):
    import sys  # noqa: F401
    from rgsim.ic import make_plummer
    return make_plummer(n, total_mass=1.0, scale_radius=1.0, seed=seed, G=1.0)


def test_vjp_formula_matches_autograd_cpu():
    s = plummer(64, 3)
    x = torch.tensor(s.pos, requires_grad=True)
    m = torch.tensor(s.mass * np.linspace(0.5, 1.5, 64), requires_grad=True)
    lam = torch.tensor(np.random.default_rng(0).normal(size=(64, 3)))
    a = dense_acc(x, m, 1.3, 1e-3)
    gx_ref, gm_ref = torch.autograd.grad((a * lam).sum(), (x, m))
    gx, gm = vjp_closed_form(x.detach(), m.detach(), lam, 1.3, 1e-3)
    torch.testing.assert_close(gx, gx_ref, rtol=1e-12, atol=1e-12)
    torch.testing.assert_close(gm, gm_ref, rtol=1e-12, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,rtol", [("fp64", 1e-11), ("fp32c", 2e-5)])
def test_gpu_vjp_vs_dense_autograd(precision, rtol):
    from rgsim.adjoint import direct_vjp
    print(f"vjp {precision} n=1024", flush=True)
    s = plummer(1024, 5)
    rng = np.random.default_rng(1)
    mass = s.mass * rng.uniform(0.5, 1.5, 1024)
    lam = rng.normal(size=(1024, 3))
    x = torch.tensor(s.pos, requires_grad=True)
    m = torch.tensor(mass, requires_grad=True)
    a = dense_acc(x, m, 1.0, 1e-4)
    gx_ref, gm_ref = torch.autograd.grad((a * torch.tensor(lam)).sum(), (x, m))
    gx, gm = direct_vjp(torch.tensor(s.pos).cuda(), torch.tensor(mass).cuda(),
                        torch.tensor(lam).cuda(), 1.0, 1e-4, precision)
    scale_x = gx_ref.abs().max()
    scale_m = gm_ref.abs().max()
    assert (gx.cpu() - gx_ref).abs().max() <= rtol * scale_x
    assert (gm.cpu() - gm_ref).abs().max() <= rtol * scale_m


@pytest.mark.gpu
def test_gpu_direct_acceleration_function():
    from rgsim.adjoint import direct_acceleration
    s = plummer(500, 2)
    x = torch.tensor(s.pos, device="cuda", requires_grad=True)
    m = torch.tensor(s.mass, device="cuda", requires_grad=True)
    a = direct_acceleration(x, m, 1.0, 0.01)
    (a.square().sum()).backward()
    xc = torch.tensor(s.pos, requires_grad=True)
    mc = torch.tensor(s.mass, requires_grad=True)
    ac = dense_acc(xc, mc, 1.0, 1e-4)
    (ac.square().sum()).backward()
    torch.testing.assert_close(a.detach().cpu(), ac.detach(), rtol=1e-12, atol=1e-12)
    torch.testing.assert_close(x.grad.cpu(), xc.grad, rtol=1e-9, atol=1e-9)
    torch.testing.assert_close(m.grad.cpu(), mc.grad, rtol=1e-9, atol=1e-9)


def dense_trajectory(x0, v0, m, steps, dt, G, eps2):
    """engine.py:108-116 in dense torch ops (autograd reference)."""
    h = 0.5 * dt
    x, v = x0, v0
    a = dense_acc(x, m, G, eps2)
    for _ in range(steps):
        v = v + a * h
        x = x + v * dt
        a = dense_acc(x, m, G, eps2)
        v = v + a * h
    return x, v


@pytest.mark.gpu
@pytest.mark.parametrize("ckpt", [7, 30])
def test_gpu_trajectory_vs_dense_autograd(ckpt):
    from rgsim.adjoint import kdk_trajectory
    print(f"trajectory ckpt={ckpt}", flush=True)
    s = plummer(256, 9)
    rng = np.random.default_rng(4)
    mass = s.mass * rng.uniform(0.8, 1.2, 256)
    target = torch.tensor(s.pos + 0.01 * rng.normal(size=s.pos.shape))
    steps, dt, eps = 30, 2e-3, 0.05
    x0 = torch.tensor(s.pos, device="cuda", requires_grad=True)
    v0 = torch.tensor(s.vel, device="cuda", requires_grad=True)
    m = torch.tensor(mass, device="cuda", requires_grad=True)
    xN, vN = kdk_trajectory(x0, v0, m, steps, dt, 1.0, eps, ckpt, "fp64")
    loss = ((xN - target.cuda()) ** 2).sum(1).mean() + 0.1 * (vN ** 2).sum()
    loss.backward()
    xc = torch.tensor(s.pos, requires_grad=True)
    vc = torch.tensor(s.vel, requires_grad=True)
    mc = torch.tensor(mass, requires_grad=True)
    xr, vr = dense_trajectory(xc, vc, mc, steps, dt, 1.0, eps ** 2)
    lr = ((xr - target) ** 2).sum(1).mean() + 0.1 * (vr ** 2).sum()
    lr.backward()
    torch.testing.assert_close(xN.detach().cpu(), xr.detach(), rtol=1e-11, atol=1e-12)
    for g, gr in ((x0.grad, xc.grad), (v0.grad, vc.grad), (m.grad, mc.grad)):
        assert (g.cpu() - gr).abs().max() <= 1e-8 * gr.abs().max()


@pytest.mark.gpu
def test_gpu_trajectory_vs_central_differences():
    """FP64 central differences (h = 1e-6) on 12 random coordinates."""
    from rgsim.adjoint import kdk_trajectory
    s = plummer(512, 11)
    rng = np.random.default_rng(12)
    target = torch.tensor(s.pos + 1e-3 * rng.normal(size=s.pos.shape), device="cuda")
    steps, dt, eps, ck = 40, 1e-3, 0.02, 9

    def loss_of(x, v, m):
        xN, _ = kdk_trajectory(x, v, m, steps, dt, 1.0, eps, ck, "fp64")
        return ((xN - target) ** 2).sum(1).mean()

    x0 = torch.tensor(s.pos, device="cuda", requires_grad=True)
    v0 = torch.tensor(s.vel, device="cuda", requires_grad=True)
    m = torch.tensor(s.mass, device="cuda", requires_grad=True)
    loss_of(x0, v0, m).backward()
    grads = {"x": x0.grad.cpu().numpy(), "v": v0.grad.cpu().numpy(), "m": m.grad.cpu().numpy()}
    base = {"x": s.pos, "v": s.vel, "m": s.mass}
    h = 1e-6
    for k in range(12):
        which = ("x", "v", "m")[k % 3]
        arr = base[which]
        idx = tuple(rng.integers(0, d) for d in arr.shape)
        vals = []
        for sgn in (1, -1):
            a = {kk: vv.copy() for kk, vv in base.items()}
            a[which][idx] += sgn * h
            with torch.no_grad():
                vals.append(loss_of(*(torch.tensor(a[kk], device="cuda") for kk in "xvm")).item())
        fd = (vals[0] - vals[1]) / (2 * h)
        g = grads[which][idx]
        assert abs(g - fd) <= 1e-5 * abs(fd) + 1e-9, (which, idx, g, fd)


@pytest.mark.gpu
def test_gpu_full_config_central_differences():
    """Differentiation config at full size: Plummer N=16384 (eps 0.01), 400
    KDK steps dt 1e-3, checkpoints every 20 steps, loss = mean squared
    distance of the final positions to a perturbed target; FP64 central
    differences (h = 1e-6) on 64 random coordinates of (x0, v0, m).  The FP64
    adjoint must match to 1e-6 relative; the FP32-compensated adjoint's
    deviation is bounded at 1e-3 relative of the gradient scale."""
    from rgsim.adjoint import kdk_trajectory
    from rgsim.ic import make_plummer
    print("full adjoint config N=16384 steps=400", flush=True)
    s = make_plummer(16384, total_mass=1.0, scale_radius=1.0, seed=7, G=1.0)
    rng = np.random.default_rng(9)
    tgt_state = make_plummer(16384, total_mass=1.0, scale_radius=1.0, seed=7, G=1.0)
    tgt0 = tgt_state.pos + 1e-3 * np.random.default_rng(8).normal(size=s.pos.shape)
    with torch.no_grad():
        tgt, _ = kdk_trajectory(torch.tensor(tgt0, device="cuda"),
                                torch.tensor(tgt_state.vel, device="cuda"),
                                torch.tensor(tgt_state.mass, device="cuda"),
                                400, 1e-3, 1.0, 0.01, 20, "fp64")
    base = {"x": s.pos, "v": s.vel, "m": s.mass}

    def loss_of(arrs, precision="fp64"):
        xN, _ = kdk_trajectory(*arrs, 400, 1e-3, 1.0, 0.01, 20, precision)
        return ((xN - tgt) ** 2).sum(1).mean()

    grads = {}
    for precision in ("fp64", "fp32c"):
        ts = [torch.tensor(base[k], device="cuda", requires_grad=True) for k in "xvm"]
        loss_of(ts, precision).backward()
        grads[precision] = {k: t.grad.cpu().numpy() for k, t in zip("xvm", ts)}
    h = 1e-6
    worst64 = worst32 = 0.0
    for c in range(64):
        which = "xvm"[c % 3]
        arr = base[which]
        idx = tuple(int(rng.integers(0, d)) for d in arr.shape)
        vals = []
        for sgn in (1, -1):
            a = {k: v.copy() for k, v in base.items()}
            a[which][idx] += sgn * h
            with torch.no_grad():
                vals.append(loss_of([torch.tensor(a[k], device="cuda") for k in "xvm"]).item())
        fd = (vals[0] - vals[1]) / (2 * h)
        scale = np.abs(grads["fp64"][which]).max()
        worst64 = max(worst64, abs(grads["fp64"][which][idx] - fd) / scale)
        worst32 = max(worst32, abs(grads["fp32c"][which][idx] - fd) / scale)
    print(f"max |adjoint - FD| / max|grad|: fp64 {worst64:.2e}, fp32c {worst32:.2e}")
    assert worst64 < 1e-6
    assert worst32 < 1e-3

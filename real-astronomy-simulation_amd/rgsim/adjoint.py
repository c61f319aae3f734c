"""Differentiation form: gradients through the direct acceleration and
through leapfrog KDK trajectories, with hand-written HIP adjoint kernels.

The reference (core/solvers/gravity.py:13-26, core/engine.py:108-116) has no
gradients.  Two ``torch.autograd.Function``s sit on the HIP C ABI:

* ``DirectAcceleration``: a = A(x, m); backward is ``rg_direct_vjp`` (the
  tidal-tensor contraction for x and the mass derivative for m).
* ``KDKTrajectory``: ``steps`` reference KDK steps from (x0, v0) with masses
  m.  Forward keeps only every ``checkpoint_every``-th state; backward
  recomputes each segment forward from its checkpoint, then runs the reverse
  steps.  One VJP per step: the acceleration at a step's end positions feeds
  both that step's closing kick and the next step's opening kick, so their
  cotangents are merged (lam = abar + vbar*dt/2).

``precision`` is "fp64" (FP64 forward and adjoint: the setting for
finite-difference checks) or "fp32c" (FP32 pair sums on hi/lo positions with
compensated accumulation, both ways).  Tensors are float64 CUDA tensors.
"""
from __future__ import annotations

import torch

from . import _lib
from .solvers import gravity
from .integrators import leapfrog


def direct_vjp(pos, mass, lam, G, eps2, precision="fp64", out=None):
    """(g_pos, g_mass) = VJP of the direct acceleration at (pos, mass) for lam."""
    n = pos.shape[0]
    g_pos = torch.empty_like(pos) if out is None else out[0]
    g_mass = torch.empty_like(mass) if out is None else out[1]
    for t, nm in ((pos, "pos"), (lam, "lam"), (g_pos, "g_pos")):
        if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous() or t.numel() != 3 * n:
            raise TypeError(f"{nm} must be a contiguous float64 CUDA tensor of shape ({n}, 3)")
    for t, nm in ((mass, "mass"), (g_mass, "g_mass")):
        if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous() or t.numel() != n:
            raise TypeError(f"{nm} must be a contiguous float64 CUDA tensor of shape ({n},)")
    _lib.require_gpu()
    _lib.check(_lib.load().rg_direct_vjp(pos.data_ptr(), mass.data_ptr(), lam.data_ptr(),
                                         g_pos.data_ptr(), g_mass.data_ptr(), n, float(G),
                                         float(eps2), _lib.precision_code(precision),
                                         _lib.stream_ptr(pos)), "rg_direct_vjp")
    return g_pos, g_mass


def _acc(pos, mass, G, eps2, precision):
    acc = torch.empty_like(pos)
    gravity.compute_acc_direct(pos, mass, acc, pos.shape[0], G, eps2, precision)
    return acc


class DirectAcceleration(torch.autograd.Function):
    """a = G sum_j m_j (x_j - x_i) / (|x_j - x_i|^2 + eps2)^{3/2}."""

    @staticmethod
    def forward(ctx, pos, mass, G, eps2, precision="fp64"):
        pos = pos.contiguous()
        mass = mass.contiguous()
        ctx.save_for_backward(pos, mass)
        ctx.G, ctx.eps2, ctx.precision = float(G), float(eps2), precision
        return _acc(pos, mass, ctx.G, ctx.eps2, precision)

    @staticmethod
    def backward(ctx, grad_acc):
        pos, mass = ctx.saved_tensors
        g_pos, g_mass = direct_vjp(pos, mass, grad_acc.contiguous(), ctx.G, ctx.eps2,
                                   ctx.precision)
        return g_pos, g_mass, None, None, None


def direct_acceleration(pos, mass, G, eps, precision="fp64"):
    return DirectAcceleration.apply(pos, mass, G, float(eps) ** 2, precision)


def _forward_segment(x, v, a, mass, steps, dt, G, eps2, precision, keep=None,
                     positions_only=False):
    """Advance (x, v, a) in place by `steps` KDK steps (the Engine's kernels);
    if `keep` is a list, append a copy of x after each drift.  With
    positions_only the last step stops after its drift: the positions of a
    segment need the forces at its first `steps` positions only, so the
    force at the segment's end (and the closing kick) is left out and v, a
    are not advanced past the last opening kick."""
    n = x.shape[0]
    half = 0.5 * dt
    for s in range(steps):
        leapfrog.kick_drift(x, v, a, n, half, dt)
        if keep is not None:
            keep.append(x.clone())
        if positions_only and s == steps - 1:
            break
        gravity.compute_acc_direct_kick(x, mass, a, v, n, G, eps2, half, precision)


class KDKTrajectory(torch.autograd.Function):
    """(x_N, v_N) after `steps` KDK steps from (x0, v0); a0 = A(x0)."""

    @staticmethod
    def forward(ctx, x0, v0, mass, steps, dt, G, eps, checkpoint_every=20, precision="fp64"):
        steps, k = int(steps), max(1, int(checkpoint_every))
        eps2 = float(eps) ** 2
        x = x0.detach().clone().contiguous()
        v = v0.detach().clone().contiguous()
        m = mass.detach().contiguous()
        a = _acc(x, m, G, eps2, precision)
        ckpts = []
        done = 0
        while done < steps:
            ckpts.append((x.clone(), v.clone(), a.clone()))
            seg = min(k, steps - done)
            _forward_segment(x, v, a, m, seg, dt, G, eps2, precision)
            done += seg
        ctx.save_for_backward(x0.detach(), m, *[t for c in ckpts for t in c])
        ctx.meta = (steps, k, float(dt), float(G), eps2, precision, len(ckpts))
        return x, v

    @staticmethod
    def backward(ctx, grad_x, grad_v):
        steps, k, dt, G, eps2, precision, nck = ctx.meta
        saved = ctx.saved_tensors
        x0, m = saved[0], saved[1]
        ckpts = [saved[2 + 3 * i: 5 + 3 * i] for i in range(nck)]
        n = x0.shape[0]
        xbar = (grad_x if grad_x is not None else torch.zeros_like(x0)).contiguous().clone()
        vbar = (grad_v if grad_v is not None else torch.zeros_like(x0)).contiguous().clone()
        mbar = torch.zeros_like(m)
        lam = vbar * (0.5 * dt)  # abar_N = 0: the final acceleration is not an output
        g_pos, g_mass = torch.empty_like(x0), torch.empty_like(m)
        L = _lib.load()
        stream = _lib.stream_ptr(x0)
        for seg in range(nck - 1, -1, -1):
            cx, cv, ca = (t.clone() for t in ckpts[seg])
            seg_steps = min(k, steps - seg * k)
            xs = [cx.clone()]
            _forward_segment(cx, cv, ca, m, seg_steps, dt, G, eps2, precision, keep=xs,
                             positions_only=True)
            for s in range(seg_steps - 1, -1, -1):
                x_end = xs[s + 1]
                direct_vjp(x_end, m, lam, G, eps2, precision, out=(g_pos, g_mass))
                # the initial acceleration A(x0) feeds only the first opening
                # kick, so its weight is vbar*dt/2 instead of vbar*dt
                coef = 0.5 * dt if (seg == 0 and s == 0) else dt
                _lib.check(L.rg_kdk_adjoint_update(xbar.data_ptr(), vbar.data_ptr(),
                                                   mbar.data_ptr(), g_pos.data_ptr(),
                                                   g_mass.data_ptr(), lam.data_ptr(), n,
                                                   dt, coef, stream), "rg_kdk_adjoint_update")
        direct_vjp(x0.contiguous(), m, lam, G, eps2, precision, out=(g_pos, g_mass))
        _lib.check(L.rg_kdk_adjoint_update(xbar.data_ptr(), vbar.data_ptr(), mbar.data_ptr(),
                                           g_pos.data_ptr(), g_mass.data_ptr(), lam.data_ptr(),
                                           n, 0.0, 0.0, stream), "rg_kdk_adjoint_update")
        return xbar, vbar, mbar, None, None, None, None, None, None


def kdk_trajectory(x0, v0, mass, steps, dt, G, eps, checkpoint_every=20, precision="fp64"):
    return KDKTrajectory.apply(x0, v0, mass, steps, dt, G, eps, checkpoint_every, precision)

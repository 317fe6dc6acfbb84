"""torch.autograd bindings of the engine (C ABI through ctypes).

DebyeIntensity: positions [N, 3] -> I(q) [n_q] (FP64); backward is the adjoint
kernel's vector-Jacobian product sum_k w_k dI_k/dx_i (rsx_debye_vjp).
Chi2Loss: positions -> chi^2 at the least-squares scale (saxs_profile.hpp:137-166
scale, chi^2 = sum (I_exp - c I)^2 / sigma^2); forward and adjoint run in one
fused call (rsx_chi2_grad) and backward scales the stored gradient.  The scale
is held fixed under differentiation (it is optimal, so dchi^2/dc = 0).

Tensors may live on any device; the C ABI takes host buffers, so values cross
to the host and back (the engine keeps its own device copies).
"""
import numpy as np
import torch

from .engine import make_bodies


def _host(pos):
    return pos.detach().to("cpu", torch.float32).contiguous().numpy()


class DebyeIntensity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, engine, types):
        bodies = make_bodies(_host(pos), types)
        iq = engine.calc_curve(bodies)
        ctx.engine = engine
        ctx.bodies = bodies
        ctx.pos_meta = (pos.device, pos.dtype)
        return torch.from_numpy(iq).to(pos.device)

    @staticmethod
    def backward(ctx, grad_iq):
        w = grad_iq.detach().to("cpu", torch.float64).contiguous().numpy()
        g = ctx.engine.vjp(ctx.bodies, w)
        dev, dt = ctx.pos_meta
        return torch.from_numpy(g).to(dev, dt), None, None


class Chi2Loss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, engine, types, i_exp, sigma):
        bodies = make_bodies(_host(pos), types)
        chi2, scale, grad, iq = engine.chi2_grad(bodies, np.asarray(i_exp, np.float32), np.asarray(sigma, np.float32))
        ctx.grad = torch.from_numpy(grad).to(pos.device, pos.dtype)
        ctx.scale = scale
        ctx.iq = iq
        return torch.tensor(chi2, dtype=torch.float64, device=pos.device)

    @staticmethod
    def backward(ctx, g):
        return ctx.grad * g.to(ctx.grad.dtype), None, None, None, None


def debye_intensity(pos, engine, types):
    """I(q) of positions pos [N, 3] (differentiable)."""
    return DebyeIntensity.apply(pos, engine, types)


def chi2_loss(pos, engine, types, i_exp, sigma):
    """chi^2 of positions against I_exp (sigma) at the optimal scale (differentiable)."""
    return Chi2Loss.apply(pos, engine, types, i_exp, sigma)


def refine(engine, types, start, i_exp, sigma, steps=100, lr=0.01):
    """Gradient-descent refinement of positions against a profile (Adam on chi^2).

    Returns (positions [N, 3] float32, chi^2 per step)."""
    pos = torch.tensor(np.asarray(start, np.float32), requires_grad=True)
    opt = torch.optim.Adam([pos], lr=lr)
    hist = []
    for _ in range(steps):
        opt.zero_grad()
        loss = chi2_loss(pos, engine, types, i_exp, sigma)
        loss.backward()
        opt.step()
        hist.append(float(loss))
    return pos.detach().numpy(), hist

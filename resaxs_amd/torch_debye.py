"""torch.autograd bindings of the engine (C ABI through ctypes).

DebyeIntensity: positions [N, 3] -> I(q) [n_q] (FP64); backward is the adjoint
kernel's vector-Jacobian product sum_k w_k dI_k/dx_i.
EnsembleIntensity: the positions of every frame of an ensemble, concatenated, and the
frame offsets -> per-frame I(q) [n_frames, n_q] (FP64) from one batched launch; backward
is the batched VJP (one adjoint launch over all frames).  The frames are laid out from a
host copy of the positions on every forward call.
Chi2Loss: positions -> chi^2 at the least-squares scale (saxs_profile.hpp:137-166
scale, chi^2 = sum (I_exp - c I)^2 / sigma^2); forward and adjoint run in one
fused call and backward scales the stored gradient.  The scale is held fixed
under differentiation (it is optimal, so dchi^2/dc = 0).

Positions on a GPU stay there: the bodies {x, y, z, type bits} are assembled on the
device and the engine's device-resident entry points (rsx_calc_curve_dev_async,
rsx_debye_vjp_dev_async, rsx_chi2_grad_dev_async) read and write the tensors' memory.
They are ordered against torch by synchronising torch's current stream before the call
and the engine's stream after it.  Those entry points reuse the engine's layout of the
atoms from a host call; the first call on an engine (and any call whose positions left
the radius the layout was made for, or after a parameter change) lays the atoms out from
a host copy of the positions and runs again.  Positions on the CPU go through the host
entry points.

Import torch before the library is loaded (before the first Engine): torch ships its own
HIP runtime under the same soname as the system one, and the library then uses torch's;
a process that loaded the library first cannot initialise torch's GPU support.
"""
import ctypes

import numpy as np
import torch

from ._lib import RSX_ARGS_NOT_SET, RSX_INVALID_ARG, RSX_NO_INITIAL_CALC, RsxError
from .engine import make_bodies

# status codes after which a device-resident call is retried once on a fresh host layout
_RELAYOUT = (RSX_ARGS_NOT_SET, RSX_NO_INITIAL_CALC, RSX_INVALID_ARG)


def _host(pos):
    return pos.detach().to("cpu", torch.float32).contiguous().numpy()


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _types_on(engine, types, device):
    """The types as an int32 tensor on device (cached on the engine per device)."""
    cache = engine.__dict__.setdefault("_torch_types", {})
    key = (str(device), id(types))
    t = cache.get(key)
    if t is None:
        t = torch.as_tensor(np.ascontiguousarray(types, dtype=np.uint32).view(np.int32), device=device)
        cache.clear()
        cache[key] = t
    return t


def _device_bodies(pos, engine, types):
    tt = _types_on(engine, types, pos.device)
    return torch.cat([pos.detach().to(torch.float32), tt.view(torch.float32).unsqueeze(1)], 1).contiguous()


def _on_device(engine, pos, types, call, layout):
    """call() enqueues device work on the engine stream; run it ordered after torch's work on
    pos's device and wait for it.  On a layout-related failure, lay out with layout() (a
    host call) and run once more."""
    for attempt in range(2):
        torch.cuda.current_stream(pos.device).synchronize()
        try:
            call()
            engine.synchronize()
            return
        except RsxError as e:
            if attempt or e.code not in _RELAYOUT:
                raise
            layout()


class DebyeIntensity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, engine, types):
        ctx.engine, ctx.types = engine, types
        if pos.is_cuda:
            bodies = _device_bodies(pos, engine, types)
            out = torch.empty(engine.nq, dtype=torch.float64, device=pos.device)

            def layout():
                engine.calc_curve(make_bodies(_host(pos), types))

            if engine.state != 4 or getattr(engine, "n", None) != len(pos):
                layout()
            _on_device(engine, pos, types, lambda: engine.calc_curve_dev_async(_ptr(bodies), len(pos), _ptr(out)),
                       layout)
            ctx.bodies_dev = bodies
            return out
        bodies = make_bodies(_host(pos), types)
        iq = engine.calc_curve(bodies)
        ctx.bodies = bodies
        ctx.pos_meta = (pos.device, pos.dtype)
        return torch.from_numpy(iq).to(pos.device)

    @staticmethod
    def backward(ctx, grad_iq):
        engine = ctx.engine
        if grad_iq.is_cuda:
            bodies = ctx.bodies_dev
            w = grad_iq.detach().to(torch.float64).contiguous()
            grad = torch.empty((len(bodies), 3), dtype=torch.float32, device=bodies.device)

            def layout():
                host = bodies.to("cpu").numpy()
                engine.calc_curve(host)

            _on_device(engine, bodies, ctx.types,
                       lambda: engine.vjp_dev_async(_ptr(bodies), len(bodies), _ptr(w), _ptr(grad)), layout)
            return grad, None, None
        w = grad_iq.detach().to("cpu", torch.float64).contiguous().numpy()
        g = engine.vjp(ctx.bodies, w)
        dev, dt = ctx.pos_meta
        return torch.from_numpy(g).to(dev, dt), None, None


class Chi2Loss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, engine, types, i_exp, sigma):
        if pos.is_cuda:
            bodies = _device_bodies(pos, engine, types)
            ie = torch.as_tensor(np.asarray(i_exp, np.float32), device=pos.device)
            sg = torch.as_tensor(np.asarray(sigma, np.float32), device=pos.device)
            out = torch.empty(2, dtype=torch.float64, device=pos.device)
            grad = torch.empty((len(pos), 3), dtype=torch.float32, device=pos.device)

            def layout():
                # a host gradient call lays the atoms out in the gradient configuration
                engine.chi2_grad(make_bodies(_host(pos), types), np.asarray(i_exp, np.float32),
                                 np.asarray(sigma, np.float32))

            if engine.state != 4 or getattr(engine, "n", None) != len(pos):
                layout()
            _on_device(engine, pos, types,
                       lambda: engine.chi2_grad_dev_async(_ptr(bodies), len(pos), _ptr(ie), _ptr(sg), _ptr(out),
                                                          _ptr(grad)), layout)
            ctx.grad = grad.to(pos.dtype)
            ctx.scale = float(out[1])
            return out[0].clone()
        bodies = make_bodies(_host(pos), types)
        chi2, scale, grad, iq = engine.chi2_grad(bodies, np.asarray(i_exp, np.float32), np.asarray(sigma, np.float32))
        ctx.grad = torch.from_numpy(grad).to(pos.device, pos.dtype)
        ctx.scale = scale
        ctx.iq = iq
        return torch.tensor(chi2, dtype=torch.float64, device=pos.device)

    @staticmethod
    def backward(ctx, g):
        return ctx.grad * g.to(ctx.grad.dtype), None, None, None, None


class EnsembleIntensity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, engine, types, offsets):
        bodies = make_bodies(_host(pos), types)
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        engine.batch_set_frames(bodies, off)
        frames, _ = engine.batch_calc()
        ctx.engine, ctx.bodies, ctx.offsets = engine, bodies, off
        ctx.gen = engine.batch_gen
        ctx.pos_meta = (pos.device, pos.dtype)
        return torch.from_numpy(frames).to(pos.device)

    @staticmethod
    def backward(ctx, grad_frames):
        engine = ctx.engine
        if engine.batch_gen != ctx.gen:  # the engine's batch holds other frames now
            engine.batch_set_frames(ctx.bodies, ctx.offsets)
            ctx.gen = engine.batch_gen
        w = grad_frames.detach().to("cpu", torch.float64).contiguous().numpy()
        g = engine.batch_vjp(w)
        dev, dt = ctx.pos_meta
        return torch.from_numpy(g).to(dev, dt), None, None, None


def ensemble_intensity(pos, engine, types, offsets):
    """Per-frame I(q) [n_frames, n_q] of an ensemble whose frame f holds the atoms
    offsets[f]..offsets[f+1] of pos [total, 3] (differentiable)."""
    return EnsembleIntensity.apply(pos, engine, types, offsets)


def debye_intensity(pos, engine, types):
    """I(q) of positions pos [N, 3] (differentiable)."""
    return DebyeIntensity.apply(pos, engine, types)


def chi2_loss(pos, engine, types, i_exp, sigma):
    """chi^2 of positions against I_exp (sigma) at the optimal scale (differentiable)."""
    return Chi2Loss.apply(pos, engine, types, i_exp, sigma)


def refine(engine, types, start, i_exp, sigma, steps=100, lr=0.01, device="cpu"):
    """Gradient-descent refinement of positions against a profile (Adam on chi^2), with the
    positions on `device` ("cuda" keeps them in HBM through the device-resident entry points).

    Returns (positions [N, 3] float32, chi^2 per step)."""
    pos = torch.tensor(np.asarray(start, np.float32), device=device, requires_grad=True)
    opt = torch.optim.Adam([pos], lr=lr)
    hist = []
    for _ in range(steps):
        opt.zero_grad()
        loss = chi2_loss(pos, engine, types, i_exp, sigma)
        loss.backward()
        opt.step()
        hist.append(float(loss))
    return pos.detach().cpu().numpy(), hist

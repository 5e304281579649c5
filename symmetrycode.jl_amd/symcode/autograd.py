"""torch.autograd bindings of the hand-written adjoint: a differentiable
RK4 / Wray3 step for training loops (the ChainRulesCore ``rrule`` a Julia
caller would add maps to the same C entry points).

Gradients follow torch's convention for complex tensors under the plain
inner product Re sum conj(a) b; the library's adjoint uses the
RFFT-weighted inner product, so cotangents are converted by the RFFT
weights (2 on interior kx columns, 1 on kx = 0 and kx = n/2) on the way in
and out.  The stored checkpoint is the step input only; the stage values
are recomputed in the backward pass.
"""
import torch

from . import solver


def rfft_weight(g, device=None):
    w = torch.ones(g.nhp, dtype=g.dtype, device=device or g.device)
    w[1:g.nh - 1] = 2.0
    w[g.nh:] = 0.0  # pad columns carry nothing
    return w


class _Step(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, dt, g, visc, scheme):
        out = u.detach().clone().contiguous()
        if scheme == "rk4":
            solver.rk4(out, dt, g, visc)
        else:
            solver.wray3(out, dt, g, visc)
        ctx.save_for_backward(u.detach())
        ctx.cfg = (dt, g, visc, scheme)
        return out

    @staticmethod
    def backward(ctx, gout):
        (u0,) = ctx.saved_tensors
        dt, g, visc, scheme = ctx.cfg
        w = rfft_weight(g, gout.device)
        ubar = (gout / w.clamp(min=1.0)).contiguous()
        ubar = ubar * (w > 0)
        vjp = solver.rk4_vjp if scheme == "rk4" else solver.wray3_vjp
        gin = vjp(u0.contiguous(), ubar, dt, g, visc) * w
        return gin, None, None, None, None


def rk4_step(u, dt, g, visc):
    """Differentiable RK4 step (fixed dt)."""
    return _Step.apply(u, dt, g, visc, "rk4")


def wray3_step(u, dt, g, visc):
    """Differentiable wray3! step (fixed dt)."""
    return _Step.apply(u, dt, g, visc, "wray3")

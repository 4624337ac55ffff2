"""PyTorch autograd over the HIP projectors.

``project(x, P)`` is y = A x with the Joseph (or Siddon) model of a
:class:`~fdk_cuda_impl_amd.core.Projector`; its backward is the matched
transpose A^T y computed by the library's deterministic gather kernels, so
gradients of any torch loss on the projections are exact up to FP32 round-off
(the dot-product identity <A x, y> = <x, A^T y> holds to ~1e-7).
``backproject`` is the transpose as a differentiable op (backward = A).
``smoothed_tv`` is TV_eps(x) = sum sqrt(|grad x|^2 + eps^2) with forward
differences over the voxel spacing (src/Reconstruction/Iterative/Common/kernels/
YkTvRegularizationKernels.cu), value and gradient from one kernel.

``least_squares_tv`` minimizes 0.5 ||A x - b||^2 + lam TV_eps(x) by gradient
descent with Barzilai-Borwein steps and an Armijo safeguard; the volume is
updated in place and never stored per iteration.
"""
from __future__ import annotations

import math

import torch

from .core import Context, Projector


class _Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, volume, projector: Projector, kind: str):
        ctx.projector, ctx.kind = projector, kind
        return projector.forward(volume.detach().contiguous(), kind=kind)

    @staticmethod
    def backward(ctx, grad_proj):
        g = ctx.projector.adjoint(grad_proj.detach().contiguous(), kind=ctx.kind)
        return g, None, None


class _Backproject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, proj, projector: Projector, kind: str):
        ctx.projector, ctx.kind = projector, kind
        return projector.adjoint(proj.detach().contiguous(), kind=kind)

    @staticmethod
    def backward(ctx, grad_vol):
        g = ctx.projector.forward(grad_vol.detach().contiguous(), kind=ctx.kind)
        return g, None, None


class _SmoothedTV(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, context: Context, grid, eps: float):
        value, grad = context.tv(x.detach().contiguous(), grid, eps)
        ctx.save_for_backward(grad)
        return torch.tensor(value, dtype=torch.float64, device=x.device)

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        return grad * grad_out.to(grad.dtype), None, None, None


def project(volume: torch.Tensor, projector: Projector, kind: str = "joseph") -> torch.Tensor:
    """y = A x (differentiable; backward is the matched transpose)."""
    return _Project.apply(volume, projector, kind)


def backproject(proj: torch.Tensor, projector: Projector, kind: str = "joseph") -> torch.Tensor:
    """x = A^T y (differentiable; backward is the forward model)."""
    return _Backproject.apply(proj, projector, kind)


def smoothed_tv(x: torch.Tensor, context: Context, grid, eps: float) -> torch.Tensor:
    """TV_eps(x) as an FP64 scalar tensor (differentiable)."""
    return _SmoothedTV.apply(x, context, grid, eps)


def operator_norm_sq(projector: Projector, kind: str = "joseph", iterations: int = 12, seed: int = 0) -> float:
    """Largest eigenvalue of A^T A by power iteration."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.rand(projector.volume_shape, device="cuda", generator=g)
    lam = 0.0
    for _ in range(iterations):
        x /= x.norm()
        y = projector.adjoint(projector.forward(x, kind=kind), kind=kind)
        lam = float(torch.dot(x.reshape(-1), y.reshape(-1)))
        x = y
    return lam


def cgls(projector: Projector, b: torch.Tensor, iterations: int = 10, kind: str = "joseph", x0=None):
    """CGLS on ||A x - b||^2 with A and A^T applied through the differentiable
    operators above (the CGLS recursion of CglsRobustBackend / AstraClassic,
    src/Reconstruction/Iterative/Flat/YkCglsBackends.hpp:22-190; dots in FP64
    on the device, without FP64 copies of the vectors).  Returns x and
    ||b - A x|| before and after every iteration."""
    dot = projector.ctx.dot
    with torch.no_grad():
        if x0 is None:
            x = torch.zeros(projector.volume_shape, device=b.device)
            r = b.clone()
        else:
            x = x0.detach().clone()
            r = b - project(x, projector, kind)
        s = backproject(r, projector, kind)
        p = s.clone()
        gamma = dot(s, s)
        history = [math.sqrt(dot(r, r))]
        for _ in range(iterations):
            if gamma == 0.0:
                break
            q = project(p, projector, kind)
            qq = dot(q, q)
            if qq == 0.0:
                break
            alpha = gamma / qq
            x.add_(p, alpha=alpha)
            r.sub_(q, alpha=alpha)
            del q
            s = backproject(r, projector, kind)
            gamma_new = dot(s, s)
            p.mul_(gamma_new / gamma).add_(s)
            gamma = gamma_new
            history.append(math.sqrt(dot(r, r)))
    return x, history


def least_squares_tv(projector: Projector, b: torch.Tensor, context: Context, x0=None, steps: int = 50,
                     lam: float = 1e-3, eps: float = 1e-6, kind: str = "joseph", nonnegative: bool = False):
    """Gradient descent on f(x) = 0.5 ||A x - b||^2 + lam TV_eps(x).

    Step sizes: Barzilai-Borwein (s.s / s.y) from the second step on, the first
    one 1 / ||A^T A||; each step is accepted only under the Armijo condition
    f(x - t g) <= f(x) - 1e-4 t ||g||^2, halving t otherwise.  Returns x and
    the objective after each step (FP64 host floats)."""
    grid = projector.grid
    x = torch.zeros(projector.volume_shape, device=b.device) if x0 is None else x0.detach().clone()

    def objective_and_grad(v):
        v = v.detach().requires_grad_(True)
        r = project(v, projector, kind) - b
        f = 0.5 * torch.sum(r.double() * r.double()) + lam * smoothed_tv(v, context, grid, eps)
        (g,) = torch.autograd.grad(f, v)
        return float(f.detach()), g

    def objective(v):
        r = projector.forward(v, kind=kind) - b
        return 0.5 * float(torch.sum(r.double() * r.double())) + lam * context.tv(v, grid, eps)[0]

    f, g = objective_and_grad(x)
    history = [f]
    t = 1.0 / max(operator_norm_sq(projector, kind), 1e-30)
    x_prev = g_prev = None
    for _ in range(steps):
        if x_prev is not None:
            s = (x - x_prev).reshape(-1).double()
            y = (g - g_prev).reshape(-1).double()
            sy = float(torch.dot(s, y))
            if sy > 0.0:
                t = float(torch.dot(s, s)) / sy
        gg = float(torch.sum(g.double() * g.double()))
        if gg == 0.0:
            break
        while True:
            trial = x - t * g
            if nonnegative:
                trial.clamp_(min=0.0)
            ft = objective(trial)
            if ft <= f - 1e-4 * t * gg or t < 1e-30:
                break
            t *= 0.5
        x_prev, g_prev = x, g
        x = trial
        f, g = objective_and_grad(x)
        if not math.isfinite(f):
            raise FloatingPointError("objective is not finite")
        history.append(f)
    return x, history

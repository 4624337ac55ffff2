"""Differentiable projectors (torch.autograd over the C ABI): gradients against
the FP64 oracle and finite differences, the dot-product identity through
autograd, and the TV-regularised least-squares solve."""
import numpy as np
import pytest
import torch

import fdk_cuda_impl_amd as yk
from fdk_cuda_impl_amd import autograd as ag
from fdk_cuda_impl_amd import phantoms
from helpers import oracle_volume, rel_rmse, small_geometry

pytestmark = pytest.mark.gpu


def setup(ctx, seed=5, **kw):
    g = small_geometry(**kw)
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, g.grid())
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, 1.0, P.volume_shape).astype(np.float32)
    b = rng.uniform(0.0, 5.0, (P.n_views, g.nv, g.nu)).astype(np.float32)
    return g, P, x, b


@pytest.mark.parametrize("kind", ["joseph", "siddon"])
def test_least_squares_gradient_matches_oracle(ctx, oracle, kind):
    """grad 0.5 ||A x - b||^2 = A^T (A x - b), FP32 autograd vs FP64 oracle."""
    g, P, x, b = setup(ctx, tilt_rad=(0.02, -0.01, 0.01))
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    bt = torch.from_numpy(b).cuda()
    r = ag.project(xt, P, kind) - bt
    (0.5 * torch.sum(r * r)).backward()
    ov = oracle.views_from_float(P.views)
    vol = oracle_volume(oracle, g)
    fp = oracle.fp_joseph if kind == "joseph" else oracle.fp_siddon
    bp = oracle.bp_joseph_adjoint if kind == "joseph" else oracle.bp_siddon_adjoint
    ref = bp(ov, g.nu, g.nv, vol, fp(ov, g.nu, g.nv, vol, x) - b)
    assert rel_rmse(xt.grad.cpu().numpy(), ref) < 1e-5


def test_finite_differences_on_64_voxels(ctx):
    """The loss is quadratic in x, so a central difference with a large step is
    exact; 64 random voxels against the autograd gradient."""
    g, P, x, b = setup(ctx, seed=8)
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    bt = torch.from_numpy(b).cuda()

    def loss(v):
        r = (ag.project(v, P) - bt).double()
        return 0.5 * torch.sum(r * r)

    loss(xt).backward()
    grad = xt.grad.cpu().numpy()
    rng = np.random.default_rng(3)
    flat = rng.choice(x.size, 64, replace=False)
    h = 0.25
    worst = 0.0
    with torch.no_grad():
        base = torch.from_numpy(x).cuda()
        for k in flat:
            idx = np.unravel_index(k, x.shape)
            xp, xm = base.clone(), base.clone()
            xp[idx] += h
            xm[idx] -= h
            fd = float(loss(xp) - loss(xm)) / (2 * h)
            worst = max(worst, abs(fd - grad[idx]))
    assert worst < 1e-3 * np.abs(grad).max(), worst


def test_dot_identity_through_autograd(ctx):
    _, P, x, b = setup(ctx, seed=9)
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    yt = torch.from_numpy(b).cuda().requires_grad_(True)
    lhs = torch.sum(ag.project(xt, P).double() * yt.double())
    rhs = torch.sum(xt.double() * ag.backproject(yt, P).double())
    lhs, rhs = float(lhs.detach()), float(rhs.detach())
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    # backward of the transpose is the forward model
    (gy,) = torch.autograd.grad(torch.sum(ag.backproject(yt, P) * xt.detach()), yt)
    assert rel_rmse(gy.cpu().numpy(), P.forward(xt.detach()).cpu().numpy()) < 1e-6


def test_tv_gradient_matches_oracle_and_fd(ctx, oracle):
    g, _, x, _ = setup(ctx, seed=10)
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    val = ag.smoothed_tv(xt, ctx, g.grid(), 0.05)
    val.backward()
    ov, og = oracle.tv_value_grad(x.astype(np.float64), oracle_volume(oracle, g), 0.05)
    assert abs(float(val.detach()) - ov) < 1e-5 * ov
    assert rel_rmse(xt.grad.cpu().numpy(), og) < 1e-5


def test_tv_regularised_solve_decreases_objective(ctx):
    g = small_geometry(total_views=36, views_per_turn=36)
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, g.grid())
    ell = phantoms.random_ellipsoids(seed=4, count=5, semi_axes=(2.0, 7.0), density=(0.01, 0.05),
                                     center_radius=4.0)
    truth = ctx.phantom(ell, g.grid())
    b = P.forward(truth)
    x, hist = ag.least_squares_tv(P, b, ctx, steps=20, lam=1e-3, eps=1e-6)
    assert all(h1 <= h0 * (1 + 1e-9) for h0, h1 in zip(hist, hist[1:]))
    assert hist[-1] < 0.05 * hist[0]
    assert rel_rmse(x.cpu().numpy(), truth.cpu().numpy()) < 0.5

"""Reference configurations as GPU parity cases, against the FP64 oracle.

* Plain FDK, FP32: modified 3-D Shepp-Logan (x 0.02 / mm) on 128^3 at
  0.5 mm, SOD 500 / SDD 1000, 180 views over 360 degrees, 256^2 at 0.8 mm,
  Joseph projections plus Gaussian noise (sigma = 0.5 % of the largest line
  integral, seed 1), cosine weight and Ram-Lak ramp.
* Precision form of the same geometry: filtered views stored as FP16 and as
  BF16 with per-row scales, FP32 accumulation with Kahan compensation, at
  180 and 720 views; relative RMSE and max-abs against the FP64 oracle.

The oracle filters and backprojects the very projections the GPU receives,
so both sides start from the same data.
"""
import numpy as np
import pytest
import torch

import fdk_cuda_impl_amd as yk
from fdk_cuda_impl_amd import phantoms
from helpers import max_abs, rel_rmse

pytestmark = pytest.mark.gpu


def plain_fdk_geometry(views):
    return yk.ScanGeometry(views, views, 500.0, 1000.0, 256, 256, 0.8, 0.8, volume=(128, 128, 128),
                           voxel_mm=(0.5, 0.5, 0.5))


def noisy_projections(ctx, geom, seed=1):
    truth = ctx.phantom(phantoms.shepp_logan(32.0), geom.grid())
    P = yk.Projector(ctx, geom.views(), geom.nu, geom.nv, geom.grid())
    p = P.forward(truth, kind="joseph").cpu().numpy()
    P.close()
    rng = np.random.default_rng(seed)
    p = p + rng.normal(0.0, 0.005 * float(p.max()), p.shape)
    return p.astype(np.float32)


def oracle_fdk(oracle, views32, proj, geom, filt="ramlak"):
    ov = oracle.views_from_float(views32)
    vol = oracle.volume(*geom.volume, *geom.voxel_mm)
    return oracle.fdk(ov, proj.astype(np.float64), vol, filt=filt)


@pytest.fixture(scope="module")
def plain180(ctx, oracle):
    g = plain_fdk_geometry(180)
    proj = noisy_projections(ctx, g)
    return g, proj, oracle_fdk(oracle, g.views(), proj, g)


def test_plain_fdk_fp32(ctx, plain180):
    g, proj, ref = plain180
    plan = yk.FdkPlan.from_geometry(ctx, g, filter="ramlak")
    got = plan.reconstruct(torch.from_numpy(proj).cuda()).cpu().numpy()
    plan.close()
    print(f"plain FDK 128^3 x 180 FP32: rel-RMSE {rel_rmse(got, ref):.3e}, "
          f"max-abs {max_abs(got, ref):.3e} (peak {np.abs(ref).max():.3e})")
    assert rel_rmse(got, ref) < 1e-5
    assert max_abs(got, ref) < 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_precision_form_180(ctx, plain180, storage):
    g, proj, ref = plain180
    plan = yk.FdkPlan.from_geometry(ctx, g, filter="ramlak", storage=storage, kahan=True)
    got = plan.reconstruct(torch.from_numpy(proj).cuda()).cpu().numpy()
    plan.close()
    err = rel_rmse(got, ref)
    print(f"precision form 128^3 x 180 {storage}+Kahan: rel-RMSE {err:.3e}, max-abs {max_abs(got, ref):.3e}")
    assert err < 1e-3
    assert err < (3e-4 if storage == "f16" else 8e-4)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_precision_form_720(ctx, oracle, storage):
    g = plain_fdk_geometry(720)
    proj = noisy_projections(ctx, g, seed=3)
    ref = oracle_fdk(oracle, g.views(), proj, g)
    plan = yk.FdkPlan.from_geometry(ctx, g, filter="ramlak", storage=storage, kahan=True)
    got = plan.reconstruct(torch.from_numpy(proj).cuda()).cpu().numpy()
    plan32 = yk.FdkPlan.from_geometry(ctx, g, filter="ramlak")
    got32 = plan32.reconstruct(torch.from_numpy(proj).cuda()).cpu().numpy()
    plan.close()
    plan32.close()
    err, err32 = rel_rmse(got, ref), rel_rmse(got32, ref)
    print(f"precision form 128^3 x 720 {storage}+Kahan: rel-RMSE {err:.3e} (FP32 {err32:.3e})")
    assert err < 1e-3
    assert err32 < 1e-5

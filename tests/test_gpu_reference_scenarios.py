"""The reference's own hot-path scenarios, restated on our C ABI:

* fdk/calibrated-geometry (Yktest/test_pipeline_smoke.cpp:2183-2248): a
  1.2 pi short scan with source / detector offsets, detector tilts and a
  per-view sinusoidal source wobble; only the per-view vectors reach the plan.
  Here checked against the FP64 oracle (the reference only checks finiteness).
* TIGRE-style pseudo-matched adjoint (unit_test/test_operator_benchmarks.cpp:
  992-1086): smooth phantom, Siddon FP, the exact Siddon transpose and the
  pseudo-matched FDK backprojection compared by inner-product error, fitted
  relative L2 and cosine.
* large/water-fdk (Yktest/test_large_water.cpp:47-97, 142-182, 476-560):
  512 x 512 x 400 water phantom, Joseph FP, FDK; central-slab correlation
  and NRMSE.
* fdk/synchronous-batch (Yktest/test_pipeline_smoke.cpp:2146-2178): after a
  synchronous return the host input may be overwritten.
"""
import math

import numpy as np
import pytest
import torch

import fdk_cuda_impl_amd as yk
from fdk_cuda_impl_amd import phantoms
from helpers import oracle_volume, rel_rmse, seeded_projection, small_geometry

pytestmark = pytest.mark.gpu


def calibrated_geometry():
    g = small_geometry(total_views=60, views_per_turn=100, start_angle_rad=-0.1,
                       source_offset_mm=(0.35, -0.20, 0.15), offset_unv_mm=(0.4, 0.0, -0.25),
                       tilt_rad=(0.004, -0.003, 0.002))
    views = g.views().copy()
    phase = 2.0 * np.pi * np.arange(views.shape[0]) / views.shape[0]
    views[:, 0] += (0.18 * np.sin(phase)).astype(np.float32)
    views[:, 1] += (0.12 * np.cos(phase)).astype(np.float32)
    views[:, 2] += (0.08 * np.sin(2.0 * phase)).astype(np.float32)
    return g, views


@pytest.mark.parametrize("storage,tol", [("f32", 2e-5), ("f16", 1e-3)])
def test_calibrated_per_view_geometry(ctx, oracle, storage, tol):
    g, views = calibrated_geometry()
    scan_range = 2.0 * np.pi * 60 / 100
    grid = g.grid()
    truth = ctx.phantom(phantoms.shepp_logan(8.0), grid)
    P = yk.Projector(ctx, views, g.nu, g.nv, grid)
    proj = P.forward(truth)
    P.close()
    opts = yk.FdkOptions(filter="ramlak", parker="auto", scan_range_rad=scan_range, start_angle_rad=-0.1,
                         storage=storage, kahan=storage != "f32")
    plan = yk.FdkPlan(ctx, views, g.nu, g.nv, grid, opts)
    got = plan.reconstruct(proj).cpu().numpy()
    plan.close()
    ov = oracle.views_from_float(views)
    ref = oracle.fdk(ov, proj.cpu().numpy().astype(np.float64), oracle_volume(oracle, g), filt="ramlak",
                     scan_range=scan_range, parker_start=-0.1, direction=1)
    err = rel_rmse(got, ref)
    print(f"calibrated short-scan geometry, {storage}: rel-RMSE vs FP64 {err:.3e}")
    assert np.all(np.isfinite(got)) and np.abs(got).max() > 0
    assert err < tol


def smooth_quality_phantom(nx, ny, nz, vox):
    """Smooth non-negative body with two inserts (the formula of
    smoothQualityPhantom, unit_test/test_operator_benchmarks.cpp:475-520)."""
    def inside(r, t):
        s = np.clip((1.0 + t - r) / (2.0 * t), 0.0, 1.0)
        return np.where(r <= 1.0 - t, 1.0, np.where(r >= 1.0 + t, 0.0, s * s * (3.0 - 2.0 * s)))
    pz = ((np.arange(nz) + 0.5 - 0.5 * nz) * vox)[:, None, None]
    py = ((np.arange(ny) + 0.5 - 0.5 * ny) * vox)[None, :, None]
    px = ((np.arange(nx) + 0.5 - 0.5 * nx) * vox)[None, None, :]
    body = np.sqrt((px / 40.0) ** 2 + (py / 34.0) ** 2 + (pz / 28.0) ** 2)
    ins = np.sqrt(((px - 15.0) / 7.0) ** 2 + ((py + 8.0) / 7.0) ** 2 + ((pz - 4.0) / 10.0) ** 2)
    low = np.sqrt(((px + 14.0) / 10.0) ** 2 + ((py - 7.0) / 8.0) ** 2 + ((pz + 5.0) / 12.0) ** 2)
    v = 0.018 * inside(body, 0.025) + 0.012 * inside(ins, 0.10) + 0.002 * inside(low, 0.08)
    return v.astype(np.float32)


def adjoint_similarity(forward_ip, cand, ref, x):
    """compareAdjointResult (unit_test/test_operator_benchmarks.cpp:107-137)."""
    cand, ref, x = (a.astype(np.float64).ravel() for a in (cand, ref, x))
    cand_ip = float(np.dot(x, cand))
    inner = abs(forward_ip - cand_ip) / max(abs(forward_ip), abs(cand_ip), 1e-30)
    cc, rr, cr = float(np.dot(cand, cand)), float(np.dot(ref, ref)), float(np.dot(cand, ref))
    scale = cr / max(cc, 1e-30)
    fitted = math.sqrt(float(np.sum((scale * cand - ref) ** 2)) / max(rr, 1e-30))
    cosine = cr / math.sqrt(max(cc * rr, 1e-30))
    return inner, fitted, cosine


def test_pseudo_matched_adjoint_against_exact_transpose(ctx):
    g = yk.ScanGeometry(90, 90, 1000.0, 1536.0, 256, 192, 0.8, 0.8, volume=(128, 128, 96),
                        voxel_mm=(1.0, 1.0, 1.0))
    x = smooth_quality_phantom(128, 128, 96, 1.0)
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, g.grid())
    y = P.forward(torch.from_numpy(x).cuda(), kind="siddon")
    forward_ip = ctx.dot(y, y)
    exact = P.adjoint(y, kind="siddon").cpu().numpy()
    pseudo = P.adjoint(y, kind="fdk_pseudo_matched").cpu().numpy()
    P.close()
    e_inner, e_fit, e_cos = adjoint_similarity(forward_ip, exact, exact, x)
    p_inner, p_fit, p_cos = adjoint_similarity(forward_ip, pseudo, exact, x)
    print(f"exact Siddon: inner {e_inner:.3e}; pseudo-matched FDK: inner {p_inner:.3e}, "
          f"fitted L2 {p_fit:.4f}, cosine {p_cos:.6f}")
    assert e_inner < 1e-4 and e_fit < 1e-6 and e_cos > 1.0 - 1e-6
    assert p_inner < 1e-2 and p_fit < 0.10 and p_cos > 0.99


def water_phantom(nx, ny, nz, vox):
    """Closed water cylinder in a shell (makeWaterPhantom,
    Yktest/test_large_water.cpp:68-97)."""
    pz = ((np.arange(nz) + 0.5 - 0.5 * nz) * vox)[:, None, None]
    py = ((np.arange(ny) + 0.5 - 0.5 * ny) * vox)[None, :, None]
    px = ((np.arange(nx) + 0.5 - 0.5 * nx) * vox)[None, None, :]
    r2 = px * px + py * py
    shell = (r2 <= 75.0 ** 2) & (np.abs(pz) <= 55.0)
    water = (r2 <= 65.0 ** 2) & (np.abs(pz) <= 45.0)
    return np.where(water, 0.020, np.where(shell, 0.040, 0.0)).astype(np.float32)


def test_large_water_fdk(ctx):
    g = yk.ScanGeometry(720, 720, 440.0, 770.0, 1024, 128, 0.417, 0.417, volume=(512, 512, 400),
                        voxel_mm=(0.3, 0.3, 0.3))
    truth = water_phantom(512, 512, 400, 0.3)
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, g.grid())
    proj = P.forward(torch.from_numpy(truth).cuda(), kind="joseph")
    P.close()
    rec = yk.FdkPlan.from_geometry(ctx, g, filter="ramlak").reconstruct(proj).cpu().numpy()
    assert np.all(np.isfinite(rec))
    pz = (np.arange(400) + 0.5 - 200.0) * 0.3
    slab = np.abs(pz) <= 12.0
    t = truth[slab].astype(np.float64).ravel()
    r = rec[slab].astype(np.float64).ravel()
    dot, tt, rr = float(t @ r), float(t @ t), float(r @ r)
    scale = dot / rr
    corr = dot / math.sqrt(tt * rr)
    nrmse = math.sqrt(float(np.sum((scale * r - t) ** 2)) / tt)
    mid = rec[200]
    yy, xx = np.meshgrid((np.arange(512) + 0.5 - 256.0) * 0.3, (np.arange(512) + 0.5 - 256.0) * 0.3,
                         indexing="ij")
    rad = np.sqrt(xx * xx + yy * yy)
    print(f"water FDK central slab: correlation {corr:.4f}, NRMSE {nrmse:.4f}; water mean "
          f"{scale * mid[rad < 55.0].mean():.5f} (0.020), shell mean "
          f"{scale * mid[(rad > 68.0) & (rad < 72.0)].mean():.5f} (0.040)")
    assert corr > 0.75 and nrmse < 0.70


def test_synchronous_batch_input_reusable(ctx):
    g = small_geometry(total_views=30, views_per_turn=30)
    plan = yk.FdkPlan.from_geometry(ctx, g)
    proj = seeded_projection(21, g.total_views, g.nv, g.nu)
    first = np.zeros(plan.volume_shape, dtype=np.float32)
    buf = proj.copy()
    plan.execute(buf, first, view_offset=0, clear_output=True)
    kept = first.copy()
    buf[:] = np.nan  # the call has returned: the input is free to reuse
    assert np.array_equal(first, kept)
    plan.reset()
    again = np.zeros_like(first)
    plan.execute(proj.copy(), again, view_offset=0, clear_output=True)
    assert np.array_equal(again, first)

"""GPU parity of the Joseph / Siddon forward projectors, their transposes, and
the iterative callers (CGLS, SART, TV) against the FP64 oracle.  All calls go
through the C ABI."""
import numpy as np
import pytest
import torch

import fdk_cuda_impl_amd as yk
from fdk_cuda_impl_amd import phantoms
from helpers import max_abs, oracle_volume, rel_rmse, small_geometry

pytestmark = pytest.mark.gpu

GEOMS = {
    "plain": dict(),
    "offset": dict(offset_unv_mm=(3.0, 0.0, -2.0), center_mm=(1.5, -1.0, 0.5), start_angle_rad=0.2),
    "aniso": dict(volume=(30, 22, 17), voxel_mm=(0.9, 1.1, 1.3), tilt_rad=(0.02, 0.01, -0.03)),
    # several 32-slice blocks of the separable (untilted) Joseph transpose
    "tall": dict(volume=(20, 16, 70), voxel_mm=(1.0, 1.0, 0.5), rows=80),
}


def random_volume(g, seed):
    rng = np.random.default_rng(seed)
    return rng.random((g.volume[2], g.volume[1], g.volume[0])).astype(np.float32)


@pytest.mark.parametrize("name", list(GEOMS))
@pytest.mark.parametrize("kind", ["joseph", "siddon"])
def test_forward_matches_oracle(ctx, oracle, name, kind):
    g = small_geometry(**GEOMS[name])
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, g.grid())
    x = random_volume(g, 11)
    got = P.forward(torch.from_numpy(x).cuda(), kind=kind).cpu().numpy()
    ov = oracle.views_from_float(P.views)
    fn = oracle.fp_joseph if kind == "joseph" else oracle.fp_siddon
    ref = fn(ov, g.nu, g.nv, oracle_volume(oracle, g), x)
    assert rel_rmse(got, ref) < 1e-5, rel_rmse(got, ref)
    assert max_abs(got, ref) < 1e-4 * np.abs(ref).max()


@pytest.mark.parametrize("name", ["plain", "aniso"])
@pytest.mark.parametrize("kind", ["joseph", "siddon"])
def test_forward_view_subset_matches_full(ctx, name, kind):
    """A view range that starts off a workgroup's view group and ends in a
    partial one (views 3..15 of 24; the Joseph kernels take eight views per
    workgroup) gives bit-identical rays to the full projection."""
    g = small_geometry(**GEOMS[name])
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, g.grid())
    x = torch.from_numpy(random_volume(g, 12)).cuda()
    full = P.forward(x, kind=kind).cpu().numpy()
    part = P.forward(x, kind=kind, first=3, count=13).cpu().numpy()
    np.testing.assert_array_equal(part, full[3:16])


@pytest.mark.parametrize("name", list(GEOMS))
@pytest.mark.parametrize("kind", ["joseph", "siddon"])
def test_adjoint_dot_product(ctx, name, kind):
    """<A x, y> = <x, A^T y> (framework/operator-matrix,
    Yktest/test_operator_properties.cpp:93-141; tolerance 1e-5)."""
    g = small_geometry(**GEOMS[name])
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, g.grid())
    rng = np.random.default_rng(12)
    x = torch.from_numpy(rng.random(P.volume_shape, dtype=np.float32)).cuda()
    y = torch.from_numpy(rng.random((g.total_views, g.nv, g.nu), dtype=np.float32)).cuda()
    ax = P.forward(x, kind=kind)
    aty = P.adjoint(y, kind=kind)
    lhs = ctx.dot(ax, y)
    rhs = ctx.dot(x, aty)
    assert abs(lhs - rhs) / abs(lhs) < 1e-5, (lhs, rhs)


@pytest.mark.parametrize("name", ["offset", "aniso", "tall"])
@pytest.mark.parametrize("kind", ["joseph", "siddon"])
def test_adjoint_matches_oracle_transpose(ctx, oracle, kind, name):
    """Untilted scans take the separable Joseph transpose (column walk), the
    tilted one the per-voxel candidate search; both against the FP64 oracle."""
    g = small_geometry(**GEOMS[name])
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, g.grid())
    rng = np.random.default_rng(13)
    y = rng.standard_normal((g.total_views, g.nv, g.nu)).astype(np.float32)
    got = P.adjoint(torch.from_numpy(y).cuda(), kind=kind).cpu().numpy()
    ov = oracle.views_from_float(P.views)
    fn = oracle.bp_joseph_adjoint if kind == "joseph" else oracle.bp_siddon_adjoint
    ref = fn(ov, g.nu, g.nv, oracle_volume(oracle, g), y.astype(np.float64))
    # zero-mean input: the FP32 sums cancel, so the relative error is a few
    # FP32 ulps of the summands rather than of the result
    assert rel_rmse(got, ref) < 2e-5, rel_rmse(got, ref)


@pytest.mark.parametrize("kind", ["fdk", "fdk_pseudo_matched", "joseph_voxel"])
@pytest.mark.parametrize("name", ["plain", "aniso"])
def test_operator_fdk_backprojection_matches_oracle(ctx, oracle, kind, name):
    """Voxel-driven backprojections used as the iterative B (FDK weight,
    pseudo-matched, voxel-driven Joseph), source lifted out of the orbit plane
    so the FDK weight uses its own denominator."""
    g = small_geometry(source_offset_mm=(0.0, 0.0, 2.0), **GEOMS[name])
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, g.grid())
    rng = np.random.default_rng(14)
    y = rng.standard_normal((g.total_views, g.nv, g.nu)).astype(np.float32)
    got = P.adjoint(torch.from_numpy(y).cuda(), kind=kind).cpu().numpy()
    ov = oracle.views_from_float(P.views)
    ref = oracle.bp_fdk_operator(ov, y.astype(np.float64), oracle_volume(oracle, g),
                                 matched={"fdk": 0, "fdk_pseudo_matched": 1, "joseph_voxel": 2}[kind])
    assert rel_rmse(got, ref) < 1e-5, rel_rmse(got, ref)
    # accumulate adds onto the volume
    acc = P.adjoint(torch.from_numpy(y).cuda(), kind=kind, out=torch.ones(P.volume_shape, device="cuda"),
                    accumulate=True).cpu().numpy()
    assert rel_rmse(acc - 1.0, ref) < 1e-5


def test_siddon_uniform_centre_ray(ctx):
    """fp/siddon-uniform-center: uniform 64^3 volume of 1 mm voxels, the centre
    ray reads 64 mm (Yktest/test_forward_projectors.cpp:64-82)."""
    g = yk.ScanGeometry(1, 4, 500.0, 1000.0, 64, 64, 1.0, 1.0, volume=(64, 64, 64), voxel_mm=(1, 1, 1))
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, g.grid())
    ones = torch.ones(P.volume_shape, device="cuda")
    for kind in ("siddon", "joseph"):
        p = P.forward(ones, kind=kind).cpu().numpy()[0]
        # detector of 64 px, even size: the four central pixels straddle the axis
        centre = p[31:33, 31:33].mean()
        assert abs(centre - 64.0) / 64.0 < 0.01, (kind, centre)


def test_forward_matches_analytic_ellipsoids(ctx, oracle):
    """Discrete projections of a voxelized ellipsoid phantom approach the
    analytic chord integrals (mean relative deviation small)."""
    g = small_geometry(volume=(64, 64, 64), voxel_mm=(0.5, 0.5, 0.5), channels=64, rows=64,
                       channel_mm=0.8, row_mm=0.8, total_views=8, views_per_turn=8)
    ell = phantoms.random_ellipsoids(3, 6, semi_axes=(4.0, 10.0), density=(0.01, 0.05),
                                     center_radius=5.0)
    vol = ctx.phantom(ell, g.grid(), clamp=False)
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, g.grid())
    ov = oracle.views_from_float(P.views)
    ref = oracle.ellipsoid_line_integrals(ell, ov, g.nu, g.nv)
    for kind in ("joseph", "siddon"):
        got = P.forward(vol, kind=kind).cpu().numpy()
        assert rel_rmse(got, ref) < 0.03, (kind, rel_rmse(got, ref))


def test_cgls_reduces_residual(ctx):
    g = small_geometry(total_views=36, views_per_turn=36)
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, g.grid())
    truth = ctx.phantom(phantoms.shepp_logan(8.0), g.grid())
    for kind in ("siddon", "joseph"):
        b = P.forward(truth, kind=kind)
        x, res = P.cgls(b, iterations=12, forward=kind, back=kind)
        assert np.all(np.diff(res) <= 1e-6 * res[0]), res
        assert res[-1] < 0.2 * res[0]


def test_cgls_exact_on_consistent_system(ctx):
    """Tiny consistent system: CGLS drives the residual essentially to zero."""
    g = small_geometry(volume=(6, 6, 6), voxel_mm=(2, 2, 2), total_views=40, views_per_turn=40)
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, g.grid())
    rng = np.random.default_rng(14)
    xt = torch.from_numpy(rng.random(P.volume_shape, dtype=np.float32)).cuda()
    b = P.forward(xt, kind="siddon")
    x, res = P.cgls(b, iterations=60, forward="siddon", back="siddon")
    assert res[-1] < 1e-3 * res[0], res[-1] / res[0]
    assert rel_rmse(x.cpu().numpy(), xt.cpu().numpy()) < 2e-2


def test_sart_converges(ctx):
    g = small_geometry(total_views=24, views_per_turn=24)
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, g.grid())
    truth = ctx.phantom(phantoms.shepp_logan(8.0), g.grid())
    b = P.forward(truth, kind="joseph")
    x1 = P.sart(b, iterations=1)
    x3 = P.sart(b, iterations=3, nonnegative=True)
    r1 = float(torch.linalg.norm(P.forward(x1) - b))
    r3 = float(torch.linalg.norm(P.forward(x3) - b))
    assert r3 < r1 < float(torch.linalg.norm(b))
    assert float(x3.min()) >= 0.0


def test_tv_matches_oracle(ctx, oracle):
    g = small_geometry(volume=(20, 17, 13), voxel_mm=(0.5, 0.7, 0.9))
    rng = np.random.default_rng(15)
    x = rng.random((13, 17, 20)).astype(np.float32)
    val, grad = ctx.tv(torch.from_numpy(x).cuda(), g.grid(), 1e-3)
    ov, og = oracle.tv_value_grad(x.astype(np.float64), oracle_volume(oracle, g), float(np.float32(1e-3)))
    assert abs(val - ov) / ov < 1e-5
    assert rel_rmse(grad.cpu().numpy(), og) < 1e-5

"""GPU parity of the unified SIRT / SART / OS-SART engine (yk_algebraic)
against the oracle's FP64 restatement, on the reference's algebraic smoke
fixture: 16 x 16 x 12 volume of 1 mm voxels, 32 x 24 detector of 1 mm pixels,
8 views over 2 pi, SID 80 mm, SDD 160 mm (Yktest/test_reconstructors_smoke.cpp:15-37),
with its four method / weight-model cases (:145-182) and its OS-SART
smoothed-TV case (:231-257), plus wider schedules.  Same measured sinogram on
both sides; tolerance on the volume: rel-RMSE 1e-4 (FP32 against FP64 through
the whole iteration)."""
import numpy as np
import pytest
import torch

import fdk_cuda_impl_amd as yk
from fdk_cuda_impl_amd import phantoms
from helpers import oracle_volume, rel_rmse, small_geometry

pytestmark = pytest.mark.gpu

FIXTURE = dict(total_views=8, views_per_turn=8, sid_mm=80.0, sdd_mm=160.0, channels=32, rows=24,
               channel_mm=1.0, row_mm=1.0, volume=(16, 16, 12), voxel_mm=(1.0, 1.0, 1.0))
TOL = 1e-4


@pytest.fixture(scope="module")
def problem(ctx, oracle):
    g = small_geometry(**FIXTURE)
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, g.grid())
    truth = ctx.phantom(phantoms.shepp_logan(7.5, scale=1.0), g.grid())
    b = P.forward(truth, kind="joseph")
    ov = oracle.views_from_float(P.views)
    return g, P, b, ov, oracle_volume(oracle, g)


def run_both(oracle, problem, **kw):
    g, P, b, ov, vol = problem
    print(f"yk_algebraic {kw}", flush=True)
    got = P.algebraic(b, **kw).cpu().numpy()
    okw = {k: v for k, v in kw.items()}
    if "tv_epsilon" in okw:
        okw["tv_epsilon"] = float(np.float32(okw["tv_epsilon"]))
    ref = oracle.algebraic(ov, g.nu, g.nv, vol, b.cpu().numpy(), **okw)
    return got, ref


@pytest.mark.parametrize("back", ["joseph_voxel", "joseph"])
@pytest.mark.parametrize("method,weights", [
    ("sirt", "detailed"), ("sart", "detailed"), ("ossart", "detailed"), ("ossart", "tigre")])
def test_reference_smoke_cases(oracle, problem, method, weights, back):
    """FP_Joseph with BP_Joseph_v3 (the fixture's pair), and with the exact
    Joseph transpose."""
    got, ref = run_both(oracle, problem, method=method, weights=weights, iterations=1, subsets=2,
                        relaxation=0.2, min_value=0.0, back=back)
    assert np.isfinite(got).all() and np.abs(got).max() > 1e-8
    assert rel_rmse(got, ref) < TOL, rel_rmse(got, ref)


def test_reference_smoothed_tv_case(oracle, problem):
    got, ref = run_both(oracle, problem, method="ossart", weights="detailed", iterations=2, subsets=2,
                        relaxation=0.2, min_value=0.0, back="joseph_voxel", tv_strength=1e-3, tv_iterations=2,
                        tv_epsilon=1e-4, tv_strength_reduction=0.95)
    assert rel_rmse(got, ref) < TOL, rel_rmse(got, ref)


@pytest.mark.parametrize("kw", [
    dict(method="sart"),  # TIGRE weights, one view per block
    dict(method="sirt", weights="tigre", relaxation=1.5),
    dict(method="ossart", subsets=4, order="golden", relaxation_reduction=0.8),
    dict(method="ossart", subsets=3, weights="tigre", relaxation_reduction=0.8, max_value=0.5),
    dict(method="ossart", subsets=4, forward="siddon", back="siddon", min_value=0.0),
    # the reference's defaults: FP_Joseph with the voxel-driven BP_Joseph_v3
    dict(method="sirt", back="joseph_voxel", min_value=0.0),
    dict(method="ossart", subsets=2, back="joseph_voxel", relaxation=0.2),
    # the Session's Automatic pair for algebraic pipelines: Siddon + FDK pseudo-matched
    dict(method="sart", forward="siddon", back="fdk_pseudo_matched", relaxation=0.5),
    dict(method="ossart", subsets=4, forward="siddon", back="fdk_pseudo_matched", min_value=0.0),
], ids=["sart-tigre", "sirt-tigre", "ossart-golden", "ossart3-tigre", "ossart-siddon", "sirt-josephv3",
        "ossart-josephv3", "sart-automatic", "ossart-automatic"])
def test_schedules(oracle, problem, kw):
    got, ref = run_both(oracle, problem, iterations=3, **kw)
    assert rel_rmse(got, ref) < TOL, rel_rmse(got, ref)


def test_sart_entry_point_is_detailed_per_view(oracle, problem):
    g, P, b, ov, vol = problem
    got = P.sart(b, iterations=2, relaxation=0.5, nonnegative=True).cpu().numpy()
    ref = oracle.algebraic(ov, g.nu, g.nv, vol, b.cpu().numpy(), method="sart", weights="detailed",
                           iterations=2, relaxation=0.5, min_value=0.0, back="joseph")
    assert rel_rmse(got, ref) < TOL, rel_rmse(got, ref)


def test_warm_start_and_convergence(problem):
    g, P, b, ov, vol = problem
    x = P.algebraic(b, method="ossart", subsets=4, iterations=2, relaxation=0.8)
    r2 = float(torch.linalg.norm(P.forward(x) - b))
    P.algebraic(b, method="ossart", subsets=4, iterations=6, relaxation=0.8, x=x)
    r8 = float(torch.linalg.norm(P.forward(x) - b))
    assert r8 < r2 < float(torch.linalg.norm(b))


def test_invalid_descriptors_are_refused(problem):
    g, P, b, ov, vol = problem
    with pytest.raises(yk.YkError):
        P.algebraic(b, method="ossart", subsets=0)
    with pytest.raises(yk.YkError):
        P.algebraic(b, method="ossart", subsets=g.total_views + 1)
    with pytest.raises(yk.YkError):
        P.algebraic(b, iterations=0)
    with pytest.raises(yk.YkError):
        P.algebraic(b, epsilon=0.0)
    with pytest.raises(yk.YkError):
        P.algebraic(b, min_value=1.0, max_value=0.5)


@pytest.mark.parametrize("kw", [
    dict(method="ossart", subsets=3, forward="joseph", back="joseph"),
    dict(method="sart", forward="siddon", back="siddon"),
    dict(method="sirt", back="joseph_voxel"),
], ids=["ossart-joseph", "sart-tigre-siddon", "sirt-josephv3"])
def test_tilted_detector_takes_the_general_operators(ctx, oracle, kw):
    """Detector tilts and an anisotropic grid: no separable view, so every
    subset runs the general projector and transpose kernels."""
    g = small_geometry(total_views=12, views_per_turn=12, volume=(20, 18, 14), voxel_mm=(0.9, 1.1, 1.3),
                       tilt_rad=(0.02, 0.01, -0.03), offset_unv_mm=(1.5, 0.0, -1.0))
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, g.grid())
    truth = ctx.phantom(phantoms.shepp_logan(8.0, scale=1.0), g.grid())
    b = P.forward(truth, kind="joseph")
    got = P.algebraic(b, iterations=2, **kw).cpu().numpy()
    ref = oracle.algebraic(oracle.views_from_float(P.views), g.nu, g.nv, oracle_volume(oracle, g),
                           b.cpu().numpy(), iterations=2, **kw)
    assert rel_rmse(got, ref) < TOL, rel_rmse(got, ref)



def test_reference_convergence_known_answer(problem):
    """Yktest/test_reconstructors_smoke.cpp:184-205: relative residual
    tolerance 2 (> 1) stops OS-SART at the first check, after one iteration."""
    g, P, b, ov, vol = problem
    kw = dict(method="ossart", subsets=2, relaxation=0.2, min_value=0.0)
    x, st = P.algebraic(b, iterations=4, relative_residual=2.0, return_stats=True, **kw)
    assert st["completed_iterations"] == 1 and st["convergence_checks"] == 1
    assert st["stopped_by_relative_residual"] == 1 and np.isfinite(st["relative_projection_residual"])
    one = P.algebraic(b, iterations=1, **kw)
    np.testing.assert_array_equal(x.cpu().numpy(), one.cpu().numpy())


def test_update_tolerance_matches_oracle(oracle, problem):
    g, P, b, ov, vol = problem
    kw = dict(method="sirt", relaxation=1.0, iterations=20, relative_update=0.05, minimum_iterations=2,
              check_interval=2, patience=1, back="joseph")
    x, st = P.algebraic(b, return_stats=True, **kw)
    ox, ost = oracle.algebraic(ov, g.nu, g.nv, vol, b.cpu().numpy(), return_stats=True, **kw)
    assert ost["stopped_by_relative_update"] == 1 and ost["completed_iterations"] < 20
    assert st["completed_iterations"] == ost["completed_iterations"]
    assert st["convergence_checks"] == ost["convergence_checks"]
    assert st["relative_volume_update"] == pytest.approx(ost["relative_volume_update"], rel=1e-3)
    assert rel_rmse(x.cpu().numpy(), ox) < TOL


def test_slice_wise_tv(oracle, problem):
    got, ref = run_both(oracle, problem, method="ossart", subsets=2, iterations=2, relaxation=0.2,
                        min_value=0.0, tv_strength=2e-3, tv_iterations=3, tv_epsilon=1e-3, tv_dimensions=2)
    assert rel_rmse(got, ref) < TOL, rel_rmse(got, ref)
    vol3, _ = run_both(oracle, problem, method="ossart", subsets=2, iterations=2, relaxation=0.2,
                       min_value=0.0, tv_strength=2e-3, tv_iterations=3, tv_epsilon=1e-3)
    assert rel_rmse(got, vol3) > 1e-4  # the z differences matter

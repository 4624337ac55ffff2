"""GPU parity of yk_cgls_ex (robust-restart and ASTRA CGLS) against the oracle's
FP64 restatement on the reference's iterative smoke fixture (16 x 16 x 12
volume, 32 x 24 detector, 8 views, SID 80 / SDD 160;
Yktest/test_reconstructors_smoke.cpp:15-37) and its convergence known answer
(:206-226).  Tolerance on the volume: rel-RMSE 1e-4."""
import numpy as np
import pytest

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
    b = P.forward(truth, kind="siddon")
    return g, P, b, oracle.views_from_float(P.views), oracle_volume(oracle, g)


def run_both(oracle, problem, **kw):
    g, P, b, ov, vol = problem
    print(f"yk_cgls_ex {kw}", flush=True)
    x, res, st = P.cgls_ex(b, **kw)
    ox, ores, ost = oracle.cgls(ov, g.nu, g.nv, vol, b.cpu().numpy(), **kw)
    return x.cpu().numpy(), res, st, ox, ores, ost


@pytest.mark.parametrize("kw", [
    dict(strategy="robust", forward="siddon", back="siddon"),       # the Session's CGLS pair
    dict(strategy="astra", forward="siddon", back="siddon"),
    dict(strategy="robust", forward="joseph", back="joseph", min_value=0.0),
    dict(strategy="astra", forward="joseph", back="joseph_voxel", min_value=0.0, max_value=1.0),
], ids=["robust-siddon", "astra-siddon", "robust-joseph-bounds", "astra-josephv3-bounds"])
def test_matches_oracle(oracle, problem, kw):
    x, res, st, ox, ores, ost = run_both(oracle, problem, iterations=8, **kw)
    assert st["completed_iterations"] == ost["completed_iterations"]
    assert st["stopped_by_divergence"] == ost["stopped_by_divergence"]
    assert rel_rmse(x, ox) < TOL, rel_rmse(x, ox)
    np.testing.assert_allclose(res, ores, rtol=1e-4)


def test_rollback_on_a_mismatched_back_model(oracle, problem):
    """FDK pseudo-matched B is not A^T: the robust strategy rolls back."""
    x, res, st, ox, ores, ost = run_both(oracle, problem, iterations=30, strategy="robust", forward="siddon",
                                         back="fdk_pseudo_matched", restart=False)
    assert ost["stopped_by_divergence"] == 1
    assert st["stopped_by_divergence"] == 1
    assert st["completed_iterations"] == ost["completed_iterations"]
    assert rel_rmse(x, ox) < TOL, rel_rmse(x, ox)


def test_reference_convergence_known_answer(problem):
    """relative_residual_tolerance 2 > 1 holds at the first check: one
    iteration, one check, stopped by the relative residual."""
    g, P, b, ov, vol = problem
    x, res, st = P.cgls_ex(b, iterations=4, min_value=0.0, relative_residual=2.0, minimum_iterations=1,
                           check_interval=1, patience=1)
    assert st["completed_iterations"] == 1 and st["convergence_checks"] == 1
    assert st["stopped_by_relative_residual"] == 1
    assert np.isfinite(st["relative_projection_residual"])


def test_breakdown_on_zero_data(problem):
    g, P, b, ov, vol = problem
    x, res, st = P.cgls_ex(b * 0, iterations=3)
    assert st["stopped_by_breakdown"] == 1 and st["completed_iterations"] == 0
    assert float(x.abs().max()) == 0.0


def test_invalid_descriptors_are_refused(problem):
    g, P, b, ov, vol = problem
    for kw in (dict(iterations=0), dict(epsilon=0.0), dict(strategy="astra", relative_residual=0.5),
               dict(check_interval=0), dict(patience=0), dict(min_value=1.0, max_value=0.0)):
        with pytest.raises(yk.YkError):
            P.cgls_ex(b, **kw)


def test_tilted_detector_robust_cgls(ctx, oracle):
    g = small_geometry(total_views=12, views_per_turn=12, volume=(20, 18, 14), voxel_mm=(0.9, 1.1, 1.3),
                       tilt_rad=(0.02, 0.01, -0.03))
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, g.grid())
    truth = ctx.phantom(phantoms.shepp_logan(8.0, scale=1.0), g.grid())
    b = P.forward(truth, kind="siddon")
    x, res, st = P.cgls_ex(b, iterations=6, forward="siddon", back="siddon")
    ox, ores, ost = oracle.cgls(oracle.views_from_float(P.views), g.nu, g.nv, oracle_volume(oracle, g),
                                b.cpu().numpy(), iterations=6, forward="siddon", back="siddon")
    assert st["completed_iterations"] == ost["completed_iterations"] == 6
    assert rel_rmse(x.cpu().numpy(), ox) < TOL, rel_rmse(x.cpu().numpy(), ox)


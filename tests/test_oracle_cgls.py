"""Pins of the oracle's CGLS restatement (CPU only).

Checked against what the recursion does not compute itself:
  * the least-squares solution from numpy's ``lstsq`` on the explicit system
    matrix: with the exact transpose both strategies reach it, and CG theory
    bounds the steps by the number of unknowns;
  * the monotone residual of CGLS with an exact transpose (no rollback);
  * the rollback: a run stopped by divergence at step k returns the iterate of
    a run of the steps before k;
  * the reference's convergence known answer (Yktest/test_reconstructors_smoke.cpp:206-226):
    tolerance 2 stops after the first check;
  * breakdown on a zero sinogram, and the bounds.
"""
import numpy as np
import pytest

NA, NU, NV = 6, 14, 10
SHAPE = (3, 4, 5)


@pytest.fixture(scope="module")
def tiny(oracle):
    views = oracle.build_views(NA, NA, 40.0, 80.0, NU, NV, 1.0, 1.0, start_angle=0.3)
    vol = oracle.volume(SHAPE[2], SHAPE[1], SHAPE[0], 1.0, 1.0, 1.0)
    n = int(np.prod(SHAPE))
    A = np.zeros((NA * NV * NU, n))
    for j in range(n):
        e = np.zeros(n, np.float32)
        e[j] = 1.0
        A[:, j] = oracle.fp_siddon(views, NU, NV, vol, e.reshape(SHAPE)).ravel()
    return views, vol, A


@pytest.mark.parametrize("strategy", ["robust", "astra"])
def test_reaches_least_squares_solution(oracle, tiny, strategy):
    views, vol, A = tiny
    rng = np.random.default_rng(3)
    b = A @ rng.random(A.shape[1]) + 0.05 * rng.standard_normal(A.shape[0])
    xs = np.linalg.lstsq(A, b, rcond=None)[0]
    x, res, st = oracle.cgls(views, NU, NV, vol, b, strategy=strategy, iterations=120, epsilon=1e-30,
                             forward="siddon", back="siddon")
    assert np.linalg.norm(x.ravel() - xs) / np.linalg.norm(xs) < 1e-6
    assert res[-1] == pytest.approx(np.linalg.norm(b - A @ xs), rel=1e-6)
    # the exact transpose: the residual never rises, so no rollback happened
    head = res[: st["completed_iterations"] + 1]
    assert np.all(np.diff(head) <= 1e-9 * head[0])


def test_rollback_returns_the_previous_iterate(oracle, tiny):
    views, vol, A = tiny
    rng = np.random.default_rng(5)
    b = A @ rng.random(A.shape[1])
    # a back model that is not the transpose makes the recursion diverge
    kw = dict(forward="siddon", back="fdk_pseudo_matched", restart=False)
    x, res, st = oracle.cgls(views, NU, NV, vol, b, iterations=40, **kw)
    assert st["stopped_by_divergence"] == 1
    k = st["completed_iterations"]
    assert 1 <= k < 40
    x_k, res_k, st_k = oracle.cgls(views, NU, NV, vol, b, iterations=k, **kw)
    assert st_k["stopped_by_divergence"] == 0
    np.testing.assert_array_equal(x, x_k)
    np.testing.assert_array_equal(res[: k + 1], res_k)
    assert np.all(res[k + 1:] == res[k])
    # with restarts it goes on past the first divergence
    _, _, st_r = oracle.cgls(views, NU, NV, vol, b, iterations=40, forward="siddon",
                             back="fdk_pseudo_matched", restart=True)
    assert st_r["completed_iterations"] > k or st_r["stopped_by_divergence"] == 1


def test_reference_convergence_known_answer(oracle, tiny):
    views, vol, A = tiny
    b = A @ np.random.default_rng(6).random(A.shape[1])
    x, res, st = oracle.cgls(views, NU, NV, vol, b, iterations=4, min_value=0.0, relative_residual=2.0,
                             minimum_iterations=1, check_interval=1, patience=1, back="joseph_voxel")
    assert st["completed_iterations"] == 1 and st["convergence_checks"] == 1
    assert st["stopped_by_relative_residual"] == 1 and np.isfinite(st["relative_projection_residual"])
    assert st["relative_projection_residual"] == pytest.approx(res[1] / np.linalg.norm(b))


def test_breakdown_and_bounds(oracle, tiny):
    views, vol, A = tiny
    x, res, st = oracle.cgls(views, NU, NV, vol, np.zeros(A.shape[0]), iterations=5)
    assert st["stopped_by_breakdown"] == 1 and st["completed_iterations"] == 0
    assert not x.any() and not res.any()
    b = A @ np.random.default_rng(7).standard_normal(A.shape[1])
    x, _, _ = oracle.cgls(views, NU, NV, vol, b, strategy="astra", iterations=6, min_value=0.0, max_value=0.4,
                          forward="siddon", back="siddon")
    assert x.min() >= 0.0 and x.max() <= 0.4

"""Pins of the oracle's algebraic reconstructor (CPU only).

``oracle.algebraic`` restates the reference's SIRT / SART / OS-SART loop over
the oracle's projector pair.  These tests check it against what the loop does
not compute itself:
  * hand-worked golden-ratio subset orders;
  * the weighted least-squares fixed point of SIRT, solved with numpy's
    ``lstsq`` on the explicit system matrix (inconsistent data, so a wrong row
    weight moves the fixed point by ~10 %);
  * the closed form of a constant volume: with exact per-subset weights one
    update maps x = a 1 to a + lambda_k (c - a), so the relaxation schedule
    fixes every iterate;
  * TIGRE row weights against ray/box chord lengths from a numpy slab
    intersection (Siddon integrates lengths exactly);
  * a second restatement on explicit matrices with explicit index sets,
    which pins the subset membership, visiting order, bounds and TV
    placement on a random volume.
The C <= 0 -> 1e30 column mask of the SIRT and TIGRE paths changes nothing
that can be observed: B has no negative entries, so B r vanishes wherever
B 1 does and both forms of the update add zero there.
"""
import numpy as np
import pytest

NA, NU, NV = 6, 14, 10
SHAPE = (3, 4, 5)  # z, y, x


@pytest.fixture(scope="module")
def tiny(oracle):
    views = oracle.build_views(NA, NA, 40.0, 80.0, NU, NV, 1.0, 1.0, start_angle=0.3)
    vol = oracle.volume(SHAPE[2], SHAPE[1], SHAPE[0], 1.0, 1.0, 1.0)
    return views, vol


def system_matrix(oracle, views, vol, kind="joseph"):
    fwd = {"joseph": oracle.fp_joseph, "siddon": oracle.fp_siddon}[kind]
    n = int(np.prod(SHAPE))
    A = np.zeros((len(views) * NV * NU, n))
    for j in range(n):
        e = np.zeros(n, np.float32)
        e[j] = 1.0
        A[:, j] = fwd(views, NU, NV, vol, e.reshape(SHAPE)).ravel()
    return A


def test_golden_subset_order_hand_worked(oracle):
    # n = 10: round(10 / phi) = 6, gcd(6, 10) = 2 -> 7
    assert oracle.golden_subset_order(10) == [0, 7, 4, 1, 8, 5, 2, 9, 6, 3]
    # n = 20: round(12.36) = 12, gcd 4 -> 13
    assert oracle.golden_subset_order(20)[:6] == [0, 13, 6, 19, 12, 5]
    # n = 4: round(2.47) = 2, gcd 2 -> 3
    assert oracle.golden_subset_order(4) == [0, 3, 2, 1]
    # n = 3: round(1.85) = 2; n = 11: round(6.80) = 7 (rounding, not truncation)
    assert oracle.golden_subset_order(3) == [0, 2, 1]
    assert oracle.golden_subset_order(11) == [0, 7, 3, 10, 6, 2, 9, 5, 1, 8, 4]
    assert oracle.golden_subset_order(2) == [0, 1]
    assert oracle.golden_subset_order(1) == [0]
    for n in range(1, 64):
        assert sorted(oracle.golden_subset_order(n)) == list(range(n))


def test_back_model_is_the_transpose(oracle, tiny):
    views, vol = tiny
    A = system_matrix(oracle, views, vol)
    p = np.random.default_rng(3).standard_normal((NA, NV, NU))
    bp = oracle.bp_joseph_adjoint(views, NU, NV, vol, p)
    np.testing.assert_allclose(bp.ravel(), A.T @ p.ravel(), rtol=1e-10, atol=1e-10)


def test_sirt_reaches_the_weighted_least_squares_solution(oracle, tiny):
    views, vol = tiny
    A = system_matrix(oracle, views, vol)
    assert np.linalg.matrix_rank(A) == A.shape[1]
    rng = np.random.default_rng(1)
    b = A @ rng.random(A.shape[1]) + 0.05 * rng.standard_normal(A.shape[0])
    R = A.sum(axis=1)
    w = np.where(R > 1e-6, 1.0 / np.where(R > 1e-6, R, 1.0), 0.0)
    xs = np.linalg.lstsq(np.sqrt(w)[:, None] * A, np.sqrt(w) * b, rcond=None)[0]
    x = oracle.algebraic(views, NU, NV, vol, b, method="sirt", iterations=3000, relaxation=1.9,
                         back="joseph")
    assert np.linalg.norm(x.ravel() - xs) / np.linalg.norm(xs) < 1e-5
    # the unweighted and the R-multiplied solutions are far from it
    for alt in (np.ones_like(R), R):
        xa = np.linalg.lstsq(np.sqrt(alt)[:, None] * A, np.sqrt(alt) * b, rcond=None)[0]
        assert np.linalg.norm(xa - xs) / np.linalg.norm(xs) > 0.05


@pytest.mark.parametrize("method,subsets,order", [
    ("sirt", 1, "sequential"), ("sart", 1, "sequential"), ("ossart", 3, "sequential"),
    ("ossart", 4, "golden")])
def test_constant_volume_closed_form(oracle, tiny, method, subsets, order):
    views, vol = tiny
    c, lam, red, iters = 0.7, 0.45, 0.6, 2
    b = oracle.fp_joseph(views, NU, NV, vol, np.full(SHAPE, c, np.float32))
    x = oracle.algebraic(views, NU, NV, vol, b, method=method, weights="detailed", order=order,
                         subsets=subsets, iterations=iters, relaxation=lam,
                         relaxation_reduction=red, back="joseph")
    n = {"sirt": 1, "sart": NA, "ossart": subsets}[method]
    # c - x_{k+1} = (1 - lambda q^k)(c - x_k), q = red^(1/n), x_0 = 0
    gap = c
    for k in range(iters * n):
        gap *= 1.0 - lam * red ** (k / n)
    np.testing.assert_allclose(x, c - gap, rtol=2e-6)


def test_constant_volume_upper_bound(oracle, tiny):
    views, vol = tiny
    b = oracle.fp_joseph(views, NU, NV, vol, np.full(SHAPE, 0.7, np.float32))
    x = oracle.algebraic(views, NU, NV, vol, b, method="ossart", subsets=2, iterations=2,
                         max_value=0.3, min_value=0.0, back="joseph")
    np.testing.assert_allclose(x, 0.3, rtol=0, atol=0)


def chord_lengths(views, lo_vol):
    """Length of src -> pixel segments inside the support box (slab method)."""
    centre = np.array([lo_vol.cx, lo_vol.cy, lo_vol.cz])
    half = np.array([lo_vol.nx * lo_vol.vx, lo_vol.ny * lo_vol.vy, lo_vol.nz * lo_vol.vz]) / 2
    out = np.zeros((len(views), NV, NU))
    for i, g in enumerate(views):
        src = g[0:3]
        for v in range(NV):
            for u in range(NU):
                d = g[3:6] + u * g[6:9] + v * g[9:12] - src
                with np.errstate(divide="ignore", invalid="ignore"):
                    t1 = (centre - half - src) / d
                    t2 = (centre + half - src) / d
                lo = np.where(d != 0, np.minimum(t1, t2), -np.inf)
                hi = np.where(d != 0, np.maximum(t1, t2), np.inf)
                a, z = max(0.0, lo.max()), min(1.0, hi.min())
                out[i, v, u] = max(0.0, z - a) * np.linalg.norm(d)
    return out


def test_tigre_row_weights_are_inverse_chords(oracle, tiny):
    views, vol = tiny
    lo = oracle.tigre_support(views, NV, vol)
    # support: 1.1 x the XY extent, max(detector rows, volume) in Z
    assert (lo.vx, lo.vy) == pytest.approx((5 * 1.1 / 2, 4 * 1.1 / 2))
    assert lo.vz == pytest.approx(max(NV * 1.0, 3.0) / 2)
    L = chord_lengths(views, lo)
    W = oracle.tigre_row_weights(views, NU, NV, vol, forward="siddon")
    expect = np.where(L <= 0.5, 1e-30, 1.0 / np.where(L > 0.5, L, 1.0))
    np.testing.assert_allclose(W, expect, rtol=1e-9, atol=1e-40)
    assert (L > 0.5).any() and (L <= 0.5).any()


def dense_reconstruction(oracle, A, b, vol, views, weights, method, subsets, order, iterations, lam,
                         red, lo=None, hi=None, tv=None):
    """The same algorithm on explicit matrices with explicit view index sets."""
    rows_per_view = NV * NU
    n = {"sirt": 1, "sart": NA, "ossart": subsets}[method]
    if weights == "tigre":
        size = -(-NA // n)
        sets = [list(range(s, min(NA, s + size))) for s in range(0, NA, size)]
        visit = range(len(sets))
        Wrow = 1.0 / np.where(chord_lengths(views, oracle.tigre_support(views, NV, vol)) <= 0.5, 1e30,
                              chord_lengths(views, oracle.tigre_support(views, NV, vol)))
        Wrow = Wrow.ravel()
    else:
        sets = [list(range(s, NA, n)) for s in range(n)]
        visit = oracle.golden_subset_order(n) if order == "golden" else range(n)
    x = np.zeros(A.shape[1])
    eps = 1e-6
    q = 1.0 if weights == "tigre" else red ** (1.0 / len(sets))
    strength = None if tv is None else tv[0]
    clamp = lambda v: np.clip(v, -np.inf if lo is None else lo, np.inf if hi is None else hi)  # noqa: E731
    for _ in range(iterations):
        for k in visit:
            idx = np.concatenate([np.arange(i * rows_per_view, (i + 1) * rows_per_view) for i in sets[k]])
            As = A[idx]
            r = b[idx] - As @ x
            if weights == "tigre":
                r = r * Wrow[idx]
            else:
                R = As.sum(axis=1)
                r = np.where(R > eps, r / np.where(R > eps, R, 1.0), 0.0)
            C = As.sum(axis=0)
            if weights == "tigre" or len(sets) == 1:
                C = np.where(C <= 0, 1e30, C)
            x = clamp(x + lam * (As.T @ r) / (C + eps))
            lam *= q
        if weights == "tigre":
            lam *= red
        if tv is not None:
            for _ in range(tv[1]):
                x = x - strength * oracle.tv_value_grad(x.reshape(SHAPE), vol, tv[2])[1].ravel()
            strength *= tv[3]
        x = clamp(x)
    return x.reshape(SHAPE)


@pytest.mark.parametrize("weights,method,subsets,order", [
    ("detailed", "sirt", 1, "sequential"), ("detailed", "sart", 1, "sequential"),
    ("detailed", "ossart", 4, "golden"), ("detailed", "ossart", 3, "sequential"),
    ("tigre", "sart", 1, "sequential"), ("tigre", "ossart", 4, "sequential"),
    ("tigre", "sirt", 1, "sequential")])
def test_matches_explicit_matrix_restatement(oracle, tiny, weights, method, subsets, order):
    views, vol = tiny
    A = system_matrix(oracle, views, vol, kind="siddon")
    rng = np.random.default_rng(7)
    b = A @ rng.random(A.shape[1]) + 0.02 * rng.standard_normal(A.shape[0])
    kw = dict(iterations=3, lam=0.8, red=0.7, lo=0.0, hi=0.9)
    tv = (2e-3, 2, 1e-3, 0.5)
    expect = dense_reconstruction(oracle, A, b, vol, views, weights, method, subsets, order, tv=tv, **kw)
    x = oracle.algebraic(views, NU, NV, vol, b, method=method, weights=weights, order=order,
                         subsets=subsets, iterations=3, relaxation=0.8, relaxation_reduction=0.7,
                         min_value=0.0, max_value=0.9, forward="siddon", back="siddon",
                         tv_strength=tv[0], tv_iterations=tv[1], tv_epsilon=tv[2],
                         tv_strength_reduction=tv[3])
    np.testing.assert_allclose(x, expect, rtol=1e-6, atol=1e-7)


def test_convergence_known_answer(oracle, tiny):
    """Yktest/test_reconstructors_smoke.cpp:184-205: OS-SART, 4 iterations, 2
    subsets, relaxation 0.2, a relative-residual tolerance of 2 (> 1) stops at
    the first check after one iteration."""
    views, vol = tiny
    A = system_matrix(oracle, views, vol)
    b = A @ np.random.default_rng(4).random(A.shape[1])
    kw = dict(method="ossart", subsets=2, relaxation=0.2, min_value=0.0, back="joseph")
    x, st = oracle.algebraic(views, NU, NV, vol, b, iterations=4, relative_residual=2.0, return_stats=True, **kw)
    assert st["completed_iterations"] == 1 and st["convergence_checks"] == 1
    assert st["stopped_by_relative_residual"] == 1
    np.testing.assert_array_equal(x, oracle.algebraic(views, NU, NV, vol, b, iterations=1, **kw))
    # the residual is |b - A x| / |b| on the explicit matrix
    assert st["relative_projection_residual"] == pytest.approx(
        np.linalg.norm(b - A @ x.ravel()) / np.linalg.norm(b), rel=1e-6)


@pytest.mark.parametrize("interval,patience", [(1, 1), (2, 1), (1, 2)])
def test_update_tolerance_stop_matches_recomputation(oracle, tiny, interval, patience):
    views, vol = tiny
    A = system_matrix(oracle, views, vol)
    b = A @ np.random.default_rng(8).random(A.shape[1])
    kw = dict(method="sirt", relaxation=1.0, back="joseph")
    xs = [np.zeros(SHAPE)] + [oracle.algebraic(views, NU, NV, vol, b, iterations=k, **kw) for k in range(1, 13)]
    tol = 0.2
    # recompute the checks: every `interval`-th iteration and the last (12)
    expect, prev, count = 12, xs[0], 0
    for k in range(1, 13):
        if k % interval and k != 12:
            continue
        ratio = np.linalg.norm(xs[k] - prev) / max(np.linalg.norm(prev), 1e-6)
        prev = xs[k]
        count = count + 1 if (k >= 2 and ratio <= tol) else 0
        if count >= patience:
            expect = k
            break
    assert expect < 12
    x, st = oracle.algebraic(views, NU, NV, vol, b, iterations=12, relative_update=tol, minimum_iterations=2,
                             check_interval=interval, patience=patience, return_stats=True, **kw)
    assert st["completed_iterations"] == expect and st["stopped_by_relative_update"] == 1
    np.testing.assert_allclose(x, xs[expect], rtol=1e-12, atol=1e-14)


def test_slice_tv_is_the_volume_tv_without_z(oracle):
    """The oracle's slice-wise TV (an infinite z spacing) against a numpy
    per-slice restatement: forward differences in x and y with zero past the
    edge, p = g / sqrt(|g|^2 + eps^2), gradient = -div p (backward)."""
    rng = np.random.default_rng(11)
    x = rng.random((4, 6, 7))
    vol = oracle.volume(7, 6, 4, 0.8, 1.3, np.inf)
    eps = 1e-2
    val, g = oracle.tv_value_grad(x, vol, eps)
    gx = np.zeros_like(x)
    gy = np.zeros_like(x)
    gx[:, :, :-1] = (x[:, :, 1:] - x[:, :, :-1]) / 0.8
    gy[:, :-1, :] = (x[:, 1:, :] - x[:, :-1, :]) / 1.3
    den = np.sqrt(gx ** 2 + gy ** 2 + eps ** 2)
    px, py = gx / den, gy / den
    div = px / 0.8 + py / 1.3
    div[:, :, 1:] -= px[:, :, :-1] / 0.8
    div[:, 1:, :] -= py[:, :-1, :] / 1.3
    np.testing.assert_allclose(g, -div, rtol=1e-12, atol=1e-12)
    assert val == pytest.approx(den.sum(), rel=1e-12)

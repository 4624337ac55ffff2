"""Pins of the FP64 CPU oracle (CPU only).

Each test checks an oracle routine against something the oracle does not
compute itself: closed forms of the ideal circular scan, numpy/scipy library
routines (FFT, windows, Bessel I0, linear interpolation), brute force on tiny
grids, physical identities (conjugate rays, adjoint dot products) and the
reference's own known-answer values.  The function names of the oracle are
those of ``oracle/yk_oracle.h``.
"""
import math

import numpy as np
import pytest
from scipy import ndimage, special
from scipy.signal import windows as sw

SID, SDD = 500.0, 1000.0


# ---------------------------------------------------------------------------
# independent numpy geometry (closed forms of the ideal circular flat scan)
# ---------------------------------------------------------------------------
def closed_form_views(thetas, sid, sdd, nu, nv, du, dv, off_u=0.0, off_v=0.0):
    """Source at Rz(t)(0,-sid,0); detector centre Rz(t)(0,sdd-sid,0) moved by
    off_u along U and off_v along V; U = Rz(t)x, V = z; detS = centre of
    pixel (0,0) (include/YKCBCT/geometry/YkSystemGeometry.hpp:61-73,
    YkDetectorGeometryBuilders.hpp:46-107)."""
    out = np.zeros((len(thetas), 13))
    for i, t in enumerate(thetas):
        c, s = math.cos(t), math.sin(t)
        src = np.array([sid * s, -sid * c, 0.0])
        U = np.array([c, s, 0.0])
        V = np.array([0.0, 0.0, 1.0])
        centre = np.array([-s * (sdd - sid), c * (sdd - sid), 0.0]) + off_u * U + off_v * V
        detS = centre - du * (nu - 1) / 2 * U - dv * (nv - 1) / 2 * V
        out[i] = np.concatenate([src, detS, du * U, dv * V, [t]])
    return out


def project_point(view, X):
    """(u, v) pixel coordinates of the ray src -> X on the detector plane,
    by a direct line-plane intersection and a 2x2 least-squares solve."""
    src, detS, dU, dV = view[0:3], view[3:6], view[6:9], view[9:12]
    m = np.cross(dU, dV)
    d = X - src
    t = np.dot(detS - src, m) / np.dot(d, m)
    P = src + t * d
    A = np.stack([dU, dV], axis=1)
    uv, *_ = np.linalg.lstsq(A, P - detS, rcond=None)
    return uv


def radial(view):
    src = view[0:3]
    r = np.array([-src[0], -src[1], 0.0])
    return r / np.linalg.norm(r), float(np.linalg.norm(r))


def interp2(plane, u, v):
    """Bilinear interpolation with zero outside the pixel centres (scipy)."""
    return ndimage.map_coordinates(plane, [np.atleast_1d(v), np.atleast_1d(u)], order=1,
                                   mode="grid-constant", cval=0.0, prefilter=False)


# ---------------------------------------------------------------------------
# geometry
# ---------------------------------------------------------------------------
def test_views_match_closed_form(oracle):
    n, nu, nv, du, dv = 12, 40, 30, 0.8, 0.7
    start = 0.3
    got = oracle.build_views(n, n, SID, SDD, nu, nv, du, dv, start_angle=start,
                             offset_unv=(0.75, 0.0, -0.35))
    thetas = start + 2 * np.pi * np.arange(n) / n
    want = closed_form_views(thetas, SID, SDD, nu, nv, du, dv, off_u=0.75, off_v=-0.35)
    np.testing.assert_allclose(got, want, atol=1e-10)
    # direction -1 reverses the angle sequence
    back = oracle.build_views(n, n, SID, SDD, nu, nv, du, dv, start_angle=start, direction=-1)
    np.testing.assert_allclose(back[:, 12], start - 2 * np.pi * np.arange(n) / n, atol=1e-12)


def test_views_invariants_tilted(oracle):
    """unit_test/test_geometry.cpp:165-204: pixel steps keep their lengths, stay
    orthogonal, and the detector faces the source."""
    v = oracle.build_views(16, 16, SID, SDD, 32, 24, 0.5, 0.6, offset_unv=(1.0, 2.0, -1.5),
                           tilt=(0.04, -0.03, 0.02), source_offset=(1.0, -2.0, 3.0))
    for g in v:
        src, detS, dU, dV = g[0:3], g[3:6], g[6:9], g[9:12]
        assert abs(np.linalg.norm(dU) - 0.5) < 1e-12
        assert abs(np.linalg.norm(dV) - 0.6) < 1e-12
        assert abs(np.dot(dU, dV)) < 1e-12
        centre = detS + 15.5 * dU + 11.5 * dV
        assert np.dot(np.cross(dU, dV), src - centre) > 0.0


def test_derived_ideal_closed_forms(oracle):
    n, nu, nv, du, dv = 90, 64, 48, 0.8, 0.6
    v = oracle.build_views(n, n, SID, SDD, nu, nv, du, dv, offset_unv=(0.75, 0.0, 0.3))
    d = oracle.derive(v, nu, nv, 2 * np.pi)
    np.testing.assert_allclose(d[:, oracle.D_SOD], SID, rtol=1e-12)
    np.testing.assert_allclose(d[:, oracle.D_SDD_RADIAL], SDD, rtol=1e-12)
    np.testing.assert_allclose(d[:, oracle.D_SDD_PLANE], SDD, rtol=1e-12)
    np.testing.assert_allclose(d[:, oracle.D_DTHETA], 2 * np.pi / n, rtol=1e-10)
    np.testing.assert_allclose(d[:, oracle.D_FSCALE], 1.0, rtol=1e-12)
    np.testing.assert_allclose(d[:, oracle.D_DU], du, rtol=1e-12)
    np.testing.assert_allclose(d[:, oracle.D_DV], dv, rtol=1e-12)
    # the panel moved, the principal point did not
    np.testing.assert_allclose(d[:, oracle.D_OFFU], -0.75 / du, atol=1e-9)
    np.testing.assert_allclose(d[:, oracle.D_OFFV], -0.3 / dv, atol=1e-9)
    for g, row in zip(v, d):
        r, _ = radial(g)
        np.testing.assert_allclose(row[oracle.D_RADIAL:oracle.D_RADIAL + 3], r, atol=1e-12)
        nrm = row[oracle.D_NORMAL:oracle.D_NORMAL + 3]
        assert np.dot(nrm, g[3:6] - g[0:3]) > 0.0  # faces away from the source
    # half scan: fscale = 2 pi / range
    d2 = oracle.derive(v[:46], nu, nv, np.pi + 0.4)
    np.testing.assert_allclose(d2[:, oracle.D_FSCALE], 2 * np.pi / (np.pi + 0.4), rtol=1e-12)


@pytest.mark.parametrize("tilted", [False, True])
def test_coefficients_match_direct_projection(oracle, tilted):
    """The affine ratios u = Cu.[X,1]/Cd.[X,1], v = Cv.[X,1]/Cd.[X,1] and the
    depth Cr.[X,1] = (X - src).r agree with a direct ray-plane intersection
    (the role of fdk/depth-denominator, Yktest/test_pipeline_smoke.cpp:2289-2353)."""
    n, nu, nv = 24, 50, 40
    kw = dict(offset_unv=(0.6, 0.0, -0.4))
    if tilted:
        kw.update(tilt=(0.05, 0.03, -0.02), source_offset=(2.0, 1.0, -3.0))
    v = oracle.build_views(n, n, SID, SDD, nu, nv, 0.9, 0.8, **kw)
    d = oracle.derive(v, nu, nv, 2 * np.pi)
    c = oracle.coeffs(v, d)
    rng = np.random.default_rng(3)
    pts = rng.uniform(-40, 40, size=(20, 3))
    for g, cf, dr in zip(v, c, d):
        r, sod = radial(g)
        np.testing.assert_allclose(dr[oracle.D_SOD], sod, rtol=1e-12)
        for X in pts:
            Xh = np.append(X, 1.0)
            den = cf[oracle.C_CD:oracle.C_CD + 4] @ Xh
            u = (cf[oracle.C_CU:oracle.C_CU + 4] @ Xh) / den
            vv = (cf[oracle.C_CV:oracle.C_CV + 4] @ Xh) / den
            np.testing.assert_allclose([u, vv], project_point(g, X), atol=1e-8)
            np.testing.assert_allclose(cf[oracle.C_CR:oracle.C_CR + 4] @ Xh, np.dot(X - g[0:3], r),
                                       rtol=1e-12)
        np.testing.assert_allclose(cf[oracle.C_WBASE], sod * sod * dr[oracle.D_DTHETA] * dr[oracle.D_FSCALE],
                                   rtol=1e-12)


# ---------------------------------------------------------------------------
# filter weights, kernels and the row filter
# ---------------------------------------------------------------------------
def library_window(kind, x):
    """Windows W(x), x in [0,1], from numpy / scipy routines (cutoff 0.5, N = 512
    puts x on a 1/256 grid, where these library windows are sampled exactly)."""
    k = np.rint(x * 256).astype(int)
    if kind == "ramlak":
        return np.ones_like(x)
    if kind == "shepplogan":
        return np.sinc(x / 2)
    if kind == "hann":
        return np.hanning(513)[256 + k]
    if kind == "hamming":
        return np.hamming(513)[256 + k]
    if kind == "blackman":
        return np.blackman(513)[256 + k]
    if kind == "kaiser":
        return np.kaiser(257, 8.6)[k]
    if kind == "tukey":
        return sw.tukey(257, 0.5)[k]
    raise KeyError(kind)


@pytest.mark.parametrize("kind", ["ramlak", "shepplogan", "hann", "hamming", "blackman", "kaiser", "tukey"])
def test_weights_match_library_windows(oracle, kind):
    """AnalyticFreq weights w[k] = gain f W(f/c) / (N du), f = k/N
    (src/Filter/YkCreateFilterKernelKernels.cu:167-192)."""
    N, du, gain = 512, 0.8, 1.3
    w = oracle.filter_weights(kind, N, du, cutoff=0.5, gain=gain, beta=8.6, tukey_alpha=0.5)
    f = np.arange(N // 2 + 1) / N
    want = gain * f * library_window(kind, f / 0.5) / (N * du)
    np.testing.assert_allclose(w, want, rtol=2e-6 if kind == "kaiser" else 1e-12, atol=1e-15)


def test_weights_special_values(oracle):
    N, du = 2048, 1.0
    f = np.arange(N // 2 + 1) / N
    base = f / N
    # cosine: W(1/2) = sqrt(2)/2 at f = c/2, W(1) = 0
    w = oracle.filter_weights("cosine", N, du, cutoff=0.5)
    k = N // 4
    assert abs(w[k] / base[k] - math.sqrt(0.5)) < 1e-12
    assert abs(w[N // 2]) < 1e-18
    # butterworth: W(1) = 1/2 for any order, W(x) = 1/(1 + x^(2n)) at x = 1/2
    w = oracle.filter_weights("butterworth", N, du, cutoff=0.5, order=3.0)
    assert abs(w[N // 2] / base[N // 2] - 0.5) < 1e-12
    assert abs(w[k] / base[k] - 1.0 / (1.0 + 0.5 ** 6)) < 1e-12
    # none: identity filter, gain / N for every bin
    w = oracle.filter_weights("none", N, 0.3, gain=2.0)
    np.testing.assert_allclose(w, 2.0 / N, rtol=1e-15)
    # cutoff: zero above c, ramp below
    w = oracle.filter_weights("ramlak", N, du, cutoff=0.3)
    above = f > 0.3
    assert np.all(w[above] == 0.0)
    np.testing.assert_allclose(w[~above], base[~above], rtol=1e-12)


def test_bessel_i0_within_published_bound(oracle):
    """Abramowitz & Stegun 9.8.1 / 9.8.2 (error < 1.6e-7 relative below 3.75,
    < 1.9e-7 of the scaled form above) against scipy's I0 and the series."""
    L = oracle.lib()
    for x in np.linspace(0.0, 20.0, 201):
        ref = special.i0(x)
        assert abs(L.yko_bessel_i0_as(x) - ref) <= 2.0e-7 * ref
        assert abs(L.yko_bessel_i0_series(x) - ref) <= 1e-13 * ref


def textbook_ramlak(N):
    """h[0] = 1/4, h[odd n] = -1/(pi^2 n^2), h[even n != 0] = 0 (du = 1),
    doc/algorithms.md:104-112."""
    n = np.arange(N)
    n = np.where(n <= N // 2, n, n - N)
    h = np.zeros(N)
    h[n == 0] = 0.25
    odd = (np.abs(n) % 2) == 1
    h[odd] = -1.0 / (np.pi ** 2 * n[odd] ** 2)
    return h


@pytest.mark.parametrize("N", [256, 2048])
def test_discrete_ramlak_dc_zero_and_textbook_kernel(oracle, N):
    """filter/discrete-ramlak-dc (test_filter_validation.cpp:17-46): DC is
    exactly 0; the other bins are the DFT of the textbook kernel / (N du);
    the effective convolution kernel of the row filter (unnormalized inverse
    transform of the weights) is the textbook one minus its mean, over du."""
    du = 0.7
    w = oracle.filter_weights("ramlak", N, du, source=1)
    assert w[0] == 0.0
    h_text = textbook_ramlak(N)
    want = np.fft.rfft(h_text).real / (N * du)
    np.testing.assert_allclose(w[1:], want[1:], rtol=1e-9, atol=1e-15)
    h = oracle.filter_kernel(w, N)
    np.testing.assert_allclose(h * du, h_text - h_text.sum() / N, atol=1e-12)


def test_padded_n_and_start_u(oracle):
    for nu, want in [(1, 2), (2, 4), (3, 8), (64, 128), (65, 256), (256, 512), (1024, 2048), (1536, 4096)]:
        assert oracle.padded_n(nu) == want
    # placement centres the (offset) principal point in the padded row
    assert oracle.start_u(256, 512, 0.0) == int(np.rint(256 - 127.5))
    assert oracle.start_u(1024, 2048, -2.5) == int(np.rint(1024 - (511.5 - 2.5)))


def test_fft_matches_numpy(oracle):
    rng = np.random.default_rng(5)
    for n in (2, 8, 64, 1024):
        re, im = rng.standard_normal(n), rng.standard_normal(n)
        fr, fi = oracle.fft(re, im)
        ref = np.fft.fft(re + 1j * im)
        np.testing.assert_allclose(fr + 1j * fi, ref, atol=1e-10 * n)
        br, bi = oracle.fft(fr, fi, inverse=True)
        np.testing.assert_allclose(br + 1j * bi, n * (re + 1j * im), atol=1e-10 * n)


@pytest.mark.parametrize("method", [0, 1])
def test_filter_rows_equal_padded_fft_convolution(oracle, method):
    """Pad to N at start_u, multiply the spectrum by the real weights,
    unnormalized inverse, crop (.../FDK/kernels/YkFDKFilterKernels.cu:32-126),
    computed here with numpy's FFT."""
    rng = np.random.default_rng(11)
    n_views, nv, nu = 3, 4, 50
    N = oracle.padded_n(nu)
    proj = rng.standard_normal((n_views, nv, nu))
    w = oracle.filter_weights("hann", N, 0.9, cutoff=0.4)
    su = [oracle.start_u(nu, N, o) for o in (0.0, 3.2, -7.6)]
    got = oracle.filter_rows(proj, w, N, su, method=method)
    for i in range(n_views):
        pad = np.zeros((nv, N))
        pad[:, su[i]:su[i] + nu] = proj[i]
        y = np.fft.irfft(np.fft.rfft(pad, axis=1) * w, N, axis=1) * N
        np.testing.assert_allclose(got[i], y[:, su[i]:su[i] + nu], atol=1e-12)


# ---------------------------------------------------------------------------
# weighting
# ---------------------------------------------------------------------------
def test_preweight_closed_form(oracle):
    """q = p SDD_r / |P - src| * SDD_r / (2 SOD), P the pixel centre
    (.../FDK/kernels/YkFDKPreWeightKernels.cu:374,401-402); SDD_r = SDD and
    SOD = SID on the ideal circle."""
    n, nu, nv, du, dv = 5, 12, 9, 1.1, 0.9
    v = oracle.build_views(n, n, SID, SDD, nu, nv, du, dv, offset_unv=(2.0, 0.0, 1.0))
    d = oracle.derive(v, nu, nv, 2 * np.pi)
    p = np.random.default_rng(2).standard_normal((n, nv, nu))
    got = oracle.preweight(v, d, p)
    uu, vv = np.meshgrid(np.arange(nu), np.arange(nv))
    for i, g in enumerate(v):
        P = g[3:6] + uu[..., None] * g[6:9] + vv[..., None] * g[9:12]
        dist = np.linalg.norm(P - g[0:3], axis=-1)
        np.testing.assert_allclose(got[i], p[i] * SDD / dist * SDD / (2 * SID), rtol=1e-12)


def test_parker_weights_sum_to_one_over_conjugate_rays(oracle):
    """Every ray of a short scan pi + 2 Gamma is measured twice with weights
    summing to 1 (times the range / pi normalization).  The conjugate of each
    ray is found geometrically (other intersection of the line with the source
    circle, then that view's detector), independent of the weight formulas."""
    nu, du = 41, 8.0
    gamma_max = math.atan((nu - 1) / 2 * du / SDD)
    rng_ = np.pi + 2 * gamma_max
    n = 1601
    start = 0.2
    thetas = start + rng_ * np.arange(n) / (n - 1)
    v = closed_form_views(thetas, SID, SDD, nu, 1, du, 1.0)
    d = oracle.derive(v, nu, 1, rng_)
    w = oracle.parker(v, d, np.ones((n, 1, nu)), start, 1, rng_)[:, 0, :] / (rng_ / np.pi)
    dbeta = rng_ / (n - 1)
    worst = 0.0
    for i in range(40, n - 40, 37):
        g = v[i]
        src = g[0:3]
        for c in range(0, nu, 4):
            P = g[3:6] + c * g[6:9]
            dxy = (P - src)[:2]
            t = -2.0 * np.dot(src[:2], dxy) / np.dot(dxy, dxy)
            s2 = src[:2] + t * dxy
            th2 = math.atan2(s2[0], -s2[1])
            beta2 = (th2 - start) % (2 * np.pi)
            view2 = closed_form_views([th2], SID, SDD, nu, 1, du, 1.0)[0]
            u2, _ = project_point(view2, src)
            u2 = min(max(u2, 0.0), nu - 1.0) if -1e-6 < u2 < nu - 1 + 1e-6 else u2
            if beta2 > rng_ or u2 < 0 or u2 > nu - 1:
                w2 = 0.0
            else:
                j = beta2 / dbeta
                w2 = float(ndimage.map_coordinates(w, [[j], [u2]], order=1, mode="nearest")[0])
            worst = max(worst, abs(w[i, c] + w2 - 1.0))
    assert worst < 2e-3, worst


# ---------------------------------------------------------------------------
# interpolation and the FDK backprojection
# ---------------------------------------------------------------------------
def test_bilinear_matches_scipy(oracle):
    rng = np.random.default_rng(8)
    plane = rng.standard_normal((7, 9))
    for _ in range(200):
        u, v = rng.uniform(-1.5, 9.5), rng.uniform(-1.5, 7.5)
        assert abs(oracle.bilinear(plane, u, v) - interp2(plane, u, v)[0]) < 1e-12
    assert oracle.bilinear(plane, 3.0, 2.0) == plane[2, 3]


@pytest.mark.parametrize("tilted", [False, True])
def test_bp_fdk_brute_force(oracle, tilted):
    """vol(X) += bilinear(q_i; u, v) SOD^2 dtheta fscale / ((X - src).r)^2
    (.../FDK/kernels/YkFDKBpKernels.cu:112-209), with (u, v) from a direct
    ray-plane intersection and scipy's interpolation."""
    n, nu, nv = 8, 20, 16
    kw = dict(tilt=(0.06, -0.04, 0.03)) if tilted else {}
    v = oracle.build_views(n, n, 60.0, 110.0, nu, nv, 1.2, 1.1, offset_unv=(0.5, 0.0, 0.4), **kw)
    d = oracle.derive(v, nu, nv, 2 * np.pi)
    c = oracle.coeffs(v, d)
    vol = oracle.volume(6, 5, 4, 1.5, 1.4, 1.3, center=(0.5, -0.3, 0.2))
    q = np.random.default_rng(4).standard_normal((n, nv, nu))
    got = oracle.bp_fdk(c, q, vol)
    want = np.zeros((4, 5, 6))
    for z in range(4):
        for y in range(5):
            for x in range(6):
                X = np.array([0.5 + (x - 2.5) * 1.5, -0.3 + (y - 2.0) * 1.4, 0.2 + (z - 1.5) * 1.3])
                acc = 0.0
                for i, g in enumerate(v):
                    r, sod = radial(g)
                    u, vv = project_point(g, X)
                    acc += interp2(q[i], u, vv)[0] * sod * sod * (2 * np.pi / n) / np.dot(X - g[0:3], r) ** 2
                want[z, y, x] = acc
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12)


def test_reference_sphere_record(oracle):
    """doc/public-api.md:46-95 example: a 10 mm sphere of mu = 0.02 on 64^3 at
    0.75 mm, 360 views, SID 200 / SDD 400, 128^2 at 0.8 mm; the reference
    records the reconstructed centre 0.0196773 (doc/testing-and-validation.md:157)."""
    v = oracle.build_views(360, 360, 200.0, 400.0, 128, 128, 0.8, 0.8)
    vol = oracle.volume(64, 64, 64, 0.75)
    idx = (np.arange(64) - 31.5) * 0.75
    zz, yy, xx = np.meshgrid(idx, idx, idx, indexing="ij")
    truth = np.where(xx * xx + yy * yy + zz * zz < 100.0, 0.02, 0.0).astype(np.float32)
    proj = oracle.fp_siddon(v, 128, 128, vol, truth).astype(np.float32)
    centre = np.array([((32 * 64) + 32) * 64 + 32])
    rec = oracle.fdk(v, proj, vol, voxel_index=centre)[0]
    assert abs(rec - 0.0196773) / 0.0196773 < 2e-3, rec


# ---------------------------------------------------------------------------
# forward projectors and their adjoints
# ---------------------------------------------------------------------------
def reference_single_view(nu, nv, du, dv):
    """makeSingleView, Yktest/test_forward_projectors.cpp:10-19."""
    g = np.zeros(13)
    g[0:3] = (0.0, -500.0, 0.0)
    g[3:6] = (-nu * 0.5 * du, 500.0, -nv * 0.5 * dv)
    g[6:9] = (du, 0.0, 0.0)
    g[9:12] = (0.0, 0.0, dv)
    return g[None, :]


def test_siddon_reference_known_answers(oracle):
    """fp/siddon-uniform-center (64 mm within 1 %) and fp/siddon-single-voxel
    (peak at (33, 33)), Yktest/test_forward_projectors.cpp:64-113."""
    vol = oracle.volume(64, 64, 64, 1.0)
    view = reference_single_view(64, 64, 1.0, 1.0)
    p = oracle.fp_siddon(view, 64, 64, vol, np.ones((64, 64, 64), np.float32))
    assert abs(p[0, 32, 32] - 64.0) <= 0.64
    one = np.zeros((64, 64, 64), np.float32)
    one[32, 32, 32] = 1.0
    p = oracle.fp_siddon(view, 64, 64, vol, one)[0]
    pv, pu = np.unravel_index(np.argmax(p), p.shape)
    assert (pu, pv) == (33, 33) and p[pv, pu] > 0


def dense_line_integral(vol_arr, vol, src, dst, samples=400000):
    """Sum of voxel values along the segment src -> dst by dense midpoint sampling."""
    t = (np.arange(samples) + 0.5) / samples
    pts = src[None, :] + t[:, None] * (dst - src)[None, :]
    lo = np.array([vol.cx - vol.nx * vol.vx / 2, vol.cy - vol.ny * vol.vy / 2, vol.cz - vol.nz * vol.vz / 2])
    idx = np.floor((pts - lo) / np.array([vol.vx, vol.vy, vol.vz])).astype(int)
    ok = np.all((idx >= 0) & (idx < np.array([vol.nx, vol.ny, vol.nz])), axis=1)
    vals = np.zeros(samples)
    vals[ok] = vol_arr[idx[ok, 2], idx[ok, 1], idx[ok, 0]]
    return vals.sum() * np.linalg.norm(dst - src) / samples


def test_siddon_ray_brute_force(oracle):
    """Exact segment lengths against dense sampling, including a segment that
    ends inside the volume (t clipped to [0, 1])."""
    rng = np.random.default_rng(21)
    vol = oracle.volume(5, 4, 3, 1.2, 0.9, 1.5, center=(0.3, -0.2, 0.1))
    x = rng.uniform(0.0, 1.0, (3, 4, 5)).astype(np.float32)
    for _ in range(12):
        src = rng.uniform(-8, 8, 3)
        dst = rng.uniform(-8, 8, 3)
        got = oracle.siddon_ray(src, dst, vol, x)
        want = dense_line_integral(x.astype(np.float64), vol, src, dst)
        assert abs(got - want) <= 2e-4 * max(1.0, abs(want)), (got, want)
    src, dst = np.array([-9.0, 0.05, 0.02]), np.array([0.0, 0.05, 0.02])
    assert abs(oracle.siddon_ray(src, dst, vol, x) - dense_line_integral(x.astype(np.float64), vol, src, dst)) < 2e-4


def joseph_brute_force(view, nu, nv, vol, x):
    """Joseph's method written in millimetres: sample the ray at every voxel
    plane of the main axis, interpolate the plane bilinearly (scipy), and scale
    by the ray length per plane step (src/FlatFpBp/FP/kernels/YkFPKernels.cu:10-202)."""
    from yk_oracle import main_axis
    m = main_axis(view, nu, nv)
    n = (vol.nx, vol.ny, vol.nz)
    vs = (vol.vx, vol.vy, vol.vz)
    cen = (vol.cx, vol.cy, vol.cz)
    arr = np.transpose(x.astype(np.float64), (2, 1, 0))  # [x][y][z]
    minor = [k for k in range(3) if k != m]
    out = np.zeros((nv, nu))
    for j in range(nv):
        for i in range(nu):
            P = view[3:6] + i * view[6:9] + j * view[9:12]
            src = view[0:3]
            dvec = P - src
            acc = 0.0
            for s in range(n[m]):
                xm = cen[m] + (s - (n[m] - 1) / 2) * vs[m]
                t = (xm - src[m]) / dvec[m]
                pt = src + t * dvec
                c = [(pt[k] - cen[k]) / vs[k] + (n[k] - 1) / 2 for k in minor]
                plane = np.take(arr, s, axis=m)
                acc += ndimage.map_coordinates(plane, [[c[0]], [c[1]]], order=1, mode="grid-constant",
                                               cval=0.0, prefilter=False)[0]
            out[j, i] = acc * vs[m] * np.linalg.norm(dvec) / abs(dvec[m])
    return out


def test_joseph_brute_force(oracle):
    rng = np.random.default_rng(9)
    vol = oracle.volume(7, 6, 5, 1.1, 1.0, 1.2, center=(0.2, -0.1, 0.3))
    x = rng.uniform(0.0, 1.0, (5, 6, 7)).astype(np.float32)
    nu, nv = 9, 7
    v = oracle.build_views(6, 6, 40.0, 80.0, nu, nv, 2.0, 2.0, offset_unv=(0.7, 0.0, -0.5),
                           start_angle=0.4)
    got = oracle.fp_joseph(v, nu, nv, vol, x)
    for k, g in enumerate(v):
        np.testing.assert_allclose(got[k], joseph_brute_force(g, nu, nv, vol, x), rtol=1e-9, atol=1e-12)


def test_joseph_axis_aligned_ray_is_a_column_sum(oracle):
    """The centre ray of an odd detector through an odd centred volume runs along
    y through voxel centres: the sum of the column times the voxel size."""
    vol = oracle.volume(9, 9, 9, 1.0, 1.5, 1.0)
    x = np.random.default_rng(1).uniform(0, 1, (9, 9, 9)).astype(np.float32)
    g = closed_form_views([0.0], 100.0, 200.0, 9, 9, 2.0, 2.0)
    p = oracle.fp_joseph(g, 9, 9, vol, x)
    assert abs(p[0, 4, 4] - 1.5 * x[4, :, 4].astype(np.float64).sum()) < 1e-9


def test_adjoint_dot_products(oracle):
    """<A x, y> = <x, A^T y> for both ray models, to FP64 round-off."""
    rng = np.random.default_rng(13)
    vol = oracle.volume(10, 9, 8, 1.0, 1.1, 0.9, center=(0.4, 0.0, -0.3))
    nu, nv = 14, 12
    v = oracle.build_views(7, 7, 30.0, 60.0, nu, nv, 1.6, 1.5, offset_unv=(0.8, 0.0, 0.3),
                           tilt=(0.02, 0.01, -0.03))
    x = rng.uniform(-1, 1, (8, 9, 10)).astype(np.float32)
    y = rng.uniform(-1, 1, (7, nv, nu))
    for fp, bp in ((oracle.fp_joseph, oracle.bp_joseph_adjoint), (oracle.fp_siddon, oracle.bp_siddon_adjoint)):
        lhs = np.sum(fp(v, nu, nv, vol, x) * y)
        rhs = np.sum(x.astype(np.float64) * bp(v, nu, nv, vol, y))
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def test_ellipsoid_line_integrals_agree_with_voxelized_phantom(oracle):
    """Analytic chords of rotated ellipsoids against Siddon through the oracle's
    own fine voxelization (pins the shared record convention), and the chord
    of an axis-aligned ellipsoid through its centre (2 a mu)."""
    from fdk_cuda_impl_amd import phantoms
    ell = np.asarray(phantoms.random_ellipsoids(seed=3, count=3, semi_axes=(4.0, 9.0), density=(0.5, 1.0),
                                                center_radius=4.0), dtype=np.float64)
    nu, nv = 8, 6
    v = oracle.build_views(3, 3, 60.0, 120.0, nu, nv, 4.0, 4.0, start_angle=0.3)
    analytic = oracle.ellipsoid_line_integrals(ell, v, nu, nv)
    vol = oracle.volume(128, 128, 128, 0.25)
    x = oracle.phantom(ell, vol, clamp=False)
    numeric = oracle.fp_siddon(v, nu, nv, vol, x)
    # centre-sampled voxelization leaves ~1 % RMS; a transposed rotation gives ~23 %
    assert np.linalg.norm(numeric - analytic) < 0.03 * np.linalg.norm(analytic)
    e = np.zeros(16)
    e[0:3] = 0.0
    e[3:6] = (7.0, 3.0, 2.0)
    e[6:15] = np.eye(3).ravel()
    e[15] = 0.3
    g = closed_form_views([np.pi / 2], 60.0, 120.0, 3, 3, 1.0, 1.0)  # source on the x axis
    li = oracle.ellipsoid_line_integrals(e, g, 3, 3)
    assert abs(li[0, 1, 1] - 2 * 7.0 * 0.3) < 1e-9


def test_tv_value_and_gradient(oracle):
    """Smoothed TV with forward differences over the voxel spacing (zero past the
    last voxel): closed forms for a constant and a linear ramp volume, and the
    gradient against central finite differences of the value."""
    eps = 1e-3
    vol = oracle.volume(5, 4, 3, 0.5, 1.0, 2.0)
    val, g = oracle.tv_value_grad(np.full((3, 4, 5), 2.0), vol, eps)
    assert abs(val - 60 * eps) < 1e-15 and np.max(np.abs(g)) == 0.0
    a = 0.7
    ramp = np.broadcast_to(a * np.arange(5)[None, None, :], (3, 4, 5)).copy()
    val, _ = oracle.tv_value_grad(ramp, vol, eps)
    assert abs(val - (4 * 12 * math.hypot(a / 0.5, eps) + 12 * eps)) < 1e-12
    rng = np.random.default_rng(6)
    x = rng.standard_normal((3, 4, 5))
    _, g = oracle.tv_value_grad(x, vol, 0.1)
    h = 1e-6
    for _ in range(12):
        k = tuple(rng.integers(0, s) for s in x.shape)
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        fd = (oracle.tv_value_grad(xp, vol, 0.1)[0] - oracle.tv_value_grad(xm, vol, 0.1)[0]) / (2 * h)
        assert abs(fd - g[k]) < 1e-6 * max(1.0, abs(fd))


@pytest.mark.parametrize("matched", [False, True, 2])
def test_bp_fdk_operator_brute_force(oracle, matched):
    """Operator-form FDK backprojections (src/FlatFpBp/BP/kernels/YkBpFdkKernel.cu:228-466)
    with a source above the orbit plane, so the FDK weight's denominator
    SOD - src.X / SOD differs from the radial depth of the FDK kernel."""
    n, nu, nv = 6, 18, 14
    v = oracle.build_views(n, n, 60.0, 110.0, nu, nv, 1.2, 1.1, offset_unv=(0.5, 0.0, 0.4),
                           source_offset=(0.0, 0.0, 3.0), tilt=(0.03, 0.0, -0.02))
    vol = oracle.volume(5, 4, 3, 1.5, 1.4, 1.3, center=(0.5, -0.3, 0.2))
    p = np.random.default_rng(12).standard_normal((n, nv, nu))
    got = oracle.bp_fdk_operator(v, p, vol, matched=matched)
    du, dv = np.linalg.norm(v[0, 6:9]), np.linalg.norm(v[0, 9:12])
    want = np.zeros((3, 4, 5))
    for z in range(3):
        for y in range(4):
            for x in range(5):
                X = np.array([0.5 + (x - 2) * 1.5, -0.3 + (y - 1.5) * 1.4, 0.2 + (z - 1) * 1.3])
                acc = 0.0
                for i, g in enumerate(v):
                    src = g[0:3]
                    u, vv = project_point(g, X)
                    if matched == 2:
                        # ray-plane intersection form (joseph_bp_v2_kernel): length of the
                        # line through one slab of the main axis in voxel units
                        dvox = (X - src) / np.array([1.5, 1.4, 1.3])
                        m = int(np.argmax(np.abs(dvox)))
                        w = np.linalg.norm(X - src) * (np.array([1.5, 1.4, 1.3])[m] / abs((X - src)[m]))
                    elif matched:
                        L = np.linalg.norm(g[3:6] + u * g[6:9] + vv * g[9:12] - src)
                        m = np.cross(g[6:9], g[9:12])
                        sdd = abs(np.dot(g[3:6] - src, m / np.linalg.norm(m)))
                        w = L ** 3 / np.dot(X - src, X - src) / sdd * (1.5 * 1.4 * 1.3) / (du * dv)
                    else:
                        sod = np.hypot(src[0], src[1])
                        w = sod * sod * (2 * np.pi / n) / (sod - np.dot(src, X) / sod) ** 2
                    acc += interp2(p[i], u, vv)[0] * w
                want[z, y, x] = acc
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12)


def test_voxel_driven_joseph_weight_is_the_joseph_matrix_entry(oracle):
    """For a ray through a voxel centre the Joseph model's entry A[ray, voxel]
    is its step length (interpolation weight 1), and the voxel-driven Joseph
    backprojection of that pixel's delta samples weight 1 at the voxel: both
    equal, with the main axis X or Y and anisotropic voxels."""
    n, nu, nv = 4, 21, 15
    v = oracle.build_views(n, n, 60.0, 110.0, nu, nv, 1.0, 1.0, start_angle=0.3)
    vol = oracle.volume(5, 5, 3, 1.0, 1.2, 0.9)
    for i in range(n):
        e = np.zeros((n, nv, nu))
        e[i, (nv - 1) // 2, (nu - 1) // 2] = 1.0
        a = oracle.bp_joseph_adjoint(v, nu, nv, vol, e)[1, 2, 2]
        b = oracle.bp_fdk_operator(v, e, vol, matched=2)[1, 2, 2]
        assert a > 0.9
        assert b == pytest.approx(a, rel=1e-12)


def test_bp_fdk_operator_equals_fdk_bp_in_the_orbit_plane(oracle):
    """With the source in z = 0 the FDK weight denominator is the radial depth,
    so the operator equals the FDK kernel's backprojection of the same data."""
    n, nu, nv = 8, 20, 16
    v = oracle.build_views(n, n, 60.0, 110.0, nu, nv, 1.2, 1.1, offset_unv=(0.5, 0.0, 0.4))
    vol = oracle.volume(6, 5, 4, 1.5, 1.4, 1.3)
    p = np.random.default_rng(2).standard_normal((n, nv, nu))
    d = oracle.derive(v, nu, nv, 2 * np.pi)
    np.testing.assert_allclose(oracle.bp_fdk_operator(v, p, vol), oracle.bp_fdk(oracle.coeffs(v, d), p, vol),
                               rtol=1e-12, atol=1e-14)

"""ctypes front end of the FP64 CPU oracle (``libyk_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py``.  The product package never imports
this module.  See ``yk_oracle.h`` for the reference citations of each routine.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libyk_oracle.so")

VIEW_STRIDE = 13
DERIVED_STRIDE = 20
COEFF_STRIDE = 17
ELLIPSOID_STRIDE = 16

# derived-record offsets
D_THETA, D_DTHETA, D_FSCALE, D_SOD, D_SDD_RADIAL, D_SDD_PLANE = 0, 1, 2, 3, 4, 5
D_OFFU, D_OFFV, D_DU, D_DV, D_RADIAL, D_NORMAL = 6, 7, 8, 9, 10, 13
# coefficient offsets
C_CU, C_CV, C_CD, C_CR, C_WBASE = 0, 4, 8, 12, 16

FILTERS = {"none": 0, "ramlak": 1, "shepplogan": 2, "cosine": 3, "hann": 4,
           "hamming": 5, "blackman": 6, "butterworth": 7, "kaiser": 8, "tukey": 9, "custom": 10}


class Volume(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("vx", ctypes.c_double), ("vy", ctypes.c_double), ("vz", ctypes.c_double),
                ("cx", ctypes.c_double), ("cy", ctypes.c_double), ("cz", ctypes.c_double)]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc -O2 -fopenmp)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "liboracle"], cwd=_HERE, check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        ip = ctypes.POINTER(ctypes.c_int)
        i64p = ctypes.POINTER(ctypes.c_int64)
        vp = ctypes.POINTER(Volume)
        c_int, c_double = ctypes.c_int, ctypes.c_double
        sig = {
            "yko_build_circular_flat_views": (c_int, [c_int, c_int, c_double, c_int, c_double, c_double,
                                                      c_double, dp, c_int, c_int, c_double, c_double, dp,
                                                      c_double, c_double, c_double, dp]),
            "yko_derive": (c_int, [dp, c_int, c_int, c_int, c_double, dp]),
            "yko_fdk_coeffs": (c_int, [dp, dp, c_int, dp]),
            "yko_preweight": (None, [dp, dp, c_int, c_int, c_int, dp, dp]),
            "yko_parker": (None, [dp, dp, c_int, c_int, c_int, c_double, c_int, c_double, dp]),
            "yko_padded_n": (c_int, [c_int]),
            "yko_start_u": (c_int, [c_int, c_int, c_double]),
            "yko_bessel_i0_as": (c_double, [c_double]),
            "yko_bessel_i0_series": (c_double, [c_double]),
            "yko_window": (c_double, [c_int, c_double, c_double, c_double, c_double, c_int]),
            "yko_filter_weights": (c_int, [c_int, c_int, c_double, c_double, c_double, c_double,
                                           c_double, c_int, c_double, dp]),
            "yko_filter_weights_ext": (c_int, [c_int, c_int, c_int, c_double, c_double, c_double, c_double,
                                               c_double, c_int, c_double, dp, c_int, dp, c_int, dp]),
            "yko_filter_kernel_from_weights": (None, [dp, c_int, dp]),
            "yko_filter_rows": (None, [dp, c_int, c_int, c_int, c_int, dp, ip, c_int, dp]),
            "yko_fft": (None, [dp, dp, c_int, c_int]),
            "yko_bilinear": (c_double, [dp, c_int, c_int, c_double, c_double]),
            "yko_bp_fdk": (None, [dp, c_int, dp, c_int, c_int, vp, i64p, ctypes.c_int64, dp]),
            "yko_main_axis": (c_int, [dp, c_int, c_int]),
            "yko_bp_fdk_operator": (None, [dp, dp, dp, c_int, dp, c_int, c_int, vp, c_int, dp]),
            "yko_fp_joseph": (None, [dp, c_int, c_int, c_int, vp, fp, dp]),
            "yko_bp_joseph_adjoint": (None, [dp, c_int, c_int, c_int, vp, dp, dp]),
            "yko_fp_siddon": (None, [dp, c_int, c_int, c_int, vp, fp, dp]),
            "yko_bp_siddon_adjoint": (None, [dp, c_int, c_int, c_int, vp, dp, dp]),
            "yko_siddon_ray": (c_double, [dp, dp, vp, fp]),
            "yko_phantom_ellipsoids": (None, [dp, c_int, vp, c_int, fp]),
            "yko_ellipsoid_line_integrals": (None, [dp, c_int, dp, c_int, c_int, c_int, dp]),
            "yko_tv_value_grad": (None, [dp, vp, c_double, dp, dp]),
            "yko_num_threads": (c_int, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def volume(nx, ny, nz, vx, vy=None, vz=None, center=(0.0, 0.0, 0.0)) -> Volume:
    vy = vx if vy is None else vy
    vz = vx if vz is None else vz
    return Volume(nx, ny, nz, vx, vy, vz, *center)


def build_views(total_views, views_per_turn, sid, sdd, channels, rows, du, dv,
                start_angle=0.0, direction=1, z_mm=0.0, source_offset=(0, 0, 0),
                offset_unv=(0, 0, 0), tilt=(0, 0, 0)) -> np.ndarray:
    out = np.zeros((total_views, VIEW_STRIDE))
    so = _d(source_offset)
    off = _d(offset_unv)
    rc = lib().yko_build_circular_flat_views(total_views, views_per_turn, start_angle, direction,
                                             sid, sdd, z_mm, _dp(so), channels, rows, du, dv,
                                             _dp(off), tilt[0], tilt[1], tilt[2], _dp(out))
    if rc != 0:
        raise ValueError("invalid circular scan")
    return out


def views_from_float(views_f32) -> np.ndarray:
    """(n, 20) float32 product views (5 x float4) -> (n, 13) float64 oracle views."""
    v = np.asarray(views_f32, dtype=np.float32).reshape(-1, 5, 4).astype(np.float64)
    out = np.zeros((v.shape[0], VIEW_STRIDE))
    out[:, 0:3] = v[:, 0, :3]
    out[:, 3:6] = v[:, 1, :3]
    out[:, 6:9] = v[:, 2, :3]
    out[:, 9:12] = v[:, 3, :3]
    out[:, 12] = v[:, 4, 0]
    return out


def derive(views, nu, nv, scan_range) -> np.ndarray:
    views = _d(views)
    out = np.zeros((views.shape[0], DERIVED_STRIDE))
    rc = lib().yko_derive(_dp(views), views.shape[0], nu, nv, scan_range, _dp(out))
    if rc != 0:
        raise ValueError("degenerate geometry")
    return out


def coeffs(views, derived) -> np.ndarray:
    views, derived = _d(views), _d(derived)
    out = np.zeros((views.shape[0], COEFF_STRIDE))
    lib().yko_fdk_coeffs(_dp(views), _dp(derived), views.shape[0], _dp(out))
    return out


def preweight(views, derived, proj) -> np.ndarray:
    views, derived, proj = _d(views), _d(derived), _d(proj)
    n, nv, nu = proj.shape
    out = np.zeros_like(proj)
    lib().yko_preweight(_dp(views), _dp(derived), n, nu, nv, _dp(proj), _dp(out))
    return out


def parker(views, derived, proj, start_angle, direction, scan_range) -> np.ndarray:
    views, derived = _d(views), _d(derived)
    data = _d(proj).copy()
    n, nv, nu = data.shape
    lib().yko_parker(_dp(views), _dp(derived), n, nu, nv, start_angle, direction, scan_range, _dp(data))
    return data


def padded_n(nu):
    return lib().yko_padded_n(nu)


def start_u(nu, n, offset_u_pix):
    return lib().yko_start_u(nu, n, offset_u_pix)


def filter_weights(kind, n, du, cutoff=0.5, gain=1.0, order=2.0, beta=8.6, tukey_alpha=0.5,
                   source=0, extract=0, ramp=None, custom=None) -> np.ndarray:
    """source 0 AnalyticFreq, 1 DiscreteRLFFT, 2 SpatialRampFFT (taps `ramp`);
    kind "custom" takes `custom` (n/2 + 1 weights); extract 0 real part, 1 magnitude."""
    k = FILTERS[kind] if isinstance(kind, str) else int(kind)
    out = np.zeros(n // 2 + 1)
    r = _d(ramp) if ramp is not None else None
    c = _d(custom) if custom is not None else None
    st = lib().yko_filter_weights_ext(k, source, extract, cutoff, gain, order, beta, tukey_alpha, n, du,
                                      _dp(r) if r is not None else None, 0 if r is None else r.size,
                                      _dp(c) if c is not None else None, 0 if c is None else c.size,
                                      _dp(out))
    if st != 0:
        raise ValueError(f"oracle filter weights rejected the input ({st})")
    return out


def filter_kernel(weights, n) -> np.ndarray:
    w = _d(weights)
    h = np.zeros(n)
    lib().yko_filter_kernel_from_weights(_dp(w), n, _dp(h))
    return h


def filter_rows(proj, weights, n, start_u_per_view, method=1) -> np.ndarray:
    proj = _d(proj)
    nviews, nv, nu = proj.shape
    su = np.ascontiguousarray(start_u_per_view, dtype=np.int32)
    w = _d(weights)
    out = np.zeros_like(proj)
    lib().yko_filter_rows(_dp(proj), nviews, nv, nu, n, _dp(w),
                          su.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), method, _dp(out))
    return out


def fft(re, im, inverse=False):
    re, im = _d(re).copy(), _d(im).copy()
    lib().yko_fft(_dp(re), _dp(im), re.size, 1 if inverse else 0)
    return re, im


def bilinear(plane, u, v) -> float:
    plane = _d(plane)
    nv, nu = plane.shape
    return lib().yko_bilinear(_dp(plane), nu, nv, u, v)


def bp_fdk(coeffs_, filtered, vol: Volume, voxel_index=None) -> np.ndarray:
    c = _d(coeffs_)
    f = _d(filtered)
    n, nv, nu = f.shape
    if voxel_index is None:
        out = np.zeros((vol.nz, vol.ny, vol.nx))
        lib().yko_bp_fdk(_dp(c), n, _dp(f), nu, nv, ctypes.byref(vol), None, 0, _dp(out))
    else:
        idx = np.ascontiguousarray(voxel_index, dtype=np.int64)
        out = np.zeros(idx.size)
        lib().yko_bp_fdk(_dp(c), n, _dp(f), nu, nv, ctypes.byref(vol),
                         idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), idx.size, _dp(out))
    return out


def bp_fdk_operator(views, proj, vol: Volume, matched=False, scan_range=None) -> np.ndarray:
    """Operator-form voxel-driven backprojection of unfiltered views: weight
    FDK (matched False/0), pseudo-matched (True/1) or the voxel-driven Joseph
    chord (2); derived geometry over the whole view sequence."""
    views = _d(views)
    p = _d(proj)
    n, nv, nu = p.shape
    d = derive(views, nu, nv, 2.0 * np.pi if scan_range is None else scan_range)
    c = coeffs(views, d)
    out = np.zeros((vol.nz, vol.ny, vol.nx))
    lib().yko_bp_fdk_operator(_dp(views), _dp(d), _dp(c), n, _dp(p), nu, nv, ctypes.byref(vol),
                              int(matched), _dp(out))
    return out


def main_axis(view, nu, nv) -> int:
    return lib().yko_main_axis(_dp(_d(view)), nu, nv)


def fp_joseph(views, nu, nv, vol: Volume, volume_f32) -> np.ndarray:
    views = _d(views)
    x = _f(volume_f32)
    out = np.zeros((views.shape[0], nv, nu))
    lib().yko_fp_joseph(_dp(views), views.shape[0], nu, nv, ctypes.byref(vol), _fp(x), _dp(out))
    return out


def bp_joseph_adjoint(views, nu, nv, vol: Volume, proj) -> np.ndarray:
    views, p = _d(views), _d(proj)
    out = np.zeros((vol.nz, vol.ny, vol.nx))
    lib().yko_bp_joseph_adjoint(_dp(views), views.shape[0], nu, nv, ctypes.byref(vol), _dp(p), _dp(out))
    return out


def fp_siddon(views, nu, nv, vol: Volume, volume_f32) -> np.ndarray:
    views = _d(views)
    x = _f(volume_f32)
    out = np.zeros((views.shape[0], nv, nu))
    lib().yko_fp_siddon(_dp(views), views.shape[0], nu, nv, ctypes.byref(vol), _fp(x), _dp(out))
    return out


def bp_siddon_adjoint(views, nu, nv, vol: Volume, proj) -> np.ndarray:
    views, p = _d(views), _d(proj)
    out = np.zeros((vol.nz, vol.ny, vol.nx))
    lib().yko_bp_siddon_adjoint(_dp(views), views.shape[0], nu, nv, ctypes.byref(vol), _dp(p), _dp(out))
    return out


def siddon_ray(src, dst, vol: Volume, volume_f32) -> float:
    return lib().yko_siddon_ray(_dp(_d(src)), _dp(_d(dst)), ctypes.byref(vol), _fp(_f(volume_f32)))


def phantom(ellipsoids, vol: Volume, clamp=True) -> np.ndarray:
    e = _d(ellipsoids).reshape(-1, ELLIPSOID_STRIDE)
    out = np.zeros((vol.nz, vol.ny, vol.nx), dtype=np.float32)
    lib().yko_phantom_ellipsoids(_dp(e), e.shape[0], ctypes.byref(vol), 1 if clamp else 0, _fp(out))
    return out


def ellipsoid_line_integrals(ellipsoids, views, nu, nv) -> np.ndarray:
    e = _d(ellipsoids).reshape(-1, ELLIPSOID_STRIDE)
    views = _d(views)
    out = np.zeros((views.shape[0], nv, nu))
    lib().yko_ellipsoid_line_integrals(_dp(e), e.shape[0], _dp(views), views.shape[0], nu, nv, _dp(out))
    return out


def num_threads() -> int:
    return lib().yko_num_threads()


def tv_value_grad(x, vol: Volume, eps):
    x = _d(x)
    g = np.zeros_like(x)
    val = ctypes.c_double(0.0)
    lib().yko_tv_value_grad(_dp(x), ctypes.byref(vol), eps, ctypes.byref(val), _dp(g))
    return val.value, g


def fdk(views, proj, vol: Volume, filt="ramlak", scan_range=None, cutoff=0.5, gain=1.0,
        parker_start=None, direction=1, voxel_index=None, method=1):
    """Whole FDK chain in FP64: derive, preweight, [Parker], filter, backproject."""
    views = _d(views)
    n, nv, nu = proj.shape
    if scan_range is None:
        scan_range = 2.0 * np.pi
    d = derive(views, nu, nv, scan_range)
    c = coeffs(views, d)
    q = preweight(views, d, proj)
    if parker_start is not None:
        q = parker(views, d, q, parker_start, direction, scan_range)
    N = padded_n(nu)
    w = filter_weights(filt, N, d[0, D_DU], cutoff=cutoff, gain=gain)
    su = [start_u(nu, N, d[i, D_OFFU]) for i in range(n)]
    f = filter_rows(q, w, N, su, method=method)
    return bp_fdk(c, f, vol, voxel_index)


# --- algebraic reconstruction (SIRT / SART / OS-SART) ----------------------

ALGEBRAIC_METHODS = ("sirt", "sart", "ossart")


def golden_subset_order(n):
    """Golden-ratio subset visiting order: step round(n / phi) made coprime
    with n by incrementing (falling back to 1), then 0, step, 2 step, ... mod n
    (AlgebraicDetailedWeightBackend::init,
    src/Reconstruction/Iterative/Flat/YkAlgebraicBackends.hpp:669-684)."""
    if n <= 1:
        return list(range(n))
    step = max(1, int(np.floor(n / 1.6180339887498949 + 0.5)))
    while np.gcd(step, n) != 1:
        step += 1
        if step >= n:
            step = 1
            break
    return [(i * step) % n for i in range(n)]


def _voxel_driven(mode):
    def back(views, nu, nv, vol, proj):
        return bp_fdk_operator(views, _d(proj).reshape(-1, nv, nu), vol, matched=mode)
    return back


def _operators(kind):
    """Forward and back models by name; the voxel-driven backs are the
    pseudo-matched FDK weight and the voxel-driven Joseph chord."""
    fwd = {"joseph": fp_joseph, "siddon": fp_siddon}[kind[0]]
    back = {"joseph": bp_joseph_adjoint, "siddon": bp_siddon_adjoint,
            "fdk_pseudo_matched": _voxel_driven(1), "joseph_voxel": _voxel_driven(2)}[kind[1]]
    return fwd, back


def tigre_support(views, nv, vol: Volume) -> Volume:
    """2 x 2 x 2 support grid of the TIGRE row weights: voxel (1.1 Nx vx / 2,
    1.1 Ny vy / 2, max(Nv dv, Nz vz) / 2), same centre
    (AlgebraicTigreBackend::init, YkAlgebraicBackends.hpp:125-134)."""
    dv = float(np.linalg.norm(_d(views)[0, 9:12]))
    return volume(2, 2, 2, vol.nx * vol.vx * 1.1 / 2, vol.ny * vol.vy * 1.1 / 2,
                  max(nv * dv, vol.nz * vol.vz) / 2, (vol.cx, vol.cy, vol.cz))


def tigre_row_weights(views, nu, nv, vol: Volume, forward="joseph") -> np.ndarray:
    """W = 1 / A_lo 1 over the support grid, A_lo 1 <= min voxel / 2 -> 1e30
    first (AlgebraicTigreBackend::iterate, YkAlgebraicBackends.hpp:178-186)."""
    fwd, _ = _operators((forward, "joseph"))
    r = fwd(_d(views), nu, nv, tigre_support(views, nv, vol), np.ones((2, 2, 2), dtype=np.float32))
    return 1.0 / np.where(r <= min(vol.vx, vol.vy, vol.vz) / 2, 1e30, r)


def algebraic(views, nu, nv, vol: Volume, b, method="sart", weights=None, order=None, iterations=10,
              subsets=1, relaxation=1.0, relaxation_reduction=1.0, epsilon=1e-6, min_value=None,
              max_value=None, forward="joseph", back="joseph_voxel", tv_strength=0.0, tv_iterations=0,
              tv_epsilon=1e-8, tv_strength_reduction=1.0, tv_dimensions=3, x0=None, relative_residual=0.0,
              relative_update=0.0, relative_improvement=0.0, minimum_iterations=1, check_interval=1,
              patience=1, return_stats=False):
    """FP64 restatement of the unified algebraic reconstructor over the
    oracle's projector pair; returns the volume [z][y][x] (and the monitor's
    statistics with ``return_stats``).

    Outer loop (AlgebraicReconstructorEx::reconstruct,
    src/Reconstruction/Iterative/Flat/YkAlgebraicReconstructorEx.hpp:137-180):
    every subset once, then ``tv_iterations`` steps x -= strength grad TV
    (SmoothedTv::apply, src/Reconstruction/Iterative/Regularization/
    RegularizerPlugins.cpp:70-79) with strength *= reduction afterwards, then
    the bounds.  Subset count: SIRT 1, SART one per view, OS-SART ``subsets``
    (resolvedSubsetCount_, :357-363).

    ``weights="detailed"`` (AlgebraicDetailedWeightBackend::iterate,
    YkAlgebraicBackends.hpp:706-790; one subset: SirtBatchBackend::iterate,
    YkSirtBatchBackend.hpp:46-78): subset s = views s, s + n, ...; R = A_s 1,
    r = R > eps ? (b_s - A_s x) / R : 0; C = B_s 1 (one subset: C <= 0 ->
    1e30); x += lambda B_s r / (C + eps), bounds; lambda *= reduction^(1/n)
    after each update (:117-121 of the reconstructor).

    ``weights="tigre"`` (AlgebraicTigreBackend::init/iterate,
    YkAlgebraicBackends.hpp:64-228): contiguous blocks of ceil(Na / n) views;
    W = 1 / A_lo 1 where A_lo projects a 2 x 2 x 2 support volume of voxel
    (1.1 Nx vx / 2, 1.1 Ny vy / 2, max(Nv dv, Nz vz) / 2) centred on the
    volume, with A_lo 1 <= min voxel / 2 -> 1e30 before the reciprocal;
    C = B_s 1 with C <= 0 -> 1e30; r = (b_s - A_s x) W; x += lambda B_s r /
    (C + eps), bounds; lambda *= reduction once per sweep.

    Convergence monitor (AlgebraicReconstructorEx::reconstruct / checkConvergence_,
    YkAlgebraicReconstructorEx.hpp:150-170, 258-330): after outer iteration k
    (every ``check_interval``-th and the last), r = |b - A x| / max(|b|, eps),
    improvement (r_prev - r) / max(r_prev, eps) since the last check, update
    |x - x_prev| / max(|x_prev|, eps) with x_prev the volume at the last check
    (or the start); satisfied when k >= minimum_iterations and any enabled
    test holds; stop after ``patience`` satisfied checks in a row.
    """
    views = _d(views)
    na = views.shape[0]
    if method not in ALGEBRAIC_METHODS:
        raise ValueError(method)
    weights = weights or ("tigre" if method == "sart" else "detailed")
    order = order or ("golden" if method == "ossart" else "sequential")
    n = 1 if method == "sirt" else na if method == "sart" else int(subsets)
    fwd, bwd = _operators((forward, back))
    A = lambda vs, x: fwd(vs, nu, nv, vol, x.astype(np.float32))  # noqa: E731
    B = lambda vs, p: bwd(vs, nu, nv, vol, p)  # noqa: E731
    b = _d(b).reshape(na, nv, nu)
    x = np.zeros((vol.nz, vol.ny, vol.nx)) if x0 is None else _d(x0).copy()
    tigre = weights == "tigre"
    if tigre:
        bs = (na + n - 1) // n
        units = [list(range(s, min(na, s + bs))) for s in range(0, na, bs)]
        visit = list(range(len(units)))
    else:
        units = [list(range(s, na, n)) for s in range(n)]
        visit = golden_subset_order(n) if order == "golden" and n > 1 else list(range(n))
    mask_cols = tigre or len(units) == 1

    rows, cols = [], []
    for u in units:
        vs = views[u]
        if tigre:
            r = tigre_row_weights(vs, nu, nv, vol, forward)
        else:
            r = A(vs, np.ones_like(x))
        c = B(vs, np.ones((len(u), nv, nu)))
        if mask_cols:
            c = np.where(c <= 0.0, 1e30, c)
        rows.append(r)
        cols.append(c)

    def bounds(v):
        if min_value is not None:
            v = np.maximum(v, min_value)
        if max_value is not None:
            v = np.minimum(v, max_value)
        return v

    lam = float(relaxation)
    per_update = 1.0 if tigre else float(relaxation_reduction) ** (1.0 / len(units))
    strength = float(tv_strength)
    watch_res = relative_residual > 0 or relative_improvement > 0
    watch_upd = relative_update > 0
    measured = float(np.linalg.norm(b)) if watch_res else 0.0
    x_prev = x.copy()
    st = dict(completed_iterations=0, convergence_checks=0, stopped_by_relative_residual=0,
              stopped_by_relative_update=0, stopped_by_stagnation=0, projection_residual_l2=np.nan,
              relative_projection_residual=np.nan, relative_volume_update=np.nan,
              relative_residual_improvement=np.nan)
    prev_res, count = np.nan, 0
    for outer in range(iterations):
        for k in visit:
            u = units[k]
            res = b[u] - A(views[u], x)
            if tigre:
                res = res * rows[k]
            else:
                res = np.where(rows[k] > epsilon, res / np.where(rows[k] > epsilon, rows[k], 1.0), 0.0)
            x = bounds(x + lam * B(views[u], res) / (cols[k] + epsilon))
            lam *= per_update
        if tigre:
            lam *= relaxation_reduction
        if tv_iterations > 0 and strength > 0:
            # Slice2D: no z differences, i.e. an infinite z spacing
            tv_vol = vol if tv_dimensions != 2 else volume(vol.nx, vol.ny, vol.nz, vol.vx, vol.vy, np.inf,
                                                           (vol.cx, vol.cy, vol.cz))
            for _ in range(tv_iterations):
                _, g = tv_value_grad(x, tv_vol, tv_epsilon)
                x = x - strength * g
            strength *= tv_strength_reduction
        x = bounds(x)
        done = outer + 1
        st["completed_iterations"] = done
        if not (watch_res or watch_upd) or (done % check_interval != 0 and done != iterations):
            continue
        st["convergence_checks"] += 1
        res_ok = upd_ok = stag_ok = False
        if watch_res:
            r = float(np.linalg.norm(b - A(views, x)))
            st["projection_residual_l2"] = r
            st["relative_projection_residual"] = r / max(measured, epsilon)
            res_ok = relative_residual > 0 and st["relative_projection_residual"] <= relative_residual
            if np.isfinite(prev_res):
                st["relative_residual_improvement"] = (prev_res - r) / max(prev_res, epsilon)
                stag_ok = (relative_improvement > 0 and st["relative_residual_improvement"] >= 0
                           and st["relative_residual_improvement"] <= relative_improvement)
            prev_res = r
        if watch_upd:
            st["relative_volume_update"] = float(np.linalg.norm(x - x_prev) / max(np.linalg.norm(x_prev), epsilon))
            upd_ok = st["relative_volume_update"] <= relative_update
            x_prev = x.copy()
        ok = done >= minimum_iterations and (res_ok or upd_ok or stag_ok)
        count = count + 1 if ok else 0
        if count >= patience:
            st["stopped_by_relative_residual"] = int(res_ok)
            st["stopped_by_relative_update"] = int(upd_ok)
            st["stopped_by_stagnation"] = int(stag_ok)
            break
    return (x, st) if return_stats else x


# --- CGLS (robust-restart and ASTRA strategies) ------------------------------

def cgls(views, nu, nv, vol: Volume, b, strategy="robust", iterations=50, epsilon=1e-8, restart=True,
         min_value=None, max_value=None, forward="joseph", back="joseph_voxel", relative_residual=0.0,
         relative_update=0.0, relative_improvement=0.0, minimum_iterations=1, check_interval=1,
         patience=1, x0=None):
    """FP64 restatement of CglsReconstructorEx over the oracle's projector
    pair; returns (x, residual norms [iterations + 1], stats dict).

    "astra" (CglsAstraBackend::run, src/Reconstruction/Iterative/Flat/
    YkCglsBackends.hpp:84-156): r = b - A x, p = B r, gamma = |p|^2; per
    iteration w = A p, stop if |w|^2 < eps, alpha = gamma / |w|^2, x += alpha p,
    r -= alpha w, z = B r, beta = |z|^2 / (gamma + eps), p = z + beta p, bounds.
    "robust" (CglsRobustBackend::run, :281-400, checkConvergence_ :454-500):
    the same step, then the true residual |b - A x|; if it rose (after the first
    iteration) x returns to its value before the step and, unless the previous
    iteration restarted too or restarts are off, the recursion restarts from x;
    then the convergence monitor on relative residual / update / improvement
    every ``check_interval`` completed iterations, ``patience`` checks in a row.
    """
    views = _d(views)
    fwd, bwd = _operators((forward, back))
    A = lambda x: fwd(views, nu, nv, vol, x.astype(np.float32))  # noqa: E731
    B = lambda p: bwd(views, nu, nv, vol, p)  # noqa: E731
    b = _d(b).reshape(views.shape[0], nv, nu)
    x = np.zeros((vol.nz, vol.ny, vol.nx)) if x0 is None else _d(x0).copy()
    eps = float(epsilon)
    monitor = relative_residual > 0 or relative_update > 0 or relative_improvement > 0
    st = dict(completed_iterations=0, convergence_checks=0, stopped_by_breakdown=0, stopped_by_divergence=0,
              stopped_by_relative_residual=0, stopped_by_relative_update=0, stopped_by_stagnation=0,
              projection_residual_l2=np.nan, relative_projection_residual=np.nan,
              relative_volume_update=np.nan, relative_residual_improvement=np.nan)

    def bounds(v):
        if min_value is not None:
            v = np.maximum(v, min_value)
        if max_value is not None:
            v = np.minimum(v, max_value)
        return v

    def init(x):
        r = b - A(x)
        p = B(r)
        return r, p, float(np.vdot(p, p))

    res = np.zeros(iterations + 1)
    r, p, gamma = init(x)
    res[0] = np.linalg.norm(r)
    if strategy == "astra":
        k = 0
        while k < iterations:
            w = A(p)
            ww = float(np.vdot(w, w))
            if ww < eps:
                st["stopped_by_breakdown"] = 1
                break
            alpha = gamma / ww
            x = x + alpha * p
            r = r - alpha * w
            z = B(r)
            g1 = float(np.vdot(z, z))
            beta = g1 / (gamma + eps)
            gamma = g1
            p = z + beta * p
            x = bounds(x)
            st["completed_iterations"] = k + 1
            res[k + 1] = np.linalg.norm(r)
            k += 1
        res[k + 1:] = res[k]
        return x, res, st

    measured = np.sqrt(max(0.0, float(np.vdot(b, b))))
    re_init_at, count, l2_prev, prev_checked = -1, 0, np.inf, np.nan
    k = 0
    while k < iterations:
        x_prev = x.copy()
        q = A(p)
        qq = float(np.vdot(q, q))
        if qq < eps:
            st["stopped_by_breakdown"] = 1
            res[k + 1:] = res[k]
            return x, res, st
        alpha = gamma / qq
        x = x + alpha * p
        l2 = float(np.linalg.norm(b - A(x)))
        if k > 0 and l2 > l2_prev:
            x = x_prev
            res[k + 1] = res[k]
            if re_init_at + 1 == k or not restart:
                st["stopped_by_divergence"] = 1
                break
            re_init_at = k
            r, p, gamma = init(x)
            l2_prev = np.inf
            k += 1
            continue
        l2_prev = l2
        r = r - alpha * q
        s = B(r)
        g1 = float(np.vdot(s, s))
        beta = g1 / (gamma + eps)
        gamma = g1
        p = s + beta * p
        x = bounds(x)
        st["completed_iterations"] = k + 1
        res[k + 1] = l2
        done = k + 1
        k += 1
        if not monitor or (done % check_interval != 0 and done != iterations):
            continue
        st["convergence_checks"] += 1
        st["projection_residual_l2"] = l2
        st["relative_projection_residual"] = l2 / max(measured, eps)
        residual_ok = relative_residual > 0 and st["relative_projection_residual"] <= relative_residual
        stagnation_ok = False
        if np.isfinite(prev_checked):
            st["relative_residual_improvement"] = (prev_checked - l2) / max(prev_checked, eps)
            stagnation_ok = (relative_improvement > 0 and st["relative_residual_improvement"] >= 0
                             and st["relative_residual_improvement"] <= relative_improvement)
        prev_checked = l2
        update_ok = False
        if relative_update > 0:
            st["relative_volume_update"] = np.linalg.norm(x - x_prev) / max(np.linalg.norm(x_prev), eps)
            update_ok = st["relative_volume_update"] <= relative_update
        satisfied = done >= minimum_iterations and (residual_ok or update_ok or stagnation_ok)
        count = count + 1 if satisfied else 0
        if count >= patience:
            st["stopped_by_relative_residual"] = int(residual_ok)
            st["stopped_by_relative_update"] = int(update_ok)
            st["stopped_by_stagnation"] = int(stagnation_ok)
            break
    last = min(k + 1, iterations) if st["stopped_by_divergence"] else min(k, iterations)
    res[last + 1:] = res[last]
    return x, res, st

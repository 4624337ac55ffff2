"""CPU tests of the C ABI library: it loads, exports every declared symbol, and
its FP64 host math (geometry builder, derived geometry, affine coefficients,
filter weights) matches the oracle.  No kernel is launched here."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import fdk_cuda_impl_amd as yk
from fdk_cuda_impl_amd import _capi as C
from helpers import oracle_views, small_geometry

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(REPO, "include", "yk_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(yk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.lib()
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []
    assert set(names) <= set(C.SIGNATURES), "binding table lacks a declared entry point"
    assert lib.yk_abi_version() == 1


def test_library_exports_the_session_factory():
    """include/yk_session.hpp: ReconstructionSessionFactory::create / destroy
    (the reference's YkReconstructionApi.hpp:95-98) are exported C++ symbols."""
    lib = C.lib()
    for name in ("_ZN2YK28ReconstructionSessionFactory6createEv",
                 "_ZN2YK28ReconstructionSessionFactory7destroyEPNS_22IReconstructionSessionE"):
        assert hasattr(lib, name), name


def test_session_facade_validation_without_device():
    """Session-API scenarios decided before any device work (struct_size,
    api_version, enum ranges, pose, unsupported combinations, scan validity,
    execute before initialize, the flat-frame tail checks):
    tests/session_check.cpp --host-only."""
    import subprocess
    from fdk_cuda_impl_amd.build import SESSION_CHECK
    assert os.path.exists(SESSION_CHECK), "build() compiles the Session check program"
    res = subprocess.run([SESSION_CHECK, "--host-only"], capture_output=True, text=True, timeout=60)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.count("PASS") == 8


def test_struct_sizes_match_reference_pod_layout():
    # SConeProjGeomVec is 5 x float4 (YkProjectionGeometry.hpp:12-18)
    assert ctypes.sizeof(C.View) == 80
    assert ctypes.sizeof(C.VolumeGrid) == 36


def test_context_without_device_reports_device_error():
    h = ctypes.c_void_p()
    st = C.lib().yk_ctx_create(-1, ctypes.byref(h))
    assert st in (C.INVALID_CONFIG, C.DEVICE_ERROR)
    assert C.lib().yk_last_error()


def test_views_match_oracle_builder(oracle):
    for kw in [dict(), dict(start_angle_rad=0.3, rotation_direction=-1),
               dict(offset_unv_mm=(2.5, 1.0, -1.5), source_offset_mm=(0.5, -1.0, 2.0), z_mm=3.0),
               dict(tilt_rad=(0.02, -0.03, 0.05))]:
        g = small_geometry(**kw)
        v = yk.build_views(g)
        o = oracle_views(oracle, g)
        got = np.concatenate([v[:, 0:3], v[:, 4:7], v[:, 8:11], v[:, 12:15], v[:, 16:17]], axis=1)
        # float rounding of the FP64 result
        np.testing.assert_allclose(got, o, rtol=2e-7, atol=2e-5)


def test_view_invariants():
    """Right-handed detector facing the source; pixel (0,0) centre at detS
    (unit_test/test_geometry.cpp:165-204 properties)."""
    g = small_geometry(tilt_rad=(0.01, 0.02, -0.01))
    v = yk.build_views(g).astype(np.float64)
    src, detS, du, dv = v[:, 0:3], v[:, 4:7], v[:, 8:11], v[:, 12:15]
    center = detS + du * (g.channels - 1) / 2 + dv * (g.rows - 1) / 2
    n = np.cross(du, dv)
    assert np.all(np.einsum("ij,ij->i", n, src - center) > 0)
    np.testing.assert_allclose(np.linalg.norm(du, axis=1), g.channel_mm, rtol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(dv, axis=1), g.row_mm, rtol=1e-6)


def test_derived_and_coefficients_match_oracle(oracle):
    g = small_geometry(offset_unv_mm=(1.5, 0.0, 0.7), tilt_rad=(0.01, 0.0, 0.02), total_views=20)
    v = yk.build_views(g)
    ov = oracle.views_from_float(v)
    d = yk.derive(v, g.nu, g.nv, g.scan_range)
    od = oracle.derive(ov, g.nu, g.nv, g.scan_range)
    idx = [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18]
    np.testing.assert_allclose(d[:, idx], od[:, idx], rtol=1e-6, atol=1e-6)
    c = yk.fdk_coefficients(v, g.nu, g.nv, g.scan_range)
    oc = oracle.coeffs(ov, od)
    np.testing.assert_allclose(c, oc, rtol=1e-6, atol=1e-4)


def test_ideal_circle_closed_forms():
    """Ideal circle: SOD = SID, SDD_r = SDD, offsetU = offset/du,
    dtheta = 2 pi / views (the oracle pinning of derived geometry)."""
    g = small_geometry(total_views=36, views_per_turn=36, offset_unv_mm=(2.5, 0.0, 0.0))
    d = yk.derive(yk.build_views(g), g.nu, g.nv, g.scan_range)
    np.testing.assert_allclose(d[:, 3], g.sid_mm, rtol=1e-6)
    np.testing.assert_allclose(d[:, 4], g.sdd_mm, rtol=1e-6)
    # the principal point stays put while the panel moves +2.5 mm along U
    np.testing.assert_allclose(d[:, 6], -2.5 / g.channel_mm, atol=1e-4)
    np.testing.assert_allclose(d[:, 1], 2 * math.pi / 36, rtol=1e-5)
    np.testing.assert_allclose(d[:, 2], 1.0, rtol=1e-6)


@pytest.mark.parametrize("kind", [k for k in C.FILTERS if k != "custom"])
@pytest.mark.parametrize("n", [512, 2048])
@pytest.mark.parametrize("cutoff", [0.5, 0.3])
def test_filter_weights_match_oracle(oracle, kind, n, cutoff):
    w = yk.filter_weights(n, 0.8, kind=kind, cutoff=cutoff, gain=1.3)
    # the parameters are float fields (SAnalyticFilterParams); give the oracle
    # the same float values so the window edge is evaluated at the same x
    f32 = lambda v: float(np.float32(v))  # noqa: E731
    o = oracle.filter_weights(kind, n, f32(0.8), cutoff=f32(cutoff), gain=f32(1.3))
    np.testing.assert_allclose(w, o, rtol=2e-6, atol=1e-12)


def _rl_taps(m):
    """the du = 1 sampled Ram-Lak kernel on taps -(m//2)..m//2"""
    n = np.arange(m) - m // 2
    t = np.zeros(m)
    t[m // 2] = 0.25
    odd = (n % 2) != 0
    t[odd] = -1.0 / (np.pi ** 2 * n[odd].astype(np.float64) ** 2)
    return t


@pytest.mark.parametrize("kind", ["ramlak", "hann", "hamming", "shepplogan", "butterworth", "kaiser", "tukey"])
@pytest.mark.parametrize("n", [512, 2048])
@pytest.mark.parametrize("extract", ["real", "magnitude"])
def test_discrete_rl_weights_match_oracle(oracle, kind, n, extract):
    """DiscreteRLFFT source: the DFT of the sampled Ram-Lak kernel, windowed
    by window_shape (Butterworth, Kaiser and Tukey give 1 there), as in
    src/Filter/YkCreateFilterKernelFFT.hpp:192-257."""
    f32 = lambda v: float(np.float32(v))  # noqa: E731
    w = yk.filter_weights(n, 0.8, kind=kind, cutoff=0.4, gain=1.3, source="discrete_rl", extract_mode=extract)
    o = oracle.filter_weights(kind, n, f32(0.8), cutoff=f32(0.4), gain=f32(1.3), source=1,
                              extract=C.RAMP_EXTRACT[extract])
    np.testing.assert_allclose(w, o, rtol=2e-6, atol=1e-12 * np.abs(o).max())


@pytest.mark.parametrize("kind", ["ramlak", "hann", "cosine", "kaiser"])
@pytest.mark.parametrize("extract", ["real", "magnitude"])
def test_spatial_ramp_weights_match_oracle(oracle, kind, extract):
    """SpatialRampFFT source with caller taps (a smoothed, truncated ramp)."""
    n = 1024
    taps = (_rl_taps(63) * np.hanning(65)[1:-1]).astype(np.float32)
    f32 = lambda v: float(np.float32(v))  # noqa: E731
    w = yk.filter_weights(n, 0.6, kind=kind, cutoff=0.45, gain=1.7, source="spatial_ramp",
                          extract_mode=extract, spatial_ramp=taps)
    o = oracle.filter_weights(kind, n, f32(0.6), cutoff=f32(0.45), gain=f32(1.7), source=2,
                              extract=C.RAMP_EXTRACT[extract], ramp=taps.astype(np.float64))
    np.testing.assert_allclose(w, o, rtol=2e-6, atol=1e-9 * np.abs(o).max())


def test_custom_weights_match_oracle(oracle):
    n = 512
    c = np.linspace(0.0, 2.0, n // 2 + 1).astype(np.float32) ** 1.5
    w = yk.filter_weights(n, 0.7, kind="custom", gain=1.25, custom_weights=c)
    o = oracle.filter_weights("custom", n, float(np.float32(0.7)), gain=1.25, custom=c.astype(np.float64))
    np.testing.assert_allclose(w, o, rtol=2e-6, atol=1e-12)


def test_oracle_filter_sources_pinned(oracle):
    """The oracle's three DFT-built paths against closed forms and each other:
    - SpatialRamp holding the sampled Ram-Lak taps of |n| < N/2 is DiscreteRL
      (the tap at N/2 is even, hence zero);
    - the SpatialRamp DFT equals numpy's rfft of the N-periodic taps / N;
    - windowed SpatialRamp carries the gain twice (the reference's order),
      the Ram-Lak one once;
    - for symmetric taps the magnitude is |real part|;
    - Custom is the caller's weights x gain / N / du."""
    n, du = 256, 0.5
    taps = _rl_taps(n - 1)
    sr = oracle.filter_weights("hann", n, du, cutoff=0.4, source=2, ramp=taps)
    drl = oracle.filter_weights("hann", n, du, cutoff=0.4, source=1)
    np.testing.assert_allclose(sr, drl, rtol=1e-12, atol=1e-15)
    h = np.zeros(n)
    h[(np.arange(taps.size) - taps.size // 2) % n] = taps
    ref = np.fft.rfft(h).real / n
    ref[0] = 0.0
    got = oracle.filter_weights("ramlak", n, 1.0, source=2, ramp=taps)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-15)
    g1 = oracle.filter_weights("hamming", n, du, source=2, ramp=taps, gain=1.0)
    g3 = oracle.filter_weights("hamming", n, du, source=2, ramp=taps, gain=3.0)
    np.testing.assert_allclose(g3, 9.0 * g1, rtol=1e-12)
    r1 = oracle.filter_weights("ramlak", n, du, source=2, ramp=taps, gain=1.0)
    r3 = oracle.filter_weights("ramlak", n, du, source=2, ramp=taps, gain=3.0)
    np.testing.assert_allclose(r3, 3.0 * r1, rtol=1e-12)
    mag = oracle.filter_weights("ramlak", n, du, source=2, ramp=taps, extract=1)
    np.testing.assert_allclose(mag, np.abs(r1), rtol=1e-9, atol=1e-15)
    c = np.arange(n // 2 + 1, dtype=np.float64)
    np.testing.assert_allclose(oracle.filter_weights("custom", n, du, gain=2.0, custom=c), c * 2.0 / n / du,
                               rtol=1e-15)


def test_spatial_ramp_of_rl_taps_is_discrete_rl():
    """The product's two DFT paths agree on the same kernel."""
    n = 2048
    taps = _rl_taps(n - 1).astype(np.float32)
    a = yk.filter_weights(n, 0.3, kind="ramlak", source="spatial_ramp", spatial_ramp=taps)
    b = yk.filter_weights(n, 0.3, kind="ramlak", source="discrete_rl")
    np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-6 * np.abs(b).max())


def test_filter_source_validation():
    taps = _rl_taps(31).astype(np.float32)
    bad = [
        dict(kind="ramlak", source="spatial_ramp", spatial_ramp=taps[:-1]),          # even tap count
        dict(kind="ramlak", source="spatial_ramp"),                                    # no taps
        dict(kind="ramlak", source="spatial_ramp", spatial_ramp=np.r_[taps[:-1], 1.0]),  # not symmetric
        dict(kind="ramlak", source="spatial_ramp", spatial_ramp=_rl_taps(1025).astype(np.float32)),  # > N
        dict(kind="custom"),                                                           # no weights
        dict(kind="custom", custom_weights=np.ones(100, np.float32)),                  # wrong count
        dict(kind="custom", custom_weights=np.full(257, np.inf, np.float32)),          # not finite
        dict(kind="ramlak", source=3),
        dict(kind="ramlak", source="discrete_rl", extract_mode=2),
    ]
    for kw in bad:
        with pytest.raises(yk.YkError) as e:
            yk.filter_weights(512, 1.0, **kw)
        assert e.value.status == C.INVALID_CONFIG, kw


def test_padded_n():
    for nu, n in [(1, 2), (128, 256), (256, 512), (1024, 2048), (1536, 4096), (1000, 2048)]:
        assert yk.padded_n(nu) == n


def test_invalid_filter_rejected():
    with pytest.raises(yk.YkError) as e:
        yk.filter_weights(512, 1.0, kind="hann", cutoff=0.7)
    assert e.value.status == C.INVALID_CONFIG
    with pytest.raises(yk.YkError):
        yk.filter_weights(500, 1.0)


def test_invalid_scan_rejected():
    g = small_geometry(sdd_mm=100.0)  # SDD < SID
    with pytest.raises(yk.YkError) as e:
        yk.build_views(g)
    assert e.value.status == C.INVALID_GEOMETRY


def test_noalloc_limit_matches_header():
    """YK_NOALLOC_MAX_ROWS of include/yk_hip.h is the binding's constant."""
    text = open(os.path.join(REPO, "include", "yk_hip.h")).read()
    m = re.search(r"#define\s+YK_NOALLOC_MAX_ROWS\s+(\d+)", text)
    assert m and int(m.group(1)) == C.NOALLOC_MAX_ROWS

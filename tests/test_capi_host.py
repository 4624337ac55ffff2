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


@pytest.mark.parametrize("kind", list(C.FILTERS))
@pytest.mark.parametrize("n", [512, 2048])
@pytest.mark.parametrize("cutoff", [0.5, 0.3])
def test_filter_weights_match_oracle(oracle, kind, n, cutoff):
    w = yk.filter_weights(n, 0.8, kind=kind, cutoff=cutoff, gain=1.3)
    # the parameters are float fields (SAnalyticFilterParams); give the oracle
    # the same float values so the window edge is evaluated at the same x
    f32 = lambda v: float(np.float32(v))  # noqa: E731
    o = oracle.filter_weights(kind, n, f32(0.8), cutoff=f32(cutoff), gain=f32(1.3))
    np.testing.assert_allclose(w, o, rtol=2e-6, atol=1e-12)


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

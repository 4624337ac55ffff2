"""CPU tests of the calibrated-geometry diagnostics and the raw file tools
(no kernel launched).  The diagnostics are pinned by what the geometry builder
puts in: an ideal scan reports SOD = SID, SDD and nothing else; a tilted
detector reports exactly the builder's U / V / N tilts (the same intrinsic
order); a detector offset reports the principal point in pixels; a lifted
source reports its height; per-view tilts report their extremes."""
import math
import os

import numpy as np
import pytest

import fdk_cuda_impl_amd as yk
from fdk_cuda_impl_amd import rawio
from fdk_cuda_impl_amd.core import geometry_diagnostics
from helpers import small_geometry


def test_ideal_scan_reports_its_distances_only():
    g = small_geometry(total_views=36, views_per_turn=36, sid_mm=500.0, sdd_mm=1000.0)
    d = geometry_diagnostics(yk.build_views(g), g.nu, g.nv)
    assert d["min_source_to_axis_mm"] == pytest.approx(500.0, rel=1e-6)
    assert d["max_source_to_axis_mm"] == pytest.approx(500.0, rel=1e-6)
    assert d["min_source_to_detector_mm"] == pytest.approx(1000.0, rel=1e-6)
    assert d["max_source_to_plane_mm"] == pytest.approx(1000.0, rel=1e-6)
    for k in ("max_abs_source_axis_mm", "max_detector_skew", "max_normal_misalignment",
              "min_relative_tilt_u_rad", "max_relative_tilt_v_rad", "max_relative_tilt_n_rad"):
        assert abs(d[k]) < 2e-6, (k, d[k])
    # float32 view vectors at 500 mm: the principal point is exact to ~1e-5 pixel
    for k in ("max_abs_principal_u_pix", "max_abs_principal_v_pix"):
        assert abs(d[k]) < 5e-5, (k, d[k])
    assert d["approximate"] is False


@pytest.mark.parametrize("tilt", [(0.02, 0.0, 0.0), (0.0, -0.015, 0.0), (0.0, 0.0, 0.03), (0.01, -0.02, 0.025)])
def test_tilts_are_recovered(tilt):
    g = small_geometry(total_views=24, views_per_turn=24, tilt_rad=tilt)
    d = geometry_diagnostics(yk.build_views(g), g.nu, g.nv)
    for axis, t in zip("uvn", tilt):
        assert d[f"min_relative_tilt_{axis}_rad"] == pytest.approx(t, abs=2e-6)
        assert d[f"max_relative_tilt_{axis}_rad"] == pytest.approx(t, abs=2e-6)
    # the radial ray still meets the detector centre, and 1 - cos(tilt) stays
    # under the 1e-3 tolerance for these tilts: not flagged approximate
    assert d["max_normal_misalignment"] == pytest.approx(
        1.0 - math.cos(tilt[0]) * math.cos(tilt[1]), abs=2e-6)
    assert d["approximate"] is False


def test_offsets_and_source_height():
    du, dv = 1.0, 1.0
    g = small_geometry(offset_unv_mm=(3.0, 0.0, -2.0), channel_mm=du, row_mm=dv,
                       source_offset_mm=(0.0, 0.0, 4.0))
    d = geometry_diagnostics(yk.build_views(g), g.nu, g.nv)
    # the principal point is the foot of the source on the detector: U offset
    # 3 mm; along V the source height minus the detector offset, 4 - (-2) mm
    assert d["max_abs_principal_u_pix"] == pytest.approx(3.0 / du, rel=1e-5)
    assert d["max_abs_principal_v_pix"] == pytest.approx((4.0 + 2.0) / dv, rel=1e-5)
    assert d["min_principal_u_pix"] == pytest.approx(d["max_principal_u_pix"], abs=1e-5)
    assert d["max_abs_source_axis_mm"] == pytest.approx(4.0, rel=1e-6)
    assert d["approximate"] is True


def test_per_view_tilts_report_their_extremes():
    tilts = [0.01, -0.02, 0.005, 0.03]
    parts = []
    for k, t in enumerate(tilts):
        g = small_geometry(total_views=1, views_per_turn=4, start_angle_rad=k * math.pi / 2, tilt_rad=(0.0, 0.0, t))
        parts.append(yk.build_views(g))
    views = np.concatenate(parts)
    d = geometry_diagnostics(views, g.nu, g.nv)
    assert d["min_relative_tilt_n_rad"] == pytest.approx(min(tilts), abs=2e-6)
    assert d["max_relative_tilt_n_rad"] == pytest.approx(max(tilts), abs=2e-6)


def test_repeated_adjacent_angles_are_refused():
    g = small_geometry()
    v = yk.build_views(g)
    v[3] = v[2]
    with pytest.raises(yk.YkError):
        geometry_diagnostics(v, g.nu, g.nv)


def test_raw_round_trip_and_names(tmp_path):
    a = np.random.default_rng(1).standard_normal((3, 4, 5)).astype(np.float32)
    p = os.path.join(tmp_path, rawio.dump_name("proj", 7))
    assert p.endswith("proj_a0007.raw")
    rawio.write_raw(p, a)
    assert os.path.getsize(p) == a.size * 4
    np.testing.assert_array_equal(rawio.read_raw(p, a.shape), a)
    with pytest.raises(ValueError):
        rawio.read_raw(p, (3, 4, 6))


def test_text_dump_and_disabled_dump(tmp_path):
    d = rawio.Dump(str(tmp_path))
    d.text("stage", "views 0-31\n", index=12)
    with open(os.path.join(tmp_path, "stage_a0012.txt")) as f:
        assert f.read() == "views 0-31\n"
    d.set_enabled(False)
    d.text("skipped.txt", "nothing")
    d.device("skipped.raw", 0, count=16)  # disabled: no device access
    assert not os.path.exists(os.path.join(tmp_path, "skipped.txt"))
    assert not os.path.exists(os.path.join(tmp_path, "skipped.raw"))
    d.set_enabled(True)
    with pytest.raises(yk.YkError):
        d.device("bad.raw", 0, count=0)
    d.close()

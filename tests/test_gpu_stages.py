"""GPU checks of the FDK stage view (FdkStageView, YkFdkPipeline.hpp:176-191)
and the raw device dump (DumpManager, src/util/YkIoDump.cpp): after a
streamed execute the last chunk's input, weighted and filtered projections
are readable through dumped raw files; the weighted stage equals the oracle's
cosine and Parker weights, the filtered stage equals the plan's own filter
stage; reset clears the view."""
import os

import numpy as np
import pytest
import torch

import fdk_cuda_impl_amd as yk
from fdk_cuda_impl_amd import rawio
from helpers import rel_rmse, seeded_projection, small_geometry

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_last_stage_of_a_short_scan(ctx, oracle, tmp_path, storage):
    g = small_geometry(total_views=200, views_per_turn=360, start_angle_rad=0.25)
    plan = yk.FdkPlan.from_geometry(ctx, g, storage=storage, batch_views=64)
    plan.set_stage_capture(True)
    proj = torch.from_numpy(seeded_projection(9, g.total_views, g.nv, g.nu)).cuda()
    vol = torch.zeros(plan.volume_shape, device="cuda")
    plan.execute(proj, vol)
    st = plan.last_stage()
    # 200 views in chunks of 64: the last chunk holds views 192..199
    assert (st["first_view"], st["count"]) == (192, 8)
    assert st["d_input"] == proj[192].data_ptr()
    first, k = st["first_view"], st["count"]
    n = k * g.nv * g.nu
    dump = rawio.Dump(str(tmp_path))
    dump.device("pre", st["d_preweighted"], count=n, stream=st["stream"], index=first)
    pre = rawio.read_raw(os.path.join(tmp_path, rawio.dump_name("pre", first)), (k, g.nv, g.nu))
    ov = oracle.views_from_float(plan.views)
    d = oracle.derive(ov, g.nu, g.nv, g.scan_range)
    q = oracle.preweight(ov, d, proj.cpu().numpy().astype(np.float64))
    q = oracle.parker(ov, d, q, g.start_angle_rad, 1, g.scan_range)
    assert rel_rmse(pre, q[first:first + k]) < 1e-6
    # the filtered stage is the plan's filter output for those views
    ref, scales = plan.filter(proj[first:first + k].contiguous(), first=first)
    if storage == "f32":
        dump.device("filtered.raw", st["d_filtered"], count=n, stream=st["stream"])
        got = rawio.read_raw(os.path.join(tmp_path, "filtered.raw"), (k, g.nv, g.nu))
        np.testing.assert_array_equal(got, ref.cpu().numpy())
        assert st["d_row_scales"] is None
    else:
        assert st["filtered_storage"] == 1 and st["d_row_scales"]
        dump.device("scales.raw", st["d_row_scales"], count=k * g.nv, stream=st["stream"])
        got = rawio.read_raw(os.path.join(tmp_path, "scales.raw"), (k, g.nv))
        np.testing.assert_array_equal(got, scales.cpu().numpy())
    plan.reset()
    assert plan.last_stage()["count"] == 0
    dump.close()


def test_capture_off_leaves_no_weighted_stage(ctx):
    g = small_geometry()
    plan = yk.FdkPlan.from_geometry(ctx, g)
    proj = seeded_projection(3, g.total_views, g.nv, g.nu)
    vol = np.zeros(plan.volume_shape, dtype=np.float32)
    plan.execute(proj, vol)
    st = plan.last_stage()
    assert st["count"] > 0 and st["d_input"] and st["d_filtered"]
    assert st["d_preweighted"] is None


def test_plan_reports_its_geometry(ctx):
    g = small_geometry(tilt_rad=(0.0, 0.0, 0.02))
    plan = yk.FdkPlan.from_geometry(ctx, g)
    d = plan.geometry_diagnostics()
    assert d["min_relative_tilt_n_rad"] == pytest.approx(0.02, abs=2e-6)
    assert d["min_source_to_axis_mm"] == pytest.approx(g.sid_mm, rel=1e-6)

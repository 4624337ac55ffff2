"""GPU test of yk_fdk_backproject_noalloc, the allocation-free variant of
yk_fdk_backproject: called through the C ABI, its volumes must equal the
existing entry point's on the same filtered views (float, half and bfloat16
storage, with and without Kahan), on two calls in a row where the second has
fewer views, and it must refuse more rows than its fixed scratch holds."""
import ctypes

import numpy as np
import pytest
import torch

import fdk_cuda_impl_amd as yk
from fdk_cuda_impl_amd import _capi as C
from fdk_cuda_impl_amd.core import _ptr
from helpers import max_abs, rel_rmse, seeded_projection, small_geometry

pytestmark = pytest.mark.gpu


def backproject(entry, plan, filtered, scales, first, count):
    st = {torch.float32: C.F32, torch.float16: C.F16, torch.bfloat16: C.BF16}[filtered.dtype]
    vol = torch.full(plan.volume_shape, float("nan"), dtype=torch.float32, device="cuda")
    C.call(entry, plan.handle, _ptr(filtered), st, _ptr(scales), first, count, _ptr(vol), 0)
    torch.cuda.synchronize()
    return vol.cpu().numpy()


@pytest.mark.parametrize("storage,kahan", [("f32", False), ("f16", True), ("f16", False), ("bf16", True)])
def test_noalloc_matches_backproject_on_two_calls(ctx, storage, kahan):
    # 48 channels: the existing entry point takes its row-normalization pass
    g = small_geometry(channels=48, offset_unv_mm=(1.5, 0.0, 0.5))
    plan = yk.FdkPlan.from_geometry(ctx, g, filter="hann", storage=storage, kahan=kahan)
    proj = torch.from_numpy(seeded_projection(7, g.total_views, g.nv, g.nu)).cuda()
    filtered, scales = plan.filter(proj)
    # second call: a smaller input, views [5, 14)
    for first, count in ((0, g.total_views), (5, 9)):
        f = filtered[first:first + count].contiguous()
        s = scales[first:first + count].contiguous() if scales is not None else None
        print(f"noalloc storage={storage} kahan={kahan} views [{first}, {first + count})", flush=True)
        ref = backproject("yk_fdk_backproject", plan, f, s, first, count)
        got = backproject("yk_fdk_backproject_noalloc", plan, f, s, first, count)
        assert np.isfinite(got).all()
        print(f"  rel-RMSE {rel_rmse(got, ref):.2e}, max-abs {max_abs(got, ref):.2e} "
              f"(peak {np.abs(ref).max():.2e})", flush=True)
        assert rel_rmse(got, ref) <= 1e-6, rel_rmse(got, ref)
        assert max_abs(got, ref) <= 1e-6 * np.abs(ref).max()
    plan.close()


def test_noalloc_refuses_rows_beyond_its_scratch(ctx):
    # 1025 views x 4096 rows = 4198400 > YK_NOALLOC_MAX_ROWS row factors
    assert C.NOALLOC_MAX_ROWS == 4194304
    g = small_geometry(total_views=1025, views_per_turn=1025, channels=8, rows=4096, row_mm=0.05,
                       volume=(8, 8, 8), voxel_mm=(0.01, 0.01, 0.01))
    plan = yk.FdkPlan.from_geometry(ctx, g, storage="f16")
    dummy = torch.zeros(16, dtype=torch.float16, device="cuda")
    scales = torch.ones(16, dtype=torch.float32, device="cuda")
    vol = torch.zeros(plan.volume_shape, dtype=torch.float32, device="cuda")
    status = C.lib().yk_fdk_backproject_noalloc(plan.handle, _ptr(dummy), C.F16, _ptr(scales), 0, 1025,
                                                _ptr(vol), 0)
    assert status == C.UNSUPPORTED
    assert "YK_NOALLOC_MAX_ROWS" in C.lib().yk_last_error().decode()
    assert float(vol.abs().max()) == 0.0
    plan.close()

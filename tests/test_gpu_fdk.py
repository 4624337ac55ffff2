"""GPU parity of the FDK path (filter stage, backprojection, end to end,
precision variants, streaming contract) against the FP64 oracle and the
reference's published numbers.  All calls go through the C ABI."""
import numpy as np
import pytest
import torch

import fdk_cuda_impl_amd as yk
from fdk_cuda_impl_amd import phantoms
from helpers import max_abs, oracle_volume, rel_rmse, seeded_projection, small_geometry

pytestmark = pytest.mark.gpu


def oracle_filtered(oracle, plan, geom, proj, filt="ramlak", cutoff=0.5, parker=None, **wkw):
    ov = oracle.views_from_float(plan.views)
    d = oracle.derive(ov, geom.nu, geom.nv, geom.scan_range)
    q = oracle.preweight(ov, d, proj.astype(np.float64))
    if parker is not None:
        q = oracle.parker(ov, d, q, parker[0], parker[1], geom.scan_range)
    n = oracle.padded_n(geom.nu)
    w = oracle.filter_weights(filt, n, float(np.float32(d[0, oracle.D_DU])),
                              cutoff=float(np.float32(cutoff)), **wkw)
    su = [oracle.start_u(geom.nu, n, d[i, oracle.D_OFFU]) for i in range(len(ov))]
    return oracle.filter_rows(q, w, n, su, method=1), oracle.coeffs(ov, d)


@pytest.mark.parametrize("case", [
    dict(filt="ramlak", kw={}),
    dict(filt="hann", kw=dict(offset_unv_mm=(2.5, 0.0, 1.0))),
    dict(filt="shepplogan", kw=dict(total_views=13, views_per_turn=24, start_angle_rad=0.4)),
])
def test_filter_stage_matches_oracle(ctx, oracle, case):
    g = small_geometry(**case["kw"])
    plan = yk.FdkPlan.from_geometry(ctx, g, filter=case["filt"], parker="disabled")
    proj = seeded_projection(1, g.total_views, g.nv, g.nu)
    out, _ = plan.filter(torch.from_numpy(proj).cuda())
    ref, _ = oracle_filtered(oracle, plan, g, proj, filt=case["filt"])
    got = out.cpu().numpy()
    assert rel_rmse(got, ref) < 2e-6
    assert max_abs(got, ref) < 1e-5 * np.abs(ref).max()


def _rl_taps(m):
    n = np.arange(m) - m // 2
    t = np.zeros(m)
    t[m // 2] = 0.25
    odd = (n % 2) != 0
    t[odd] = -1.0 / (np.pi ** 2 * n[odd].astype(np.float64) ** 2)
    return t


@pytest.mark.parametrize("source", ["discrete_rl", "spatial_ramp", "custom"])
def test_filter_sources_stage_matches_oracle(ctx, oracle, source):
    """The weight sources other than the analytic one through the filter
    kernel (N = 1024): DiscreteRLFFT with a Hann window and magnitude
    extraction, SpatialRampFFT with caller taps and a Hamming window (gain
    1.2, applied twice as the reference does), Custom weights."""
    g = small_geometry(channels=300, rows=5, channel_mm=0.1, row_mm=0.1, total_views=4, views_per_turn=24,
                       volume=(24, 20, 18), voxel_mm=(0.05, 0.05, 0.05), offset_unv_mm=(0.15, 0.0, 0.0))
    n = 1024
    taps = (_rl_taps(41) * np.hanning(43)[1:-1]).astype(np.float32)
    custom = (np.linspace(0.0, 0.5, n // 2 + 1) * np.cos(np.linspace(0.0, 1.4, n // 2 + 1))).astype(np.float32)
    opts, wkw, filt = {
        "discrete_rl": (dict(filter="hann", cutoff=0.45, filter_source="discrete_rl", extract_mode="magnitude"),
                        dict(source=1, extract=1), "hann"),
        "spatial_ramp": (dict(filter="hamming", gain=1.2, filter_source="spatial_ramp", spatial_ramp=taps),
                         dict(source=2, ramp=taps.astype(np.float64), gain=float(np.float32(1.2))), "hamming"),
        "custom": (dict(filter="custom", gain=0.9, custom_weights=custom),
                   dict(custom=custom.astype(np.float64), gain=float(np.float32(0.9))), "custom"),
    }[source]
    cutoff = opts.get("cutoff", 0.5)
    plan = yk.FdkPlan.from_geometry(ctx, g, parker="disabled", **opts)
    assert plan.padded_n == n
    proj = seeded_projection(5, g.total_views, g.nv, g.nu)
    ref, _ = oracle_filtered(oracle, plan, g, proj, filt=filt, cutoff=cutoff, **wkw)
    out, _ = plan.filter(torch.from_numpy(proj).cuda())
    got = out.cpu().numpy()
    print(f"filter source {source}: FP32 rel-RMSE {rel_rmse(got, ref):.2e}")
    assert rel_rmse(got, ref) < 2e-6, rel_rmse(got, ref)
    assert max_abs(got, ref) < 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("nu", [256, 512, 1024, 1536])
@pytest.mark.parametrize("case", ["full", "short-offset"])
def test_filter_stage_lengths_match_oracle(ctx, oracle, nu, case):
    """The filter kernel's compile-time lengths N = 512, 1024, 2048 and 4096
    (NPT = 2, 4, 8 and 16 points per lane; the register first pass from
    N = 2048), with FP32 and FP16 outputs, against the oracle's FP64 rows.
    "short-offset": a 234-degree short scan with Parker weights, a 2.5-pixel
    detector offset (the row is not centred in its padded length) and an odd
    row count."""
    kw = dict(channels=nu, rows=7, channel_mm=0.1, row_mm=0.1, volume=(24, 20, 18),
              voxel_mm=(0.05, 0.05, 0.05))
    if case == "full":
        kw.update(total_views=3, views_per_turn=24)
        parker = None
    else:
        kw.update(total_views=13, views_per_turn=20, start_angle_rad=0.3, offset_unv_mm=(0.25, 0.0, 0.0))
        parker = (0.3, 1)
    g = small_geometry(**kw)
    plan = yk.FdkPlan.from_geometry(ctx, g, filter="hann", parker="auto" if parker else "disabled")
    assert plan.padded_n == {256: 512, 512: 1024, 1024: 2048, 1536: 4096}[nu]
    proj = seeded_projection(11, g.total_views, g.nv, g.nu)
    ref, _ = oracle_filtered(oracle, plan, g, proj, filt="hann", parker=parker)
    dproj = torch.from_numpy(proj).cuda()
    out32, _ = plan.filter(dproj, storage="f32")
    got = out32.cpu().numpy()
    assert rel_rmse(got, ref) < 2e-6, rel_rmse(got, ref)
    assert max_abs(got, ref) < 1e-5 * np.abs(ref).max()
    out16, scales = plan.filter(dproj, storage="f16")
    got16 = out16.float().cpu().numpy() * scales.cpu().numpy()[:, :, None]
    # rows a Parker weight zeroes (or nearly: FP32 and FP64 angles differ at the
    # ends of the short scan) are measured against the global peak
    row_peak = np.maximum(np.abs(ref).max(axis=2), 1e-3 * np.abs(ref).max())
    row_err = np.abs(got16 - ref).max(axis=2) / row_peak
    print(f"filter nu={nu} ({case}): FP32 rel-RMSE {rel_rmse(got, ref):.2e}, FP16 rel-RMSE "
          f"{rel_rmse(got16, ref):.2e}, worst FP16 row error / row peak {row_err.max():.2e}")
    assert rel_rmse(got16, ref) < 3e-4
    assert row_err.max() < 2.0 ** -10
    plan.close()


@pytest.mark.parametrize("kw", [
    dict(),
    dict(offset_unv_mm=(2.5, 0.0, -1.5), source_offset_mm=(0.0, 0.0, 2.0), center_mm=(1.0, -2.0, 0.5)),
    dict(volume=(40, 33, 70), voxel_mm=(0.6, 0.6, 0.5)),
])
def test_backprojection_matches_oracle(ctx, oracle, kw):
    g = small_geometry(**kw)
    plan = yk.FdkPlan.from_geometry(ctx, g)
    proj = seeded_projection(2, g.total_views, g.nv, g.nu)
    ref_f, coeffs = oracle_filtered(oracle, plan, g, proj)
    filt32 = ref_f.astype(np.float32)
    vol = plan.backproject(torch.from_numpy(filt32).cuda()).cpu().numpy()
    ref = oracle.bp_fdk(coeffs, filt32.astype(np.float64), oracle_volume(oracle, g))
    assert rel_rmse(vol, ref) < 1e-5, rel_rmse(vol, ref)
    assert max_abs(vol, ref) < 2e-5 * np.abs(ref).max()


def test_backprojection_accumulate_and_view_ranges(ctx, oracle):
    g = small_geometry(total_views=30, views_per_turn=30)
    plan = yk.FdkPlan.from_geometry(ctx, g)
    proj = torch.from_numpy(seeded_projection(3, g.total_views, g.nv, g.nu)).cuda()
    filt, _ = plan.filter(proj)
    full = plan.backproject(filt)
    part = plan.backproject(filt[:11].contiguous(), first=0)
    part = plan.backproject(filt[11:].contiguous(), first=11, volume=part, accumulate=True)
    assert max_abs(part.cpu().numpy(), full.cpu().numpy()) < 1e-5 * float(full.abs().max())


@pytest.mark.parametrize("kw", [
    dict(),
    # tilted detectors: the depth denominator changes by 2.9e-4 per slice
    # (a reciprocal per sample) and by ~4e-5 (Newton steps)
    dict(tilt_rad=(0.03, -0.02, 0.04), offset_unv_mm=(1.0, 0.0, 2.0)),
    dict(tilt_rad=(0.004, -0.003, 0.005), offset_unv_mm=(1.0, 0.0, 2.0)),
])
@pytest.mark.parametrize("storage,tol", [("f32", 2e-5), ("f16", 1e-3)])
def test_fdk_end_to_end_matches_oracle(ctx, oracle, kw, storage, tol):
    """FP32, and FP16 + Kahan (white-noise input: the worst case for half
    storage, 5.5e-4 measured), against the oracle's FDK."""
    g = small_geometry(**kw)
    plan = yk.FdkPlan.from_geometry(ctx, g, filter="hann", storage=storage, kahan=storage != "f32")
    proj = seeded_projection(4, g.total_views, g.nv, g.nu)
    vol = plan.reconstruct(torch.from_numpy(proj).cuda()).cpu().numpy()
    ov = oracle.views_from_float(plan.views)
    ref = oracle.fdk(ov, proj.astype(np.float64), oracle_volume(oracle, g), filt="hann")
    print(f"FDK {storage} {kw}: rel-RMSE {rel_rmse(vol, ref):.2e}")
    assert rel_rmse(vol, ref) < tol, rel_rmse(vol, ref)


def test_fdk_short_scan_parker_matches_oracle(ctx, oracle):
    # 200 degrees of a 360-view turn: Parker switched on by Auto
    g = small_geometry(total_views=200, views_per_turn=360, start_angle_rad=0.25)
    plan = yk.FdkPlan.from_geometry(ctx, g)
    proj = seeded_projection(5, g.total_views, g.nv, g.nu)
    out, _ = plan.filter(torch.from_numpy(proj).cuda())
    ref, _ = oracle_filtered(oracle, plan, g, proj, parker=(g.start_angle_rad, 1))
    assert rel_rmse(out.cpu().numpy(), ref) < 2e-6


def test_parker_rejects_insufficient_range(ctx):
    g = small_geometry(total_views=150, views_per_turn=360)  # 150 deg < 180 + fan
    with pytest.raises(yk.YkError) as e:
        yk.FdkPlan.from_geometry(ctx, g)
    assert e.value.status == 1


def shepp_logan_projection(ctx, g, kind="siddon"):
    grid = g.grid()
    ell = phantoms.shepp_logan(0.4 * g.volume[0] * g.voxel_mm[0])
    vol = ctx.phantom(ell, grid)
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, grid)
    return vol, P.forward(vol, kind=kind)


@pytest.mark.parametrize("storage,tol", [("f16", 1e-3), ("bf16", 1e-3)])
@pytest.mark.parametrize("kahan", [False, True])
@pytest.mark.parametrize("channels", [44, 48])
def test_reduced_precision_storage(ctx, oracle, storage, tol, kahan, channels):
    """44 channels: the backprojection applies the row factors while staging;
    48 (a multiple of 8): one pass normalizes the rows first.  Both 16-bit
    forms within the 1e-3 relative-RMSE bar of the FP64 oracle (measured:
    FP16 1.4e-4, BF16 4.7e-4: bfloat16 keeps 8 significand bits to half's
    11, so its rounding is ~8x coarser)."""
    g = small_geometry(total_views=90, views_per_turn=90, volume=(32, 32, 24), voxel_mm=(0.8, 0.8, 0.8),
                       channels=channels)
    _, proj = shepp_logan_projection(ctx, g, "joseph")
    plan32 = yk.FdkPlan.from_geometry(ctx, g)
    plan = yk.FdkPlan.from_geometry(ctx, g, storage=storage, kahan=kahan)
    v32 = plan32.reconstruct(proj).cpu().numpy()
    vlo = plan.reconstruct(proj).cpu().numpy()
    ov = oracle.views_from_float(plan.views)
    ref = oracle.fdk(ov, proj.cpu().numpy().astype(np.float64), oracle_volume(oracle, g))
    assert rel_rmse(v32, ref) < 2e-5
    err = rel_rmse(vlo, ref)
    print(f"32x32x24, 90 views, {channels} channels, {storage}{'+Kahan' if kahan else ''}: rel-RMSE {err:.3e}")
    assert err < tol, (storage, err)


def test_fp32_kahan_matches_plain(ctx):
    g = small_geometry(total_views=64, views_per_turn=64)
    proj = torch.from_numpy(seeded_projection(6, g.total_views, g.nv, g.nu)).cuda()
    a = yk.FdkPlan.from_geometry(ctx, g).reconstruct(proj)
    b = yk.FdkPlan.from_geometry(ctx, g, kahan=True).reconstruct(proj)
    assert rel_rmse(b.cpu().numpy(), a.cpu().numpy()) < 1e-6


def test_streaming_batches_match_full(ctx):
    """fdk/batch-consistency and the DLL 47-view streaming check
    (Yktest/test_pipeline_smoke.cpp:2077-2120; dlltest/test_shepp_logan.cpp:118-158)."""
    g = small_geometry(total_views=120, views_per_turn=120)
    plan = yk.FdkPlan.from_geometry(ctx, g)
    proj = seeded_projection(7, g.total_views, g.nv, g.nu)
    full = np.zeros(plan.volume_shape, dtype=np.float32)
    plan.execute(proj, full, view_offset=0, clear_output=True)
    assert plan.complete
    plan.reset()
    streamed = np.zeros_like(full)
    for off in range(0, g.total_views, 47):
        cnt = min(47, g.total_views - off)
        plan.execute(np.ascontiguousarray(proj[off:off + cnt]), streamed, view_offset=off,
                     clear_output=(off == 0))
    assert max_abs(streamed, full) < 1e-5
    # device projection + device volume through the same contract
    plan.reset()
    dvol = torch.zeros(plan.volume_shape, device="cuda")
    plan.execute(torch.from_numpy(proj).cuda(), dvol, 0, True)
    assert max_abs(dvol.cpu().numpy(), full) < 1e-6 * np.abs(full).max()


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_staged_host_chunks_match_device_path(ctx, storage):
    """Host projections move through the two pinned staging slots in chunks
    of batch_views views (here 7: 18 chunks, each slot reused 9 times, the
    copy stream waiting on the filter of the chunk two back); a host volume
    comes back through the same slots.  Chunked backprojection adds the
    chunks' sums to the volume one after another, so the result equals the
    device-resident single batch up to FP32 summation order."""
    g = small_geometry(total_views=120, views_per_turn=120)
    proj = seeded_projection(11, g.total_views, g.nv, g.nu)
    ref_plan = yk.FdkPlan.from_geometry(ctx, g, storage=storage)
    ref = torch.zeros(ref_plan.volume_shape, device="cuda")
    ref_plan.execute(torch.from_numpy(proj).cuda(), ref, 0, True)
    ref = ref.cpu().numpy()
    ref_plan.close()
    plan = yk.FdkPlan.from_geometry(ctx, g, storage=storage, batch_views=7)
    host_vol = np.full(plan.volume_shape, np.nan, dtype=np.float32)
    plan.execute(proj, host_vol, view_offset=0, clear_output=True)
    assert plan.complete
    peak = np.abs(ref).max()
    assert max_abs(host_vol, ref) < 1e-5 * peak
    # streamed in uneven host batches into a device volume, twice (slots and
    # events carry over between calls and plans' resets)
    for _ in range(2):
        plan.reset()
        dvol = torch.zeros(plan.volume_shape, device="cuda")
        for off, cnt in ((0, 50), (50, 3), (53, 67)):
            plan.execute(np.ascontiguousarray(proj[off:off + cnt]), dvol, view_offset=off, clear_output=(off == 0))
        assert max_abs(dvol.cpu().numpy(), ref) < 1e-5 * peak
    plan.close()


def test_streaming_state_errors(ctx):
    g = small_geometry(total_views=20, views_per_turn=20)
    plan = yk.FdkPlan.from_geometry(ctx, g)
    proj = seeded_projection(8, 10, g.nv, g.nu)
    vol = np.zeros(plan.volume_shape, dtype=np.float32)
    with pytest.raises(yk.YkError) as e:
        plan.execute(proj, vol, view_offset=5, clear_output=True)
    assert e.value.status == 6
    plan.execute(proj, vol, view_offset=0, clear_output=True)
    with pytest.raises(yk.YkError) as e:
        plan.execute(proj, vol, view_offset=10, clear_output=True)
    assert e.value.status == 6
    with pytest.raises(yk.YkError) as e:
        plan.execute(proj[:1], np.zeros(10, dtype=np.float32), view_offset=10, clear_output=False)
    assert e.value.status == 5


def test_reference_sphere_example(ctx):
    """doc/public-api.md:46-95: r = 10 mm, mu = 0.02 sphere, 64^3 @ 0.75 mm,
    128^2 @ 0.8 mm, 360 views, SID 200 / SDD 400, Ram-Lak, Siddon FP; the
    reference records centre = 0.0196773 (doc/testing-and-validation.md:157)."""
    g = yk.ScanGeometry(360, 360, 200.0, 400.0, 128, 128, 0.8, 0.8, volume=(64, 64, 64),
                        voxel_mm=(0.75, 0.75, 0.75))
    grid = g.grid()
    vol = ctx.phantom(phantoms.sphere(10.0, 0.02), grid)
    # the example uses the strict '< 100' test on voxel centres; no centre lies on r = 10
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, grid)
    proj = P.forward(vol, kind="siddon")
    rec = yk.FdkPlan.from_geometry(ctx, g).reconstruct(proj).cpu().numpy()
    centre = float(rec[32, 32, 32])
    assert abs(centre - 0.0196773) / 0.0196773 < 2e-3, centre


def test_reference_shepp_logan_quality(ctx):
    """dlltest/test_shepp_logan.cpp:32-158 thresholds (cosine > 0.85, NRMSE < 0.65)."""
    g = yk.ScanGeometry(360, 360, 200.0, 400.0, 128, 128, 0.8, 0.8, volume=(64, 64, 64),
                        voxel_mm=(0.75, 0.75, 0.75))
    grid = g.grid()
    truth = ctx.phantom(phantoms.shepp_logan(20.0), grid)
    P = yk.Projector(ctx, g.views(), g.nu, g.nv, grid)
    proj = P.forward(truth, kind="siddon")
    rec = yk.FdkPlan.from_geometry(ctx, g).reconstruct(proj).cpu().numpy().astype(np.float64).ravel()
    t = truth.cpu().numpy().astype(np.float64).ravel()
    cos = t @ rec / np.sqrt((t @ t) * (rec @ rec))
    nrmse = np.linalg.norm(rec - t) / np.linalg.norm(t)
    assert cos > 0.85 and nrmse < 0.65, (cos, nrmse)


def test_single_channel_detector_is_refused(ctx):
    """The backprojection stages texel pairs of a row, so an FDK plan needs at
    least two detector channels; one channel is UnsupportedCombination."""
    g = small_geometry(channels=1)
    with pytest.raises(yk.YkError) as e:
        yk.FdkPlan.from_geometry(ctx, g)
    assert e.value.status == 3  # UnsupportedCombination

"""FDK backprojection throughput on one MI355X.

Workload (the reference's 512^3 x 720 configuration): 512^3 voxels at 0.25 mm,
SOD 600 mm, SDD 1100 mm, 720 views over 360 degrees, flat detector 1024 x 768
at 0.3 mm with a 2.5 px (0.75 mm) horizontal offset, Hann-windowed ramp.  The
phantom is 64 random ellipsoids (semi-axes 5-40 mm, densities 0-0.05 / mm,
seed 2), voxelized and forward-projected on the GPU by the Joseph projector.

A step is one backprojection of all 720 views into the 512^3 volume from
filtered projections resident in HBM.  The headline value is the GUPS of the
FP16-stored path (half rows with per-row scales, FP32 + Kahan accumulation);
the FP32 path is timed beside it.  Accuracy of both is measured against the
FP64 oracle on 64 evenly spaced z-slices, whose oracle backprojection is also
the timed CPU baseline (about 1.2e10 updates, 10-30 s on the host cores).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
METRIC = "FDK backprojection GUPS at 512³×720; FP16-stored vs FP32 at ≤1e-3 rel-RMSE"
# MI355X FP32 vector peak; plain wave64 FP32 FMAs issue in 2 cycles on gfx950
# and reach it, as do packed half FMAs / dot2 (4 cycles, two FMAs per lane)
# (profiles/valu_instruction_costs_gfx950.txt)
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
FLOP_PER_UPDATE = {"f32": 25, "f16_kahan": 29}  # algorithmic FLOPs per voxel-view update


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle accuracy/CPU leg")
    ap.add_argument("--no-projectors", action="store_true",
                    help="skip the forward-projector / transpose timings")
    ap.add_argument("--size", type=int, default=512, help="volume edge (default 512)")
    ap.add_argument("--views", type=int, default=720)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one 1 GiB device copy (reference dispatch for rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE scaling)")
    return ap.parse_args()


def measured_traffic():
    """HBM bytes per FP16 backprojection launch at this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated);
    None when the summary is absent."""
    path = os.path.join(REPO, "profiles", "fdk_bp_512cube_720views_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["fdk_bp_f16_kahan"]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def measured_counters():
    """VALU figures of the backprojection kernels at this workload from the
    committed rocprofv3 SQ counter summary: VALU instructions per update and
    SQ_INSTS_VALU x 4 cycles / (1024 SIMDs x GRBM_GUI_ACTIVE per XCD), the
    busy fraction those counters imply when every VALU instruction takes 4
    cycles (FP32 FMA / add / mul take 2 on gfx950); None when the summary is
    absent."""
    path = os.path.join(REPO, "profiles", "fdk_bp_512cube_720views_counters.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def time_launches(fn, steps, warmup, stream):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn()
    end.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps


def projector_leg(yk, phantoms, ctx, stream):
    """Forward projectors and their transposes at the iterative workload
    (256^3 @ 0.5 mm, SOD 500 / SDD 1000, 360 views, 512 x 384 @ 0.6 mm,
    random ellipsoids seed 4): milliseconds per call with HIP events on the
    launch stream, and Joseph samples (rays x main-axis planes) per second."""
    import torch
    geom = yk.ScanGeometry(360, 360, 500.0, 1000.0, 512, 384, 0.6, 0.6, volume=(256, 256, 256),
                           voxel_mm=(0.5, 0.5, 0.5))
    grid = geom.grid()
    ell = phantoms.random_ellipsoids(4, 32, semi_axes=(5.0, 30.0), density=(0.0, 0.05),
                                     center_radius=25.0)
    vol = ctx.phantom(ell, grid, clamp=False)
    P = yk.Projector(ctx, geom.views(), geom.nu, geom.nv, grid)
    proj = torch.empty((360, geom.nv, geom.nu), device="cuda")
    back = torch.empty_like(vol)
    out = {}
    for kind in ("joseph", "siddon"):
        out[f"{kind}_fp_ms"] = round(time_launches(lambda: P.forward(vol, kind=kind, out=proj), 3, 1, stream), 3)
        out[f"{kind}_adjoint_ms"] = round(time_launches(lambda: P.adjoint(proj, kind=kind, out=back), 3, 1,
                                                        stream), 3)
    # the voxel-driven backprojections the iterative solvers use as B
    for kind in ("fdk", "fdk_pseudo_matched", "joseph_voxel"):
        out[f"{kind}_bp_ms"] = round(time_launches(lambda: P.adjoint(proj, kind=kind, out=back), 3, 1,
                                                   stream), 3)
    P.close()
    # the general (non-separable) kernels: the same scan with the detector
    # tilted by (2, -1, 3) mrad about U, V, N
    tgeom = yk.ScanGeometry(360, 360, 500.0, 1000.0, 512, 384, 0.6, 0.6, volume=(256, 256, 256),
                            voxel_mm=(0.5, 0.5, 0.5), tilt_rad=(0.002, -0.001, 0.003))
    PT = yk.Projector(ctx, tgeom.views(), tgeom.nu, tgeom.nv, grid)
    for kind in ("joseph", "siddon"):
        out[f"tilted_{kind}_fp_ms"] = round(time_launches(lambda: PT.forward(vol, kind=kind, out=proj), 3, 1,
                                                          stream), 3)
        out[f"tilted_{kind}_adjoint_ms"] = round(time_launches(lambda: PT.adjoint(proj, kind=kind, out=back), 2, 1,
                                                               stream), 3)
    PT.close()
    samples = 360 * geom.nu * geom.nv * 256
    out["joseph_fp_gsamples_per_s"] = round(samples / (out["joseph_fp_ms"] * 1e-3) / 1e9, 1)
    out["workload"] = "256^3 @ 0.5 mm, 360 views, 512x384 @ 0.6 mm, SOD 500 / SDD 1000"
    return out


def oracle_check(views32, proj_cpu, geom, vols, args):
    """FP64 oracle FDK on 64 evenly spaced z-slices; returns accuracy per
    volume and the timed oracle backprojection (CPU baseline)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import yk_oracle as O
    ov = O.views_from_float(views32)
    nu, nv = geom.channels, geom.rows
    d = O.derive(ov, nu, nv, geom.scan_range)
    c = O.coeffs(ov, d)
    q = O.preweight(ov, d, proj_cpu.astype(np.float64))
    n = O.padded_n(nu)
    w = O.filter_weights("hann", n, float(np.float32(d[0, O.D_DU])), cutoff=0.5)
    su = [O.start_u(nu, n, d[i, O.D_OFFU]) for i in range(len(ov))]
    f = O.filter_rows(q, w, n, su, method=1)
    del q
    nx, ny, nz = geom.volume
    slices = np.linspace(0, nz - 1, 64).round().astype(np.int64)
    idx = (slices[:, None] * (nx * ny) + np.arange(nx * ny, dtype=np.int64)[None, :]).ravel()
    vol = O.volume(nx, ny, nz, *geom.voxel_mm, center=geom.center_mm)
    t0 = time.perf_counter()
    ref = O.bp_fdk(c, f, vol, idx)
    cpu_s = time.perf_counter() - t0
    acc = {}
    for name, v in vols.items():
        got = v.reshape(-1)[idx].astype(np.float64)
        acc[name] = {"rel_rmse": float(np.linalg.norm(got - ref) / np.linalg.norm(ref)),
                     "max_abs": float(np.abs(got - ref).max()),
                     "ref_max_abs": float(np.abs(ref).max())}
    updates = idx.size * len(ov)
    cpu = {"value": updates / cpu_s / 1e9, "unit": "GUPS", "cores": O.num_threads(),
           "kind": "port",
           "sample": f"FP64 oracle backprojection of {len(slices)} z-slices "
                     f"({idx.size} voxels) x {len(ov)} views = {updates:.3e} updates "
                     f"in {cpu_s:.1f} s"}
    return acc, cpu


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)
    import torch
    sys.path.insert(0, REPO)
    import fdk_cuda_impl_amd as yk
    from fdk_cuda_impl_amd import phantoms

    torch.cuda.set_device(0)
    if args.copy_calibration:
        src = torch.ones(256 * 1024 * 1024, device="cuda")
        torch.empty_like(src).copy_(src)
        torch.cuda.synchronize()
        del src
    s = args.size
    vox = 0.25 * 512 / s
    geom = yk.ScanGeometry(args.views, args.views, 600.0, 1100.0, 1024, 768, 0.3, 0.3,
                           volume=(s, s, s), voxel_mm=(vox, vox, vox), offset_unv_mm=(0.75, 0.0, 0.0))
    ctx = yk.Context(0)
    stream = torch.cuda.current_stream()
    grid = geom.grid()
    views = geom.views()
    ell = phantoms.random_ellipsoids(2, 64, semi_axes=(5.0, 40.0), density=(0.0, 0.05),
                                     center_radius=50.0)
    truth = ctx.phantom(ell, grid, clamp=False)
    P = yk.Projector(ctx, views, geom.nu, geom.nv, grid)
    t0 = time.perf_counter()
    proj = P.forward(truth, kind="joseph")
    torch.cuda.synchronize()
    fp_s = time.perf_counter() - t0
    del truth
    print(f"[bench] projections ready ({fp_s:.2f} s Joseph FP)", file=sys.stderr, flush=True)

    plan32 = yk.FdkPlan.from_geometry(ctx, geom, filter="hann")
    plan16 = yk.FdkPlan.from_geometry(ctx, geom, filter="hann", storage="f16", kahan=True)
    ctx.set_synchronous(False)
    f32, _ = plan32.filter(proj)
    f16, s16 = plan16.filter(proj)
    v32 = torch.empty(plan32.volume_shape, device="cuda")
    v16 = torch.empty(plan16.volume_shape, device="cuda")
    filt_ms = time_launches(lambda: plan16.filter(proj, out=f16, scales=s16), 3, 1, stream)
    filt32_ms = time_launches(lambda: plan32.filter(proj, out=f32), 3, 1, stream)
    ms32 = time_launches(lambda: plan32.backproject(f32, None, volume=v32), args.steps, args.warmup, stream)
    print(f"[bench] fp32 backprojection {ms32:.2f} ms", file=sys.stderr, flush=True)
    ms16 = time_launches(lambda: plan16.backproject(f16, s16, volume=v16), args.steps, args.warmup, stream)
    print(f"[bench] fp16 backprojection {ms16:.2f} ms", file=sys.stderr, flush=True)
    # the general (per-sample perspective) path: the same scan with a detector
    # tilted by a few mrad, as a calibrated scanner reports it
    tgeom = yk.ScanGeometry(args.views, args.views, 600.0, 1100.0, 1024, 768, 0.3, 0.3,
                            volume=(s, s, s), voxel_mm=(vox, vox, vox), offset_unv_mm=(0.75, 0.0, 0.0),
                            tilt_rad=(0.002, -0.001, 0.003))
    general = {"workload": "as above, detector tilted (0.002, -0.001, 0.003) rad about U, V, N"}
    vt = torch.empty_like(v16)
    for storage, kahan in (("f32", False), ("f16", True)):
        tplan = yk.FdkPlan.from_geometry(ctx, tgeom, filter="hann", storage=storage, kahan=kahan)
        tf, ts = tplan.filter(proj)
        general[f"{storage}{'_kahan' if kahan else ''}_ms"] = round(
            time_launches(lambda: tplan.backproject(tf, ts, volume=vt), 3, 1, stream), 3)
        tplan.close()
        del tf, ts
    del vt
    print(f"[bench] general path {general}", file=sys.stderr, flush=True)
    # Host -> device end to end: the whole FDK (filter + backprojection of all
    # views) from pageable host projections into a device volume through
    # yk_fdk_execute, the path Session::Execute takes (pinned staging, H2D on a
    # copy stream overlapping the GPU work).  Best of two runs per storage.
    proj_cpu = proj.cpu().numpy()
    host_e2e = {"workload": "as above, projections in pageable host memory (%.2f GB FP32), volume on the device"
                            % (proj_cpu.nbytes / 1e9)}
    vh = torch.empty(plan32.volume_shape, device="cuda")
    for storage, kahan in (("f32", False), ("f16", True)):
        hplan = yk.FdkPlan.from_geometry(ctx, geom, filter="hann", storage=storage, kahan=kahan)
        best = float("inf")
        for _ in range(2):
            hplan.reset()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            hplan.execute(proj_cpu, vh, view_offset=0, clear_output=True)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        hplan.close()
        host_e2e[f"{storage}{'_kahan' if kahan else ''}_ms"] = round(best * 1e3, 2)
    host_e2e["pcie_gb_per_s"] = round(proj_cpu.nbytes / (host_e2e["f32_ms"] * 1e-3) / 1e9, 1)
    del vh
    print(f"[bench] host-to-device end to end {host_e2e}", file=sys.stderr, flush=True)
    proj_leg = None
    if not args.no_projectors:
        proj_leg = projector_leg(yk, phantoms, ctx, stream)
        print(f"[bench] projectors {proj_leg}", file=sys.stderr, flush=True)
    ctx.set_synchronous(True)

    nvox = s * s * s
    updates = nvox * args.views
    gups16 = updates / (ms16 * 1e-3) / 1e9
    gups32 = updates / (ms32 * 1e-3) / 1e9
    achieved = updates * FLOP_PER_UPDATE["f16_kahan"] / (ms16 * 1e-3) / 1e12
    alg_bytes16 = args.views * geom.nu * geom.nv * 2 + args.views * geom.nv * 4 + nvox * 4

    counters = measured_counters() or {}
    c16 = counters.get("fdk_bp_f16_kahan", {})
    roofline = {
        "bound": "valu", "achieved": round(achieved, 2), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": measured_traffic(),
        "valu_insts_per_update": c16.get("valu_insts_per_update"),
        "sq_valu_quadcycle_frac": c16.get("valu_busy"),
        "hbm_frac": round(alg_bytes16 / (ms16 * 1e-3) / 1e9 / PEAK_HBM_GBS, 5),
        "note": (f"VALU-bound kernel (no MFMA): {FLOP_PER_UPDATE['f16_kahan']} algorithmic FLOP per update "
                 f"against the FP32 vector peak, which plain FP32 FMAs and packed half FMAs both reach on "
                 f"gfx950; sq_valu_quadcycle_frac counts 4 cycles per VALU instruction (FP32 FMA / add / mul "
                 f"take 2); HBM {alg_bytes16 / (ms16 * 1e-3) / 1e9:.0f} GB/s algorithmic"),
    }
    result = {
        "metric": METRIC, "value": round(gups16, 2), "unit": "GUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms16, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f16",
        "data": "synthetic (64 seeded random ellipsoids, GPU Joseph projections)",
        "config": {"workload": f"FDK backprojection {s}^3 voxels @ {vox} mm x {args.views} views, "
                               f"1024x768 @ 0.3 mm detector (+2.5 px offset), SOD 600 / SDD 1100, Hann",
                   "storage": "fp16 rows + per-row fp32 scales; bilinear interpolation in fp16 (packed "
                              "half FMA lerps), weighting and accumulation in fp32 (dot2), Kahan "
                              "compensation across views",
                   "views_per_launch": args.views},
        "roofline": roofline,
        "fp32_gups": round(gups32, 2), "fp32_ms_per_step": round(ms32, 3),
        "fp16_speedup_vs_fp32": round(ms32 / ms16, 3),
        "filter_ms_fp16": round(filt_ms, 3),
        "filter_ms_fp32": round(filt32_ms, 3),
        "fdk_host_to_device_end_to_end": host_e2e,
        "joseph_fp_s": round(fp_s, 3),
        "projectors": proj_leg,
        "general_path_backprojection": general,
    }
    if not args.no_cpu:
        print("[bench] FP64 oracle check", file=sys.stderr, flush=True)
        acc, cpu = oracle_check(views, proj_cpu, geom,
                                {"f16": v16.cpu().numpy(), "f32": v32.cpu().numpy()}, args)
        result["accuracy_vs_fp64"] = acc
        result["cpu_baseline"] = cpu
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result))


if __name__ == "__main__":
    main()

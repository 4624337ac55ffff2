"""Shared small geometries and error metrics for the tests."""
import numpy as np


def rel_rmse(x, ref):
    x = np.asarray(x, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    return float(np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300))


def max_abs(x, ref):
    return float(np.max(np.abs(np.asarray(x, dtype=np.float64) - np.asarray(ref, dtype=np.float64))))


def small_geometry(**kw):
    """A small flat circular scan the FP64 oracle handles in well under a second."""
    from fdk_cuda_impl_amd import ScanGeometry
    args = dict(total_views=24, views_per_turn=24, sid_mm=120.0, sdd_mm=240.0, channels=44, rows=36,
                channel_mm=1.0, row_mm=1.0, volume=(24, 20, 18), voxel_mm=(1.0, 1.0, 1.0))
    args.update(kw)
    return ScanGeometry(**args)


def oracle_views(oracle, geom):
    """FP64 oracle view vectors built from the same scan description."""
    return oracle.build_views(geom.total_views, geom.views_per_turn, geom.sid_mm, geom.sdd_mm,
                              geom.channels, geom.rows, geom.channel_mm, geom.row_mm,
                              start_angle=geom.start_angle_rad, direction=geom.rotation_direction,
                              z_mm=geom.z_mm, source_offset=geom.source_offset_mm,
                              offset_unv=geom.offset_unv_mm, tilt=geom.tilt_rad)


def oracle_volume(oracle, geom):
    return oracle.volume(*geom.volume, *geom.voxel_mm, center=geom.center_mm)


def seeded_projection(seed, n, nv, nu):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, nv, nu)).astype(np.float32)

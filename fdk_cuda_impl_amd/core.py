"""Python host side over the C ABI: contexts, FDK plans and projection operators
working on device tensors (PyTorch supplies device memory and streams).

Names and argument meanings mirror the reference's FdkPipeline
(src/Reconstruction/Analytic/Circular/Flat/FDK/YkFdkPipeline.hpp) and
IForwardOperator / IBackOperator (src/common/YkProjectionOperators.hpp:24-40).
"""
from __future__ import annotations

import atexit
import ctypes
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _capi as C

VIEW_FLOATS = 20


# ---------------------------------------------------------------------------
# geometry
# ---------------------------------------------------------------------------
@dataclass
class ScanGeometry:
    """Flat circular scan (SRegularCircularScanSpec + SFlatDetectorSpec + SVolumeGridSpec)."""
    total_views: int
    views_per_turn: int
    sid_mm: float
    sdd_mm: float
    channels: int
    rows: int
    channel_mm: float
    row_mm: float
    volume: tuple = (64, 64, 64)
    voxel_mm: tuple = (1.0, 1.0, 1.0)
    center_mm: tuple = (0.0, 0.0, 0.0)
    start_angle_rad: float = 0.0
    rotation_direction: int = 1
    z_mm: float = 0.0
    source_offset_mm: tuple = (0.0, 0.0, 0.0)
    offset_unv_mm: tuple = (0.0, 0.0, 0.0)
    tilt_rad: tuple = (0.0, 0.0, 0.0)

    def system(self) -> C.System:
        s = C.System()
        sc = s.scan
        sc.total_views, sc.views_per_turn = self.total_views, self.views_per_turn
        sc.start_angle_rad, sc.rotation_direction = self.start_angle_rad, self.rotation_direction
        sc.sid_mm, sc.sdd_mm, sc.z_mm = self.sid_mm, self.sdd_mm, self.z_mm
        sc.source_offset_mm[:] = list(self.source_offset_mm)
        d = s.detector
        d.channels, d.rows = self.channels, self.rows
        d.channel_size_mm, d.row_size_mm = self.channel_mm, self.row_mm
        d.pose.offset_unv_mm[:] = list(self.offset_unv_mm)
        d.pose.tilt_u_rad, d.pose.tilt_v_rad, d.pose.tilt_n_rad = self.tilt_rad
        s.volume = self.grid()
        return s

    def grid(self) -> C.VolumeGrid:
        g = C.VolumeGrid()
        g.nx, g.ny, g.nz = self.volume
        g.voxel_x_mm, g.voxel_y_mm, g.voxel_z_mm = self.voxel_mm
        g.center_mm[:] = list(self.center_mm)
        return g

    @property
    def scan_range(self) -> float:
        return float(np.float32(2.0 * np.pi) * np.float32(self.total_views) / np.float32(self.views_per_turn))

    def views(self) -> np.ndarray:
        return build_views(self)

    @property
    def nu(self):
        return self.channels

    @property
    def nv(self):
        return self.rows


def build_views(geom: ScanGeometry) -> np.ndarray:
    """(n, 20) float32 cone vectors (src, detS, detU, detV, angle as float4)."""
    out = np.zeros((geom.total_views, VIEW_FLOATS), dtype=np.float32)
    n = ctypes.c_int(0)
    sysd = geom.system()
    C.call("yk_geometry_build", ctypes.byref(sysd), _views_ptr(out), geom.total_views, ctypes.byref(n))
    return out


def _views_ptr(views: np.ndarray):
    if views.dtype != np.float32 or not views.flags.c_contiguous or views.shape[-1] != VIEW_FLOATS:
        raise ValueError("views must be a C-contiguous float32 (n, 20) array")
    return ctypes.cast(views.ctypes.data, ctypes.POINTER(C.View))


def _diagnostics_dict(d) -> dict:
    out = {name: float(getattr(d, name)) for name in C._DIAG_FLOATS}
    out["approximate"] = bool(d.approximate)
    return out


def geometry_diagnostics(views: np.ndarray, nu: int, nv: int) -> dict:
    """Calibrated-geometry report of a view sequence (host only; raises
    YkError for geometry FDK cannot use, e.g. repeated adjacent angles)."""
    views = np.ascontiguousarray(views, dtype=np.float32)
    d = C.GeometryDiagnostics()
    d.struct_size = ctypes.sizeof(d)
    C.call("yk_geometry_diagnose", _views_ptr(views), views.shape[0], nu, nv, ctypes.byref(d))
    return _diagnostics_dict(d)


def grid_of(volume, voxel, center=(0.0, 0.0, 0.0)) -> C.VolumeGrid:
    g = C.VolumeGrid()
    g.nx, g.ny, g.nz = volume
    g.voxel_x_mm, g.voxel_y_mm, g.voxel_z_mm = voxel
    g.center_mm[:] = list(center)
    return g


def derive(views: np.ndarray, nu: int, nv: int, scan_range: float = 0.0) -> np.ndarray:
    views = np.ascontiguousarray(views, dtype=np.float32)
    out = np.zeros((views.shape[0], 20), dtype=np.float32)
    C.call("yk_geometry_derive", _views_ptr(views), views.shape[0], nu, nv, scan_range,
           out.ctypes.data_as(C._FP))
    return out


def fdk_coefficients(views: np.ndarray, nu: int, nv: int, scan_range: float = 0.0) -> np.ndarray:
    views = np.ascontiguousarray(views, dtype=np.float32)
    out = np.zeros((views.shape[0], 17), dtype=np.float32)
    C.call("yk_geometry_fdk_coefficients", _views_ptr(views), views.shape[0], nu, nv, scan_range,
           out.ctypes.data_as(C._FP))
    return out


def make_filter(kind="ramlak", cutoff=0.5, gain=1.0, butterworth_order=2.0, kaiser_beta=8.6,
                tukey_alpha=0.5, source="analytic", extract_mode="real", spatial_ramp=None,
                custom_weights=None) -> C.Filter:
    """yk_filter.  source: "analytic" | "discrete_rl" | "spatial_ramp" (taps in
    spatial_ramp); kind "custom" takes N/2 + 1 frequency weights.  The arrays
    are kept on the returned struct (attribute ``keep``) for the C call."""
    f = C.Filter()
    f.type = C.FILTERS[kind] if isinstance(kind, str) else int(kind)
    f.cutoff, f.gain = cutoff, gain
    f.butterworth_order, f.kaiser_beta, f.tukey_alpha = butterworth_order, kaiser_beta, tukey_alpha
    f.source = C.FILTER_SOURCES[source] if isinstance(source, str) else int(source)
    f.extract_mode = C.RAMP_EXTRACT[extract_mode] if isinstance(extract_mode, str) else int(extract_mode)
    keep = []
    if spatial_ramp is not None:
        a = np.ascontiguousarray(spatial_ramp, dtype=np.float32)
        keep.append(a)
        f.spatial_ramp, f.spatial_ramp_size = a.ctypes.data_as(C._FP), a.size
    if custom_weights is not None:
        a = np.ascontiguousarray(custom_weights, dtype=np.float32)
        keep.append(a)
        f.custom_weights, f.custom_weights_size = a.ctypes.data_as(C._FP), a.size
    f.keep = keep
    return f


def filter_weights(n: int, du: float, **kw) -> np.ndarray:
    f = make_filter(**kw)
    out = np.zeros(n // 2 + 1, dtype=np.float32)
    C.call("yk_filter_weights", ctypes.byref(f), n, du, out.ctypes.data_as(C._FP))
    return out


def padded_n(nu: int) -> int:
    return C.lib().yk_filter_padded_n(nu)


# ---------------------------------------------------------------------------
# context
# ---------------------------------------------------------------------------
def _torch():
    import torch
    return torch


# Live library objects, closed by an exit hook registered with the first of
# them (so it runs before exit hooks registered earlier, such as torch's),
# plans and projectors before their contexts: their handles are then never
# released by __del__ during interpreter teardown.
_live = weakref.WeakSet()
_exit_hook = False


def _close_all():
    for obj in sorted(list(_live), key=lambda o: isinstance(o, Context)):
        try:
            obj.close()
        except Exception:
            pass


def _track(obj):
    global _exit_hook
    _live.add(obj)
    if not _exit_hook:
        atexit.register(_close_all)
        _exit_hook = True


class Context:
    """One HIP stream on one device (yk_ctx).  By default the context adopts
    PyTorch's current stream of that device so library work is ordered with the
    caller's tensor work."""

    def __init__(self, device: int = 0, use_torch_stream: bool = True, synchronous: bool = True):
        self.device = device
        self.handle = ctypes.c_void_p()
        _track(self)
        s = _torch().cuda.current_stream(device) if use_torch_stream else None
        C.call("yk_ctx_create", device, ctypes.byref(self.handle))
        if s is not None:
            C.call("yk_ctx_set_stream", self.handle, ctypes.c_void_p(s.cuda_stream))
        C.call("yk_ctx_set_synchronous", self.handle, 1 if synchronous else 0)

    @property
    def stream(self) -> int:
        return C.lib().yk_ctx_stream(self.handle) or 0

    def set_synchronous(self, flag: bool):
        C.call("yk_ctx_set_synchronous", self.handle, 1 if flag else 0)

    def synchronize(self):
        C.call("yk_ctx_synchronize", self.handle)

    def close(self):
        if self.handle:
            C.call("yk_ctx_destroy", self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def dot(self, a, b) -> float:
        out = ctypes.c_double(0.0)
        C.call("yk_dot", self.handle, _ptr(a), _ptr(b), a.numel(), ctypes.byref(out))
        return out.value

    def tv(self, x, grid: C.VolumeGrid, eps: float, grad=None):
        """Smoothed TV value and gradient of a [z][y][x] device volume."""
        torch = _torch()
        if grad is None:
            grad = torch.empty_like(x)
        val = ctypes.c_double(0.0)
        C.call("yk_tv_gradient", self.handle, ctypes.byref(grid), _ptr(x), eps, _ptr(grad),
               ctypes.byref(val))
        return val.value, grad

    def phantom(self, ellipsoids: np.ndarray, grid: C.VolumeGrid, clamp=True, out=None):
        torch = _torch()
        e = np.ascontiguousarray(ellipsoids, dtype=np.float32).reshape(-1, 16)
        if out is None:
            out = torch.empty((grid.nz, grid.ny, grid.nx), dtype=torch.float32,
                              device=f"cuda:{self.device}")
        C.call("yk_phantom_ellipsoids", self.handle, e.ctypes.data_as(C._FP), e.shape[0],
               ctypes.byref(grid), 1 if clamp else 0, _ptr(out))
        return out


def _ptr(t):
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _check_cuda(t, dtype, shape=None, name="tensor"):
    torch = _torch()
    if not (isinstance(t, torch.Tensor) and t.is_cuda):
        raise TypeError(f"{name} must be a CUDA tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")


# ---------------------------------------------------------------------------
# FDK
# ---------------------------------------------------------------------------
@dataclass
class FdkOptions:
    filter: str = "ramlak"
    cutoff: float = 0.5
    gain: float = 1.0
    butterworth_order: float = 2.0
    kaiser_beta: float = 8.6
    tukey_alpha: float = 0.5
    filter_source: str = "analytic"
    extract_mode: str = "real"
    spatial_ramp: np.ndarray | None = None
    custom_weights: np.ndarray | None = None
    parker: str = "auto"
    scan_range_rad: float = 0.0
    start_angle_rad: float = 0.0
    rotation_direction: int = 1
    storage: str = "f32"
    kahan: bool = False
    batch_views: int = 0

    def desc(self) -> C.FdkDesc:
        d = C.FdkDesc()
        d.struct_size = ctypes.sizeof(C.FdkDesc)
        f = make_filter(self.filter, self.cutoff, self.gain, self.butterworth_order,
                        self.kaiser_beta, self.tukey_alpha, self.filter_source, self.extract_mode,
                        self.spatial_ramp, self.custom_weights)
        d.filter = f
        d.keep = f.keep  # the arrays d.filter points at
        d.parker_mode = {"auto": C.PARKER_AUTO, "disabled": C.PARKER_DISABLED,
                         "enabled": C.PARKER_ENABLED}[self.parker]
        d.scan_range_rad = self.scan_range_rad
        d.start_angle_rad = self.start_angle_rad
        d.rotation_direction = self.rotation_direction
        d.storage = C.STORAGE[self.storage]
        d.kahan = 1 if self.kahan else 0
        d.batch_views = self.batch_views
        return d


class FdkPlan:
    """FDK plan over a complete view sequence (FdkPipeline::prepareWithGeometry)."""

    def __init__(self, ctx: Context, views: np.ndarray, nu: int, nv: int, grid: C.VolumeGrid,
                 options: FdkOptions | None = None, system: C.System | None = None):
        self.ctx = ctx
        self.options = options or FdkOptions()
        self.handle = ctypes.c_void_p()
        _track(self)
        desc = self.options.desc()
        if system is not None:
            C.call("yk_fdk_create_system", ctx.handle, ctypes.byref(desc), ctypes.byref(system),
                   ctypes.byref(self.handle))
            views = build_views_from_system(system)
        else:
            views = np.ascontiguousarray(views, dtype=np.float32)
            C.call("yk_fdk_create", ctx.handle, ctypes.byref(desc), _views_ptr(views), views.shape[0],
                   nu, nv, ctypes.byref(grid), ctypes.byref(self.handle))
        self.views = views
        self.nu, self.nv = nu, nv
        self.grid = grid
        self.n_views = views.shape[0]

    @classmethod
    def from_geometry(cls, ctx: Context, geom: ScanGeometry, **opts) -> "FdkPlan":
        o = FdkOptions(**opts)
        return cls(ctx, None, geom.channels, geom.rows, geom.grid(), o, system=geom.system())

    @property
    def volume_shape(self):
        return (self.grid.nz, self.grid.ny, self.grid.nx)

    @property
    def padded_n(self) -> int:
        return C.lib().yk_fdk_padded_n(self.handle)

    def weights(self) -> np.ndarray:
        out = np.zeros(self.padded_n // 2 + 1, dtype=np.float32)
        C.call("yk_fdk_get_weights", self.handle, out.ctypes.data_as(C._FP))
        return out

    def derived(self) -> np.ndarray:
        out = np.zeros((self.n_views, 20), dtype=np.float32)
        C.call("yk_fdk_get_derived", self.handle, out.ctypes.data_as(C._FP))
        return out

    def coefficients(self) -> np.ndarray:
        out = np.zeros((self.n_views, 17), dtype=np.float32)
        C.call("yk_fdk_get_coefficients", self.handle, out.ctypes.data_as(C._FP))
        return out

    def geometry_diagnostics(self) -> dict:
        """How far the plan's per-view geometry is from the ideal circular scan
        (FdkGeometryDiagnostics, YkFdkPipeline.hpp:35-58)."""
        d = C.GeometryDiagnostics()
        d.struct_size = ctypes.sizeof(d)
        C.call("yk_fdk_geometry_diagnostics", self.handle, ctypes.byref(d))
        return _diagnostics_dict(d)

    def set_stage_capture(self, on: bool = True):
        """Keep the weighted (pre-filter) projections of each execute chunk."""
        C.call("yk_fdk_set_stage_capture", self.handle, 1 if on else 0)

    def last_stage(self) -> dict:
        """Device pointers of the last chunk yk_fdk_execute processed (input,
        weighted when captured, filtered, row scales; FdkStageView,
        YkFdkPipeline.hpp:176-191), valid until the next execute / reset."""
        st = C.FdkStage()
        C.call("yk_fdk_last_stage", self.handle, ctypes.byref(st))
        return {name: getattr(st, name) for name, _ in C.FdkStage._fields_}

    def filter(self, proj, first: int = 0, storage: str | None = None, out=None, scales=None):
        """Preweight [+ Parker] + ramp filter of device views [first, first + len(proj))."""
        torch = _torch()
        count = proj.shape[0]
        _check_cuda(proj, torch.float32, (count, self.nv, self.nu), "projection")
        st = C.STORAGE[storage or self.options.storage]
        dt = {C.F32: torch.float32, C.F16: torch.float16, C.BF16: torch.bfloat16}[st]
        if out is None:
            out = torch.empty((count, self.nv, self.nu), dtype=dt, device=proj.device)
        if st != C.F32 and scales is None:
            scales = torch.empty((count, self.nv), dtype=torch.float32, device=proj.device)
        C.call("yk_fdk_filter", self.handle, _ptr(proj), first, count, _ptr(out), st,
               _ptr(scales) if st != C.F32 else None)
        return out, (scales if st != C.F32 else None)

    def backproject(self, filtered, scales=None, first: int = 0, volume=None, accumulate=False,
                    no_alloc=False):
        """Backprojection of filtered views [first, first + len(filtered)).
        no_alloc: the allocation-free entry point (yk_fdk_backproject_noalloc)."""
        torch = _torch()
        st = {torch.float32: C.F32, torch.float16: C.F16, torch.bfloat16: C.BF16}[filtered.dtype]
        count = filtered.shape[0]
        if volume is None:
            volume = torch.empty(self.volume_shape, dtype=torch.float32, device=filtered.device)
            accumulate = False
        _check_cuda(volume, torch.float32, self.volume_shape, "volume")
        C.call("yk_fdk_backproject_noalloc" if no_alloc else "yk_fdk_backproject", self.handle,
               _ptr(filtered), st, _ptr(scales), first, count, _ptr(volume), 1 if accumulate else 0)
        return volume

    def execute(self, projection, volume, view_offset=0, clear_output=True):
        """Streaming batch with the Session contract; numpy arrays are Host buffers,
        CUDA tensors Device buffers."""
        torch = _torch()

        def loc_ptr(a):
            if isinstance(a, np.ndarray):
                if a.dtype != np.float32 or not a.flags.c_contiguous:
                    raise ValueError("host buffers must be contiguous float32")
                return C.HOST, ctypes.c_void_p(a.ctypes.data), a.size
            if isinstance(a, torch.Tensor) and a.is_cuda and a.dtype == torch.float32:
                return C.DEVICE, _ptr(a), a.numel()
            raise TypeError("buffers must be float32 numpy arrays or CUDA tensors")

        pl, pp, pc = loc_ptr(projection)
        vl, vp, vc = loc_ptr(volume)
        count = pc // (self.nu * self.nv)
        C.call("yk_fdk_execute", self.handle, pp, pl, pc, view_offset, count, vp, vl, vc,
               1 if clear_output else 0)
        return volume

    def reset(self):
        C.call("yk_fdk_reset", self.handle)

    @property
    def received_views(self) -> int:
        return C.lib().yk_fdk_received_views(self.handle)

    @property
    def complete(self) -> bool:
        return bool(C.lib().yk_fdk_complete(self.handle))

    def reconstruct(self, proj):
        """Filter + backproject a complete device projection stack; returns the volume."""
        filt, scales = self.filter(proj)
        return self.backproject(filt, scales)

    def close(self):
        if self.handle:
            C.call("yk_fdk_destroy", self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_views_from_system(system: C.System) -> np.ndarray:
    n = system.scan.total_views
    out = np.zeros((n, VIEW_FLOATS), dtype=np.float32)
    cnt = ctypes.c_int(0)
    C.call("yk_geometry_build", ctypes.byref(system), _views_ptr(out), n, ctypes.byref(cnt))
    return out


# ---------------------------------------------------------------------------
# projection operators
# ---------------------------------------------------------------------------
FORWARD = {"joseph": C.FP_JOSEPH, "siddon": C.FP_SIDDON}
BACK = {"joseph": C.BP_JOSEPH_ADJOINT, "siddon": C.BP_SIDDON_ADJOINT,
        "fdk": C.BP_FDK, "fdk_pseudo_matched": C.BP_FDK_PSEUDO_MATCHED, "joseph_voxel": C.BP_JOSEPH_VOXEL}
ALGEBRAIC = {"sirt": 0, "sart": 1, "ossart": 2}  # yk_algebraic_method
WEIGHT_MODELS = {"detailed": 0, "tigre": 1}     # yk_weight_model
SUBSET_ORDERS = {"sequential": 0, "golden": 1}  # yk_subset_order
CGLS_STRATEGIES = {"robust": 0, "astra": 1}     # yk_cgls_strategy


class Projector:
    """Flat forward projectors and their exact transposes (IForwardOperator /
    IBackOperator) over a fixed view sequence."""

    def __init__(self, ctx: Context, views: np.ndarray, nu: int, nv: int, grid: C.VolumeGrid,
                 scan_range: float = 0.0):
        self.ctx = ctx
        self.views = np.ascontiguousarray(views, dtype=np.float32)
        self.nu, self.nv, self.grid = nu, nv, grid
        self.n_views = self.views.shape[0]
        self.handle = ctypes.c_void_p()
        _track(self)
        C.call("yk_projector_create", ctx.handle, _views_ptr(self.views), self.n_views, nu, nv,
               ctypes.byref(grid), scan_range, ctypes.byref(self.handle))

    @property
    def volume_shape(self):
        return (self.grid.nz, self.grid.ny, self.grid.nx)

    def forward(self, volume, kind="joseph", first=0, count=None, out=None, accumulate=False):
        torch = _torch()
        count = self.n_views - first if count is None else count
        _check_cuda(volume, torch.float32, self.volume_shape, "volume")
        if out is None:
            out = torch.empty((count, self.nv, self.nu), dtype=torch.float32, device=volume.device)
            accumulate = False
        _check_cuda(out, torch.float32, (count, self.nv, self.nu), "projection")
        C.call("yk_project", self.handle, FORWARD[kind], _ptr(volume), _ptr(out), first, count,
               1 if accumulate else 0)
        return out

    def adjoint(self, proj, kind="joseph", first=0, out=None, accumulate=False):
        torch = _torch()
        count = proj.shape[0]
        _check_cuda(proj, torch.float32, (count, self.nv, self.nu), "projection")
        if out is None:
            out = torch.empty(self.volume_shape, dtype=torch.float32, device=proj.device)
            accumulate = False
        _check_cuda(out, torch.float32, self.volume_shape, "volume")
        C.call("yk_backproject", self.handle, BACK[kind], _ptr(proj), _ptr(out), first, count,
               1 if accumulate else 0)
        return out

    def cgls(self, b, x=None, iterations=10, forward="siddon", back="siddon"):
        torch = _torch()
        _check_cuda(b, torch.float32, (self.n_views, self.nv, self.nu), "projection")
        if x is None:
            x = torch.zeros(self.volume_shape, dtype=torch.float32, device=b.device)
        res = np.zeros(iterations + 1, dtype=np.float64)
        C.call("yk_cgls", self.handle, FORWARD[forward], BACK[back], _ptr(b), _ptr(x), iterations,
               res.ctypes.data_as(C._DP))
        return x, res

    def cgls_ex(self, b, x=None, iterations=50, strategy="robust", epsilon=1e-8, restart=True,
                min_value=None, max_value=None, forward="joseph", back="joseph_voxel",
                relative_residual=0.0, relative_update=0.0, relative_improvement=0.0,
                minimum_iterations=1, check_interval=1, patience=1):
        """CGLS with the reference's strategies: "robust" (true residual each
        step, rollback and one restart on divergence, convergence monitor) or
        "astra" (plain recursion); defaults follow CglsReconstructionConfig
        (YkCglsReconstructorEx.hpp:15-27).  Returns (x, residual norms, stats)."""
        torch = _torch()
        _check_cuda(b, torch.float32, (self.n_views, self.nv, self.nu), "projection")
        if x is None:
            x = torch.zeros(self.volume_shape, dtype=torch.float32, device=b.device)
        _check_cuda(x, torch.float32, self.volume_shape, "volume")
        d = C.CglsDesc()
        d.struct_size = ctypes.sizeof(C.CglsDesc)
        d.strategy = CGLS_STRATEGIES[strategy]
        d.iterations = iterations
        d.epsilon = epsilon
        d.restart_on_divergence = 1 if restart else 0
        d.use_min = 0 if min_value is None else 1
        d.min_value = 0.0 if min_value is None else min_value
        d.use_max = 0 if max_value is None else 1
        d.max_value = 0.0 if max_value is None else max_value
        d.forward_kind = FORWARD[forward]
        d.back_kind = BACK[back]
        d.relative_residual_tolerance = relative_residual
        d.relative_update_tolerance = relative_update
        d.relative_improvement_tolerance = relative_improvement
        d.minimum_iterations = minimum_iterations
        d.check_interval = check_interval
        d.patience = patience
        st = C.IterativeStats()
        st.struct_size = ctypes.sizeof(C.IterativeStats)
        res = np.zeros(max(iterations, 0) + 1, dtype=np.float64)
        C.call("yk_cgls_ex", self.handle, ctypes.byref(d), _ptr(b), _ptr(x), ctypes.byref(st),
               res.ctypes.data_as(C._DP))
        stats = {name: getattr(st, name) for name, _ in C.IterativeStats._fields_ if name != "struct_size"}
        return x, res, stats

    def sart(self, b, x=None, iterations=1, relaxation=1.0, nonnegative=False, forward="joseph",
             back="joseph"):
        torch = _torch()
        _check_cuda(b, torch.float32, (self.n_views, self.nv, self.nu), "projection")
        if x is None:
            x = torch.zeros(self.volume_shape, dtype=torch.float32, device=b.device)
        C.call("yk_sart", self.handle, FORWARD[forward], BACK[back], _ptr(b), _ptr(x), iterations,
               relaxation, 1 if nonnegative else 0)
        return x

    def algebraic(self, b, method="sart", x=None, iterations=10, relaxation=1.0, subsets=1,
                  weights=None, order=None, relaxation_reduction=1.0, epsilon=1e-6, min_value=None,
                  max_value=None, forward="joseph", back="joseph_voxel", tv_strength=0.0, tv_iterations=0,
                  tv_epsilon=1e-8, tv_strength_reduction=1.0, tv_dimensions=3, relative_residual=0.0,
                  relative_update=0.0,
                  relative_improvement=0.0, minimum_iterations=1, check_interval=1, patience=1,
                  return_stats=False):
        """SIRT / SART / OS-SART on the device, updating x (zeros when None) in place.

        ``weights`` is "detailed" (exact per-subset row and column sums) or
        "tigre" (TIGRE-style approximate row weights over contiguous view
        blocks); the defaults follow the reference's public entry points:
        SART uses "tigre", SIRT and OS-SART "detailed", OS-SART with the
        golden-ratio subset order, and the model pair defaults to Joseph with
        the voxel-driven Joseph backprojection (FP_Joseph / BP_Joseph_v3)
        (YkAlgebraicReconstructorEx.hpp:15-180, YkAlgebraicBackends.hpp:32-310,
        597-800).
        """
        torch = _torch()
        if method not in ALGEBRAIC:
            raise ValueError(f"method must be one of {sorted(ALGEBRAIC)}")
        _check_cuda(b, torch.float32, (self.n_views, self.nv, self.nu), "projection")
        if x is None:
            x = torch.zeros(self.volume_shape, dtype=torch.float32, device=b.device)
        _check_cuda(x, torch.float32, self.volume_shape, "volume")
        if weights is None:
            weights = "tigre" if method == "sart" else "detailed"
        if order is None:
            order = "golden" if method == "ossart" else "sequential"
        d = C.AlgebraicDesc()
        d.struct_size = ctypes.sizeof(C.AlgebraicDesc)
        d.method = ALGEBRAIC[method]
        d.weight_model = WEIGHT_MODELS[weights]
        d.subset_order = SUBSET_ORDERS[order]
        d.iterations = iterations
        d.subsets = subsets
        d.relaxation = relaxation
        d.relaxation_reduction = relaxation_reduction
        d.epsilon = epsilon
        d.use_min = 0 if min_value is None else 1
        d.min_value = 0.0 if min_value is None else min_value
        d.use_max = 0 if max_value is None else 1
        d.max_value = 0.0 if max_value is None else max_value
        d.forward_kind = FORWARD[forward]
        d.back_kind = BACK[back]
        d.tv_strength = tv_strength
        d.tv_iterations = tv_iterations
        d.tv_epsilon = tv_epsilon
        d.tv_strength_reduction = tv_strength_reduction
        d.tv_dimensions = tv_dimensions
        d.relative_residual_tolerance = relative_residual
        d.relative_update_tolerance = relative_update
        d.relative_improvement_tolerance = relative_improvement
        d.minimum_iterations = minimum_iterations
        d.check_interval = check_interval
        d.patience = patience
        st = C.IterativeStats()
        st.struct_size = ctypes.sizeof(C.IterativeStats)
        C.call("yk_algebraic", self.handle, ctypes.byref(d), _ptr(b), _ptr(x), ctypes.byref(st))
        if return_stats:
            return x, {name: getattr(st, name) for name, _ in C.IterativeStats._fields_ if name != "struct_size"}
        return x

    def close(self):
        if self.handle:
            C.call("yk_projector_destroy", self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

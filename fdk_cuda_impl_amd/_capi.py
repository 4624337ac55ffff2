"""ctypes binding of ``libyk_hip.so`` (the C ABI declared in include/yk_hip.h).

The library is loaded from this package directory.  Nothing here falls back to
another implementation: a missing or unloadable library raises immediately.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libyk_hip.so")

# yk_status (mirrors YK::EApiErrorCode)
OK, INVALID_CONFIG, INVALID_GEOMETRY, UNSUPPORTED, ALGORITHM_UNDER_TEST, INVALID_BUFFER, \
    STREAMING_STATE, DEVICE_ERROR = range(8)
STATUS_NAMES = {0: "None", 1: "InvalidConfig", 2: "InvalidGeometry", 3: "UnsupportedCombination",
                4: "AlgorithmUnderTest", 5: "InvalidBuffer", 6: "StreamingStateError",
                7: "DeviceError"}

HOST, DEVICE = 0, 1
F32, F16, BF16 = 0, 1, 2
FP_JOSEPH, FP_SIDDON = 1, 2
BP_JOSEPH_ADJOINT, BP_SIDDON_ADJOINT, BP_FDK, BP_FDK_PSEUDO_MATCHED, BP_JOSEPH_VOXEL = 8, 4, 6, 7, 10
PARKER_AUTO, PARKER_DISABLED, PARKER_ENABLED = 0, 1, 2
FILTERS = {"none": 0, "ramlak": 1, "shepplogan": 2, "cosine": 3, "hann": 4, "hamming": 5,
           "blackman": 6, "butterworth": 7, "kaiser": 8, "tukey": 9, "custom": 10}
# yk_filter_source / yk_ramp_extract
FILTER_SOURCES = {"analytic": 0, "discrete_rl": 1, "spatial_ramp": 2}
RAMP_EXTRACT = {"real": 0, "magnitude": 1}
# YK_NOALLOC_MAX_ROWS: row factors (views x detector rows) one call of
# yk_fdk_backproject_noalloc may use
NOALLOC_MAX_ROWS = 4194304
# yk_progress_fn: (user data, completed, total)
PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_int)
STORAGE = {"f32": F32, "float32": F32, "f16": F16, "float16": F16, "bf16": BF16, "bfloat16": BF16}


class YkError(RuntimeError):
    """A yk_status other than YK_OK, with the library's message."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


class View(ctypes.Structure):
    _fields_ = [("src", ctypes.c_float * 4), ("detS", ctypes.c_float * 4),
                ("detU", ctypes.c_float * 4), ("detV", ctypes.c_float * 4),
                ("angle", ctypes.c_float * 4)]


class CircularScan(ctypes.Structure):
    _fields_ = [("total_views", ctypes.c_int), ("views_per_turn", ctypes.c_int),
                ("start_angle_rad", ctypes.c_float), ("rotation_direction", ctypes.c_int),
                ("sid_mm", ctypes.c_float), ("sdd_mm", ctypes.c_float), ("z_mm", ctypes.c_float),
                ("source_offset_mm", ctypes.c_float * 3)]


class DetectorPose(ctypes.Structure):
    _fields_ = [("offset_unv_mm", ctypes.c_float * 3), ("tilt_u_rad", ctypes.c_float),
                ("tilt_v_rad", ctypes.c_float), ("tilt_n_rad", ctypes.c_float)]


class FlatDetector(ctypes.Structure):
    _fields_ = [("channels", ctypes.c_int), ("rows", ctypes.c_int),
                ("channel_size_mm", ctypes.c_float), ("row_size_mm", ctypes.c_float),
                ("pose", DetectorPose)]


class VolumeGrid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("voxel_x_mm", ctypes.c_float), ("voxel_y_mm", ctypes.c_float),
                ("voxel_z_mm", ctypes.c_float), ("center_mm", ctypes.c_float * 3)]


class System(ctypes.Structure):
    _fields_ = [("scan", CircularScan), ("detector", FlatDetector), ("volume", VolumeGrid)]


class Filter(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("cutoff", ctypes.c_float), ("gain", ctypes.c_float),
                ("butterworth_order", ctypes.c_float), ("kaiser_beta", ctypes.c_float),
                ("tukey_alpha", ctypes.c_float), ("source", ctypes.c_int), ("extract_mode", ctypes.c_int),
                ("spatial_ramp", ctypes.POINTER(ctypes.c_float)), ("spatial_ramp_size", ctypes.c_int),
                ("custom_weights", ctypes.POINTER(ctypes.c_float)), ("custom_weights_size", ctypes.c_int)]


class FdkDesc(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("filter", Filter), ("parker_mode", ctypes.c_int),
                ("scan_range_rad", ctypes.c_float), ("start_angle_rad", ctypes.c_float),
                ("rotation_direction", ctypes.c_int), ("storage", ctypes.c_int),
                ("kahan", ctypes.c_int), ("batch_views", ctypes.c_int)]


class AlgebraicDesc(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("method", ctypes.c_int),
                ("weight_model", ctypes.c_int), ("subset_order", ctypes.c_int),
                ("iterations", ctypes.c_int), ("subsets", ctypes.c_int),
                ("relaxation", ctypes.c_float), ("relaxation_reduction", ctypes.c_float),
                ("epsilon", ctypes.c_float), ("use_min", ctypes.c_int), ("min_value", ctypes.c_float),
                ("use_max", ctypes.c_int), ("max_value", ctypes.c_float),
                ("forward_kind", ctypes.c_int), ("back_kind", ctypes.c_int),
                ("tv_strength", ctypes.c_float), ("tv_iterations", ctypes.c_int),
                ("tv_epsilon", ctypes.c_float), ("tv_strength_reduction", ctypes.c_float),
                ("relative_residual_tolerance", ctypes.c_float),
                ("relative_update_tolerance", ctypes.c_float),
                ("relative_improvement_tolerance", ctypes.c_float),
                ("minimum_iterations", ctypes.c_int), ("check_interval", ctypes.c_int),
                ("patience", ctypes.c_int), ("tv_dimensions", ctypes.c_int),
                ("progress_callback", ctypes.c_void_p), ("progress_user_data", ctypes.c_void_p)]


class FdkStage(ctypes.Structure):
    _fields_ = [("first_view", ctypes.c_int), ("count", ctypes.c_int), ("d_input", ctypes.c_void_p),
                ("d_preweighted", ctypes.c_void_p), ("d_filtered", ctypes.c_void_p),
                ("filtered_storage", ctypes.c_int), ("d_row_scales", ctypes.c_void_p),
                ("stream", ctypes.c_void_p)]


_DIAG_FLOATS = ["min_source_to_axis_mm", "max_source_to_axis_mm", "min_source_to_detector_mm",
                "max_source_to_detector_mm", "min_source_to_plane_mm", "max_source_to_plane_mm",
                "max_abs_source_axis_mm", "max_detector_skew", "max_abs_principal_u_pix",
                "max_abs_principal_v_pix", "min_principal_u_pix", "max_principal_u_pix",
                "min_principal_v_pix", "max_principal_v_pix", "min_relative_tilt_u_rad",
                "max_relative_tilt_u_rad", "min_relative_tilt_v_rad", "max_relative_tilt_v_rad",
                "min_relative_tilt_n_rad", "max_relative_tilt_n_rad", "max_normal_misalignment"]


class GeometryDiagnostics(ctypes.Structure):
    _fields_ = ([("struct_size", ctypes.c_uint32)] + [(n, ctypes.c_float) for n in _DIAG_FLOATS]
                + [("approximate", ctypes.c_int)])


class CglsDesc(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("strategy", ctypes.c_int),
                ("iterations", ctypes.c_int), ("epsilon", ctypes.c_float),
                ("restart_on_divergence", ctypes.c_int), ("use_min", ctypes.c_int),
                ("min_value", ctypes.c_float), ("use_max", ctypes.c_int), ("max_value", ctypes.c_float),
                ("forward_kind", ctypes.c_int), ("back_kind", ctypes.c_int),
                ("relative_residual_tolerance", ctypes.c_float),
                ("relative_update_tolerance", ctypes.c_float),
                ("relative_improvement_tolerance", ctypes.c_float),
                ("minimum_iterations", ctypes.c_int), ("check_interval", ctypes.c_int),
                ("patience", ctypes.c_int),
                ("progress_callback", ctypes.c_void_p), ("progress_user_data", ctypes.c_void_p)]


class IterativeStats(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("completed_iterations", ctypes.c_int),
                ("convergence_checks", ctypes.c_int), ("stopped_by_breakdown", ctypes.c_int),
                ("stopped_by_divergence", ctypes.c_int), ("stopped_by_relative_residual", ctypes.c_int),
                ("stopped_by_relative_update", ctypes.c_int), ("stopped_by_stagnation", ctypes.c_int),
                ("projection_residual_l2", ctypes.c_double),
                ("relative_projection_residual", ctypes.c_double),
                ("relative_volume_update", ctypes.c_double),
                ("relative_residual_improvement", ctypes.c_double)]


_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_SZ = ctypes.c_size_t
_FP = ctypes.POINTER(ctypes.c_float)
_DP = ctypes.POINTER(ctypes.c_double)
_VP = ctypes.POINTER(View)

SIGNATURES = {
    "yk_abi_version": (_I, []),
    "yk_last_error": (ctypes.c_char_p, []),
    "yk_ctx_create": (_I, [_I, ctypes.POINTER(_P)]),
    "yk_ctx_destroy": (_I, [_P]),
    "yk_ctx_set_stream": (_I, [_P, _P]),
    "yk_ctx_reset_stream": (_I, [_P]),
    "yk_ctx_stream": (_P, [_P]),
    "yk_ctx_set_synchronous": (_I, [_P, _I]),
    "yk_ctx_synchronize": (_I, [_P]),
    "yk_geometry_build": (_I, [ctypes.POINTER(System), _VP, _I, ctypes.POINTER(_I)]),
    "yk_geometry_derive": (_I, [_VP, _I, _I, _I, _F, _FP]),
    "yk_geometry_fdk_coefficients": (_I, [_VP, _I, _I, _I, _F, _FP]),
    "yk_filter_weights": (_I, [ctypes.POINTER(Filter), _I, _F, _FP]),
    "yk_filter_padded_n": (_I, [_I]),
    "yk_fdk_create": (_I, [_P, ctypes.POINTER(FdkDesc), _VP, _I, _I, _I, ctypes.POINTER(VolumeGrid),
                           ctypes.POINTER(_P)]),
    "yk_fdk_create_system": (_I, [_P, ctypes.POINTER(FdkDesc), ctypes.POINTER(System),
                                  ctypes.POINTER(_P)]),
    "yk_fdk_destroy": (_I, [_P]),
    "yk_fdk_execute": (_I, [_P, _P, _I, _SZ, _I, _I, _P, _I, _SZ, _I]),
    "yk_fdk_reset": (_I, [_P]),
    "yk_fdk_received_views": (_I, [_P]),
    "yk_fdk_complete": (_I, [_P]),
    "yk_fdk_filter": (_I, [_P, _P, _I, _I, _P, _I, _P]),
    "yk_fdk_backproject": (_I, [_P, _P, _I, _P, _I, _I, _P, _I]),
    "yk_fdk_backproject_noalloc": (_I, [_P, _P, _I, _P, _I, _I, _P, _I]),
    "yk_fdk_padded_n": (_I, [_P]),
    "yk_fdk_get_weights": (_I, [_P, _FP]),
    "yk_fdk_get_derived": (_I, [_P, _FP]),
    "yk_fdk_get_coefficients": (_I, [_P, _FP]),
    "yk_projector_create": (_I, [_P, _VP, _I, _I, _I, ctypes.POINTER(VolumeGrid), _F,
                                 ctypes.POINTER(_P)]),
    "yk_projector_destroy": (_I, [_P]),
    "yk_project": (_I, [_P, _I, _P, _P, _I, _I, _I]),
    "yk_backproject": (_I, [_P, _I, _P, _P, _I, _I, _I]),
    "yk_cgls": (_I, [_P, _I, _I, _P, _P, _I, _DP]),
    "yk_cgls_ex": (_I, [_P, ctypes.POINTER(CglsDesc), _P, _P, ctypes.POINTER(IterativeStats), _DP]),
    "yk_fdk_set_stage_capture": (_I, [_P, _I]),
    "yk_fdk_last_stage": (_I, [_P, ctypes.POINTER(FdkStage)]),
    "yk_fdk_geometry_diagnostics": (_I, [_P, ctypes.POINTER(GeometryDiagnostics)]),
    "yk_geometry_diagnose": (_I, [_VP, _I, _I, _I, ctypes.POINTER(GeometryDiagnostics)]),
    "yk_dump_create": (_I, [ctypes.c_char_p, _I, ctypes.POINTER(_P)]),
    "yk_dump_destroy": (_I, [_P]),
    "yk_dump_set_enabled": (_I, [_P, _I]),
    "yk_dump_device_f32": (_I, [_P, ctypes.c_char_p, _P, _SZ, _P]),
    "yk_dump_device_f32_a": (_I, [_P, ctypes.c_char_p, _I, _P, _SZ, _P]),
    "yk_dump_text": (_I, [_P, ctypes.c_char_p, ctypes.c_char_p]),
    "yk_dump_text_a": (_I, [_P, ctypes.c_char_p, _I, ctypes.c_char_p]),
    "yk_sart": (_I, [_P, _I, _I, _P, _P, _I, _F, _I]),
    "yk_algebraic": (_I, [_P, ctypes.POINTER(AlgebraicDesc), _P, _P, ctypes.POINTER(IterativeStats)]),
    "yk_tv_gradient": (_I, [_P, ctypes.POINTER(VolumeGrid), _P, _F, _P, _DP]),
    "yk_dot": (_I, [_P, _P, _P, _SZ, _DP]),
    "yk_phantom_ellipsoids": (_I, [_P, _FP, _I, ctypes.POINTER(VolumeGrid), _I, _P]),
}

_lib = None


def lib():
    """Load libyk_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run fdk_cuda_impl_amd.build.build() first")
        # PyTorch ships its own HIP runtime: load it before this library so the
        # process holds one runtime (a HIP runtime initialized here first
        # leaves torch.cuda without devices)
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != OK:
        msg = lib().yk_last_error()
        raise YkError(status, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))

/*
 * yk_hip.h -- C ABI of the MI355X (gfx950) flat-panel circular-orbit
 * reconstruction path: FDK filter + backprojection, Joseph and Siddon forward
 * projection, their matched transposes, and the CGLS / SART / TV callers.
 *
 * This is the drop-in boundary behind the YKCBCT Session API
 * (include/YKCBCT/interface/YkReconstructionApi.hpp in the reference).
 * Every entry point is extern "C", takes plain pointers and sizes, returns a
 * yk_status code and never aborts.  Status values mirror YK::EApiErrorCode
 * (YkReconstructionApi.hpp:72-81) one for one; YK_ERR_DEVICE stands where the
 * reference says CudaError.
 *
 * Data conventions are the reference's, unchanged
 * (doc/geometry-and-data.md:39,74,105-110):
 *   projections [view][v][u] float32, volumes [z][y][x] float32,
 *   millimetres, attenuation in mm^-1, p = integral of mu dl;
 *   yk_view.detS is the centre of pixel (0,0), detU/detV one-pixel steps;
 *   voxel (i,j,k) centre = centre_mm + ((i,j,k) - (N-1)/2) * voxel_mm.
 *
 * Threading and lifetime (reference: doc/public-api.md:12-16,137,183-195):
 *   a yk_ctx owns one HIP stream (or adopts a caller stream) on one device and
 *   is not thread-safe; separate contexts may be used from separate threads.
 *   Handles (yk_fdk, yk_projector) are library-owned and must be destroyed
 *   through this API.  All buffer pointers are caller-owned and borrowed for
 *   the call.  Device pointers must belong to the context's device.  A call
 *   is synchronous on success unless yk_ctx_set_synchronous(ctx, 0) was used,
 *   in which case work is only enqueued on the context stream.
 */
#ifndef YK_HIP_H
#define YK_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YK_HIP_ABI_VERSION 1

/* ------------------------------------------------------------------------ */
/* status, enums                                                             */
/* ------------------------------------------------------------------------ */
typedef enum yk_status {
    YK_OK = 0,
    YK_ERR_INVALID_CONFIG = 1,
    YK_ERR_INVALID_GEOMETRY = 2,
    YK_ERR_UNSUPPORTED = 3,
    YK_ERR_ALGORITHM_UNDER_TEST = 4,
    YK_ERR_INVALID_BUFFER = 5,
    YK_ERR_STREAMING_STATE = 6,
    YK_ERR_DEVICE = 7
} yk_status;

/* YK::EMemoryLocation (YkReconstructionTypes.hpp:132-135) */
typedef enum yk_location { YK_HOST = 0, YK_DEVICE = 1 } yk_location;

/* storage type of filtered projections */
typedef enum yk_dtype { YK_F32 = 0, YK_F16 = 1, YK_BF16 = 2 } yk_dtype;

/* YK::EFdkFilter (algorithms/analytic/YkAnalyticFilterParams.hpp:5-8) */
typedef enum yk_filter_type {
    YK_FILTER_NONE = 0, YK_FILTER_RAMLAK = 1, YK_FILTER_SHEPPLOGAN = 2,
    YK_FILTER_COSINE = 3, YK_FILTER_HANN = 4, YK_FILTER_HAMMING = 5,
    YK_FILTER_BLACKMAN = 6, YK_FILTER_BUTTERWORTH = 7, YK_FILTER_KAISER = 8,
    YK_FILTER_TUKEY = 9,
    /* caller-given frequency weights (EFilterKernel::Custom, src/global/YkFilterTypes.hpp:8-20) */
    YK_FILTER_CUSTOM = 10
} yk_filter_type;

/* YK::EWeightsBuildSource (src/global/YkFilterTypes.hpp:22-26): how the
 * frequency weights of a windowed ramp are built */
typedef enum yk_filter_source {
    YK_FILTER_SOURCE_ANALYTIC_FREQ = 0,     /* gain |f| W(f / cutoff) */
    YK_FILTER_SOURCE_DISCRETE_RL_FFT = 1,   /* DFT of the sampled Ram-Lak kernel, windowed */
    YK_FILTER_SOURCE_SPATIAL_RAMP_FFT = 2   /* DFT of caller-given spatial taps, windowed */
} yk_filter_source;

/* YK::ERampExtractMode (src/global/YkFilterTypes.hpp:28-31): the weight a
 * DFT bin gives */
typedef enum yk_ramp_extract {
    YK_RAMP_EXTRACT_REAL = 0, YK_RAMP_EXTRACT_MAGNITUDE = 1
} yk_ramp_extract;

/* YK::EParkerMode (algorithms/analytic/YkFdkParams.hpp) */
typedef enum yk_parker_mode {
    YK_PARKER_AUTO = 0, YK_PARKER_DISABLED = 1, YK_PARKER_ENABLED = 2
} yk_parker_mode;

/* forward models, YK::ETask FP_* (src/common/YkOperatorTypes.hpp:5-22) */
typedef enum yk_forward_kind { YK_FP_JOSEPH = 1, YK_FP_SIDDON = 2 } yk_forward_kind;

/* back models: exact transposes of the forward models, and the operator-form
 * FDK backprojectors of unfiltered views used as B by the iterative solvers
 * (fdk_bp_launch / fdk_pseudo_matched_bp_launch,
 * src/FlatFpBp/BP/kernels/YkBpFdkKernel.cu:228-557): weight
 * SOD^2 dtheta fscale / (SOD - src.X / SOD)^2, or the pseudo-matched
 * L^3 / (|X - src|^2 SDD_plane) * vx vy vz / (du dv).  Derived geometry
 * (dtheta) is taken over the projector's whole view sequence.
 * YK_BP_JOSEPH_VOXEL is the voxel-driven Joseph backprojection the reference's
 * algebraic solvers use as B by default (BP_Joseph_v3 / v2,
 * src/FlatFpBp/BP/kernels/YkBPJosephKernel.cu:430-735): the bilinear sample at
 * the projected voxel centre times the ray's length through one voxel along
 * its main axis; not the exact transpose of YK_FP_JOSEPH. */
typedef enum yk_back_kind {
    YK_BP_JOSEPH_ADJOINT = 8, YK_BP_SIDDON_ADJOINT = 4,
    YK_BP_FDK = 6, YK_BP_FDK_PSEUDO_MATCHED = 7, YK_BP_JOSEPH_VOXEL = 10
} yk_back_kind;

/* ------------------------------------------------------------------------ */
/* geometry PODs (mirror the reference's public structs field for field)      */
/* ------------------------------------------------------------------------ */
/* SConeProjGeomVec (geometry/YkProjectionGeometry.hpp:12-18), 80 bytes */
typedef struct yk_view {
    float src[4];
    float detS[4];
    float detU[4];
    float detV[4];
    float angle[4];  /* angle[0] = acquisition angle, rad */
} yk_view;

/* SRegularCircularScanSpec (geometry/YkGeometryBuilderTypes.hpp:173-183) */
typedef struct yk_circular_scan {
    int total_views;
    int views_per_turn;
    float start_angle_rad;
    int rotation_direction;
    float sid_mm;
    float sdd_mm;
    float z_mm;
    float source_offset_mm[3];
} yk_circular_scan;

/* SDetectorPoseSpec (YkGeometryBuilderTypes.hpp:216-223): offset is U, N, V */
typedef struct yk_detector_pose {
    float offset_unv_mm[3];
    float tilt_u_rad;
    float tilt_v_rad;
    float tilt_n_rad;
} yk_detector_pose;

/* SFlatDetectorSpec (YkGeometryBuilderTypes.hpp:232-238) */
typedef struct yk_flat_detector {
    int channels;
    int rows;
    float channel_size_mm;
    float row_size_mm;
    yk_detector_pose pose;
} yk_flat_detector;

/* SVolumeGridSpec (YkGeometryBuilderTypes.hpp:252-260) */
typedef struct yk_volume_grid {
    int nx, ny, nz;
    float voxel_x_mm, voxel_y_mm, voxel_z_mm;
    float center_mm[3];
} yk_volume_grid;

/* Flat + circular subset of SSystemConfig (geometry/YkSystemGeometry.hpp:23-32) */
typedef struct yk_system {
    yk_circular_scan scan;
    yk_flat_detector detector;
    yk_volume_grid volume;
} yk_system;

/* SAnalyticFilterParams (algorithms/analytic/YkAnalyticFilterParams.hpp:9-14),
 * with the weight sources of SFilterKernelDesc (src/global/YkFilterTypes.hpp:34-44).
 * A zero tail (source, extract_mode, null arrays) is the analytic filter.
 * The host arrays are read when the weights are built (plan creation,
 * yk_filter_weights) and not kept. */
typedef struct yk_filter {
    int type;  /* yk_filter_type */
    float cutoff;
    float gain;
    float butterworth_order;
    float kaiser_beta;
    float tukey_alpha;
    int source;                  /* yk_filter_source */
    int extract_mode;            /* yk_ramp_extract (DiscreteRL / SpatialRamp sources) */
    const float* spatial_ramp;   /* SpatialRamp source: odd number of taps, centre tap
                                    at index size / 2, symmetric to 1e-4 */
    int spatial_ramp_size;       /* <= padded N */
    const float* custom_weights; /* YK_FILTER_CUSTOM: padded N / 2 + 1 weights of bins 0..N/2 */
    int custom_weights_size;
} yk_filter;

/* FDK plan options (SFdkAlgoParams + scan range, plus the MI355X storage
 * choices of this build) */
typedef struct yk_fdk_desc {
    uint32_t struct_size;       /* sizeof(yk_fdk_desc) */
    yk_filter filter;
    int parker_mode;            /* yk_parker_mode */
    float scan_range_rad;       /* total angular range; 0 = 2 pi */
    float start_angle_rad;      /* Parker angle base */
    int rotation_direction;     /* +1 / -1 */
    int storage;                /* yk_dtype of filtered projections */
    int kahan;                  /* compensated summation across views */
    int batch_views;            /* views filtered per workspace pass; 0 = auto */
} yk_fdk_desc;

typedef struct yk_ctx yk_ctx;
typedef struct yk_fdk yk_fdk;
typedef struct yk_projector yk_projector;

/* ------------------------------------------------------------------------ */
/* library and context                                                       */
/* ------------------------------------------------------------------------ */
/* ABI version of this header */
int yk_abi_version(void);
/* message of the last failure on this thread (never NULL) */
const char* yk_last_error(void);

/* One HIP stream on one device (ResourceContext::initialize,
 * src/common/YkExecutionContext.hpp:164-170). */
yk_status yk_ctx_create(int device, yk_ctx** out);
yk_status yk_ctx_destroy(yk_ctx* ctx);
/* Adopt a caller-owned HIP stream (hipStream_t, default priority); NULL
 * means the device's null stream.  yk_ctx_reset_stream returns to the
 * context's own stream. */
yk_status yk_ctx_set_stream(yk_ctx* ctx, void* hip_stream);
yk_status yk_ctx_reset_stream(yk_ctx* ctx);
void* yk_ctx_stream(yk_ctx* ctx);
/* 1 (default): every call waits for its GPU work; 0: enqueue only */
yk_status yk_ctx_set_synchronous(yk_ctx* ctx, int synchronous);
yk_status yk_ctx_synchronize(yk_ctx* ctx);

/* ------------------------------------------------------------------------ */
/* geometry                                                                  */
/* ------------------------------------------------------------------------ */
/* Per-view cone vectors of a regular flat circular scan (buildSystemGeometry
 * for Flat + Circular, geometry/YkSystemGeometry.hpp:115-131).  Computed in
 * FP64 and rounded to float.  *count receives total_views. */
yk_status yk_geometry_build(const yk_system* system, yk_view* out, int capacity,
    int* count);

/* FDK derived geometry, 20 floats per view in the order theta, dtheta,
 * fscale, SOD, SDD_radial, SDD_plane, offsetU_pix, offsetV_pix, du, dv,
 * radial[3], normal[3], UU, VV, UV, 0 (SFDKGeoParamPerView,
 * src/global/YkFdkKernelTypes.hpp:7-32). */
yk_status yk_geometry_derive(const yk_view* views, int n_views, int nu, int nv,
    float scan_range_rad, float* out);

/* FDK affine coefficients, 17 floats per view: Cu[4], Cv[4], Cd[4], Cr[4],
 * wbase = SOD^2 dtheta 2pi/range (FdkAffineCoeff,
 * src/global/YkFdkKernelTypes.hpp:35-55), computed in FP64. */
yk_status yk_geometry_fdk_coefficients(const yk_view* views, int n_views, int nu,
    int nv, float scan_range_rad, float* out);

/* Ramp-filter frequency weights w[0..N/2] with 1/N baked in
 * (CreateFilterKernelFromFFT::build_weights, src/Filter/YkCreateFilterKernelFFT.hpp:114-258:
 * every source and the Custom and None kinds).  N must be a power of two. */
yk_status yk_filter_weights(const yk_filter* filter, int padded_n, float du_mm,
    float* out);
/* smallest power of two >= 2 nu (.../FDK/kernels/YkFDKFilterHelpers.cuh:29-35) */
int yk_filter_padded_n(int nu);

/* ------------------------------------------------------------------------ */
/* FDK (FdkPipeline, src/Reconstruction/Analytic/Circular/Flat/FDK/YkFdkPipeline.hpp) */
/* ------------------------------------------------------------------------ */
/* Plan for a complete view sequence (prepareWithGeometry, YkFdkPipeline.hpp:344-366). */
yk_status yk_fdk_create(yk_ctx* ctx, const yk_fdk_desc* desc, const yk_view* views,
    int n_views, int nu, int nv, const yk_volume_grid* volume, yk_fdk** out);
/* Plan from a regular scan (prepare(SSystemConfig...), YkFdkPipeline.hpp:247-272);
 * scan range, start angle, direction and Parker come from the scan. */
yk_status yk_fdk_create_system(yk_ctx* ctx, const yk_fdk_desc* desc,
    const yk_system* system, yk_fdk** out);
yk_status yk_fdk_destroy(yk_fdk* fdk);

/* Streaming batch (Session::execute FDK branch, YkReconstructionApi.cpp:353-370,
 * YkSession.cpp:693-728): view_offset must equal the views received so far,
 * clear_output is only allowed on the first batch, capacities are in floats.
 * The projection may be Host or Device; the volume may be Host or Device.  A
 * Host volume is written once the last view has been received. */
yk_status yk_fdk_execute(yk_fdk* fdk, const float* projection, int projection_location,
    size_t projection_capacity, int view_offset, int view_count, float* volume,
    int volume_location, size_t volume_capacity, int clear_output);
yk_status yk_fdk_reset(yk_fdk* fdk);
int yk_fdk_received_views(const yk_fdk* fdk);

/* Read-only device view of the last chunk yk_fdk_execute processed
 * (FdkStageView / lastStage, YkFdkPipeline.hpp:176-191, 549): the input on the
 * device (the caller's device buffer, or the library's staging copy of a host
 * one), the cosine- (and Parker-) weighted projections when stage capture is
 * on (else null), and the filtered projections the backprojection consumed,
 * in the plan's storage with their per-row scales for F16 / BF16.  Valid
 * until the next execute, reset or destroy; the work is ordered on `stream`.
 * No counterpart fields: filtered_storage, d_row_scales. */
typedef struct yk_fdk_stage {
    int first_view;
    int count;                       /* 0: no chunk yet */
    const float* d_input;            /* [count][nv][nu] */
    const float* d_preweighted;      /* [count][nv][nu] or null */
    const void* d_filtered;          /* [count][nv][nu] in filtered_storage */
    int filtered_storage;            /* yk_storage */
    const float* d_row_scales;       /* [count][nv] for F16 / BF16, else null */
    void* stream;                    /* hipStream_t */
} yk_fdk_stage;
/* on: yk_fdk_execute also writes the weighted projections of each chunk
 * (one more pass over the chunk; off by default) */
yk_status yk_fdk_set_stage_capture(yk_fdk* fdk, int on);
yk_status yk_fdk_last_stage(const yk_fdk* fdk, yk_fdk_stage* out);

/* Distance of calibrated per-view geometry from the ideal circular scan
 * (FdkGeometryDiagnostics, YkFdkPipeline.hpp:35-58, computed as
 * validateApproximateFdkGeometry_, :664-866): source-to-axis, source-to-
 * detector (radial) and source-to-plane ranges, the largest |source z|,
 * detector skew |U.V| / (|U||V|), principal-point ranges in pixels, detector
 * tilts about U, V, N relative to the ideal detector of each source, the
 * largest 1 - |radial . normal|, and `approximate` when any of them (or a
 * pixel pitch) departs by more than 1e-3 relative.  Geometry FDK cannot use
 * (non-positive pitches or distances, non-finite angles, repeated adjacent
 * angles) is INVALID_GEOMETRY, as in the plan constructors. */
typedef struct yk_geometry_diagnostics {
    uint32_t struct_size;            /* caller sets sizeof(yk_geometry_diagnostics) */
    float min_source_to_axis_mm, max_source_to_axis_mm;
    float min_source_to_detector_mm, max_source_to_detector_mm;
    float min_source_to_plane_mm, max_source_to_plane_mm;
    float max_abs_source_axis_mm;
    float max_detector_skew;
    float max_abs_principal_u_pix, max_abs_principal_v_pix;
    float min_principal_u_pix, max_principal_u_pix;
    float min_principal_v_pix, max_principal_v_pix;
    float min_relative_tilt_u_rad, max_relative_tilt_u_rad;
    float min_relative_tilt_v_rad, max_relative_tilt_v_rad;
    float min_relative_tilt_n_rad, max_relative_tilt_n_rad;
    float max_normal_misalignment;
    int approximate;
} yk_geometry_diagnostics;
yk_status yk_geometry_diagnose(const yk_view* views, int n_views, int nu, int nv,
    yk_geometry_diagnostics* out);
yk_status yk_fdk_geometry_diagnostics(const yk_fdk* fdk, yk_geometry_diagnostics* out);
int yk_fdk_complete(const yk_fdk* fdk);

/* Stage entry points on device buffers.
 * Cosine preweight [+ Parker] + ramp filter of views [first, first+count):
 * d_projection holds those views, d_filtered receives them in `storage`
 * (float / half / bfloat16), d_row_scales receives one float per detector row
 * for half/bfloat16 (value = stored * scale) and may be NULL for float. */
yk_status yk_fdk_filter(yk_fdk* fdk, const float* d_projection, int first_view,
    int count, void* d_filtered, int storage, float* d_row_scales);
/* Voxel-driven FDK backprojection of filtered views [first, first+count) into
 * d_volume ([z][y][x]); accumulate = 0 overwrites. */
yk_status yk_fdk_backproject(yk_fdk* fdk, const void* d_filtered, int storage,
    const float* d_row_scales, int first_view, int count, float* d_volume,
    int accumulate);
/* yk_fdk_backproject without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Results
 * equal yk_fdk_backproject's.  For half / bfloat16 storage the normalized row
 * factors and the global row scale live in a fixed scratch array of the
 * library, one per GPU (YK_NOALLOC_MAX_ROWS factors, 16 MiB, held until the
 * process ends), and the kernel applies the factors while it stages texels
 * (yk_fdk_backproject instead normalizes the rows into a workspace first).
 * count * nv above YK_NOALLOC_MAX_ROWS is YK_ERR_UNSUPPORTED.  Calls that
 * overlap on one GPU (other contexts or streams) race on that array; every
 * call rewrites what it reads. */
#define YK_NOALLOC_MAX_ROWS 4194304
yk_status yk_fdk_backproject_noalloc(yk_fdk* fdk, const void* d_filtered, int storage,
    const float* d_row_scales, int first_view, int count, float* d_volume, int accumulate);
/* Introspection for tests: padded row length, weights (N/2+1 floats),
 * derived geometry (20 floats/view), coefficients (17 floats/view). */
int yk_fdk_padded_n(const yk_fdk* fdk);
yk_status yk_fdk_get_weights(const yk_fdk* fdk, float* out);
yk_status yk_fdk_get_derived(const yk_fdk* fdk, float* out);
yk_status yk_fdk_get_coefficients(const yk_fdk* fdk, float* out);

/* ------------------------------------------------------------------------ */
/* projection operators (IForwardOperator / IBackOperator,                   */
/* src/common/YkProjectionOperators.hpp:24-40)                               */
/* ------------------------------------------------------------------------ */
yk_status yk_projector_create(yk_ctx* ctx, const yk_view* views, int n_views,
    int nu, int nv, const yk_volume_grid* volume, float scan_range_rad,
    yk_projector** out);
yk_status yk_projector_destroy(yk_projector* projector);
/* d_projection holds views [first, first+count); accumulate = 0 overwrites */
yk_status yk_project(yk_projector* projector, int forward_kind, const float* d_volume,
    float* d_projection, int first_view, int count, int accumulate);
yk_status yk_backproject(yk_projector* projector, int back_kind, const float* d_projection,
    float* d_volume, int first_view, int count, int accumulate);

/* ------------------------------------------------------------------------ */
/* iterative callers and vector helpers                                      */
/* ------------------------------------------------------------------------ */
/* CGLS (doc/algorithms.md:269; CglsRobustBackend::run,
 * src/Reconstruction/Iterative/Flat/YkCglsBackends.hpp:282-415), x in/out.
 * residual_norms (optional, iterations+1 doubles) receives ||b - A x_k||. */
yk_status yk_cgls(yk_projector* projector, int forward_kind, int back_kind,
    const float* d_b, float* d_x, int iterations, double* residual_norms);
/* CGLS with the reference's strategies (CglsReconstructorEx,
 * src/Reconstruction/Iterative/Flat/YkCglsReconstructorEx.hpp:10-100):
 *   ROBUST_RESTART (CglsRobustBackend, YkCglsBackends.hpp:222-560): after each
 *     step the true residual ||b - A x||; a step that raises it is rolled back
 *     and the recursion restarts from there once (a second divergence in a row,
 *     or any with restart_on_divergence = 0, stops); optional convergence
 *     monitor (IterativeConvergenceConfig, YkIterativeConvergence.hpp:9-47);
 *   ASTRA (CglsAstraBackend, :22-190): the plain recursion.
 * Both stop when ||A p||^2 < epsilon, use beta = gamma1 / (gamma + epsilon) and
 * clamp x to the bounds after each iteration.  x is the start and the result.
 * residual_norms (optional, iterations + 1 doubles): [0] = ||b - A x0||, then
 * per completed iteration the true (ROBUST) or recursive (ASTRA) residual,
 * repeated past an early stop. */
/* progress callback of the iterative entry points: (user data, completed
 * units, total units) */
typedef void (*yk_progress_fn)(void* user_data, int completed, int total);
typedef enum yk_cgls_strategy { YK_CGLS_ROBUST_RESTART = 0, YK_CGLS_ASTRA = 1 } yk_cgls_strategy;
typedef struct yk_cgls_desc {
    uint32_t struct_size;            /* sizeof(yk_cgls_desc) */
    int strategy;                    /* yk_cgls_strategy */
    int iterations;                  /* >= 1 */
    float epsilon;                   /* > 0, 1e-8 in the reference */
    int restart_on_divergence;       /* ROBUST_RESTART only */
    int use_min;
    float min_value;
    int use_max;
    float max_value;
    int forward_kind;                /* yk_forward_kind */
    int back_kind;                   /* yk_back_kind */
    /* convergence monitor (ROBUST_RESTART; all tolerances 0: off) */
    float relative_residual_tolerance;     /* ||b - A x|| / ||b|| */
    float relative_update_tolerance;       /* ||x - x_prev|| / ||x_prev|| */
    float relative_improvement_tolerance;  /* residual improvement between checks */
    int minimum_iterations;          /* >= 0 */
    int check_interval;              /* >= 1 */
    int patience;                    /* >= 1 checks in a row */
    /* optional tail, read when struct_size covers it (reportExecutionProgress,
     * src/common/YkExecutionProgress.hpp:18-25; YkCglsBackends.hpp:147,402):
     * called on the calling thread once the GPU work of each completed
     * iteration has finished, with (completed, iterations); a rolled-back
     * step reports nothing */
    yk_progress_fn progress_callback;
    void* progress_user_data;
} yk_cgls_desc;
/* IterativeConvergenceStatistics (src/Reconstruction/Iterative/Common/
 * YkIterativeConvergence.hpp:18-30) plus the CGLS stop reasons */
typedef struct yk_iterative_stats {
    uint32_t struct_size;            /* caller sets sizeof(yk_iterative_stats) */
    int completed_iterations;
    int convergence_checks;
    int stopped_by_breakdown;
    int stopped_by_divergence;
    int stopped_by_relative_residual;
    int stopped_by_relative_update;
    int stopped_by_stagnation;
    double projection_residual_l2;          /* at the last check (NaN: none) */
    double relative_projection_residual;
    double relative_volume_update;
    double relative_residual_improvement;
} yk_iterative_stats;
yk_status yk_cgls_ex(yk_projector* projector, const yk_cgls_desc* desc, const float* d_b, float* d_x,
    yk_iterative_stats* stats, double* residual_norms);
/* SART with view-by-view updates (AlgebraicReconstructorEx,
 * src/Reconstruction/Iterative/Flat/YkAlgebraicReconstructorEx.hpp:137-180) */
yk_status yk_sart(yk_projector* projector, int forward_kind, int back_kind,
    const float* d_b, float* d_x, int iterations, float relaxation,
    int nonnegative);
/* Unified algebraic reconstruction (AlgebraicReconstructorEx,
 * src/Reconstruction/Iterative/Flat/YkAlgebraicReconstructorEx.hpp:15-180):
 *   x <- P[x + lambda C_s^-1 B_s R_s^-1 (b_s - A_s x)] over subsets s, then
 *   after each outer iteration smoothed-TV steps and the bounds again.
 * Subsets: SIRT one (all views), SART one per view, OS-SART `subsets`.
 * Weight models:
 *   DETAILED (AlgebraicDetailedWeightBackend, YkAlgebraicBackends.hpp:597-800;
 *   SirtBatchBackend, YkSirtBatchBackend.hpp:11-80): interleaved subsets
 *   (views s, s+n, ...), R = A_s 1 with R > eps ? r / R : 0, C = B_s 1
 *   (SIRT: C <= 0 -> 1e30), sequential or golden-ratio subset order, lambda
 *   reduced by reduction^(1/subsets) after every subset update;
 *   TIGRE (AlgebraicTigreBackend, YkAlgebraicBackends.hpp:32-310): contiguous
 *   blocks of ceil(n/subsets) views in order, R^-1 from the projection of a
 *   2 x 2 x 2 support volume (1.1 x the XY extent, max(detector, volume) Z
 *   extent; <= half the smallest voxel -> 0), C = B_s 1 with C <= 0 -> 1e30,
 *   lambda reduced once per outer iteration.
 * The update is x += lambda bp / (C + eps).  d_x holds the initial volume and
 * receives the result; d_b is the whole sinogram [view][v][u]. */
typedef enum yk_algebraic_method { YK_ALG_SIRT = 0, YK_ALG_SART = 1, YK_ALG_OSSART = 2 } yk_algebraic_method;
typedef enum yk_weight_model { YK_WEIGHTS_DETAILED = 0, YK_WEIGHTS_TIGRE = 1 } yk_weight_model;
typedef enum yk_subset_order { YK_ORDER_SEQUENTIAL = 0, YK_ORDER_GOLDEN = 1 } yk_subset_order;
typedef struct yk_algebraic_desc {
    uint32_t struct_size;        /* sizeof(yk_algebraic_desc) */
    int method;                  /* yk_algebraic_method */
    int weight_model;            /* yk_weight_model */
    int subset_order;            /* yk_subset_order (DETAILED OS-SART only) */
    int iterations;              /* outer iterations, >= 1 (every subset once each) */
    int subsets;                 /* OS-SART subset count in [1, views] */
    float relaxation;            /* lambda */
    float relaxation_reduction;  /* per outer iteration */
    float epsilon;               /* weight guard, > 0 (1e-6 in the reference) */
    int use_min;
    float min_value;
    int use_max;
    float max_value;
    int forward_kind;            /* yk_forward_kind */
    int back_kind;               /* yk_back_kind */
    float tv_strength;           /* smoothed-TV step x -= strength grad (0: off) */
    int tv_iterations;           /* TV steps after each outer iteration */
    float tv_epsilon;
    float tv_strength_reduction; /* strength *= this after each outer iteration */
    /* convergence monitor after each checked outer iteration (all tolerances
     * 0: off; AlgebraicReconstructorEx::checkConvergence_,
     * YkAlgebraicReconstructorEx.hpp:258-330): ||b - A x|| / ||b||, the
     * residual's relative improvement since the last check, ||x - x_prev|| /
     * ||x_prev|| since the last check; stop after `patience` satisfied checks
     * in a row at or past minimum_iterations */
    float relative_residual_tolerance;
    float relative_update_tolerance;
    float relative_improvement_tolerance;
    int minimum_iterations;          /* >= 0 */
    int check_interval;              /* >= 1 */
    int patience;                    /* >= 1 */
    int tv_dimensions;               /* 3 (Volume3D; 0 means 3) or 2 (Slice2D: no z differences) */
    /* optional tail, read when struct_size covers it
     * (AlgebraicReconstructorEx::run, YkAlgebraicReconstructorEx.hpp:164):
     * called on the calling thread once the GPU work of each outer iteration
     * (its subset updates, TV steps and bounds) has finished, with
     * (completed, iterations) */
    yk_progress_fn progress_callback;
    void* progress_user_data;
} yk_algebraic_desc;
/* stats (optional) receives the completed outer iterations and the monitor's
 * last readings */
yk_status yk_algebraic(yk_projector* projector, const yk_algebraic_desc* desc,
    const float* d_b, float* d_x, yk_iterative_stats* stats);
/* smoothed TV value and gradient (YkTvRegularizationKernels.cu:17-73) */
yk_status yk_tv_gradient(yk_ctx* ctx, const yk_volume_grid* volume, const float* d_x,
    float epsilon, float* d_gradient, double* value);
/* FP64-accumulated dot product of two device float arrays */
yk_status yk_dot(yk_ctx* ctx, const float* d_a, const float* d_b, size_t n, double* out);
/* Point-sampled ellipsoid phantom into a device volume.  Each ellipsoid is 16
 * floats: centre[3], semi_axes[3], R[9] (body = R (p - centre)), value. */
yk_status yk_phantom_ellipsoids(yk_ctx* ctx, const float* ellipsoids, int count,
    const yk_volume_grid* volume, int clamp_nonnegative, float* d_volume);

/* ------------------------------------------------------------------------ */
/* raw dumps (DumpManager, src/util/YkIoDump.hpp:9-72)                       */
/* ------------------------------------------------------------------------ */
/* Float32 device buffers and text into files under output_dir, little-endian
 * with no header: projections [view][v][u], volumes [z][y][x].  The *_a forms
 * name the file <prefix>_a<4-digit index>.raw / .txt.  A dump copies through
 * a pinned host buffer after synchronizing `stream` (hipStream_t, null: the
 * default stream).  While disabled every call succeeds without writing. */
typedef struct yk_dump yk_dump;
yk_status yk_dump_create(const char* output_dir, int enabled, yk_dump** out);
yk_status yk_dump_destroy(yk_dump* dump);
yk_status yk_dump_set_enabled(yk_dump* dump, int enabled);
yk_status yk_dump_device_f32(yk_dump* dump, const char* file_name, const float* d_src,
    size_t count, void* stream);
yk_status yk_dump_device_f32_a(yk_dump* dump, const char* prefix, int index, const float* d_src,
    size_t count, void* stream);
yk_status yk_dump_text(yk_dump* dump, const char* file_name, const char* text);
yk_status yk_dump_text_a(yk_dump* dump, const char* prefix, int index, const char* text);

#ifdef __cplusplus
}
#endif
#endif

// C++ Session API of the flat circular path, source-compatible with the
// reference's public interface for the FDK and ForwardProjection pipelines:
//   include/YKCBCT/interface/YkReconstructionApi.hpp:12-98 (SSystemSpec,
//     SExecutionRequest, EApiErrorCode, IReconstructionSession,
//     ReconstructionSessionFactory),
//   include/YKCBCT/interface/YkReconstructionTypes.hpp:14-44 (EPipeline,
//     EMemoryLocation, Buffer),
//   include/YKCBCT/interface/YkSystemReconstruction.hpp:8-18 (SReconstructionSpec),
//   include/YKCBCT/geometry/YkSystemGeometry.hpp:12-31 and
//     YkGeometryBuilderTypes.hpp:11-117 (SSystemConfig and its scan, detector,
//     volume and pose records),
//   include/YKCBCT/algorithms/analytic/YkFdkParams.hpp, YkAnalyticFilterParams.hpp,
//   YkWfbpParams.hpp, include/YKCBCT/algorithms/projection/YkProjectionParams.hpp,
//   include/YKCBCT/algorithms/iterative/YkIterationParams.hpp, YkCglsParams.hpp,
//   YkPwlsParams.hpp, YkTigreParams.hpp.
// Field names, order, defaults and meanings follow those headers.  Pipelines:
// FDK, ForwardProjection, SART, SIRT, OSSART and CGLS on flat detectors and
// circular scans.  Records of what this library does not provide (helical
// and cylindrical geometry, statistical and plugin tails) are omitted; the
// PWLS, WFBP and TIGRE-gradient blocks are kept so code that fills them
// compiles, and those pipelines are refused with UnsupportedCombination.
// float3 is HIP's.
//
// The implementation (ReconstructionSessionFactory in libyk_hip.so) sits on
// the C ABI of yk_hip.h.  Like the reference, a session is not thread-safe,
// execute() is synchronous on success, buffers are borrowed for the call and
// element_count is a float count.
#pragma once

#include <cstddef>
#include <cstdint>

#include <hip/hip_vector_types.h>

namespace YK {

// --- geometry ---------------------------------------------------------------
enum class EDetectorKind : int32_t { Flat, Cylindrical };
enum class ETrajectoryKind : int32_t { Circular, Helical, PlanarSourceOrbit, NikonTilted };

// Regular static circular scan: total_views < views_per_turn is a short scan.
struct SRegularCircularScanSpec {
    int total_views = 0;
    int views_per_turn = 0;
    float start_angle_rad = 0.f;
    int rotation_direction = 1;  // +1 / -1
    float sid_mm = 0.f;
    float sdd_mm = 0.f;
    float z_mm = 0.f;
    float3 source_offset_mm = make_float3(0.f, 0.f, 0.f);
};

// Detector pose relative to the nominal principal point: offsets along
// U, N, V; tilts (rad) about U, then V, then N.  U x V = N, N toward the source.
struct SDetectorPoseSpec {
    float3 offset_unv_mm = make_float3(0.f, 0.f, 0.f);
    float tilt_u_rad = 0.f;
    float tilt_v_rad = 0.f;
    float tilt_n_rad = 0.f;
};

struct SFlatDetectorSpec {
    int channels = 0;
    int rows = 0;
    float channel_size_mm = 0.f;
    float row_size_mm = 0.f;
    SDetectorPoseSpec pose{};
};

struct SVolumeGridSpec {
    int nx = 0;
    int ny = 0;
    int nz = 0;
    float voxel_x_mm = 0.f;
    float voxel_y_mm = 0.f;
    float voxel_z_mm = 0.f;
    float3 center_mm = make_float3(0.f, 0.f, 0.f);
};

// Volume rotation of the standalone forward projection: fixed world-axis X,
// then Y, then Z rotations about the volume centre.
struct SForwardProjectionPose {
    bool enabled = false;
    float rotation_x_rad = 0.f;
    float rotation_y_rad = 0.f;
    float rotation_z_rad = 0.f;
};

// Per-view cone geometry (SConeProjGeomVec,
// include/YKCBCT/geometry/YkProjectionGeometry.hpp:12-18), 80 bytes: source,
// centre of detector pixel (0,0), one-pixel steps along U and V (mm);
// angle.x is the acquisition angle (rad).
struct SConeProjGeomVec {
    float4 src;
    float4 detS;
    float4 detU;
    float4 detV;
    float4 angle;
};

struct SSystemConfig {
    EDetectorKind detector = EDetectorKind::Flat;
    ETrajectoryKind trajectory = ETrajectoryKind::Circular;
    SRegularCircularScanSpec circular{};
    SFlatDetectorSpec flat_detector{};
    SVolumeGridSpec volume{};
    SForwardProjectionPose forward_projection_pose{};
};

// --- algorithm --------------------------------------------------------------
enum class EPipeline : int32_t {
    FDK, ForwardProjection, SIRT, OSSART, CGLS, PWLS, XFDK, CFDK, WFBP, TigreGradient, SART, Statistical,
};

enum class EForwardProjector : int32_t { Automatic, Siddon, Joseph, SeparableFootprint };
enum class EBackProjector : int32_t {
    Automatic, FdkPseudoMatched, SiddonRayDriven, Joseph, JosephV3, SeparableFootprint,
};

enum class EParkerMode : int32_t { Auto, Disabled, Enabled };
struct SParkerScanSpec {
    EParkerMode mode = EParkerMode::Auto;
};

enum class EFdkFilter : int32_t {
    None = 0, RamLak = 1, SheppLogan = 2, Cosine = 3, Hann = 4, Hamming = 5,
    Blackman = 6, Butterworth = 7, Kaiser = 8, Tukey = 9
};
struct SAnalyticFilterParams {
    EFdkFilter type = EFdkFilter::RamLak;
    float cutoff = .5f, gain = 1.f;
    float butterworth_order = 2.f, kaiser_beta = 8.6f, tukey_alpha = .5f;
};

struct SFdkAlgoParams {
    SParkerScanSpec parker{};
    SAnalyticFilterParams filter{};
};

// SART / SIRT / OSSART / CGLS: outer iterations, relaxation, OS-SART subsets
struct SIterAlgoParams {
    int iterations = 10;
    float relaxation = 1.f;
    int subsets = 1;
};

// convergence monitor of the cylindrical CGLS (not read on flat scans)
struct SCglsAlgoParams {
    int minimum_iterations = 0;
    int check_interval = 1;
    int patience = 1;
    float relative_residual = 0.f;
    bool robust_restart = true;
};

enum class EPwlsRegularizerSpec : int32_t { None, Quadratic, Huber };
struct SPwlsAlgoParams {
    EPwlsRegularizerSpec regularizer = EPwlsRegularizerSpec::Quadratic;
    float regularization = 1e-3f;
    float huber_delta = 3e-3f;
    float epsilon = 1e-6f;
    float lower_bound = 0.f;
    float upper_bound = 3.40282347e+38f;
    bool use_projection_weights = false;
};

enum class ETigreGradientMethodSpec : int32_t {
    Sart, OsSart, Sirt, AsdPocs, OsAsdPocs, BAsdPocsBeta, Pcsd, OsPcsd, AwPcsd, OsAwPcsd, AwAsdPocs, OsAwAsdPocs,
};
// Of this block the algebraic pipelines read non_negative (clamp at 0).
struct STigreGradientAlgoParams {
    ETigreGradientMethodSpec method = ETigreGradientMethodSpec::OsSart;
    int block_size = 20;
    float lambda_reduction = 0.f;
    bool nesterov_relaxation = false;
    bool fdk_initialization = false;
    bool non_negative = true;
    int tv_iterations = 20;
    float tv_alpha = 0.002f;
    float tv_alpha_reduction = 0.95f;
    float maximum_update_ratio = 0.95f;
    float max_l2_error = -1.f;
};

struct SWfbpAlgoParams {
    float redundancy_flat = 0.6f;
    SAnalyticFilterParams filter{};
};

struct SReconstructionSpec {
    EPipeline pipeline = EPipeline::FDK;
    EForwardProjector forward_projector = EForwardProjector::Automatic;
    EBackProjector back_projector = EBackProjector::Automatic;
    SFdkAlgoParams fdk{};
    SIterAlgoParams iterative{};
    SCglsAlgoParams cgls{};
    SPwlsAlgoParams pwls{};
    STigreGradientAlgoParams tigre{};
    SWfbpAlgoParams wfbp{};
};

// --- session ----------------------------------------------------------------
enum class EMemoryLocation : int32_t { Host, Device };

// Contiguous float arrays: projection [view][v][u], volume [z][y][x].
// The iterative pipelines start from the volume's contents and need it in
// device memory (CGLS the projection too).
struct Buffer {
    float* data = nullptr;
    EMemoryLocation location = EMemoryLocation::Device;
    size_t element_count = 0;  // capacity in floats
};

struct SSystemSpec {
    uint32_t struct_size = sizeof(SSystemSpec);
    uint32_t api_version = 2;
    SSystemConfig geometry{};
    SReconstructionSpec reconstruction{};
    int device = 0;
    // Optional flat-frame tail: calibrated per-view geometry, one frame per
    // view of the scan, used in place of the builder's views by
    // ForwardProjection, SART, SIRT, OSSART and CGLS.  FDK accepts frames
    // only within 1e-3 of the nominal circular trajectory.  Host memory
    // borrowed during initialize only (the session copies it).
    const SConeProjGeomVec* flat_frames = nullptr;
    int flat_frame_count = 0;
};

struct SExecutionRequest {
    uint32_t struct_size = sizeof(SExecutionRequest);
    uint32_t api_version = 2;
    Buffer projection{};
    Buffer volume{};
    int view_offset = 0;
    int view_count = 0;  // 0: every prepared view
    bool clear_output = true;
    int iteration_count = 0;
    // Optional tail: called on execute's thread once the GPU work of each
    // outer iteration of SART / SIRT / OSSART / CGLS has finished
    // (completed, total = iterations); FDK and ForwardProjection report nothing.
    void (*progress_callback)(void*, int completed, int total) = nullptr;
    void* progress_user_data = nullptr;
};

enum class EApiErrorCode : int32_t {
    None = 0,
    InvalidConfig,
    InvalidGeometry,
    UnsupportedCombination,
    AlgorithmUnderTest,
    InvalidBuffer,
    StreamingStateError,
    CudaError,  // a HIP runtime failure on this build
};

class IReconstructionSession {
public:
    virtual ~IReconstructionSession() = default;
    virtual bool initialize(const SSystemSpec& system) = 0;
    virtual bool execute(const SExecutionRequest& request) = 0;
    virtual bool isInitialized() const = 0;
    virtual EApiErrorCode lastError() const = 0;
    // valid until the next call on this session
    virtual const char* lastErrorMessage() const = 0;
    virtual void reset() = 0;
    virtual void release() = 0;
};

struct ReconstructionSessionFactory {
    static IReconstructionSession* create();
    static void destroy(IReconstructionSession* session);
};

}  // namespace YK

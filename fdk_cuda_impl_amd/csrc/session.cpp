// C++ Session facade (include/yk_session.hpp) over this library's C ABI: the
// reference's ReconstructionSession (src/interface/YkReconstructionApi.cpp:
// 133-418) with its FDK, ForwardProjection and flat algebraic / CGLS backends
// (src/interface/YkSession.cpp:290-350, 693-728, 881-1030,
// src/common/YkProjectionOperators.hpp:138-180) for flat detectors on
// circular scans.  Validation order and error codes follow the reference:
// struct_size / api_version, enum ranges, projector selection, pose, then
// the backend; execute checks the streaming contract and the buffer
// capacities before any GPU work.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/yk_hip.h"
#include "../../include/yk_session.hpp"

namespace YK {
namespace {

bool has_bytes(uint32_t actual, size_t required) { return actual >= required; }

bool valid_location(EMemoryLocation l) { return l == EMemoryLocation::Host || l == EMemoryLocation::Device; }

// 3 x 3 rotation R = Rz Ry Rx (fixed world axes X, then Y, then Z), FP64
void euler_rotation(double rx, double ry, double rz, double R[3][3])
{
    const double cx = std::cos(rx), sx = std::sin(rx), cy = std::cos(ry), sy = std::sin(ry);
    const double cz = std::cos(rz), sz = std::sin(rz);
    const double X[3][3] = { { 1, 0, 0 }, { 0, cx, -sx }, { 0, sx, cx } };
    const double Y[3][3] = { { cy, 0, sy }, { 0, 1, 0 }, { -sy, 0, cy } };
    const double Z[3][3] = { { cz, -sz, 0 }, { sz, cz, 0 }, { 0, 0, 1 } };
    double ZY[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            ZY[i][j] = 0.0;
            for (int k = 0; k < 3; ++k) ZY[i][j] += Z[i][k] * Y[k][j];
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            R[i][j] = 0.0;
            for (int k = 0; k < 3; ++k) R[i][j] += ZY[i][k] * X[k][j];
        }
}

// Views in the frame of a volume rotated by R about c: points map by
// R^T (p - c) + c, directions by R^T (the inverse rigid transform applied to
// the cone vectors, as transformProjectionGeometry does with local_from_world).
void rotate_views(std::vector<yk_view>& views, const SForwardProjectionPose& pose, const float c[3])
{
    double R[3][3];
    euler_rotation(pose.rotation_x_rad, pose.rotation_y_rad, pose.rotation_z_rad, R);
    auto vec = [&](float* v) {
        double o[3];
        for (int i = 0; i < 3; ++i) o[i] = R[0][i] * v[0] + R[1][i] * v[1] + R[2][i] * v[2];
        for (int i = 0; i < 3; ++i) v[i] = float(o[i]);
    };
    auto point = [&](float* p) {
        double d[3], o[3];
        for (int i = 0; i < 3; ++i) d[i] = double(p[i]) - c[i];
        for (int i = 0; i < 3; ++i) o[i] = R[0][i] * d[0] + R[1][i] * d[1] + R[2][i] * d[2] + c[i];
        for (int i = 0; i < 3; ++i) p[i] = float(o[i]);
    };
    for (auto& v : views) {
        point(v.src);
        point(v.detS);
        vec(v.detU);
        vec(v.detV);
    }
}

// The flat circular part of SSystemConfig as the C ABI's yk_system.
yk_system c_system(const SSystemConfig& g)
{
    yk_system sys{};
    const SRegularCircularScanSpec& c = g.circular;
    sys.scan.total_views = c.total_views;
    sys.scan.views_per_turn = c.views_per_turn;
    sys.scan.start_angle_rad = c.start_angle_rad;
    sys.scan.rotation_direction = c.rotation_direction;
    sys.scan.sid_mm = c.sid_mm;
    sys.scan.sdd_mm = c.sdd_mm;
    sys.scan.z_mm = c.z_mm;
    sys.scan.source_offset_mm[0] = c.source_offset_mm.x;
    sys.scan.source_offset_mm[1] = c.source_offset_mm.y;
    sys.scan.source_offset_mm[2] = c.source_offset_mm.z;
    const SFlatDetectorSpec& d = g.flat_detector;
    sys.detector.channels = d.channels;
    sys.detector.rows = d.rows;
    sys.detector.channel_size_mm = d.channel_size_mm;
    sys.detector.row_size_mm = d.row_size_mm;
    sys.detector.pose.offset_unv_mm[0] = d.pose.offset_unv_mm.x;
    sys.detector.pose.offset_unv_mm[1] = d.pose.offset_unv_mm.y;
    sys.detector.pose.offset_unv_mm[2] = d.pose.offset_unv_mm.z;
    sys.detector.pose.tilt_u_rad = d.pose.tilt_u_rad;
    sys.detector.pose.tilt_v_rad = d.pose.tilt_v_rad;
    sys.detector.pose.tilt_n_rad = d.pose.tilt_n_rad;
    const SVolumeGridSpec& v = g.volume;
    sys.volume.nx = v.nx;
    sys.volume.ny = v.ny;
    sys.volume.nz = v.nz;
    sys.volume.voxel_x_mm = v.voxel_x_mm;
    sys.volume.voxel_y_mm = v.voxel_y_mm;
    sys.volume.voxel_z_mm = v.voxel_z_mm;
    sys.volume.center_mm[0] = v.center_mm.x;
    sys.volume.center_mm[1] = v.center_mm.y;
    sys.volume.center_mm[2] = v.center_mm.z;
    return sys;
}

// nominal views of the scan; false when the builder refuses it
bool nominal_views(const yk_system& sys, std::vector<yk_view>& out)
{
    int n = 0;
    const yk_status st = yk_geometry_build(&sys, nullptr, 0, &n);
    if ((st != YK_OK && st != YK_ERR_INVALID_BUFFER) || n <= 0) return false;
    out.resize(size_t(n));
    return yk_geometry_build(&sys, out.data(), n, &n) == YK_OK;
}

yk_view to_view(const SConeProjGeomVec& f)
{
    yk_view v{};
    const float4* in[5] = { &f.src, &f.detS, &f.detU, &f.detV, &f.angle };
    float* out[5] = { v.src, v.detS, v.detU, v.detV, v.angle };
    for (int k = 0; k < 5; ++k) {
        out[k][0] = in[k]->x;
        out[k][1] = in[k]->y;
        out[k][2] = in[k]->z;
        out[k][3] = in[k]->w;
    }
    return v;
}

// the frame checks of PreparedGeometry::initialize
// (src/common/YkExecutionContext.hpp:22-36): finite vectors, non-degenerate
// and orthogonal detector steps (|U.V| <= 1e-5 |U||V|), a finite angle
bool usable_frame(const SConeProjGeomVec& f)
{
    for (const float4* a : { &f.src, &f.detS, &f.detU, &f.detV })
        if (!std::isfinite(a->x) || !std::isfinite(a->y) || !std::isfinite(a->z)) return false;
    const double uu = double(f.detU.x) * f.detU.x + double(f.detU.y) * f.detU.y + double(f.detU.z) * f.detU.z;
    const double vv = double(f.detV.x) * f.detV.x + double(f.detV.y) * f.detV.y + double(f.detV.z) * f.detV.z;
    const double uv = double(f.detU.x) * f.detV.x + double(f.detU.y) * f.detV.y + double(f.detU.z) * f.detV.z;
    return uu > 0.0 && vv > 0.0 && std::abs(uv) <= 1e-5 * std::sqrt(uu * vv) && std::isfinite(f.angle.x);
}

// Scoped switch to the session's device, restoring the caller's.
class DeviceScope {
public:
    explicit DeviceScope(int device)
    {
        ok_ = hipGetDevice(&previous_) == hipSuccess && hipSetDevice(device) == hipSuccess;
    }
    ~DeviceScope()
    {
        if (ok_) (void)hipSetDevice(previous_);
    }
    bool ok() const { return ok_; }

private:
    int previous_ = 0;
    bool ok_ = false;
};

class Session final : public IReconstructionSession {
public:
    ~Session() override { release(); }

    bool initialize(const SSystemSpec& spec) override
    {
        release();
        clear();
        if (!has_bytes(spec.struct_size, offsetof(SSystemSpec, device) + sizeof(spec.device)))
            return fail(EApiErrorCode::InvalidConfig, "SSystemSpec is truncated or from an older layout");
        if (spec.api_version != 2) return fail(EApiErrorCode::InvalidConfig, "unsupported SSystemSpec api_version");
        const SSystemConfig& g = spec.geometry;
        const SReconstructionSpec& rec = spec.reconstruction;
        // flat-frame tail (YkReconstructionApi.cpp:146-167)
        const bool frame_tail =
            has_bytes(spec.struct_size, offsetof(SSystemSpec, flat_frame_count) + sizeof(spec.flat_frame_count));
        if (spec.struct_size > offsetof(SSystemSpec, flat_frames) && !frame_tail)
            return fail(EApiErrorCode::InvalidConfig, "truncated flat-frame tail");
        const SConeProjGeomVec* frames = frame_tail ? spec.flat_frames : nullptr;
        const int frame_count = frame_tail ? spec.flat_frame_count : 0;
        if ((frames == nullptr) != (frame_count == 0) || frame_count < 0)
            return fail(EApiErrorCode::InvalidConfig, "invalid flat-frame buffer / count");
        if (frames) {
            const EPipeline p = rec.pipeline;
            if (g.detector != EDetectorKind::Flat ||
                (g.trajectory != ETrajectoryKind::Circular && g.trajectory != ETrajectoryKind::Helical) ||
                (p != EPipeline::ForwardProjection && p != EPipeline::SART && p != EPipeline::SIRT &&
                 p != EPipeline::OSSART && p != EPipeline::CGLS && p != EPipeline::FDK && p != EPipeline::WFBP))
                return fail(EApiErrorCode::UnsupportedCombination, "external flat frames are not supported by this pipeline");
            if ((p == EPipeline::FDK || p == EPipeline::WFBP) && g.trajectory == ETrajectoryKind::Circular) {
                // analytic reconstruction keeps its circular-orbit model: the
                // frames must be the nominal trajectory within 1e-3
                std::vector<yk_view> expected;
                if (!nominal_views(c_system(g), expected) || int(expected.size()) != frame_count)
                    return fail(EApiErrorCode::InvalidGeometry,
                                "analytic reconstruction with external frames needs the matching nominal trajectory");
                for (int i = 0; i < frame_count; ++i) {
                    const yk_view a = expected[size_t(i)], b = to_view(frames[i]);
                    const float* av[5] = { a.src, a.detS, a.detU, a.detV, a.angle };
                    const float* bv[5] = { b.src, b.detS, b.detU, b.detV, b.angle };
                    for (int k = 0; k < 5; ++k)
                        for (int j = 0; j < 3; ++j)
                            if (std::abs(av[k][j] - bv[k][j]) > 1e-3f)
                                return fail(EApiErrorCode::UnsupportedCombination,
                                            "external frames depart from the circular trajectory FDK assumes; "
                                            "use an iterative reconstruction");
                }
            }
        }
        const int det = int(g.detector), traj = int(g.trajectory), pipe = int(rec.pipeline);
        if (det < 0 || det > 1 || traj < 0 || traj > 3 || pipe < 0 || pipe > int(EPipeline::Statistical))
            return fail(EApiErrorCode::InvalidConfig, "pipeline, detector or trajectory enum out of range");
        const int fwd = int(rec.forward_projector), back = int(rec.back_projector);
        if (fwd < 0 || fwd > int(EForwardProjector::SeparableFootprint) || back < 0 ||
            back > int(EBackProjector::SeparableFootprint))
            return fail(EApiErrorCode::InvalidConfig, "forward or back projector enum out of range");
        if (g.detector != EDetectorKind::Flat || g.trajectory != ETrajectoryKind::Circular)
            return fail(EApiErrorCode::UnsupportedCombination,
                        "this library provides the flat-detector circular-scan path only");
        const bool iterative = rec.pipeline == EPipeline::SART || rec.pipeline == EPipeline::SIRT ||
                               rec.pipeline == EPipeline::OSSART || rec.pipeline == EPipeline::CGLS;
        if (rec.pipeline != EPipeline::FDK && rec.pipeline != EPipeline::ForwardProjection && !iterative)
            return fail(EApiErrorCode::UnsupportedCombination,
                        "the Session facade runs the FDK, ForwardProjection, SART, SIRT, OSSART and CGLS pipelines");
        const SForwardProjectionPose& pose = g.forward_projection_pose;
        if (!std::isfinite(pose.rotation_x_rad) || !std::isfinite(pose.rotation_y_rad) ||
            !std::isfinite(pose.rotation_z_rad))
            return fail(EApiErrorCode::InvalidGeometry, "forward-projection volume rotation must be finite radians");
        if (rec.pipeline == EPipeline::ForwardProjection) {
            // Automatic is Siddon for flat detectors (resolveForwardProjector,
            // src/interface/YkProjectorSelection.hpp:11-23)
            if (rec.forward_projector == EForwardProjector::SeparableFootprint)
                return fail(EApiErrorCode::UnsupportedCombination, "the separable-footprint projector is not provided");
            forward_kind_ = rec.forward_projector == EForwardProjector::Joseph ? YK_FP_JOSEPH : YK_FP_SIDDON;
        } else if (iterative) {
            // resolveForwardProjector / resolveBackProjector
            // (src/interface/YkProjectorSelection.hpp:11-44): Automatic is
            // Siddon forward; backward the Siddon transpose where the solver
            // needs A^T (CGLS), else the FDK pseudo-matched backprojection
            if (rec.forward_projector == EForwardProjector::SeparableFootprint ||
                rec.back_projector == EBackProjector::SeparableFootprint)
                return fail(EApiErrorCode::UnsupportedCombination, "the separable-footprint projector is not provided");
            forward_kind_ = rec.forward_projector == EForwardProjector::Joseph ? YK_FP_JOSEPH : YK_FP_SIDDON;
            switch (rec.back_projector) {
            case EBackProjector::Automatic:
                back_kind_ = rec.pipeline == EPipeline::CGLS ? YK_BP_SIDDON_ADJOINT : YK_BP_FDK_PSEUDO_MATCHED;
                break;
            case EBackProjector::FdkPseudoMatched: back_kind_ = YK_BP_FDK_PSEUDO_MATCHED; break;
            case EBackProjector::SiddonRayDriven: back_kind_ = YK_BP_SIDDON_ADJOINT; break;
            case EBackProjector::Joseph: back_kind_ = YK_BP_JOSEPH_ADJOINT; break;
            default: back_kind_ = YK_BP_JOSEPH_VOXEL; break;  // JosephV3
            }
            // the backends' own checks (AlgebraicReconstructorEx::validate_,
            // CglsReconstructorEx::prepare); a failed backend initialisation
            // is reported as InvalidGeometry
            const SIterAlgoParams& it = rec.iterative;
            subsets_ = it.subsets > 1 ? it.subsets : 1;
            if (it.iterations <= 0 || (rec.pipeline != EPipeline::CGLS && !(it.relaxation > 0.f)) ||
                !std::isfinite(it.relaxation) ||
                (rec.pipeline == EPipeline::OSSART && subsets_ > g.circular.total_views))
                return fail(EApiErrorCode::InvalidGeometry,
                            "iterative backend initialisation failed: iterations, relaxation or subset count");
            iterations_ = it.iterations;
            relaxation_ = it.relaxation;
            non_negative_ = rec.tigre.non_negative;
        } else {
            const int p = int(rec.fdk.parker.mode), f = int(rec.fdk.filter.type);
            if (p < 0 || p > 2 || f < 0 || f > 9)
                return fail(EApiErrorCode::InvalidConfig, "FDK Parker mode or filter enum out of range");
        }

        const yk_system sys = c_system(g);
        const SRegularCircularScanSpec& c = g.circular;
        const SFlatDetectorSpec& d = g.flat_detector;
        const SVolumeGridSpec& v = g.volume;
        if (v.nx <= 0 || v.ny <= 0 || v.nz <= 0 || d.channels <= 0 || d.rows <= 0)
            return fail(EApiErrorCode::InvalidGeometry, "volume and detector sizes must be positive");

        // first call: validate the scan and count its views
        int n = 0;
        if (yk_status st = yk_geometry_build(&sys, nullptr, 0, &n); st != YK_OK && st != YK_ERR_INVALID_BUFFER)
            return fail_status(st, "geometry");
        if (n <= 0) return fail(EApiErrorCode::InvalidGeometry, "the scan has no views");
        std::vector<yk_view> views(static_cast<size_t>(n));
        if (yk_status st = yk_geometry_build(&sys, views.data(), n, &n); st != YK_OK)
            return fail_status(st, "geometry");
        if (frames) {
            // PreparedGeometry::initialize with external frames
            // (src/common/YkExecutionContext.hpp:22-36); its failure is the
            // backend's, reported as InvalidGeometry
            if (frame_count != c.total_views)
                return fail(EApiErrorCode::InvalidGeometry, "one external frame per view of the scan is required");
            for (int i = 0; i < frame_count; ++i)
                if (!usable_frame(frames[i]))
                    return fail(EApiErrorCode::InvalidGeometry,
                                "external frame " + std::to_string(i) +
                                    ": non-finite, degenerate or non-orthogonal detector vectors");
            for (int i = 0; i < frame_count; ++i) views[size_t(i)] = to_view(frames[i]);
        }

        device_ = spec.device;
        if (yk_status st = yk_ctx_create(device_, &ctx_); st != YK_OK) return fail_status(st, "device context");
        if (rec.pipeline == EPipeline::FDK) {
            yk_fdk_desc desc{};
            desc.struct_size = sizeof(desc);
            const SAnalyticFilterParams& f = rec.fdk.filter;
            desc.filter.type = int(f.type);
            desc.filter.cutoff = f.cutoff;
            desc.filter.gain = f.gain;
            desc.filter.butterworth_order = f.butterworth_order;
            desc.filter.kaiser_beta = f.kaiser_beta;
            desc.filter.tukey_alpha = f.tukey_alpha;
            desc.parker_mode = int(rec.fdk.parker.mode);
            desc.rotation_direction = c.rotation_direction;
            desc.storage = YK_F32;
            yk_status st;
            if (frames) {
                // prepareWithGeometry (YkSession.cpp:131-163): the frames are
                // the geometry; angles from angle.x, the direction from the
                // first two, the range of the scan
                desc.scan_range_rad = float(2.0 * 3.14159265358979323846 * c.total_views / c.views_per_turn);
                desc.start_angle_rad = views[0].angle[0];
                if (n >= 2) desc.rotation_direction = views[1].angle[0] >= views[0].angle[0] ? 1 : -1;
                st = yk_fdk_create(ctx_, &desc, views.data(), n, d.channels, d.rows, &sys.volume, &fdk_);
            } else {
                st = yk_fdk_create_system(ctx_, &desc, &sys, &fdk_);
            }
            if (st != YK_OK) {
                const bool r = fail_status(st, "FDK plan");
                release_handles();
                return r;
            }
        } else {
            if (pose.enabled && rec.pipeline == EPipeline::ForwardProjection)
                rotate_views(views, pose, sys.volume.center_mm);
            const float range = float(2.0 * 3.14159265358979323846 * c.total_views / c.views_per_turn);
            if (yk_status st = yk_projector_create(ctx_, views.data(), n, d.channels, d.rows, &sys.volume, range,
                                                   &proj_);
                st != YK_OK) {
                const bool r = fail_status(st, "projector");
                release_handles();
                return r;
            }
        }
        pipeline_ = rec.pipeline;
        views_ = n;
        view_elems_ = size_t(d.channels) * d.rows;
        vol_elems_ = size_t(v.nx) * v.ny * v.nz;
        received_ = 0;
        initialized_ = true;
        return true;
    }

    bool execute(const SExecutionRequest& r) override
    {
        clear();
        if (!initialized_) return fail(EApiErrorCode::StreamingStateError, "the session is not initialized");
        if (!has_bytes(r.struct_size, offsetof(SExecutionRequest, iteration_count) + sizeof(r.iteration_count)))
            return fail(EApiErrorCode::InvalidConfig, "SExecutionRequest is truncated or from an older layout");
        if (r.struct_size > offsetof(SExecutionRequest, progress_callback) &&
            !has_bytes(r.struct_size, offsetof(SExecutionRequest, progress_user_data) + sizeof(r.progress_user_data)))
            return fail(EApiErrorCode::InvalidConfig, "SExecutionRequest progress tail is truncated");
        if (r.api_version != 2) return fail(EApiErrorCode::InvalidConfig, "unsupported SExecutionRequest api_version");
        if (!r.projection.data || !r.volume.data || !valid_location(r.projection.location) ||
            !valid_location(r.volume.location))
            return fail(EApiErrorCode::InvalidBuffer, "projection / volume buffer is null or its location is invalid");
        const int offset = r.view_offset;
        const int count = r.view_count > 0 ? r.view_count : views_;
        if (offset < 0 || count <= 0 || offset > views_ - count)
            return fail(EApiErrorCode::InvalidConfig, "view_offset / view_count outside the prepared views");
        if (pipeline_ == EPipeline::FDK) {
            if (offset != received_)
                return fail(EApiErrorCode::StreamingStateError,
                            "FDK batches must arrive in order: view_offset must equal the views received");
            if (received_ > 0 && r.clear_output)
                return fail(EApiErrorCode::StreamingStateError, "only the first FDK batch may clear the output");
        } else if (offset != 0 || count != views_) {
            return fail(EApiErrorCode::StreamingStateError,
                        "this pipeline takes the whole view sequence from view_offset 0");
        }
        if (r.projection.element_count < view_elems_ * size_t(count) || r.volume.element_count < vol_elems_)
            return fail(EApiErrorCode::InvalidBuffer, "projection / volume element_count below what this call needs");

        // progress (ScopedExecutionProgress, YkReconstructionApi.cpp:383-384):
        // the iterative backends report each outer iteration; FDK and the
        // forward projection report nothing
        const bool tail = has_bytes(r.struct_size, offsetof(SExecutionRequest, progress_user_data) +
                                                       sizeof(r.progress_user_data));
        yk_progress_fn progress = tail ? r.progress_callback : nullptr;
        void* progress_user = tail ? r.progress_user_data : nullptr;
        bool ok;
        if (pipeline_ == EPipeline::FDK)
            ok = execute_fdk(r, offset, count);
        else if (pipeline_ == EPipeline::ForwardProjection)
            ok = execute_fp(r);
        else
            ok = execute_iterative(r, progress, progress_user);
        if (!ok) return false;
        if (pipeline_ == EPipeline::FDK) received_ += count;
        return true;
    }

    bool isInitialized() const override { return initialized_; }
    EApiErrorCode lastError() const override { return code_; }
    const char* lastErrorMessage() const override { return message_.c_str(); }

    void reset() override
    {
        if (fdk_) (void)yk_fdk_reset(fdk_);
        received_ = 0;
    }

    void release() override
    {
        release_handles();
        views_ = received_ = 0;
        view_elems_ = vol_elems_ = 0;
        initialized_ = false;
    }

private:
    bool execute_fdk(const SExecutionRequest& r, int offset, int count)
    {
        const yk_status st = yk_fdk_execute(fdk_, r.projection.data, int(r.projection.location),
            r.projection.element_count, offset, count, r.volume.data, int(r.volume.location),
            r.volume.element_count, r.clear_output ? 1 : 0);
        return st == YK_OK ? true : fail_status(st, "FDK");
    }

    // staging buffers for host-side FP inputs and outputs
    float* device_buffer(void*& p, size_t& have, size_t n)
    {
        if (have < n) {
            if (p) (void)hipFree(p);
            p = nullptr;
            have = 0;
            if (hipMalloc(&p, n * sizeof(float)) != hipSuccess) return nullptr;
            have = n;
        }
        return static_cast<float*>(p);
    }

    bool execute_fp(const SExecutionRequest& r)
    {
        DeviceScope scope(device_);
        if (!scope.ok()) return fail(EApiErrorCode::CudaError, "cannot select the session's device");
        const size_t np = view_elems_ * size_t(views_);
        const float* d_vol = r.volume.data;
        if (r.volume.location == EMemoryLocation::Host) {
            float* s = device_buffer(d_vol_, vol_have_, vol_elems_);
            if (!s || hipMemcpy(s, r.volume.data, vol_elems_ * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
                return fail(EApiErrorCode::CudaError, "volume upload failed");
            d_vol = s;
        }
        float* d_proj = r.projection.data;
        if (r.projection.location == EMemoryLocation::Host) {
            d_proj = device_buffer(d_proj_, proj_have_, np);
            if (!d_proj) return fail(EApiErrorCode::CudaError, "projection staging allocation failed");
        }
        if (yk_status st = yk_project(proj_, forward_kind_, d_vol, d_proj, 0, views_, 0); st != YK_OK)
            return fail_status(st, "forward projection");
        if (r.projection.location == EMemoryLocation::Host &&
            hipMemcpy(r.projection.data, d_proj, np * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            return fail(EApiErrorCode::CudaError, "projection download failed");
        return true;
    }

    // SART / SIRT / OSSART / CGLS (YkSession.cpp:290-350 prepare, :940-1030
    // executeIterative_): the volume holds the start and must be in device
    // memory; SART uses the TIGRE-approximate weights, SIRT the exact ones,
    // OSSART the exact ones in golden-ratio subset order, CGLS the
    // robust-restart strategy; non_negative clamps at zero (not for CGLS).
    // iteration_count > 0 overrides the outer iterations of SIRT and OSSART.
    bool execute_iterative(const SExecutionRequest& r, yk_progress_fn progress, void* progress_user)
    {
        if (r.volume.location != EMemoryLocation::Device ||
            (pipeline_ == EPipeline::CGLS && r.projection.location != EMemoryLocation::Device))
            return fail(EApiErrorCode::CudaError,
                        "iterative pipelines need a device volume (and CGLS a device projection)");
        DeviceScope scope(device_);
        if (!scope.ok()) return fail(EApiErrorCode::CudaError, "cannot select the session's device");
        const size_t np = view_elems_ * size_t(views_);
        const float* d_b = r.projection.data;
        if (r.projection.location == EMemoryLocation::Host) {
            float* s = device_buffer(d_proj_, proj_have_, np);
            if (!s || hipMemcpy(s, r.projection.data, np * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
                return fail(EApiErrorCode::CudaError, "projection upload failed");
            d_b = s;
        }
        const bool sirt_or_os = pipeline_ == EPipeline::SIRT || pipeline_ == EPipeline::OSSART;
        const int iterations = sirt_or_os && r.iteration_count > 0 ? r.iteration_count : iterations_;
        if (pipeline_ == EPipeline::CGLS) {
            yk_cgls_desc d{};
            d.struct_size = sizeof(d);
            d.strategy = YK_CGLS_ROBUST_RESTART;
            d.iterations = iterations;
            d.epsilon = 1e-8f;
            d.restart_on_divergence = 1;
            d.forward_kind = forward_kind_;
            d.back_kind = back_kind_;
            d.check_interval = 1;
            d.patience = 1;
            d.progress_callback = progress;
            d.progress_user_data = progress_user;
            if (yk_status e = yk_cgls_ex(proj_, &d, d_b, r.volume.data, nullptr, nullptr); e != YK_OK)
                return fail(EApiErrorCode::CudaError, std::string("CGLS: ") + yk_last_error());
        } else {
            yk_algebraic_desc d{};
            d.struct_size = sizeof(d);
            d.method = pipeline_ == EPipeline::SART ? YK_ALG_SART : pipeline_ == EPipeline::SIRT ? YK_ALG_SIRT : YK_ALG_OSSART;
            d.weight_model = pipeline_ == EPipeline::SART ? YK_WEIGHTS_TIGRE : YK_WEIGHTS_DETAILED;
            d.subset_order = pipeline_ == EPipeline::OSSART ? YK_ORDER_GOLDEN : YK_ORDER_SEQUENTIAL;
            d.iterations = iterations;
            d.subsets = subsets_;
            d.relaxation = relaxation_;
            d.relaxation_reduction = 1.f;
            d.epsilon = 1e-6f;
            d.use_min = non_negative_ ? 1 : 0;
            d.min_value = 0.f;
            d.forward_kind = forward_kind_;
            d.back_kind = back_kind_;
            d.tv_strength_reduction = 1.f;
            d.check_interval = 1;
            d.patience = 1;
            d.progress_callback = progress;
            d.progress_user_data = progress_user;
            if (yk_status e = yk_algebraic(proj_, &d, d_b, r.volume.data, nullptr); e != YK_OK)
                return fail(EApiErrorCode::CudaError, std::string("algebraic reconstruction: ") + yk_last_error());
        }
        return true;
    }

    void release_handles()
    {
        if (fdk_) (void)yk_fdk_destroy(fdk_);
        if (proj_) (void)yk_projector_destroy(proj_);
        if (d_vol_ || d_proj_) {
            DeviceScope scope(device_);
            if (d_vol_) (void)hipFree(d_vol_);
            if (d_proj_) (void)hipFree(d_proj_);
        }
        if (ctx_) (void)yk_ctx_destroy(ctx_);
        fdk_ = nullptr;
        proj_ = nullptr;
        ctx_ = nullptr;
        d_vol_ = d_proj_ = nullptr;
        vol_have_ = proj_have_ = 0;
    }

    bool fail(EApiErrorCode c, const std::string& m)
    {
        code_ = c;
        message_ = m;
        return false;
    }
    bool fail_status(yk_status st, const char* what)
    {
        return fail(EApiErrorCode(int(st)), std::string(what) + ": " + yk_last_error());
    }
    void clear()
    {
        code_ = EApiErrorCode::None;
        message_.clear();
    }

    yk_ctx* ctx_ = nullptr;
    yk_fdk* fdk_ = nullptr;
    yk_projector* proj_ = nullptr;
    EPipeline pipeline_ = EPipeline::FDK;
    int forward_kind_ = YK_FP_SIDDON;
    int back_kind_ = YK_BP_FDK_PSEUDO_MATCHED;
    int iterations_ = 0, subsets_ = 1;
    float relaxation_ = 1.f;
    bool non_negative_ = true;
    int device_ = 0;
    int views_ = 0;
    int received_ = 0;
    size_t view_elems_ = 0, vol_elems_ = 0;
    void* d_vol_ = nullptr;
    void* d_proj_ = nullptr;
    size_t vol_have_ = 0, proj_have_ = 0;
    bool initialized_ = false;
    EApiErrorCode code_ = EApiErrorCode::None;
    std::string message_;
};

}  // namespace

IReconstructionSession* ReconstructionSessionFactory::create() { return new (std::nothrow) Session(); }

void ReconstructionSessionFactory::destroy(IReconstructionSession* session) { delete session; }

}  // namespace YK

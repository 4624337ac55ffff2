// Scenarios of the reference's Session-API tests run against the Session
// facade of this library (include/yk_session.hpp, libyk_hip.so):
//   dlltest/test_Task_dll.cpp:102-206, 427-505, 522-556 (validation and error
//   codes, forward projection, FDK offline vs streaming, streaming state
//   machine, reset / reuse) and dlltest/test_shepp_logan.cpp:32-158 (DLL FP
//   then DLL FDK of a modified 3-D Shepp-Logan, 47-view streaming;
//   cosine > 0.85, NRMSE < 0.65, streaming max difference < 1e-5), plus the
//   iterative ones (:150-168 pose ignored outside forward projection,
//   :207-244 SIRT smoke, :246-316 SIRT reset / reuse) and checks that each
//   iterative pipeline runs the C-ABI solver with the reference's mapping
//   (src/interface/YkSession.cpp:290-350, 940-1030,
//   src/interface/YkProjectorSelection.hpp:11-44).
// Prints one "PASS <name>" / "FAIL <name>: <why>" line per scenario and
// exits non-zero if any failed.  --host-only runs the scenarios that are
// decided before any device work (they need no GPU).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "yk_hip.h"
#include "yk_session.hpp"

namespace {

using SessionPtr = std::unique_ptr<YK::IReconstructionSession, decltype(&YK::ReconstructionSessionFactory::destroy)>;

SessionPtr make_session()
{
    return SessionPtr(YK::ReconstructionSessionFactory::create(), &YK::ReconstructionSessionFactory::destroy);
}

constexpr int kNu = 32, kNv = 24, kNx = 16, kNy = 16, kNz = 12;

YK::SSystemSpec small_system(YK::EPipeline pipeline, int views)
{
    YK::SSystemSpec s{};
    s.geometry.detector = YK::EDetectorKind::Flat;
    s.geometry.trajectory = YK::ETrajectoryKind::Circular;
    s.geometry.circular.total_views = views;
    s.geometry.circular.views_per_turn = views;
    s.geometry.circular.sid_mm = 500.f;
    s.geometry.circular.sdd_mm = 1000.f;
    s.geometry.flat_detector.channels = kNu;
    s.geometry.flat_detector.rows = kNv;
    s.geometry.flat_detector.channel_size_mm = 1.f;
    s.geometry.flat_detector.row_size_mm = 1.f;
    s.geometry.volume = { kNx, kNy, kNz, 1.f, 1.f, 1.f, make_float3(0.f, 0.f, 0.f) };
    s.reconstruction.pipeline = pipeline;
    return s;
}

// smooth, non-zero, reproducible projections
std::vector<float> synthetic_projection(int views)
{
    std::vector<float> p(size_t(views) * kNu * kNv);
    for (int i = 0; i < views; ++i)
        for (int v = 0; v < kNv; ++v)
            for (int u = 0; u < kNu; ++u)
                p[(size_t(i) * kNv + v) * kNu + u] = 0.25f + 0.5f * float(u + 1) / kNu +
                    0.125f * std::sin(6.2831853f * float(i + 1) / views) * float(v + 1) / kNv;
    return p;
}

// device buffer with RAII
struct DeviceBuffer {
    float* p = nullptr;
    size_t n = 0;
    explicit DeviceBuffer(size_t count) : n(count)
    {
        if (hipMalloc(&p, count * sizeof(float)) != hipSuccess) p = nullptr;
        else (void)hipMemset(p, 0, count * sizeof(float));
    }
    ~DeviceBuffer() { if (p) (void)hipFree(p); }
    std::vector<float> download() const
    {
        std::vector<float> h(n);
        if (p) (void)hipMemcpy(h.data(), p, n * sizeof(float), hipMemcpyDeviceToHost);
        return h;
    }
    void upload(const std::vector<float>& h) { (void)hipMemcpy(p, h.data(), n * sizeof(float), hipMemcpyHostToDevice); }
};

YK::Buffer host(std::vector<float>& v) { return { v.data(), YK::EMemoryLocation::Host, v.size() }; }
YK::Buffer dev(DeviceBuffer& b) { return { b.p, YK::EMemoryLocation::Device, b.n }; }

double max_abs_diff(const std::vector<float>& a, const std::vector<float>& b)
{
    double m = 0.0;
    for (size_t i = 0; i < a.size(); ++i) m = std::max(m, std::abs(double(a[i]) - double(b[i])));
    return m;
}

bool finite_nonzero(const std::vector<float>& v)
{
    double e = 0.0;
    for (float x : v) {
        if (!std::isfinite(x)) return false;
        e += std::abs(double(x));
    }
    return e > 0.0;
}

struct Result {
    bool ok;
    std::string why;
};

Result expect_error(YK::IReconstructionSession& s, bool call_ok, YK::EApiErrorCode code)
{
    if (call_ok) return { false, "call succeeded" };
    if (s.lastError() != code) return { false, "error code " + std::to_string(int(s.lastError())) + ": " + s.lastErrorMessage() };
    if (!s.lastErrorMessage() || !s.lastErrorMessage()[0]) return { false, "empty error message" };
    return { true, "" };
}

// ---------------------------------------------------------------------------
// decided before device work
// ---------------------------------------------------------------------------
Result truncated_system_spec()
{
    auto s = make_session();
    auto spec = small_system(YK::EPipeline::ForwardProjection, 2);
    spec.struct_size = sizeof(std::uint32_t);
    return expect_error(*s, s->initialize(spec), YK::EApiErrorCode::InvalidConfig);
}

Result unsupported_api_version()
{
    auto s = make_session();
    auto spec = small_system(YK::EPipeline::FDK, 32);
    spec.api_version = 999;
    return expect_error(*s, s->initialize(spec), YK::EApiErrorCode::InvalidConfig);
}

Result enum_out_of_range()
{
    auto s = make_session();
    auto spec = small_system(YK::EPipeline::FDK, 32);
    spec.reconstruction.pipeline = static_cast<YK::EPipeline>(99);
    Result r = expect_error(*s, s->initialize(spec), YK::EApiErrorCode::InvalidConfig);
    if (!r.ok) return r;
    spec = small_system(YK::EPipeline::ForwardProjection, 2);
    spec.reconstruction.forward_projector = static_cast<YK::EForwardProjector>(-1);
    return expect_error(*s, s->initialize(spec), YK::EApiErrorCode::InvalidConfig);
}

Result nan_pose_rejected()
{
    auto s = make_session();
    auto spec = small_system(YK::EPipeline::ForwardProjection, 2);
    spec.geometry.forward_projection_pose.enabled = true;
    spec.geometry.forward_projection_pose.rotation_x_rad = NAN;
    return expect_error(*s, s->initialize(spec), YK::EApiErrorCode::InvalidGeometry);
}

Result outside_this_path()
{
    auto s = make_session();
    auto spec = small_system(YK::EPipeline::FDK, 32);
    spec.geometry.trajectory = YK::ETrajectoryKind::Helical;
    Result r = expect_error(*s, s->initialize(spec), YK::EApiErrorCode::UnsupportedCombination);
    if (!r.ok) return r;
    spec = small_system(YK::EPipeline::FDK, 32);
    spec.geometry.detector = YK::EDetectorKind::Cylindrical;
    r = expect_error(*s, s->initialize(spec), YK::EApiErrorCode::UnsupportedCombination);
    if (!r.ok) return r;
    spec = small_system(YK::EPipeline::PWLS, 32);
    r = expect_error(*s, s->initialize(spec), YK::EApiErrorCode::UnsupportedCombination);
    if (!r.ok) return r;
    spec = small_system(YK::EPipeline::SART, 32);
    spec.reconstruction.back_projector = YK::EBackProjector::SeparableFootprint;
    r = expect_error(*s, s->initialize(spec), YK::EApiErrorCode::UnsupportedCombination);
    if (!r.ok) return r;
    spec = small_system(YK::EPipeline::ForwardProjection, 2);
    spec.reconstruction.forward_projector = YK::EForwardProjector::SeparableFootprint;
    return expect_error(*s, s->initialize(spec), YK::EApiErrorCode::UnsupportedCombination);
}

Result invalid_scan_rejected()
{
    auto s = make_session();
    auto spec = small_system(YK::EPipeline::FDK, 32);
    spec.geometry.circular.sdd_mm = 100.f;  // detector inside the source circle
    return expect_error(*s, s->initialize(spec), YK::EApiErrorCode::InvalidGeometry);
}

Result execute_before_initialize()
{
    auto s = make_session();
    float x = 0.f;
    YK::SExecutionRequest r{};
    r.projection = { &x, YK::EMemoryLocation::Host, 1 };
    r.volume = { &x, YK::EMemoryLocation::Host, 1 };
    return expect_error(*s, s->execute(r), YK::EApiErrorCode::StreamingStateError);
}

// ---------------------------------------------------------------------------
// device scenarios
// ---------------------------------------------------------------------------
Result forward_projection(YK::EForwardProjector kind)
{
    auto s = make_session();
    auto spec = small_system(YK::EPipeline::ForwardProjection, 8);
    spec.reconstruction.forward_projector = kind;
    if (!s->initialize(spec)) return { false, s->lastErrorMessage() };
    std::vector<float> vol(size_t(kNx) * kNy * kNz, 0.01f);
    DeviceBuffer dvol(vol.size()), dproj(size_t(8) * kNu * kNv);
    dvol.upload(vol);
    YK::SExecutionRequest r{};
    r.view_count = 8;
    r.volume = dev(dvol);
    r.projection = dev(dproj);
    if (!s->execute(r)) return { false, s->lastErrorMessage() };
    const auto p = dproj.download();
    // the same call with host buffers gives the same projections
    std::vector<float> hp(p.size(), -1.f);
    r.volume = host(vol);
    r.projection = host(hp);
    if (!s->execute(r)) return { false, s->lastErrorMessage() };
    if (!finite_nonzero(p)) return { false, "projection is zero or non-finite" };
    if (max_abs_diff(p, hp) != 0.0) return { false, "host and device paths differ" };
    return { true, "" };
}

Result execution_validation()
{
    auto s = make_session();
    if (!s->initialize(small_system(YK::EPipeline::ForwardProjection, 2))) return { false, s->lastErrorMessage() };
    YK::SExecutionRequest r{};
    r.projection = { reinterpret_cast<float*>(1), YK::EMemoryLocation::Device, 1 };
    r.volume = { reinterpret_cast<float*>(1), YK::EMemoryLocation::Device, 1 };
    Result res = expect_error(*s, s->execute(r), YK::EApiErrorCode::InvalidBuffer);
    if (!res.ok) return res;
    r.projection.element_count = size_t(2) * kNu * kNv;
    r.volume.element_count = size_t(kNx) * kNy * kNz;
    r.view_offset = 1;
    r.view_count = 1;
    return expect_error(*s, s->execute(r), YK::EApiErrorCode::StreamingStateError);
}

Result fdk_offline_vs_streaming()
{
    constexpr int views = 32, first = 13;
    const auto spec = small_system(YK::EPipeline::FDK, views);
    auto offline = make_session(), streaming = make_session();
    if (!offline->initialize(spec)) return { false, offline->lastErrorMessage() };
    if (!streaming->initialize(spec)) return { false, streaming->lastErrorMessage() };
    auto proj = synthetic_projection(views);
    const size_t ve = size_t(kNu) * kNv, vn = size_t(kNx) * kNy * kNz;
    DeviceBuffer a(vn), b(vn);
    YK::SExecutionRequest r{};
    r.view_count = views;
    r.projection = host(proj);
    r.volume = dev(a);
    if (!offline->execute(r)) return { false, offline->lastErrorMessage() };
    YK::SExecutionRequest r1{};
    r1.view_offset = 0;
    r1.view_count = first;
    r1.clear_output = true;
    r1.projection = { proj.data(), YK::EMemoryLocation::Host, size_t(first) * ve };
    r1.volume = dev(b);
    YK::SExecutionRequest r2 = r1;
    r2.view_offset = first;
    r2.view_count = views - first;
    r2.clear_output = false;
    r2.projection = { proj.data() + size_t(first) * ve, YK::EMemoryLocation::Host, size_t(views - first) * ve };
    if (!streaming->execute(r1) || !streaming->execute(r2)) return { false, streaming->lastErrorMessage() };
    const auto va = a.download(), vb = b.download();
    if (!finite_nonzero(va)) return { false, "zero or non-finite volume" };
    const double d = max_abs_diff(va, vb);
    if (!(d <= 1e-5)) return { false, "offline / streaming max difference " + std::to_string(d) };
    return { true, "" };
}

Result fdk_streaming_state_machine()
{
    constexpr int views = 32;
    auto s = make_session();
    if (!s->initialize(small_system(YK::EPipeline::FDK, views))) return { false, s->lastErrorMessage() };
    auto proj = synthetic_projection(views);
    const size_t ve = size_t(kNu) * kNv;
    std::vector<float> vol(size_t(kNx) * kNy * kNz);
    YK::SExecutionRequest r{};
    r.projection = { proj.data(), YK::EMemoryLocation::Host, 8 * ve };
    r.volume = host(vol);
    r.view_offset = 4;  // skips views 0..3
    r.view_count = 8;
    Result res = expect_error(*s, s->execute(r), YK::EApiErrorCode::StreamingStateError);
    if (!res.ok) return res;
    r.view_offset = 0;
    if (!s->execute(r)) return { false, s->lastErrorMessage() };
    r.view_offset = 8;  // a later batch may not clear the output
    r.clear_output = true;
    res = expect_error(*s, s->execute(r), YK::EApiErrorCode::StreamingStateError);
    if (!res.ok) return res;
    r.clear_output = false;
    r.projection.element_count = 8 * ve - 1;  // undersized projection
    res = expect_error(*s, s->execute(r), YK::EApiErrorCode::InvalidBuffer);
    if (!res.ok) return res;
    r.projection.element_count = 8 * ve;
    r.view_count = 40;  // beyond the prepared views
    res = expect_error(*s, s->execute(r), YK::EApiErrorCode::InvalidConfig);
    if (!res.ok) return res;
    // a host volume is written once the last view has arrived
    std::fill(vol.begin(), vol.end(), -7.f);
    r.view_count = views - 8;
    r.projection = { proj.data() + 8 * ve, YK::EMemoryLocation::Host, size_t(views - 8) * ve };
    int progress_done = -1, progress_total = -1;
    auto cb = [](void* user, int done, int total) {
        int* p = static_cast<int*>(user);
        p[0] = done;
        p[1] = total;
    };
    int pr[2] = { -1, -1 };
    r.progress_callback = cb;
    r.progress_user_data = pr;
    if (!s->execute(r)) return { false, s->lastErrorMessage() };
    progress_done = pr[0];
    progress_total = pr[1];
    // FDK has no iterations: the reference's backends report nothing
    if (progress_done != -1 || progress_total != -1) return { false, "FDK called the progress callback" };
    if (!finite_nonzero(vol) || vol[0] == -7.f) return { false, "host volume not written at completion" };
    // complete: a further batch is out of range
    r.view_offset = views;
    r.view_count = 1;
    return expect_error(*s, s->execute(r), YK::EApiErrorCode::InvalidConfig);
}

Result session_reset_and_reuse()
{
    constexpr int views = 32;
    auto s = make_session();
    const auto spec = small_system(YK::EPipeline::FDK, views);
    if (!s->initialize(spec)) return { false, s->lastErrorMessage() };
    auto proj = synthetic_projection(views);
    std::vector<float> v1(size_t(kNx) * kNy * kNz), v2(v1.size()), v3(v1.size());
    YK::SExecutionRequest r{};
    r.projection = host(proj);
    r.volume = host(v1);
    if (!s->execute(r)) return { false, s->lastErrorMessage() };
    s->reset();
    r.volume = host(v2);
    if (!s->execute(r)) return { false, s->lastErrorMessage() };
    s->release();
    if (s->isInitialized()) return { false, "still initialized after release" };
    if (!s->initialize(spec)) return { false, s->lastErrorMessage() };
    r.volume = host(v3);
    if (!s->execute(r)) return { false, s->lastErrorMessage() };
    if (max_abs_diff(v1, v2) != 0.0 || max_abs_diff(v1, v3) != 0.0) return { false, "repeated runs differ" };
    return { true, "" };
}

// modified 3-D Shepp-Logan (Kak & Slaney table), head half-extent 20 mm
std::vector<float> shepp_logan(const YK::SVolumeGridSpec& g)
{
    struct E { double x, y, z, a, b, c, deg, val; };
    const E table[] = {
        { 0, 0, 0, .69, .92, .81, 0, 1 }, { 0, -.0184, 0, .6624, .874, .78, 0, -.8 },
        { .22, 0, 0, .11, .31, .22, -18, -.2 }, { -.22, 0, 0, .16, .41, .28, 18, -.2 },
        { 0, .35, 0, .21, .25, .41, 0, .1 }, { 0, .1, 0, .046, .046, .05, 0, .1 },
        { 0, -.1, 0, .046, .046, .05, 0, .1 }, { -.08, -.605, 0, .046, .023, .05, 0, .1 },
        { 0, -.605, 0, .023, .023, .02, 0, .1 }, { .06, -.605, 0, .023, .046, .02, 0, .1 },
    };
    std::vector<float> out(size_t(g.nx) * g.ny * g.nz);
    for (int z = 0; z < g.nz; ++z)
        for (int y = 0; y < g.ny; ++y)
            for (int x = 0; x < g.nx; ++x) {
                const double px = (x + .5 - .5 * g.nx) * g.voxel_x_mm / 20.0;
                const double py = (y + .5 - .5 * g.ny) * g.voxel_y_mm / 20.0;
                const double pz = (z + .5 - .5 * g.nz) * g.voxel_z_mm / 20.0;
                double val = 0.0;
                for (const E& e : table) {
                    const double t = e.deg * 3.14159265358979323846 / 180.0;
                    const double dx = px - e.x, dy = py - e.y, dz = pz - e.z;
                    const double rx = dx * std::cos(t) - dy * std::sin(t), ry = dx * std::sin(t) + dy * std::cos(t);
                    if (rx * rx / (e.a * e.a) + ry * ry / (e.b * e.b) + dz * dz / (e.c * e.c) <= 1.0) val += e.val;
                }
                out[(size_t(z) * g.ny + y) * g.nx + x] = float(std::max(val, 0.0) * 0.02);
            }
    return out;
}

Result shepp_logan_fp_then_fdk()
{
    YK::SSystemSpec spec{};
    auto& g = spec.geometry;
    g.circular.total_views = 360;
    g.circular.views_per_turn = 360;
    g.circular.sid_mm = 200.f;
    g.circular.sdd_mm = 400.f;
    g.flat_detector.channels = 128;
    g.flat_detector.rows = 128;
    g.flat_detector.channel_size_mm = 0.8f;
    g.flat_detector.row_size_mm = 0.8f;
    g.volume = { 64, 64, 64, 0.75f, 0.75f, 0.75f, make_float3(0.f, 0.f, 0.f) };
    spec.reconstruction.pipeline = YK::EPipeline::FDK;
    spec.reconstruction.fdk.filter.type = YK::EFdkFilter::RamLak;
    auto truth = shepp_logan(g.volume);
    const int views = 360;
    const size_t ve = size_t(128) * 128;
    std::vector<float> proj(ve * views), rec(truth.size()), streamed(truth.size());

    auto fp = make_session();
    auto fp_spec = spec;
    fp_spec.reconstruction.pipeline = YK::EPipeline::ForwardProjection;
    if (!fp->initialize(fp_spec)) return { false, fp->lastErrorMessage() };
    YK::SExecutionRequest r{};
    r.volume = host(truth);
    r.projection = host(proj);
    r.view_count = views;
    if (!fp->execute(r)) return { false, fp->lastErrorMessage() };
    fp.reset();

    auto fdk = make_session();
    if (!fdk->initialize(spec)) return { false, fdk->lastErrorMessage() };
    r.volume = host(rec);
    if (!fdk->execute(r)) return { false, fdk->lastErrorMessage() };
    fdk->reset();
    for (int off = 0; off < views; off += 47) {
        const int n = std::min(47, views - off);
        YK::SExecutionRequest b{};
        b.projection = { proj.data() + size_t(off) * ve, YK::EMemoryLocation::Host, size_t(n) * ve };
        b.volume = host(streamed);
        b.view_offset = off;
        b.view_count = n;
        b.clear_output = off == 0;
        if (!fdk->execute(b)) return { false, fdk->lastErrorMessage() };
    }
    double tt = 0, rr = 0, tr = 0, e2 = 0;
    for (size_t i = 0; i < truth.size(); ++i) {
        if (!std::isfinite(rec[i]) || !std::isfinite(streamed[i])) return { false, "non-finite reconstruction" };
        const double t = truth[i], x = rec[i];
        tt += t * t; rr += x * x; tr += t * x; e2 += (x - t) * (x - t);
    }
    const double cosine = tr / std::sqrt(std::max(tt * rr, 1e-30)), nrmse = std::sqrt(e2 / std::max(tt, 1e-30));
    const double sd = max_abs_diff(rec, streamed);
    std::printf("shepp-logan: cosine %.6f nrmse %.6f streaming max difference %.3g\n", cosine, nrmse, sd);
    if (!(cosine > 0.85 && nrmse < 0.65 && sd < 1e-5)) return { false, "quality gates" };
    return { true, "" };
}

// The pose rotates the volume about its centre: a quarter turn about Z maps
// voxel (x, y) of the rotated volume to (y, N-1-x) of the stored one, so the
// projections of the stored volume with the pose equal those of the
// index-permuted volume without it.
Result forward_projection_pose_quarter_turn()
{
    constexpr int n = 16, views = 12;
    auto spec = small_system(YK::EPipeline::ForwardProjection, views);
    spec.geometry.volume = { n, n, n, 1.f, 1.f, 1.f, make_float3(0.f, 0.f, 0.f) };
    std::vector<float> v(size_t(n) * n * n), rot(v.size());
    unsigned s = 12345u;
    for (float& x : v) {
        s = s * 1664525u + 1013904223u;
        x = float((s >> 8) & 0xffff) / 65536.f;
    }
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x)
                rot[(size_t(z) * n + y) * n + x] = v[(size_t(z) * n + (n - 1 - x)) * n + y];
    std::vector<float> pa(size_t(views) * kNu * kNv), pb(pa.size());
    auto a = make_session(), b = make_session();
    auto posed = spec;
    posed.geometry.forward_projection_pose.enabled = true;
    posed.geometry.forward_projection_pose.rotation_z_rad = 1.5707963267948966f;
    if (!a->initialize(posed)) return { false, a->lastErrorMessage() };
    if (!b->initialize(spec)) return { false, b->lastErrorMessage() };
    YK::SExecutionRequest r{};
    r.volume = host(v);
    r.projection = host(pa);
    if (!a->execute(r)) return { false, a->lastErrorMessage() };
    r.volume = host(rot);
    r.projection = host(pb);
    if (!b->execute(r)) return { false, b->lastErrorMessage() };
    double num = 0, den = 0;
    for (size_t i = 0; i < pa.size(); ++i) {
        num += (double(pa[i]) - pb[i]) * (double(pa[i]) - pb[i]);
        den += double(pb[i]) * pb[i];
    }
    const double rel = std::sqrt(num / std::max(den, 1e-30));
    if (!(den > 0 && rel < 1e-4)) return { false, "posed vs permuted projections differ: rel " + std::to_string(rel) };
    return { true, "" };
}

// ---------------------------------------------------------------------------
// iterative pipelines
// ---------------------------------------------------------------------------
YK::SSystemSpec iterative_system(YK::EPipeline pipeline, int views, int iterations, float relaxation, int subsets)
{
    auto spec = small_system(pipeline, views);
    spec.geometry.circular.sid_mm = 80.f;
    spec.geometry.circular.sdd_mm = 160.f;
    spec.reconstruction.iterative.iterations = iterations;
    spec.reconstruction.iterative.relaxation = relaxation;
    spec.reconstruction.iterative.subsets = subsets;
    return spec;
}

// dlltest/test_Task_dll.cpp:150-168: the forward-projection pose is ignored
// by other pipelines, and a NaN pose is still an InvalidGeometry
Result iterative_pose_ignored()
{
    auto s = make_session();
    auto spec = iterative_system(YK::EPipeline::SIRT, 32, 1, 0.2f, 1);
    spec.geometry.forward_projection_pose.enabled = true;
    spec.geometry.forward_projection_pose.rotation_x_rad = 0.7f;
    if (!s->initialize(spec)) return { false, s->lastErrorMessage() };
    spec.geometry.forward_projection_pose.rotation_x_rad = NAN;
    return expect_error(*s, s->initialize(spec), YK::EApiErrorCode::InvalidGeometry);
}

// dlltest/test_Task_dll.cpp:207-244: SIRT, one iteration, zero data
Result sirt_zero_data()
{
    auto s = make_session();
    auto spec = iterative_system(YK::EPipeline::SIRT, 32, 1, 0.2f, 1);
    if (!s->initialize(spec)) return { false, s->lastErrorMessage() };
    DeviceBuffer vol(size_t(kNx) * kNy * kNz), proj(size_t(32) * kNu * kNv);
    YK::SExecutionRequest r{};
    r.view_count = 32;
    r.volume = dev(vol);
    r.projection = dev(proj);
    if (!s->execute(r)) return { false, s->lastErrorMessage() };
    for (float x : vol.download())
        if (x != 0.f) return { false, "zero data gave a non-zero volume" };
    return { true, "" };
}

// dlltest/test_Task_dll.cpp:246-316: SIRT twice with reset in between, from
// a cleared volume each time, gives the same result (max difference 1e-5)
Result sirt_reset_reuse()
{
    constexpr int views = 24;
    auto s = make_session();
    auto spec = iterative_system(YK::EPipeline::SIRT, views, 2, 0.2f, 1);
    if (!s->initialize(spec)) return { false, s->lastErrorMessage() };
    auto hp = synthetic_projection(views);
    DeviceBuffer vol(size_t(kNx) * kNy * kNz), proj(hp.size());
    proj.upload(hp);
    YK::SExecutionRequest r{};
    r.volume = dev(vol);
    r.projection = dev(proj);
    if (!s->execute(r)) return { false, s->lastErrorMessage() };
    const auto first = vol.download();
    s->reset();
    (void)hipMemset(vol.p, 0, vol.n * sizeof(float));
    if (!s->execute(r)) return { false, s->lastErrorMessage() };
    const auto second = vol.download();
    if (!finite_nonzero(first)) return { false, "empty reconstruction" };
    const double d = max_abs_diff(first, second);
    if (!(d <= 1e-5)) return { false, "reset changed the result by " + std::to_string(d) };
    return { true, "" };
}

// Each pipeline equals the C-ABI solver called with the reference's mapping:
// SART -> TIGRE weights, SIRT -> exact weights, OSSART -> exact weights in
// golden-ratio order, CGLS -> robust restart; Automatic -> Siddon forward and
// the FDK pseudo-matched back (CGLS: the Siddon transpose); non_negative -> a
// lower bound of 0 (not for CGLS).  Bit-identical: the same code runs.
Result iterative_mapping()
{
    constexpr int views = 12;
    const size_t nvol = size_t(kNx) * kNy * kNz;
    auto hp = synthetic_projection(views);
    DeviceBuffer proj(hp.size());
    proj.upload(hp);
    yk_ctx* ctx = nullptr;
    if (yk_ctx_create(0, &ctx) != YK_OK) return { false, "context" };
    std::unique_ptr<yk_ctx, yk_status (*)(yk_ctx*)> ctx_guard(ctx, yk_ctx_destroy);
    yk_system sys{};
    sys.scan.total_views = views;
    sys.scan.views_per_turn = views;
    sys.scan.rotation_direction = 1;
    sys.scan.sid_mm = 80.f;
    sys.scan.sdd_mm = 160.f;
    sys.detector.channels = kNu;
    sys.detector.rows = kNv;
    sys.detector.channel_size_mm = 1.f;
    sys.detector.row_size_mm = 1.f;
    sys.volume = { kNx, kNy, kNz, 1.f, 1.f, 1.f, { 0.f, 0.f, 0.f } };
    std::vector<yk_view> vw(views);
    int n = views;
    if (yk_geometry_build(&sys, vw.data(), views, &n) != YK_OK) return { false, yk_last_error() };
    yk_projector* pr = nullptr;
    if (yk_projector_create(ctx, vw.data(), views, kNu, kNv, &sys.volume, 6.2831853f, &pr) != YK_OK)
        return { false, yk_last_error() };
    std::unique_ptr<yk_projector, yk_status (*)(yk_projector*)> pr_guard(pr, yk_projector_destroy);

    struct Case {
        YK::EPipeline pipeline;
        YK::EBackProjector back;
        int method, weights, order, bk;
    };
    const Case cases[] = {
        { YK::EPipeline::SART, YK::EBackProjector::Automatic, YK_ALG_SART, YK_WEIGHTS_TIGRE, YK_ORDER_SEQUENTIAL,
          YK_BP_FDK_PSEUDO_MATCHED },
        { YK::EPipeline::SIRT, YK::EBackProjector::JosephV3, YK_ALG_SIRT, YK_WEIGHTS_DETAILED, YK_ORDER_SEQUENTIAL,
          YK_BP_JOSEPH_VOXEL },
        { YK::EPipeline::OSSART, YK::EBackProjector::SiddonRayDriven, YK_ALG_OSSART, YK_WEIGHTS_DETAILED,
          YK_ORDER_GOLDEN, YK_BP_SIDDON_ADJOINT },
        { YK::EPipeline::CGLS, YK::EBackProjector::Automatic, -1, 0, 0, YK_BP_SIDDON_ADJOINT },
    };
    for (const Case& c : cases) {
        auto s = make_session();
        auto spec = iterative_system(c.pipeline, views, 3, 0.7f, 4);
        spec.reconstruction.back_projector = c.back;
        if (!s->initialize(spec)) return { false, s->lastErrorMessage() };
        DeviceBuffer a(nvol), b(nvol);
        YK::SExecutionRequest r{};
        r.volume = dev(a);
        r.projection = dev(proj);
        if (!s->execute(r)) return { false, s->lastErrorMessage() };
        yk_status st;
        if (c.method < 0) {
            yk_cgls_desc d{};
            d.struct_size = sizeof(d);
            d.strategy = YK_CGLS_ROBUST_RESTART;
            d.iterations = 3;
            d.epsilon = 1e-8f;
            d.restart_on_divergence = 1;
            d.forward_kind = YK_FP_SIDDON;
            d.back_kind = c.bk;
            d.check_interval = 1;
            d.patience = 1;
            st = yk_cgls_ex(pr, &d, proj.p, b.p, nullptr, nullptr);
        } else {
            yk_algebraic_desc d{};
            d.struct_size = sizeof(d);
            d.method = c.method;
            d.weight_model = c.weights;
            d.subset_order = c.order;
            d.iterations = 3;
            d.subsets = 4;
            d.relaxation = 0.7f;
            d.relaxation_reduction = 1.f;
            d.epsilon = 1e-6f;
            d.use_min = 1;
            d.forward_kind = YK_FP_SIDDON;
            d.back_kind = c.bk;
            d.tv_strength_reduction = 1.f;
            d.check_interval = 1;
            d.patience = 1;
            st = yk_algebraic(pr, &d, proj.p, b.p, nullptr);
        }
        if (st != YK_OK) return { false, yk_last_error() };
        const auto va = a.download(), vb = b.download();
        if (!finite_nonzero(va)) return { false, "empty reconstruction" };
        if (max_abs_diff(va, vb) != 0.0)
            return { false, "pipeline " + std::to_string(int(c.pipeline)) + " differs from the C-ABI call" };
    }
    return { true, "" };
}

// buffer rules: device volume required (CudaError otherwise), CGLS also a
// device projection; SART / SIRT / OSSART accept a host projection with the
// same result; the whole sequence from view 0; iteration_count overrides the
// SIRT / OSSART iterations; bad parameters fail initialisation
Result iterative_requests()
{
    constexpr int views = 12;
    const size_t nvol = size_t(kNx) * kNy * kNz;
    auto hp = synthetic_projection(views);
    DeviceBuffer proj(hp.size()), a(nvol), b(nvol);
    proj.upload(hp);
    std::vector<float> hv(nvol, 0.f);
    auto s = make_session();
    if (!s->initialize(iterative_system(YK::EPipeline::OSSART, views, 2, 0.5f, 3))) return { false, s->lastErrorMessage() };
    YK::SExecutionRequest r{};
    r.projection = dev(proj);
    r.volume = host(hv);
    Result e = expect_error(*s, s->execute(r), YK::EApiErrorCode::CudaError);
    if (!e.ok) return e;
    r.volume = dev(a);
    r.view_offset = 1;
    r.view_count = views - 1;
    e = expect_error(*s, s->execute(r), YK::EApiErrorCode::StreamingStateError);
    if (!e.ok) return e;
    r.view_offset = 0;
    r.view_count = 0;
    if (!s->execute(r)) return { false, s->lastErrorMessage() };
    r.projection = host(hp);
    r.volume = dev(b);
    if (!s->execute(r)) return { false, s->lastErrorMessage() };
    if (max_abs_diff(a.download(), b.download()) != 0.0) return { false, "host projection changed the result" };
    // iteration_count = 5 equals a session built with 5 iterations
    (void)hipMemset(a.p, 0, nvol * sizeof(float));
    (void)hipMemset(b.p, 0, nvol * sizeof(float));
    r.projection = dev(proj);
    r.volume = dev(a);
    r.iteration_count = 5;
    if (!s->execute(r)) return { false, s->lastErrorMessage() };
    auto s5 = make_session();
    if (!s5->initialize(iterative_system(YK::EPipeline::OSSART, views, 5, 0.5f, 3))) return { false, s5->lastErrorMessage() };
    r.volume = dev(b);
    r.iteration_count = 0;
    if (!s5->execute(r)) return { false, s5->lastErrorMessage() };
    if (max_abs_diff(a.download(), b.download()) != 0.0) return { false, "iteration_count did not override" };
    // CGLS needs a device projection
    auto c = make_session();
    if (!c->initialize(iterative_system(YK::EPipeline::CGLS, views, 2, 1.f, 1))) return { false, c->lastErrorMessage() };
    r.projection = host(hp);
    e = expect_error(*c, c->execute(r), YK::EApiErrorCode::CudaError);
    if (!e.ok) return e;
    // bad parameters
    auto bad = make_session();
    e = expect_error(*bad, bad->initialize(iterative_system(YK::EPipeline::SIRT, views, 0, 0.5f, 1)),
                     YK::EApiErrorCode::InvalidGeometry);
    if (!e.ok) return e;
    return expect_error(*bad, bad->initialize(iterative_system(YK::EPipeline::OSSART, views, 2, 0.5f, views + 1)),
                        YK::EApiErrorCode::InvalidGeometry);
}

// ---------------------------------------------------------------------------
// external flat frames and iteration progress
// (example/multispectrum_sim/tests/recon/backend/test_reconstruction_progress.cpp,
// src/interface/YkReconstructionApi.cpp:146-167, src/common/YkExecutionContext.hpp:22-36)
// ---------------------------------------------------------------------------
// the builder's views of `spec` as SConeProjGeomVec frames
std::vector<YK::SConeProjGeomVec> nominal_frames(const YK::SSystemSpec& spec)
{
    const YK::SSystemConfig& g = spec.geometry;
    yk_system sys{};
    sys.scan.total_views = g.circular.total_views;
    sys.scan.views_per_turn = g.circular.views_per_turn;
    sys.scan.start_angle_rad = g.circular.start_angle_rad;
    sys.scan.rotation_direction = g.circular.rotation_direction;
    sys.scan.sid_mm = g.circular.sid_mm;
    sys.scan.sdd_mm = g.circular.sdd_mm;
    sys.scan.z_mm = g.circular.z_mm;
    sys.detector.channels = g.flat_detector.channels;
    sys.detector.rows = g.flat_detector.rows;
    sys.detector.channel_size_mm = g.flat_detector.channel_size_mm;
    sys.detector.row_size_mm = g.flat_detector.row_size_mm;
    sys.volume.nx = g.volume.nx;
    sys.volume.ny = g.volume.ny;
    sys.volume.nz = g.volume.nz;
    sys.volume.voxel_x_mm = g.volume.voxel_x_mm;
    sys.volume.voxel_y_mm = g.volume.voxel_y_mm;
    sys.volume.voxel_z_mm = g.volume.voxel_z_mm;
    int n = 0;
    (void)yk_geometry_build(&sys, nullptr, 0, &n);
    std::vector<yk_view> v(size_t(std::max(n, 0)));
    if (n <= 0 || yk_geometry_build(&sys, v.data(), n, &n) != YK_OK) return {};
    std::vector<YK::SConeProjGeomVec> f(v.size());
    for (size_t i = 0; i < v.size(); ++i) {
        f[i].src = make_float4(v[i].src[0], v[i].src[1], v[i].src[2], v[i].src[3]);
        f[i].detS = make_float4(v[i].detS[0], v[i].detS[1], v[i].detS[2], v[i].detS[3]);
        f[i].detU = make_float4(v[i].detU[0], v[i].detU[1], v[i].detU[2], v[i].detU[3]);
        f[i].detV = make_float4(v[i].detV[0], v[i].detV[1], v[i].detV[2], v[i].detV[3]);
        f[i].angle = make_float4(v[i].angle[0], v[i].angle[1], v[i].angle[2], v[i].angle[3]);
    }
    return f;
}

Result flat_frame_tail_validation()
{
    auto s = make_session();
    auto spec = iterative_system(YK::EPipeline::SART, 12, 2, 0.5f, 1);
    auto frames = nominal_frames(spec);
    if (frames.size() != 12) return { false, "nominal frames" };
    // the tail is cut inside the count
    spec.flat_frames = frames.data();
    spec.flat_frame_count = 12;
    spec.struct_size = uint32_t(offsetof(YK::SSystemSpec, flat_frames) + sizeof(void*));
    Result r = expect_error(*s, s->initialize(spec), YK::EApiErrorCode::InvalidConfig);
    if (!r.ok) return { false, "step 1: " + r.why };
    spec.struct_size = sizeof(spec);
    // buffer / count mismatches
    spec.flat_frames = nullptr;
    r = expect_error(*s, s->initialize(spec), YK::EApiErrorCode::InvalidConfig);
    if (!r.ok) return { false, "step 2: " + r.why };
    spec.flat_frames = frames.data();
    spec.flat_frame_count = 0;
    r = expect_error(*s, s->initialize(spec), YK::EApiErrorCode::InvalidConfig);
    if (!r.ok) return { false, "step 3: " + r.why };
    spec.flat_frame_count = -1;
    r = expect_error(*s, s->initialize(spec), YK::EApiErrorCode::InvalidConfig);
    if (!r.ok) return { false, "step 4: " + r.why };
    // pipelines without external-frame support
    spec.flat_frame_count = 12;
    spec.reconstruction.pipeline = YK::EPipeline::PWLS;
    r = expect_error(*s, s->initialize(spec), YK::EApiErrorCode::UnsupportedCombination);
    if (!r.ok) return { false, "step 5: " + r.why };
    // one frame per view of the scan
    spec.reconstruction.pipeline = YK::EPipeline::SART;
    spec.flat_frame_count = 11;
    r = expect_error(*s, s->initialize(spec), YK::EApiErrorCode::InvalidGeometry);
    if (!r.ok) return { false, "step 6: " + r.why };
    // non-orthogonal detector steps
    spec.flat_frame_count = 12;
    frames[3].detV.x += 0.01f;
    frames[3].detV.y += 0.01f;
    r = expect_error(*s, s->initialize(spec), YK::EApiErrorCode::InvalidGeometry);
    if (!r.ok) return { false, "step 7: " + r.why };
    frames[3].detU.y = NAN;
    r = expect_error(*s, s->initialize(spec), YK::EApiErrorCode::InvalidGeometry);
    if (!r.ok) return { false, "step 8: " + r.why };
    // FDK: the frames must be the nominal trajectory within 1e-3
    auto fspec = small_system(YK::EPipeline::FDK, 12);
    auto fframes = nominal_frames(fspec);
    fspec.flat_frames = fframes.data();
    fspec.flat_frame_count = 11;
    r = expect_error(*s, s->initialize(fspec), YK::EApiErrorCode::InvalidGeometry);
    if (!r.ok) return { false, "step 9: " + r.why };
    fspec.flat_frame_count = 12;
    fframes[0].src.z += 2.f;  // the reference's "noncircular frame"
    r = expect_error(*s, s->initialize(fspec), YK::EApiErrorCode::UnsupportedCombination);
    if (!r.ok) return { false, "step 10: " + r.why };
    fframes[0].src.z -= 2.f;
    fframes[5].detS.x += 0.002f;
    return expect_error(*s, s->initialize(fspec), YK::EApiErrorCode::UnsupportedCombination);
}

// projection of 12 views for the invariance runs
std::vector<float> progress_projection(int views)
{
    std::vector<float> p(size_t(views) * kNu * kNv, 0.2f);
    for (size_t i = 0; i < p.size(); ++i) p[i] += 0.003f * float(i % 17);
    return p;
}

// run `spec` on the given projection into a fresh volume: iterative
// pipelines take a device projection, FDK streams single-view host batches
bool run_pipeline(const YK::SSystemSpec& spec, std::vector<float>& proj, std::vector<float>& out, std::string& why)
{
    const int views = spec.geometry.circular.total_views;
    const size_t ve = size_t(kNu) * kNv, vn = size_t(kNx) * kNy * kNz;
    auto s = make_session();
    if (!s->initialize(spec)) {
        why = std::string("initialize: ") + s->lastErrorMessage();
        return false;
    }
    DeviceBuffer vol(vn), dproj(proj.size());
    dproj.upload(proj);
    YK::SExecutionRequest r{};
    r.volume = dev(vol);
    if (spec.reconstruction.pipeline == YK::EPipeline::FDK) {
        for (int f = 0; f < views; ++f) {
            r.projection = { proj.data() + size_t(f) * ve, YK::EMemoryLocation::Host, ve };
            r.view_offset = f;
            r.view_count = 1;
            r.clear_output = f == 0;
            if (!s->execute(r)) {
                why = std::string("single-view FDK: ") + s->lastErrorMessage();
                return false;
            }
        }
    } else {
        r.view_count = views;
        r.projection = dev(dproj);
        if (!s->execute(r)) {
            why = std::string("execute: ") + s->lastErrorMessage();
            return false;
        }
    }
    out = vol.download();
    return true;
}

// explicit frames equal to the nominal trajectory leave every pipeline's
// result unchanged (the reference's 1e-5 bound)
Result external_frames_invariance()
{
    constexpr int views = 12;
    for (auto p : { YK::EPipeline::SART, YK::EPipeline::SIRT, YK::EPipeline::OSSART, YK::EPipeline::CGLS,
                    YK::EPipeline::FDK }) {
        auto spec = iterative_system(p, views, 4, 0.5f, p == YK::EPipeline::OSSART ? 3 : 1);
        if (p == YK::EPipeline::FDK) spec = small_system(p, views);
        auto proj = progress_projection(views);
        std::vector<float> base, observed;
        std::string why;
        if (!run_pipeline(spec, proj, base, why)) return { false, why };
        auto frames = nominal_frames(spec);
        spec.flat_frames = frames.data();
        spec.flat_frame_count = int(frames.size());
        if (!run_pipeline(spec, proj, observed, why)) return { false, why };
        if (!finite_nonzero(observed)) return { false, "non-finite or zero result with frames" };
        if (max_abs_diff(base, observed) >= 1e-5)
            return { false, "explicit nominal frames changed pipeline " + std::to_string(int(p)) };
    }
    return { true, "" };
}

// calibrated frames drive the forward projection: the Session's projections
// equal the C ABI's projection with the same views
Result external_frames_forward_projection()
{
    constexpr int views = 8;
    auto spec = small_system(YK::EPipeline::ForwardProjection, views);
    spec.reconstruction.forward_projector = YK::EForwardProjector::Joseph;
    auto frames = nominal_frames(spec);
    for (int i = 0; i < views; ++i) {
        frames[size_t(i)].src.x += 0.3f * std::sin(1.7f * float(i));  // per-view source wobble
        frames[size_t(i)].detS.z += 0.2f * float(i % 3);
    }
    spec.flat_frames = frames.data();
    spec.flat_frame_count = views;
    auto s = make_session();
    if (!s->initialize(spec)) return { false, s->lastErrorMessage() };
    std::vector<float> vol(size_t(kNx) * kNy * kNz);
    for (size_t i = 0; i < vol.size(); ++i) vol[i] = 0.01f + 0.001f * float(i % 13);
    DeviceBuffer dvol(vol.size()), dproj(size_t(views) * kNu * kNv), ref(size_t(views) * kNu * kNv);
    dvol.upload(vol);
    YK::SExecutionRequest r{};
    r.view_count = views;
    r.volume = dev(dvol);
    r.projection = dev(dproj);
    if (!s->execute(r)) return { false, s->lastErrorMessage() };
    // the same views through the C ABI
    std::vector<yk_view> v(static_cast<size_t>(views));
    for (int i = 0; i < views; ++i) {
        const YK::SConeProjGeomVec& f = frames[size_t(i)];
        const float4* in[5] = { &f.src, &f.detS, &f.detU, &f.detV, &f.angle };
        float* out[5] = { v[size_t(i)].src, v[size_t(i)].detS, v[size_t(i)].detU, v[size_t(i)].detV, v[size_t(i)].angle };
        for (int k = 0; k < 5; ++k) {
            out[k][0] = in[k]->x;
            out[k][1] = in[k]->y;
            out[k][2] = in[k]->z;
            out[k][3] = in[k]->w;
        }
    }
    yk_ctx* ctx = nullptr;
    yk_projector* pj = nullptr;
    yk_volume_grid grid{ kNx, kNy, kNz, 1.f, 1.f, 1.f, { 0.f, 0.f, 0.f } };
    if (yk_ctx_create(0, &ctx) != YK_OK) return { false, yk_last_error() };
    bool ok = yk_projector_create(ctx, v.data(), views, kNu, kNv, &grid, 6.2831853f, &pj) == YK_OK &&
              yk_project(pj, YK_FP_JOSEPH, dvol.p, ref.p, 0, views, 0) == YK_OK;
    const std::string err = ok ? "" : yk_last_error();
    if (pj) (void)yk_projector_destroy(pj);
    (void)yk_ctx_destroy(ctx);
    if (!ok) return { false, err };
    const auto a = dproj.download(), b = ref.download();
    if (!finite_nonzero(a)) return { false, "projection is zero or non-finite" };
    if (max_abs_diff(a, b) != 0.0) return { false, "Session and C ABI projections differ" };
    // and they differ from the nominal views' projections
    auto nominal = make_session();
    auto nspec = small_system(YK::EPipeline::ForwardProjection, views);
    nspec.reconstruction.forward_projector = YK::EForwardProjector::Joseph;
    if (!nominal->initialize(nspec)) return { false, nominal->lastErrorMessage() };
    if (!nominal->execute(r)) return { false, nominal->lastErrorMessage() };
    if (max_abs_diff(dproj.download(), b) < 1e-4) return { false, "the frames did not change the projection" };
    return { true, "" };
}

// an older SSystemSpec (struct_size ending before the tail) never reads it
Result old_system_spec_ignores_frame_tail()
{
    constexpr int views = 12;
    auto spec = iterative_system(YK::EPipeline::SIRT, views, 2, 0.5f, 1);
    auto proj = progress_projection(views);
    std::vector<float> base, observed;
    std::string why;
    if (!run_pipeline(spec, proj, base, why)) return { false, why };
    spec.struct_size = uint32_t(offsetof(YK::SSystemSpec, flat_frames));
    spec.flat_frames = reinterpret_cast<const YK::SConeProjGeomVec*>(16);  // never dereferenced
    spec.flat_frame_count = 5;
    if (!run_pipeline(spec, proj, observed, why)) return { false, why };
    if (max_abs_diff(base, observed) != 0.0) return { false, "old-size spec changed the result" };
    return { true, "" };
}

struct Progress {
    int calls = 0, last = 0, total = 0;
    bool monotone = true;
};

void progress_cb(void* data, int completed, int total)
{
    auto& p = *static_cast<Progress*>(data);
    if (completed < p.last || completed > total || total <= 0) p.monotone = false;
    ++p.calls;
    p.last = completed;
    p.total = total;
}

// per outer iteration progress of SIRT and SART; an old-size request never
// reads the callback tail; the callback does not change the result
Result iteration_progress()
{
    constexpr int views = 12;
    for (auto p : { YK::EPipeline::SIRT, YK::EPipeline::SART, YK::EPipeline::OSSART, YK::EPipeline::CGLS }) {
        auto spec = iterative_system(p, views, 4, 0.5f, p == YK::EPipeline::OSSART ? 3 : 1);
        auto s = make_session();
        if (!s->initialize(spec)) return { false, s->lastErrorMessage() };
        auto hp = progress_projection(views);
        DeviceBuffer vol(size_t(kNx) * kNy * kNz), proj(hp.size());
        proj.upload(hp);
        YK::SExecutionRequest r{};
        r.view_count = views;
        r.iteration_count = 4;
        r.projection = dev(proj);
        r.volume = dev(vol);
        Progress progress;
        r.struct_size = uint32_t(offsetof(YK::SExecutionRequest, progress_callback));
        r.progress_callback = progress_cb;
        r.progress_user_data = &progress;
        if (!s->execute(r)) return { false, s->lastErrorMessage() };
        if (progress.calls != 0) return { false, "old-size request called the callback" };
        const auto base = vol.download();
        s->reset();
        (void)hipMemset(vol.p, 0, vol.n * sizeof(float));
        r.struct_size = sizeof(r);
        if (!s->execute(r)) return { false, s->lastErrorMessage() };
        const std::string name = "pipeline " + std::to_string(int(p));
        if (!progress.monotone) return { false, name + ": progress not monotone or beyond total" };
        if (p == YK::EPipeline::CGLS) {
            // a rolled-back step reports nothing
            if (progress.calls < 1 || progress.total != 4) return { false, name + ": no CGLS progress" };
        } else if (progress.calls != 4 || progress.last != 4 || progress.total != 4) {
            return { false, name + ": expected progress (1..4 of 4), got " + std::to_string(progress.calls) +
                                " calls, last " + std::to_string(progress.last) };
        }
        if (max_abs_diff(base, vol.download()) > 1e-6) return { false, name + ": the callback changed the result" };
    }
    return { true, "" };
}

}  // namespace

int main(int argc, char** argv)
{
    const bool host_only = argc > 1 && std::strcmp(argv[1], "--host-only") == 0;
    struct Case {
        const char* name;
        std::function<Result()> run;
        bool device;
    };
    const std::vector<Case> cases = {
        { "truncated_system_spec", truncated_system_spec, false },
        { "unsupported_api_version", unsupported_api_version, false },
        { "enum_out_of_range", enum_out_of_range, false },
        { "nan_pose_rejected", nan_pose_rejected, false },
        { "outside_this_path", outside_this_path, false },
        { "invalid_scan_rejected", invalid_scan_rejected, false },
        { "execute_before_initialize", execute_before_initialize, false },
        { "flat_frame_tail_validation", flat_frame_tail_validation, false },
        { "forward_projection_siddon", [] { return forward_projection(YK::EForwardProjector::Siddon); }, true },
        { "forward_projection_joseph", [] { return forward_projection(YK::EForwardProjector::Joseph); }, true },
        { "forward_projection_automatic", [] { return forward_projection(YK::EForwardProjector::Automatic); }, true },
        { "execution_validation", execution_validation, true },
        { "fdk_offline_vs_streaming", fdk_offline_vs_streaming, true },
        { "fdk_streaming_state_machine", fdk_streaming_state_machine, true },
        { "session_reset_and_reuse", session_reset_and_reuse, true },
        { "forward_projection_pose_quarter_turn", forward_projection_pose_quarter_turn, true },
        { "shepp_logan_fp_then_fdk", shepp_logan_fp_then_fdk, true },
        { "iterative_pose_ignored", iterative_pose_ignored, true },
        { "sirt_zero_data", sirt_zero_data, true },
        { "sirt_reset_reuse", sirt_reset_reuse, true },
        { "iterative_mapping", iterative_mapping, true },
        { "iterative_requests", iterative_requests, true },
        { "external_frames_invariance", external_frames_invariance, true },
        { "external_frames_forward_projection", external_frames_forward_projection, true },
        { "old_system_spec_ignores_frame_tail", old_system_spec_ignores_frame_tail, true },
        { "iteration_progress", iteration_progress, true },
    };
    int failed = 0;
    for (const Case& c : cases) {
        if (host_only && c.device) continue;
        std::fflush(stdout);
        const Result r = c.run();
        if (r.ok) std::printf("PASS %s\n", c.name);
        else {
            std::printf("FAIL %s: %s\n", c.name, r.why.c_str());
            ++failed;
        }
        std::fflush(stdout);
    }
    return failed == 0 ? 0 : 1;
}

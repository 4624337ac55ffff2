// extern "C" boundary of libyk_hip.so: argument checks, exception-to-status
// translation (the reference aborts through YK_CUDA_CHECK,
// src/global/YkMacro.hpp:35-57; this layer returns codes instead), context
// management, and the iterative callers.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "fdk.hpp"

namespace yk {
namespace {
thread_local std::string g_last_error;
}

void fail(yk_status status, const std::string& message) { throw Error(status, message); }

void check_hip(hipError_t err, const char* what)
{
    if (err != hipSuccess)
        throw Error(YK_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(err));
}

void report_progress(yk_progress_fn cb, void* user, int completed, int total, hipStream_t s)
{
    if (!cb) return;
    // the report describes finished GPU work, not enqueued work
    YK_HIP(hipStreamSynchronize(s));
    cb(user, completed, total);
}

void PinnedBuffer::ensure(size_t bytes)
{
    if (bytes <= bytes_) return;
    reset();
    YK_HIP(hipHostMalloc(&ptr_, bytes, hipHostMallocDefault));
    bytes_ = bytes;
}

void PinnedBuffer::reset()
{
    if (ptr_) (void)hipHostFree(ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
}

template <typename F>
yk_status guard(F&& f)
{
    try {
        f();
        g_last_error.clear();
        return YK_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return YK_ERR_DEVICE;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return YK_ERR_INVALID_CONFIG;
    } catch (...) {
        g_last_error = "unknown error";
        return YK_ERR_INVALID_CONFIG;
    }
}

void require(bool ok, yk_status s, const char* msg)
{
    if (!ok) fail(s, msg);
}

}  // namespace yk

using namespace yk;

void yk_ctx::activate() const { YK_HIP(hipSetDevice(device)); }

void yk_ctx::finish()
{
    YK_HIP(hipGetLastError());
    if (synchronous) YK_HIP(hipStreamSynchronize(stream));
}

extern "C" {

int yk_abi_version(void) { return YK_HIP_ABI_VERSION; }

const char* yk_last_error(void) { return g_last_error.c_str(); }

yk_status yk_ctx_create(int device, yk_ctx** out)
{
    return guard([&] {
        require(out != nullptr, YK_ERR_INVALID_BUFFER, "null output handle");
        *out = nullptr;
        int n = 0;
        YK_HIP(hipGetDeviceCount(&n));
        require(device >= 0 && device < n, YK_ERR_INVALID_CONFIG, "invalid device index");
        auto* c = new yk_ctx();
        c->device = device;
        try {
            YK_HIP(hipSetDevice(device));
            YK_HIP(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
            // set-up of the allocation-free backprojection: the device
            // address of this GPU's fixed scratch array
            (void)bp_noalloc_scratch(device);
        } catch (...) {
            delete c;
            throw;
        }
        c->stream = c->own_stream;
        *out = c;
    });
}

yk_status yk_ctx_destroy(yk_ctx* ctx)
{
    return guard([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        ctx->reduce_scratch.reset();
        ctx->pinned.reset();
        if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
        delete ctx;
    });
}

yk_status yk_ctx_set_stream(yk_ctx* ctx, void* stream)
{
    return guard([&] {
        require(ctx != nullptr, YK_ERR_INVALID_CONFIG, "null context");
        ctx->activate();
        YK_HIP(hipStreamSynchronize(ctx->stream));
        ctx->stream = static_cast<hipStream_t>(stream);
    });
}

yk_status yk_ctx_reset_stream(yk_ctx* ctx)
{
    return guard([&] {
        require(ctx != nullptr, YK_ERR_INVALID_CONFIG, "null context");
        ctx->activate();
        YK_HIP(hipStreamSynchronize(ctx->stream));
        ctx->stream = ctx->own_stream;
    });
}

void* yk_ctx_stream(yk_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

yk_status yk_ctx_set_synchronous(yk_ctx* ctx, int synchronous)
{
    return guard([&] {
        require(ctx != nullptr, YK_ERR_INVALID_CONFIG, "null context");
        ctx->synchronous = synchronous != 0;
    });
}

yk_status yk_ctx_synchronize(yk_ctx* ctx)
{
    return guard([&] {
        require(ctx != nullptr, YK_ERR_INVALID_CONFIG, "null context");
        ctx->activate();
        YK_HIP(hipStreamSynchronize(ctx->stream));
        YK_HIP(hipGetLastError());
    });
}

yk_status yk_geometry_build(const yk_system* system, yk_view* out, int capacity, int* count)
{
    return guard([&] {
        require(system != nullptr, YK_ERR_INVALID_CONFIG, "null system");
        std::vector<yk_view> v;
        build_views(*system, v);
        if (count) *count = int(v.size());
        require(out != nullptr && capacity >= int(v.size()), YK_ERR_INVALID_BUFFER,
            "view buffer smaller than total_views");
        std::copy(v.begin(), v.end(), out);
    });
}

yk_status yk_geometry_derive(const yk_view* views, int n_views, int nu, int nv, float range, float* out)
{
    return guard([&] {
        require(views && out && n_views > 0, YK_ERR_INVALID_BUFFER, "null geometry buffer");
        std::vector<yk_view> v(views, views + n_views);
        std::vector<DerivedD> d;
        derive_geometry(v, nu, nv, range > 0.f ? double(range) : 6.28318530717958647692, d);
        for (int i = 0; i < n_views; ++i) {
            float* o = out + size_t(i) * 20;
            const DerivedD& g = d[i];
            const double rec[20] = { g.theta, g.dtheta, g.fscale, g.sod, g.sdd_radial, g.sdd_plane,
                g.off_u, g.off_v, g.du, g.dv, g.radial[0], g.radial[1], g.radial[2], g.normal[0],
                g.normal[1], g.normal[2], g.uu, g.vv, g.uv, 0.0 };
            for (int k = 0; k < 20; ++k) o[k] = float(rec[k]);
        }
    });
}

yk_status yk_geometry_fdk_coefficients(const yk_view* views, int n_views, int nu, int nv, float range,
    float* out)
{
    return guard([&] {
        require(views && out && n_views > 0, YK_ERR_INVALID_BUFFER, "null geometry buffer");
        std::vector<yk_view> v(views, views + n_views);
        std::vector<DerivedD> d;
        std::vector<CoeffD> c;
        derive_geometry(v, nu, nv, range > 0.f ? double(range) : 6.28318530717958647692, d);
        fdk_coefficients(v, d, c);
        for (int i = 0; i < n_views; ++i) {
            float* o = out + size_t(i) * 17;
            for (int k = 0; k < 4; ++k) {
                o[k] = float(c[i].cu[k]);
                o[4 + k] = float(c[i].cv[k]);
                o[8 + k] = float(c[i].cd[k]);
                o[12 + k] = float(c[i].cr[k]);
            }
            o[16] = float(c[i].wbase);
        }
    });
}

yk_status yk_filter_weights(const yk_filter* filter, int n, float du, float* out)
{
    return guard([&] {
        require(filter && out, YK_ERR_INVALID_BUFFER, "null filter buffer");
        std::vector<float> w;
        filter_weights(*filter, n, du, w);
        std::copy(w.begin(), w.end(), out);
    });
}

int yk_filter_padded_n(int nu) { return nu > 0 ? padded_n(nu) : 0; }

// ---------------------------------------------------------------------------
// FDK
// ---------------------------------------------------------------------------
yk_status yk_fdk_create(yk_ctx* ctx, const yk_fdk_desc* desc, const yk_view* views, int n_views,
    int nu, int nv, const yk_volume_grid* volume, yk_fdk** out)
{
    return guard([&] {
        require(out != nullptr, YK_ERR_INVALID_BUFFER, "null output handle");
        *out = nullptr;
        require(ctx && desc && volume, YK_ERR_INVALID_CONFIG, "null context, descriptor or volume");
        require(desc->struct_size >= sizeof(yk_fdk_desc), YK_ERR_INVALID_CONFIG,
            "yk_fdk_desc.struct_size too small");
        *out = new yk_fdk(ctx, *desc, views, n_views, nu, nv, *volume);
    });
}

yk_status yk_fdk_create_system(yk_ctx* ctx, const yk_fdk_desc* desc, const yk_system* system, yk_fdk** out)
{
    return guard([&] {
        require(out != nullptr, YK_ERR_INVALID_BUFFER, "null output handle");
        *out = nullptr;
        require(ctx && desc && system, YK_ERR_INVALID_CONFIG, "null context, descriptor or system");
        require(desc->struct_size >= sizeof(yk_fdk_desc), YK_ERR_INVALID_CONFIG,
            "yk_fdk_desc.struct_size too small");
        std::vector<yk_view> v;
        build_views(*system, v);
        yk_fdk_desc d = *desc;
        // regularScanRangeRad (include/YKCBCT/interface/YkSystemReconstruction.hpp:167-173)
        d.scan_range_rad = 6.28318530717958647692f * float(system->scan.total_views) /
            float(system->scan.views_per_turn);
        d.start_angle_rad = system->scan.start_angle_rad;
        d.rotation_direction = system->scan.rotation_direction;
        *out = new yk_fdk(ctx, d, v.data(), int(v.size()), system->detector.channels,
            system->detector.rows, system->volume);
    });
}

yk_status yk_fdk_destroy(yk_fdk* fdk)
{
    return guard([&] {
        if (!fdk) return;
        (void)hipSetDevice(fdk->ctx->device);
        (void)hipStreamSynchronize(fdk->ctx->stream);
        delete fdk;
    });
}

yk_status yk_fdk_execute(yk_fdk* fdk, const float* projection, int ploc, size_t pcap, int offset,
    int count, float* volume, int vloc, size_t vcap, int clear_output)
{
    return guard([&] {
        require(fdk != nullptr, YK_ERR_INVALID_CONFIG, "null FDK plan");
        fdk->ctx->activate();
        fdk->execute(projection, ploc, pcap, offset, count, volume, vloc, vcap, clear_output);
        fdk->ctx->finish();
    });
}

yk_status yk_fdk_reset(yk_fdk* fdk)
{
    return guard([&] {
        require(fdk != nullptr, YK_ERR_INVALID_CONFIG, "null FDK plan");
        fdk->received = 0;
        fdk->last_stage = yk_fdk_stage{};
    });
}

yk_status yk_fdk_set_stage_capture(yk_fdk* fdk, int on)
{
    return guard([&] {
        require(fdk != nullptr, YK_ERR_INVALID_CONFIG, "null FDK plan");
        fdk->capture_stages = on != 0;
    });
}

yk_status yk_fdk_last_stage(const yk_fdk* fdk, yk_fdk_stage* out)
{
    return guard([&] {
        require(fdk != nullptr, YK_ERR_INVALID_CONFIG, "null FDK plan");
        require(out != nullptr, YK_ERR_INVALID_BUFFER, "null output");
        *out = fdk->last_stage;
    });
}

yk_status yk_fdk_geometry_diagnostics(const yk_fdk* fdk, yk_geometry_diagnostics* out)
{
    return guard([&] {
        require(fdk != nullptr, YK_ERR_INVALID_CONFIG, "null FDK plan");
        require(out != nullptr && out->struct_size >= sizeof(yk_geometry_diagnostics), YK_ERR_INVALID_BUFFER,
            "null or truncated yk_geometry_diagnostics");
        *out = fdk->diagnostics;
    });
}

yk_status yk_geometry_diagnose(const yk_view* views, int n_views, int nu, int nv, yk_geometry_diagnostics* out)
{
    return guard([&] {
        require(views != nullptr && n_views > 0, YK_ERR_INVALID_BUFFER, "null geometry buffer");
        require(out != nullptr && out->struct_size >= sizeof(yk_geometry_diagnostics), YK_ERR_INVALID_BUFFER,
            "null or truncated yk_geometry_diagnostics");
        std::vector<yk_view> v(views, views + n_views);
        std::vector<DerivedD> d;
        derive_geometry(v, nu, nv, 6.28318530717958647692, d);
        geometry_diagnostics(v, d, *out);
    });
}

yk_status yk_dump_create(const char* output_dir, int enabled, yk_dump** out)
{
    return guard([&] {
        require(out != nullptr, YK_ERR_INVALID_BUFFER, "null output handle");
        *out = nullptr;
        *out = dump_create(output_dir, enabled);
    });
}

yk_status yk_dump_destroy(yk_dump* dump)
{
    return guard([&] { dump_destroy(dump); });
}

yk_status yk_dump_set_enabled(yk_dump* dump, int enabled)
{
    return guard([&] {
        require(dump != nullptr, YK_ERR_INVALID_CONFIG, "null dump");
        dump_set_enabled(dump, enabled);
    });
}

yk_status yk_dump_device_f32(yk_dump* dump, const char* file_name, const float* d_src, size_t count, void* stream)
{
    return guard([&] {
        require(dump != nullptr, YK_ERR_INVALID_CONFIG, "null dump");
        dump_device_f32(dump, file_name ? file_name : "", d_src, count, static_cast<hipStream_t>(stream));
    });
}

yk_status yk_dump_device_f32_a(yk_dump* dump, const char* prefix, int index, const float* d_src, size_t count,
    void* stream)
{
    return guard([&] {
        require(dump != nullptr, YK_ERR_INVALID_CONFIG, "null dump");
        require(prefix != nullptr, YK_ERR_INVALID_BUFFER, "null prefix");
        dump_device_f32(dump, dump_name_a(prefix, index, ".raw"), d_src, count, static_cast<hipStream_t>(stream));
    });
}

yk_status yk_dump_text(yk_dump* dump, const char* file_name, const char* text)
{
    return guard([&] {
        require(dump != nullptr, YK_ERR_INVALID_CONFIG, "null dump");
        dump_text(dump, file_name ? file_name : "", text);
    });
}

yk_status yk_dump_text_a(yk_dump* dump, const char* prefix, int index, const char* text)
{
    return guard([&] {
        require(dump != nullptr, YK_ERR_INVALID_CONFIG, "null dump");
        require(prefix != nullptr, YK_ERR_INVALID_BUFFER, "null prefix");
        dump_text(dump, dump_name_a(prefix, index, ".txt"), text);
    });
}

int yk_fdk_received_views(const yk_fdk* fdk) { return fdk ? fdk->received : 0; }

int yk_fdk_complete(const yk_fdk* fdk) { return fdk && fdk->received == fdk->n_views ? 1 : 0; }

yk_status yk_fdk_filter(yk_fdk* fdk, const float* d_projection, int first, int count, void* d_filtered,
    int storage, float* d_row_scales)
{
    return guard([&] {
        require(fdk != nullptr, YK_ERR_INVALID_CONFIG, "null FDK plan");
        fdk->ctx->activate();
        fdk->filter(d_projection, first, count, d_filtered, storage, d_row_scales);
        fdk->ctx->finish();
    });
}

yk_status yk_fdk_backproject(yk_fdk* fdk, const void* d_filtered, int storage, const float* d_row_scales,
    int first, int count, float* d_volume, int accumulate)
{
    return guard([&] {
        require(fdk != nullptr, YK_ERR_INVALID_CONFIG, "null FDK plan");
        fdk->ctx->activate();
        fdk->backproject(d_filtered, storage, d_row_scales, first, count, d_volume, accumulate);
        fdk->ctx->finish();
    });
}

yk_status yk_fdk_backproject_noalloc(yk_fdk* fdk, const void* d_filtered, int storage,
    const float* d_row_scales, int first, int count, float* d_volume, int accumulate)
{
    return guard([&] {
        require(fdk != nullptr, YK_ERR_INVALID_CONFIG, "null FDK plan");
        fdk->ctx->activate();
        fdk->backproject_noalloc(d_filtered, storage, d_row_scales, first, count, d_volume, accumulate);
        fdk->ctx->finish();
    });
}

int yk_fdk_padded_n(const yk_fdk* fdk) { return fdk ? fdk->n_pad : 0; }

yk_status yk_fdk_get_weights(const yk_fdk* fdk, float* out)
{
    return guard([&] {
        require(fdk && out, YK_ERR_INVALID_BUFFER, "null argument");
        std::copy(fdk->weights.begin(), fdk->weights.end(), out);
    });
}

yk_status yk_fdk_get_derived(const yk_fdk* fdk, float* out)
{
    return guard([&] {
        require(fdk && out, YK_ERR_INVALID_BUFFER, "null argument");
        for (int i = 0; i < fdk->n_views; ++i) {
            const DerivedD& g = fdk->derived[i];
            const double rec[20] = { g.theta, g.dtheta, g.fscale, g.sod, g.sdd_radial, g.sdd_plane,
                g.off_u, g.off_v, g.du, g.dv, g.radial[0], g.radial[1], g.radial[2], g.normal[0],
                g.normal[1], g.normal[2], g.uu, g.vv, g.uv, 0.0 };
            for (int k = 0; k < 20; ++k) out[size_t(i) * 20 + k] = float(rec[k]);
        }
    });
}

yk_status yk_fdk_get_coefficients(const yk_fdk* fdk, float* out)
{
    return guard([&] {
        require(fdk && out, YK_ERR_INVALID_BUFFER, "null argument");
        for (int i = 0; i < fdk->n_views; ++i) {
            float* o = out + size_t(i) * 17;
            const CoeffD& c = fdk->coeffs[i];
            for (int k = 0; k < 4; ++k) {
                o[k] = float(c.cu[k]);
                o[4 + k] = float(c.cv[k]);
                o[8 + k] = float(c.cd[k]);
                o[12 + k] = float(c.cr[k]);
            }
            o[16] = float(c.wbase);
        }
    });
}

// ---------------------------------------------------------------------------
// projection operators
// ---------------------------------------------------------------------------
yk_status yk_projector_create(yk_ctx* ctx, const yk_view* views, int n_views, int nu, int nv,
    const yk_volume_grid* volume, float scan_range, yk_projector** out)
{
    return guard([&] {
        require(out != nullptr, YK_ERR_INVALID_BUFFER, "null output handle");
        *out = nullptr;
        require(ctx && volume, YK_ERR_INVALID_CONFIG, "null context or volume");
        *out = new yk_projector(ctx, views, n_views, nu, nv, *volume, scan_range);
    });
}

yk_status yk_projector_destroy(yk_projector* p)
{
    return guard([&] {
        if (!p) return;
        (void)hipSetDevice(p->ctx->device);
        (void)hipStreamSynchronize(p->ctx->stream);
        delete p;
    });
}

yk_status yk_project(yk_projector* p, int kind, const float* d_volume, float* d_projection, int first,
    int count, int accumulate)
{
    return guard([&] {
        require(p != nullptr, YK_ERR_INVALID_CONFIG, "null projector");
        p->ctx->activate();
        p->project(kind, d_volume, d_projection, first, count, accumulate);
        p->ctx->finish();
    });
}

yk_status yk_backproject(yk_projector* p, int kind, const float* d_projection, float* d_volume, int first,
    int count, int accumulate)
{
    return guard([&] {
        require(p != nullptr, YK_ERR_INVALID_CONFIG, "null projector");
        p->ctx->activate();
        p->backproject(kind, d_projection, d_volume, first, count, accumulate);
        p->ctx->finish();
    });
}

// CGLS, doc/algorithms.md:269 (AstraClassic recursion of
// src/Reconstruction/Iterative/Flat/YkCglsBackends.hpp:22-190)
yk_status yk_cgls(yk_projector* p, int fk, int bk, const float* d_b, float* d_x, int iterations,
    double* residual_norms)
{
    return guard([&] {
        require(p != nullptr, YK_ERR_INVALID_CONFIG, "null projector");
        require(d_b && d_x, YK_ERR_INVALID_BUFFER, "null device buffer");
        require(iterations >= 0, YK_ERR_INVALID_CONFIG, "negative iteration count");
        p->ctx->activate();
        hipStream_t s = p->ctx->stream;
        const size_t np = size_t(p->n_views) * p->nu * p->nv;
        const size_t nx = size_t(p->grid.nx) * p->grid.ny * p->grid.nz;
        DeviceArray<float> r(np), q(np), sv(nx), pv(nx);
        const int n = p->n_views;
        p->project(fk, d_x, q.get(), 0, n, 0);
        launch_scale_copy(r.get(), d_b, 1.f, np, s);
        launch_axpy(r.get(), q.get(), -1.f, np, s);
        p->backproject(bk, r.get(), sv.get(), 0, n, 0);
        launch_scale_copy(pv.get(), sv.get(), 1.f, nx, s);
        double gamma = p->dot(sv.get(), sv.get(), nx);
        if (residual_norms) residual_norms[0] = std::sqrt(p->dot(r.get(), r.get(), np));
        for (int k = 0; k < iterations; ++k) {
            p->project(fk, pv.get(), q.get(), 0, n, 0);
            const double delta = p->dot(q.get(), q.get(), np);
            if (!(delta > 0.0) || !(gamma > 0.0)) {
                for (int j = k; j < iterations && residual_norms; ++j)
                    residual_norms[j + 1] = residual_norms[j];
                break;
            }
            const double alpha = gamma / delta;
            launch_axpy(d_x, pv.get(), float(alpha), nx, s);
            launch_axpy(r.get(), q.get(), float(-alpha), np, s);
            p->backproject(bk, r.get(), sv.get(), 0, n, 0);
            const double gnew = p->dot(sv.get(), sv.get(), nx);
            const double beta = gnew / gamma;
            launch_xpby(pv.get(), sv.get(), float(beta), nx, s);
            gamma = gnew;
            if (residual_norms) residual_norms[k + 1] = std::sqrt(p->dot(r.get(), r.get(), np));
        }
        p->ctx->finish();
    });
}

yk_status yk_cgls_ex(yk_projector* p, const yk_cgls_desc* desc, const float* d_b, float* d_x, yk_iterative_stats* stats,
    double* residual_norms)
{
    return guard([&] {
        require(p != nullptr && desc != nullptr, YK_ERR_INVALID_CONFIG, "null projector or descriptor");
        require(desc->struct_size >= offsetof(yk_cgls_desc, progress_callback), YK_ERR_INVALID_CONFIG,
            "yk_cgls_desc is truncated or from an older layout");
        require(d_b && d_x, YK_ERR_INVALID_BUFFER, "null device buffer");
        // an older caller's struct ends before the progress tail: read only
        // struct_size bytes of it
        yk_cgls_desc d{};
        std::memcpy(&d, desc, std::min<size_t>(desc->struct_size, sizeof(d)));
        run_cgls(p, d, d_b, d_x, stats, residual_norms);
        p->ctx->finish();
    });
}

// SART: view-by-view updates with exact per-view weights, the DETAILED
// SART schedule of yk_algebraic (AlgebraicReconstructorEx per-view order,
// src/Reconstruction/Iterative/Flat/YkAlgebraicReconstructorEx.hpp:137-180)
yk_status yk_sart(yk_projector* p, int fk, int bk, const float* d_b, float* d_x, int iterations,
    float relaxation, int nonnegative)
{
    return guard([&] {
        require(p != nullptr, YK_ERR_INVALID_CONFIG, "null projector");
        require(d_b && d_x, YK_ERR_INVALID_BUFFER, "null device buffer");
        require(iterations >= 0 && std::isfinite(relaxation) && relaxation > 0.f, YK_ERR_INVALID_CONFIG,
            "invalid SART parameters");
        if (iterations == 0) return;
        yk_algebraic_desc d{};
        d.struct_size = sizeof(d);
        d.method = YK_ALG_SART;
        d.weight_model = YK_WEIGHTS_DETAILED;
        d.iterations = iterations;
        d.relaxation = relaxation;
        d.relaxation_reduction = 1.f;
        d.epsilon = 1e-6f;
        d.use_min = nonnegative ? 1 : 0;
        d.min_value = 0.f;
        d.forward_kind = fk;
        d.back_kind = bk;
        d.tv_strength_reduction = 1.f;
        d.check_interval = 1;
        d.patience = 1;
        run_algebraic(p, d, d_b, d_x, nullptr);
        p->ctx->finish();
    });
}

yk_status yk_algebraic(yk_projector* p, const yk_algebraic_desc* desc, const float* d_b, float* d_x,
    yk_iterative_stats* stats)
{
    return guard([&] {
        require(p != nullptr && desc != nullptr, YK_ERR_INVALID_CONFIG, "null projector or descriptor");
        require(desc->struct_size >= offsetof(yk_algebraic_desc, progress_callback), YK_ERR_INVALID_CONFIG,
            "yk_algebraic_desc is truncated or from an older layout");
        require(d_b && d_x, YK_ERR_INVALID_BUFFER, "null device buffer");
        yk_algebraic_desc d{};
        std::memcpy(&d, desc, std::min<size_t>(desc->struct_size, sizeof(d)));
        run_algebraic(p, d, d_b, d_x, stats);
        p->ctx->finish();
    });
}

yk_status yk_tv_gradient(yk_ctx* ctx, const yk_volume_grid* volume, const float* d_x, float eps,
    float* d_gradient, double* value)
{
    return guard([&] {
        require(ctx && volume && d_x, YK_ERR_INVALID_BUFFER, "null argument");
        validate_volume(*volume);
        require(std::isfinite(eps) && eps > 0.f, YK_ERR_INVALID_CONFIG, "epsilon must be positive");
        ctx->activate();
        const int parts = 1024;
        ctx->reduce_scratch.resize(parts);
        launch_tv_gradient(d_x, d_gradient, ctx->reduce_scratch.get(), parts, volume_geom(*volume), eps,
            ctx->stream);
        std::vector<double> h(parts);
        YK_HIP(hipMemcpyAsync(h.data(), ctx->reduce_scratch.get(), parts * sizeof(double),
            hipMemcpyDeviceToHost, ctx->stream));
        YK_HIP(hipStreamSynchronize(ctx->stream));
        double acc = 0.0;
        for (double v : h) acc += v;
        if (value) *value = acc;
        ctx->finish();
    });
}

yk_status yk_dot(yk_ctx* ctx, const float* a, const float* b, size_t n, double* out)
{
    return guard([&] {
        require(ctx && a && b && out, YK_ERR_INVALID_BUFFER, "null argument");
        ctx->activate();
        const int parts = int(std::min<size_t>(1024, std::max<size_t>(1, (n + 255) / 256)));
        ctx->reduce_scratch.resize(1024);
        launch_dot(a, b, n, ctx->reduce_scratch.get(), parts, ctx->stream);
        std::vector<double> h(parts);
        YK_HIP(hipMemcpyAsync(h.data(), ctx->reduce_scratch.get(), parts * sizeof(double),
            hipMemcpyDeviceToHost, ctx->stream));
        YK_HIP(hipStreamSynchronize(ctx->stream));
        double acc = 0.0;
        for (double v : h) acc += v;
        *out = acc;
    });
}

yk_status yk_phantom_ellipsoids(yk_ctx* ctx, const float* ellipsoids, int count,
    const yk_volume_grid* volume, int clamp, float* d_volume)
{
    return guard([&] {
        require(ctx && volume && d_volume, YK_ERR_INVALID_BUFFER, "null argument");
        require(count >= 0 && (count == 0 || ellipsoids), YK_ERR_INVALID_BUFFER, "null ellipsoid table");
        validate_volume(*volume);
        ctx->activate();
        DeviceArray<float> e;
        if (count) e.upload(ellipsoids, size_t(count) * 16, ctx->stream);
        launch_phantom(e.get(), count, volume_geom(*volume), clamp, d_volume, ctx->stream);
        YK_HIP(hipStreamSynchronize(ctx->stream));
        ctx->finish();
    });
}

}  // extern "C"

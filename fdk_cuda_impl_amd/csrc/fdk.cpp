// FDK plan: geometry precompute (FP64), device records, the per-(block, view)
// footprint table, workspaces and the streaming execute contract of the
// reference FdkPipeline / Session FDK branch.
#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "fdk.hpp"

namespace yk {
namespace {

constexpr double kTwoPi = 6.28318530717958647692;

size_t storage_bytes(int storage)
{
    switch (storage) {
    case YK_F32: return 4;
    case YK_F16:
    case YK_BF16: return 2;
    default: fail(YK_ERR_INVALID_CONFIG, "unknown storage type");
    }
}

// Parker weight table w(view, u) * range / pi (parker_weight_kernel /
// pk_uploadAngles, .../FDK/kernels/YkFDKParkerKernels.cu:46-145)
void parker_table(const std::vector<yk_view>& views, const std::vector<DerivedD>& derived,
    int nu, int nv, double start, int dir, double range, std::vector<float>& out)
{
    const double pi = kTwoPi / 2.0;
    const double fscale = range / pi;
    out.assign(views.size() * size_t(nu), 0.f);
    for (size_t i = 0; i < views.size(); ++i) {
        const ViewD g = widen(views[i]);
        const double rx = derived[i].radial[0], ry = derived[i].radial[1];
        const double cv = 0.5 * (nv - 1);
        auto fan = [&](double u) {
            const double x = g.detS[0] + u * g.detU[0] + cv * g.detV[0] - g.src[0];
            const double y = g.detS[1] + u * g.detU[1] + cv * g.detV[1] - g.src[1];
            return std::atan2(x * ry - y * rx, x * rx + y * ry);
        };
        double beta = dir * (g.angle - start);
        while (beta < 0.0) beta += kTwoPi;
        while (beta >= kTwoPi) beta -= kTwoPi;
        const double Gamma = std::max(std::fabs(fan(0.0)), std::fabs(fan(double(nu - 1))));
        for (int u = 0; u < nu; ++u) {
            const double gamma = dir * fan(double(u));
            const double t1 = 2.0 * (Gamma + gamma), t2 = pi + 2.0 * gamma, t3 = pi + 2.0 * Gamma;
            double w;
            if (beta <= 0.0) w = 0.0;
            else if (beta < t1) {
                const double d = Gamma + gamma;
                const double s = std::sin(0.25 * pi * beta / (std::fabs(d) > 1e-6 ? d : 1e-6));
                w = s * s;
            } else if (beta <= t2) w = 1.0;
            else if (beta < t3) {
                const double d = Gamma - gamma;
                const double s = std::sin(0.25 * pi * (t3 - beta) / (std::fabs(d) > 1e-6 ? d : 1e-6));
                w = s * s;
            } else w = 0.0;
            out[i * nu + u] = float(w * fscale);
        }
    }
}

// validateParkerCoverage_ (YkFdkPipeline.hpp:597-635): range >= pi + 2 max fan
void check_parker_coverage(const std::vector<yk_view>& views, int nu, int nv, double range)
{
    if (range >= kTwoPi - 1e-5) fail(YK_ERR_INVALID_CONFIG, "Parker weighting needs a short scan (< 2 pi)");
    double max_fan = 0.0;
    for (const yk_view& v : views) {
        const ViewD g = widen(v);
        const double cu = 0.5 * (nu - 1), cv = 0.5 * (nv - 1);
        const double px = g.detS[0] + cu * g.detU[0] + cv * g.detV[0] - g.src[0];
        const double py = g.detS[1] + cu * g.detU[1] + cv * g.detV[1] - g.src[1];
        for (double u : { 0.0, double(nu - 1) }) {
            const double x = g.detS[0] + u * g.detU[0] + cv * g.detV[0] - g.src[0];
            const double y = g.detS[1] + u * g.detU[1] + cv * g.detV[1] - g.src[1];
            max_fan = std::max(max_fan, std::fabs(std::atan2(x * py - y * px, x * px + y * py)));
        }
    }
    if (range + 1e-5 < kTwoPi / 2.0 + 2.0 * max_fan)
        fail(YK_ERR_INVALID_CONFIG, "short-scan range below pi + 2 * fan angle");
}

// Geometry mode of the tilted-detector backprojection: 2 (a Newton step from
// the previous slice's reciprocal of the depth denominator) when the
// denominator changes by at most 2e-4 relative per slice anywhere in the
// volume, so the step's relative error (the square of that change) stays
// below FP32 resolution; else 1 (a reciprocal per sample).  The denominator
// is affine in the voxel centre, so its minimum over the volume lies at a
// corner.
int general_mode(const std::vector<CoeffD>& coeffs, const VolumeGeom& g)
{
    double worst = 0.0;
    for (const CoeffD& c : coeffs) {
        double den_min = 1e300;
        for (int a = 0; a < 8; ++a) {
            const double X = g.ox + ((a & 1) ? (g.nx - 1) * double(g.vx) : 0.0);
            const double Y = g.oy + ((a & 2) ? (g.ny - 1) * double(g.vy) : 0.0);
            const double Z = g.oz + ((a & 4) ? (g.nz - 1) * double(g.vz) : 0.0);
            den_min = std::min(den_min, c.cd[0] * X + c.cd[1] * Y + c.cd[2] * Z + c.cd[3]);
        }
        if (!(den_min > 0.0)) return 1;
        worst = std::max(worst, std::fabs(c.cd[2]) * g.vz / den_min);
    }
    return worst <= 2e-4 ? 2 : 1;
}

}  // namespace
}  // namespace yk

using namespace yk;

yk_fdk::yk_fdk(yk_ctx* c, const yk_fdk_desc& d, const yk_view* v, int n, int nu_, int nv_,
    const yk_volume_grid& g)
    : ctx(c), desc(d), nu(nu_), nv(nv_), n_views(n), grid(g)
{
    if (!v || n <= 0) fail(YK_ERR_INVALID_GEOMETRY, "FDK needs at least one view");
    if (nu <= 0 || nv <= 0) fail(YK_ERR_INVALID_GEOMETRY, "invalid detector size");
    // the backprojection stages texel pairs (c, c + 1) of a row
    if (nu < 2) fail(YK_ERR_UNSUPPORTED, "FDK needs a detector of at least two channels");
    validate_volume(grid);
    validate_filter(desc.filter);
    storage_bytes(desc.storage);
    if (desc.rotation_direction != 1 && desc.rotation_direction != -1)
        fail(YK_ERR_INVALID_CONFIG, "rotation direction must be +1 or -1");
    if (desc.parker_mode < YK_PARKER_AUTO || desc.parker_mode > YK_PARKER_ENABLED)
        fail(YK_ERR_INVALID_CONFIG, "invalid Parker mode");
    scan_range = desc.scan_range_rad > 0.f ? double(desc.scan_range_rad) : kTwoPi;
    if (!std::isfinite(scan_range)) fail(YK_ERR_INVALID_CONFIG, "scan range must be finite");
    views.assign(v, v + n);
    vol = volume_geom(grid);

    ctx->activate();
    hipStream_t s = ctx->stream;
    derive_geometry(views, nu, nv, scan_range, derived);
    geometry_diagnostics(views, derived, diagnostics);
    fdk_coefficients(views, derived, coeffs);

    // resolveParkerEnabled (include/YKCBCT/interface/YkSystemReconstruction.hpp:175-185)
    parker = desc.parker_mode == YK_PARKER_ENABLED ||
        (desc.parker_mode == YK_PARKER_AUTO && scan_range < kTwoPi - 1e-5);
    if (parker) check_parker_coverage(views, nu, nv, scan_range);

    // filter: weights use du of the first view (YkFDKFilterProcessor.hpp:111)
    n_pad = padded_n(nu);
    filter_weights(desc.filter, n_pad, derived[0].du, weights);
    response_l1 = float(filter_response_l1(weights, n_pad));
    // the caller's spatial ramp / custom weights are read only here
    desc.filter.spatial_ramp = nullptr;
    desc.filter.custom_weights = nullptr;
    d_weights.upload(weights.data(), weights.size(), s);
    std::vector<float2> tw;
    fft_twiddles(n_pad, tw);
    d_twiddles.upload(tw.data(), tw.size(), s);

    std::vector<FilterView> fv(n);
    std::vector<BpCoeff> bc(n);
    std::vector<double> c64(size_t(n) * 12);
    general = false;
    for (int i = 0; i < n; ++i) {
        FilterView& f = fv[i];
        std::memset(&f, 0, sizeof(f));
        std::memcpy(f.src, views[i].src, sizeof(f.src));
        std::memcpy(f.detS, views[i].detS, sizeof(f.detS));
        std::memcpy(f.detU, views[i].detU, sizeof(f.detU));
        std::memcpy(f.detV, views[i].detV, sizeof(f.detV));
        f.sdd_radial = float(derived[i].sdd_radial);
        f.wgeom = float(derived[i].sdd_radial / derived[i].sod / 2.0);
        f.start_u = start_u(nu, n_pad, derived[i].off_u);
        BpCoeff& b = bc[i];
        std::memset(&b, 0, sizeof(b));
        for (int k = 0; k < 4; ++k) {
            b.cu[k] = float(coeffs[i].cu[k]);
            b.cv[k] = float(coeffs[i].cv[k]);
            b.cd[k] = float(coeffs[i].cd[k]);
            b.cr[k] = float(coeffs[i].cr[k]);
            c64[size_t(i) * 12 + k] = coeffs[i].cu[k];
            c64[size_t(i) * 12 + 4 + k] = coeffs[i].cv[k];
            c64[size_t(i) * 12 + 8 + k] = coeffs[i].cd[k];
        }
        b.wbase = float(coeffs[i].wbase);
        if (b.cu[2] != 0.f || b.cd[2] != 0.f) general = true;
        if (b.cr[2] != 0.f) fail(YK_ERR_INVALID_GEOMETRY, "radial ray with a z component");
    }
    geo_mode = general ? general_mode(coeffs, vol) : 0;
    d_fviews.upload(fv.data(), fv.size(), s);
    d_coeffs.upload(bc.data(), bc.size(), s);
    d_c64.upload(c64.data(), c64.size(), s);
    if (parker) {
        std::vector<float> pt;
        parker_table(views, derived, nu, nv, desc.start_angle_rad, desc.rotation_direction,
            scan_range, pt);
        d_parker.upload(pt.data(), pt.size(), s);
    }
    build_tiles();
    YK_HIP(hipStreamSynchronize(s));

    batch = desc.batch_views > 0 ? std::min(desc.batch_views, n) : n;
    // keep the filtered workspace below 16 GiB by default
    const size_t per_view = size_t(nu) * nv * 4;
    if (desc.batch_views <= 0) {
        const size_t cap = size_t(16) << 30;
        batch = int(std::max<size_t>(1, std::min<size_t>(size_t(n), cap / per_view)));
    }
}

void yk_fdk::build_tiles()
{
    hipStream_t s = ctx->stream;
    const size_t max_lds = 96 * 1024;
    const size_t quad = bp_quad_bytes(desc.storage, general ? 1 : 0);
    const int first_bz = bp_block_z(general ? 1 : 0, desc.storage, desc.kahan);
    for (bz = first_bz; bz >= 8; bz = bz > 16 ? 16 : bz / 2) {
        nbx = (grid.nx + 15) / 16;
        nby = (grid.ny + 15) / 16;
        nbz = (grid.nz + bz - 1) / bz;
        const size_t entries = size_t(nbx) * nby * nbz * n_views;
        d_tiles.resize(entries);
        d_stage.resize(entries);
        d_maxes.resize(8);
        TileLaunch t{};
        t.coeff64 = d_c64.get();
        t.n_views = n_views;
        t.vol = vol;
        t.bx = 16; t.by = 16; t.bz = bz;
        t.nbx = nbx; t.nby = nby; t.nbz = nbz;
        t.nu = nu; t.nv = nv;
        t.tiles = d_tiles.get();
        t.stage = d_stage.get();
        t.max_quads = d_maxes.get();
        launch_bp_tiles(t, s);
        int maxes[4] = { 0, 0, 0, 0 };
        YK_HIP(hipMemcpyAsync(maxes, d_maxes.get(), sizeof(maxes), hipMemcpyDeviceToHost, s));
        YK_HIP(hipStreamSynchronize(s));
        if (maxes[3]) fail(YK_ERR_INVALID_GEOMETRY, "volume reaches the source plane of some view");
        lds_quads = std::max(1, maxes[2]);
        if (size_t(lds_quads) * quad <= max_lds) return;
    }
    fail(YK_ERR_UNSUPPORTED, "detector footprint of a voxel block exceeds the LDS budget");
}

void yk_fdk::filter(const float* d_proj, int first, int count, void* d_out, int storage,
    float* d_scales, float* d_preweighted)
{
    if (first < 0 || count <= 0 || first + count > n_views) fail(YK_ERR_INVALID_CONFIG, "view range outside the plan");
    if (!d_proj || !d_out) fail(YK_ERR_INVALID_BUFFER, "null device buffer");
    storage_bytes(storage);
    if (storage != YK_F32 && !d_scales) fail(YK_ERR_INVALID_BUFFER, "half / bfloat16 storage needs row scales");
    FilterLaunch p{};
    p.proj = d_proj;
    p.out = d_out;
    p.row_scales = d_scales;
    p.views = d_fviews.get() + first;
    p.parker = parker ? d_parker.get() + size_t(first) * nu : nullptr;
    p.weights = d_weights.get();
    p.twiddles = d_twiddles.get();
    p.nu = nu; p.nv = nv; p.count = count; p.n = n_pad;
    p.response_l1 = response_l1;
    p.storage = storage;
    launch_fdk_filter(p, ctx->stream);
    if (d_preweighted) launch_preweight(p, d_preweighted, ctx->stream);
}

void yk_fdk::backproject(const void* d_filtered, int storage, const float* d_scales, int first,
    int count, float* d_volume, int accumulate)
{
    if (first < 0 || count <= 0 || first + count > n_views) fail(YK_ERR_INVALID_CONFIG, "view range outside the plan");
    if (!d_filtered || !d_volume) fail(YK_ERR_INVALID_BUFFER, "null device buffer");
    storage_bytes(storage);
    if (storage != YK_F32 && !d_scales) fail(YK_ERR_INVALID_BUFFER, "half / bfloat16 storage needs row scales");
    const size_t quad = bp_quad_bytes(storage, general ? 1 : 0);
    if (size_t(lds_quads) * quad > 160 * 1024) fail(YK_ERR_UNSUPPORTED, "footprint exceeds LDS");
    hipStream_t s = ctx->stream;
    int kernel_storage = storage;
    const void* data = d_filtered;
    if (storage != YK_F32) {
        // one pass applies the normalized row factors, so the backprojection
        // stages plain half texels (when nu is a multiple of 8: 16-byte rows)
        d_rowfac.resize(std::max(d_rowfac.size(), size_t(count) * nv));
        launch_row_factors(d_scales, count * nv, storage, d_maxes.get() + 4, d_rowfac.get(), s);
        if (nu % 8 == 0 && reinterpret_cast<uintptr_t>(d_filtered) % 16 == 0) {
            d_norm_ws.resize(std::max(d_norm_ws.size(), size_t(count) * nv * nu));
            launch_normalize_rows(d_filtered, d_rowfac.get(), storage, count * nv, nu, d_norm_ws.get(), s);
            data = d_norm_ws.get();
            kernel_storage = kBpHalfNormalized;
        }
    }
    BpLaunch p{};
    p.proj = data;
    p.row_factors = storage == YK_F32 ? nullptr : d_rowfac.get();
    p.gscale_bits = d_maxes.get() + 4;
    p.coeffs = d_coeffs.get() + first;
    p.tiles = d_tiles.get();
    p.stage = d_stage.get();
    p.tile_view_stride = n_views;
    p.tile_view_offset = first;
    p.volume = d_volume;
    p.vol = vol;
    p.nu = nu; p.nv = nv; p.count = count;
    p.storage = kernel_storage;
    p.accumulate = accumulate;
    p.kahan = desc.kahan;
    p.general = geo_mode;
    p.nbx = nbx; p.nby = nby; p.nbz = nbz;
    p.lds_quads = lds_quads;
    p.bz = bz;
    launch_fdk_bp(p, s);
}

void yk_fdk::backproject_noalloc(const void* d_filtered, int storage, const float* d_scales, int first,
    int count, float* d_volume, int accumulate)
{
    // the checks of backproject(), then the same kernels with every buffer
    // beyond the inputs, the outputs and the plan's own tables in the
    // library's fixed per-GPU scratch array
    if (first < 0 || count <= 0 || first + count > n_views) fail(YK_ERR_INVALID_CONFIG, "view range outside the plan");
    if (!d_filtered || !d_volume) fail(YK_ERR_INVALID_BUFFER, "null device buffer");
    storage_bytes(storage);
    if (storage != YK_F32 && !d_scales) fail(YK_ERR_INVALID_BUFFER, "half / bfloat16 storage needs row scales");
    const size_t quad = bp_quad_bytes(storage, general ? 1 : 0);
    if (size_t(lds_quads) * quad > 160 * 1024) fail(YK_ERR_UNSUPPORTED, "footprint exceeds LDS");
    const size_t rows = size_t(count) * nv;
    if (storage != YK_F32 && rows > kNoAllocRowFactors)
        fail(YK_ERR_UNSUPPORTED, "view_count x detector rows exceeds YK_NOALLOC_MAX_ROWS (" +
            std::to_string(kNoAllocRowFactors) + ") of the allocation-free backprojection");
    unsigned* scratch = bp_noalloc_scratch(ctx->device);
    int* gbits = reinterpret_cast<int*>(scratch);
    unsigned* factors = scratch + kNoAllocHeader;
    hipStream_t s = ctx->stream;
    if (storage != YK_F32) launch_row_factors(d_scales, int(rows), storage, gbits, factors, s);
    BpLaunch p{};
    p.proj = d_filtered;
    p.row_factors = storage == YK_F32 ? nullptr : factors;
    p.gscale_bits = gbits;
    p.coeffs = d_coeffs.get() + first;
    p.tiles = d_tiles.get();
    p.stage = d_stage.get();
    p.tile_view_stride = n_views;
    p.tile_view_offset = first;
    p.volume = d_volume;
    p.vol = vol;
    p.nu = nu; p.nv = nv; p.count = count;
    p.storage = storage;
    p.accumulate = accumulate;
    p.kahan = desc.kahan;
    p.general = geo_mode;
    p.nbx = nbx; p.nby = nby; p.nbz = nbz;
    p.lds_quads = lds_quads;
    p.bz = bz;
    launch_fdk_bp(p, s);
}

namespace {

// memcpy with up to 8 host threads (pageable <-> pinned staging runs at
// several times the one-thread rate); the threads join before it returns
void copy_parallel(void* dst, const void* src, size_t bytes)
{
    const size_t min_part = size_t(8) << 20;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>(std::min<size_t>(8, hw), std::max<size_t>(1, bytes / min_part));
    char* d = static_cast<char*>(dst);
    const char* sp = static_cast<const char*>(src);
    if (nt <= 1) {
        std::memcpy(d, sp, bytes);
        return;
    }
    const size_t part = (bytes / nt + 4095) & ~size_t(4095);
    std::vector<std::thread> threads;
    size_t done = part;  // the calling thread copies [0, part)
    for (size_t off = part; off < bytes; off += part) {
        const size_t len = std::min(part, bytes - off);
        try {
            threads.emplace_back([d, sp, off, len] { std::memcpy(d + off, sp + off, len); });
            done = off + len;
        } catch (...) {
            break;  // no thread available: the rest is copied here
        }
    }
    std::memcpy(d, sp, std::min(part, bytes));
    if (done < bytes) std::memcpy(d + done, sp + done, bytes - done);
    for (auto& t : threads) t.join();
}

}  // namespace

yk_fdk::~yk_fdk()
{
    for (int k = 0; k < 2; ++k) {
        if (ev_pinned[k]) (void)hipEventDestroy(ev_pinned[k]);
        if (ev_consumed[k]) (void)hipEventDestroy(ev_consumed[k]);
    }
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
}

int yk_fdk::stage_chunk_views() const
{
    // about 128 MiB of FP32 projections per staged chunk (two pinned slots)
    const size_t view_bytes = size_t(nu) * nv * 4;
    const size_t target = size_t(128) << 20;
    return int(std::max<size_t>(1, std::min<size_t>(size_t(batch), target / view_bytes)));
}

void yk_fdk::ensure_staging(size_t slot_bytes)
{
    for (int k = 0; k < 2; ++k) {
        pinned[k].ensure(slot_bytes);
        if (!ev_pinned[k]) YK_HIP(hipEventCreateWithFlags(&ev_pinned[k], hipEventDisableTiming));
        if (!ev_consumed[k]) YK_HIP(hipEventCreateWithFlags(&ev_consumed[k], hipEventDisableTiming));
    }
    if (!copy_stream) YK_HIP(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking));
}

void yk_fdk::upload_host(float* d_dst, const float* host, size_t bytes)
{
    // host -> pinned slot -> device on the plan's stream, two slots in flight
    const size_t slot = size_t(stage_chunk_views()) * nu * nv * 4;
    ensure_staging(slot);
    hipStream_t s = ctx->stream;
    const char* src = reinterpret_cast<const char*>(host);
    char* dst = reinterpret_cast<char*>(d_dst);
    for (size_t off = 0, i = 0; off < bytes; off += slot, ++i) {
        const int k = int(i & 1);
        const size_t len = std::min(slot, bytes - off);
        if (pinned_used[k]) YK_HIP(hipEventSynchronize(ev_pinned[k]));
        copy_parallel(pinned[k].get(), src + off, len);
        YK_HIP(hipMemcpyAsync(dst + off, pinned[k].get(), len, hipMemcpyHostToDevice, s));
        YK_HIP(hipEventRecord(ev_pinned[k], s));
        pinned_used[k] = true;
    }
}

void yk_fdk::download_host(float* host, const float* d_src, size_t bytes)
{
    // device -> pinned slot (plan's stream) -> host, the next slot's copy in
    // flight while the host drains the current one
    const size_t slot = size_t(stage_chunk_views()) * nu * nv * 4;
    ensure_staging(slot);
    hipStream_t s = ctx->stream;
    const char* src = reinterpret_cast<const char*>(d_src);
    char* dst = reinterpret_cast<char*>(host);
    const size_t n = (bytes + slot - 1) / slot;
    auto issue = [&](size_t i) {
        const int k = int(i & 1);
        const size_t off = i * slot, len = std::min(slot, bytes - off);
        if (pinned_used[k]) YK_HIP(hipEventSynchronize(ev_pinned[k]));
        YK_HIP(hipMemcpyAsync(pinned[k].get(), src + off, len, hipMemcpyDeviceToHost, s));
        YK_HIP(hipEventRecord(ev_pinned[k], s));
        pinned_used[k] = true;
    };
    for (size_t i = 0; i < n && i < 2; ++i) issue(i);
    for (size_t i = 0; i < n; ++i) {
        const int k = int(i & 1);
        const size_t off = i * slot, len = std::min(slot, bytes - off);
        YK_HIP(hipEventSynchronize(ev_pinned[k]));
        copy_parallel(dst + off, pinned[k].get(), len);
        if (i + 2 < n) issue(i + 2);
    }
}

void yk_fdk::execute(const float* projection, int ploc, size_t pcap, int offset, int count,
    float* volume, int vloc, size_t vcap, int clear_output)
{
    // YkReconstructionApi.cpp:353-370 and YkSession.cpp:693-728
    if (!projection || !volume) fail(YK_ERR_INVALID_BUFFER, "projection and volume are required");
    if ((ploc != YK_HOST && ploc != YK_DEVICE) || (vloc != YK_HOST && vloc != YK_DEVICE))
        fail(YK_ERR_INVALID_BUFFER, "invalid memory location");
    if (count <= 0) fail(YK_ERR_INVALID_CONFIG, "view_count must be positive");
    if (offset != received) fail(YK_ERR_STREAMING_STATE, "view_offset must equal the views received so far");
    if (offset + count > n_views) fail(YK_ERR_STREAMING_STATE, "batch runs past the prepared view sequence");
    if (clear_output && received != 0) fail(YK_ERR_STREAMING_STATE, "clear_output is only valid for the first batch");
    const size_t view_elems = size_t(nu) * nv;
    const size_t vol_elems = size_t(grid.nx) * grid.ny * grid.nz;
    if (pcap < view_elems * count) fail(YK_ERR_INVALID_BUFFER, "projection buffer smaller than view_count views");
    if (vcap < vol_elems) fail(YK_ERR_INVALID_BUFFER, "volume buffer smaller than the grid");
    hipStream_t s = ctx->stream;
    float* d_vol = volume;
    if (vloc == YK_HOST) {
        d_volume_ws.resize(vol_elems);
        d_vol = d_volume_ws.get();
        if (received == 0 && !clear_output) upload_host(d_vol, volume, vol_elems * 4);
    }
    // Device projections run in batches of `batch` views.  Host projections
    // run in staged chunks: chunk c is copied into pinned slot c mod 2 by the
    // host, moved to device slot c mod 2 on the copy stream (once the filter
    // of chunk c - 2 has read that device slot), and filtered and
    // backprojected on the plan's stream (once its copy has landed), so the
    // transfers of later chunks overlap the GPU work of earlier ones.
    const bool staged = ploc == YK_HOST;
    const int chunk = std::min(staged ? stage_chunk_views() : batch, count);
    const size_t chunk_elems = size_t(chunk) * view_elems;
    const size_t eb = storage_bytes(desc.storage);
    d_filtered_ws.resize(chunk_elems * eb);
    if (desc.storage != YK_F32) d_scales_ws.resize(size_t(chunk) * nv);
    if (staged) {
        ensure_staging(chunk_elems * 4);
        d_proj_ws.resize(2 * chunk_elems);
    }
    if (capture_stages) d_preweighted_ws.resize(chunk_elems);
    last_stage = yk_fdk_stage{};
    for (int base = 0, c = 0; base < count; base += chunk, ++c) {
        const int k = std::min(chunk, count - base);
        const int slot = c & 1;
        const float* d_proj = projection + size_t(base) * view_elems;
        if (staged) {
            const size_t bytes = size_t(k) * view_elems * 4;
            float* d_in = d_proj_ws.get() + size_t(slot) * chunk_elems;
            if (pinned_used[slot]) YK_HIP(hipEventSynchronize(ev_pinned[slot]));
            copy_parallel(pinned[slot].get(), d_proj, bytes);
            if (consumed_used[slot]) YK_HIP(hipStreamWaitEvent(copy_stream, ev_consumed[slot], 0));
            YK_HIP(hipMemcpyAsync(d_in, pinned[slot].get(), bytes, hipMemcpyHostToDevice, copy_stream));
            YK_HIP(hipEventRecord(ev_pinned[slot], copy_stream));
            pinned_used[slot] = true;
            YK_HIP(hipStreamWaitEvent(s, ev_pinned[slot], 0));
            d_proj = d_in;
        }
        filter(d_proj, offset + base, k, d_filtered_ws.get(), desc.storage,
            desc.storage == YK_F32 ? nullptr : d_scales_ws.get(),
            capture_stages ? d_preweighted_ws.get() : nullptr);
        if (staged) {
            YK_HIP(hipEventRecord(ev_consumed[slot], s));
            consumed_used[slot] = true;
        }
        last_stage.first_view = offset + base;
        last_stage.count = k;
        last_stage.d_input = d_proj;
        last_stage.d_preweighted = capture_stages ? d_preweighted_ws.get() : nullptr;
        last_stage.d_filtered = d_filtered_ws.get();
        last_stage.filtered_storage = desc.storage;
        last_stage.d_row_scales = desc.storage == YK_F32 ? nullptr : d_scales_ws.get();
        last_stage.stream = s;
        const int acc = (clear_output && base == 0) ? 0 : 1;
        backproject(d_filtered_ws.get(), desc.storage,
            desc.storage == YK_F32 ? nullptr : d_scales_ws.get(), offset + base, k, d_vol, acc);
    }
    received += count;
    if (vloc == YK_HOST && received == n_views) download_host(volume, d_vol, vol_elems * 4);
    // host buffers: the call returns with the GPU work done (the caller may
    // reuse or read them at once)
    if (vloc == YK_HOST || staged) YK_HIP(hipStreamSynchronize(s));
}

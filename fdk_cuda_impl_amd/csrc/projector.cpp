// Projection-operator plan: per-view device records for the Joseph and
// Siddon forward models and their transposes, and the iterative callers
// (CGLS, SART) built on them.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "fdk.hpp"

using namespace yk;

namespace {

// scratch that only grows: the separable paths alternate between layouts of
// different sizes in one array, and a reallocation would synchronize
template <typename T>
void grow(DeviceArray<T>& a, size_t n)
{
    if (a.size() < n) a.resize(n);
}

}  // namespace

yk_projector::yk_projector(yk_ctx* c, const yk_view* v, int n, int nu_, int nv_,
    const yk_volume_grid& g, float range)
    : ctx(c), nu(nu_), nv(nv_), n_views(n), grid(g),
      scan_range(range > 0.f ? double(range) : 6.28318530717958647692)
{
    if (!v || n <= 0) fail(YK_ERR_INVALID_GEOMETRY, "projector needs at least one view");
    if (nu <= 0 || nv <= 0) fail(YK_ERR_INVALID_GEOMETRY, "invalid detector size");
    validate_volume(grid);
    views.assign(v, v + n);
    vol = volume_geom(grid);
    ctx->activate();
    std::vector<ProjViewDev> jv(n);
    std::vector<SiddonViewDev> sv(n);
    std::vector<double> c64(size_t(n) * 12);
    const double vox[3] = { grid.voxel_x_mm, grid.voxel_y_mm, grid.voxel_z_mm };
    const double cen[3] = { grid.center_mm[0], grid.center_mm[1], grid.center_mm[2] };
    for (int i = 0; i < n; ++i) {
        const yk_view& w = views[i];
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(w.src[k]) || !std::isfinite(w.detS[k]) || !std::isfinite(w.detU[k]) ||
                !std::isfinite(w.detV[k]))
                fail(YK_ERR_INVALID_GEOMETRY, "non-finite view vector");
        // normalizeToVoxel (src/common/cuda/operators/YkOperatorKernelTypes.cuh:148-168)
        ProjViewDev& p = jv[i];
        std::memset(&p, 0, sizeof(p));
        for (int k = 0; k < 3; ++k) {
            p.src[k] = float((double(w.src[k]) - cen[k]) / vox[k]);
            p.detS[k] = float((double(w.detS[k]) - cen[k]) / vox[k]);
            p.detU[k] = float(double(w.detU[k]) / vox[k]);
            p.detV[k] = float(double(w.detV[k]) / vox[k]);
        }
        p.axis = main_axis(w, nu, nv);
        if (p.axis == 0) any_x_axis = true;
        // d0 and d1 of every ray independent of v (exact zeros, so bit-exact)
        separable.push_back(p.axis != 2 && p.detV[p.axis == 0 ? 0 : 1] == 0.f &&
                            p.detV[p.axis == 0 ? 1 : 0] == 0.f);
        siddon_separable.push_back(w.detV[0] == 0.f && w.detV[1] == 0.f);
        SiddonViewDev& s = sv[i];
        std::memset(&s, 0, sizeof(s));
        std::memcpy(s.src, w.src, sizeof(s.src));
        std::memcpy(s.detS, w.detS, sizeof(s.detS));
        std::memcpy(s.detU, w.detU, sizeof(s.detU));
        std::memcpy(s.detV, w.detV, sizeof(s.detV));
        projection_map(w, &c64[size_t(i) * 12]);
    }
    // operator-form FDK backprojection records (FP64 geometry over the whole
    // view sequence, like FdkAffineCoeff in .../FDK/kernels/YkFDKBpPrecompute.cu:14-138)
    std::vector<DerivedD> der;
    std::vector<CoeffD> cf;
    derive_geometry(views, nu, nv, scan_range, der);
    fdk_coefficients(views, der, cf);
    std::vector<OpBpViewDev> ov(n);
    op_scale.resize(n);
    for (int i = 0; i < n; ++i) {
        const ViewD w = widen(views[i]);
        OpBpViewDev& o = ov[i];
        std::memset(&o, 0, sizeof(o));
        for (int k = 0; k < 4; ++k) {
            o.cu[k] = float(cf[i].cu[k]);
            o.cv[k] = float(cf[i].cv[k]);
            o.cd[k] = float(cf[i].cd[k]);
        }
        const double sod = der[i].sod;
        for (int k = 0; k < 3; ++k) {
            o.cc[k] = float(-w.src[k] / sod);
            o.src[k] = float(w.src[k]);
        }
        o.cc[3] = float(sod);
        double r0[3];
        for (int k = 0; k < 3; ++k) r0[k] = w.detS[k] - w.src[k];
        auto dot = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
        o.l2[0] = float(dot(r0, r0));
        o.l2[1] = float(2.0 * dot(r0, w.detU));
        o.l2[2] = float(2.0 * dot(r0, w.detV));
        o.l2[3] = float(dot(w.detU, w.detU));
        o.l2[4] = float(2.0 * dot(w.detU, w.detV));
        o.l2[5] = float(dot(w.detV, w.detV));
        o.wbase = float(cf[i].wbase);
        o.inv_sdd_plane = std::fabs(der[i].sdd_plane) > 1e-12 ? float(1.0 / std::fabs(der[i].sdd_plane)) : 0.f;
        op_scale[i] = vox[0] * vox[1] * vox[2] / der[i].du / der[i].dv;
    }
    hipStream_t s = ctx->stream;
    d_opviews.upload(ov.data(), ov.size(), s);
    d_jviews.upload(jv.data(), jv.size(), s);
    d_sviews.upload(sv.data(), sv.size(), s);
    std::vector<float> c32(c64.begin(), c64.end());
    d_c32.upload(c32.data(), c32.size(), s);
    d_partials.resize(1024);
    YK_HIP(hipStreamSynchronize(s));
}

void yk_projector::project(int kind, const float* d_volume, float* d_proj, int first, int count,
    int accumulate)
{
    if (first < 0 || count <= 0 || first + count > n_views) fail(YK_ERR_INVALID_CONFIG, "view range outside the plan");
    if (!d_volume || !d_proj) fail(YK_ERR_INVALID_BUFFER, "null device buffer");
    hipStream_t s = ctx->stream;
    yk_projector& w = scratch_owner ? *scratch_owner : *this;  // scratch arrays
    bool all_sep = true, any_x = false, any_y = false, any_z = false;
    for (int i = first; i < first + count; ++i) {
        all_sep = all_sep && separable[i];
        const int ax = main_axis(views[i], nu, nv);
        any_x = any_x || ax == 0;
        any_y = any_y || ax == 1;
        any_z = any_z || ax == 2;
    }
    if (kind == YK_FP_JOSEPH && all_sep) {
        // paired guarded copies (every call rewrites the copy it reads)
        if (any_x) {
            grow(w.d_vol_zx, joseph_pair_size(grid.nx, grid.ny, grid.nz, true, false));
            launch_joseph_pair_copy(d_volume, w.d_vol_zx.get(), grid.nx, grid.ny, grid.nz, true, false, s);
        }
        if (any_y) {
            grow(w.d_vol_zy, joseph_pair_size(grid.nx, grid.ny, grid.nz, false, false));
            launch_joseph_pair_copy(d_volume, w.d_vol_zy.get(), grid.nx, grid.ny, grid.nz, false, false, s);
        }
        JosephSepFpLaunch p{};
        p.vol_zx = any_x ? w.d_vol_zx.get() : nullptr;
        p.vol_zy = any_y ? w.d_vol_zy.get() : nullptr;
        p.proj = d_proj;
        p.views = d_jviews.get() + first;
        p.vol = vol;
        p.nu = nu; p.nv = nv; p.count = count;
        p.accumulate = accumulate;
        launch_joseph_fp_sep(p, s);
    } else if (kind == YK_FP_JOSEPH && !any_z) {
        // X/Y-main views of any orientation: paired slab copies
        if (any_x) {
            grow(w.d_vol_zx, joseph_pair_size(grid.nx, grid.ny, grid.nz, true, true));
            launch_joseph_pair_copy(d_volume, w.d_vol_zx.get(), grid.nx, grid.ny, grid.nz, true, true, s);
        }
        if (any_y) {
            grow(w.d_vol_zy, joseph_pair_size(grid.nx, grid.ny, grid.nz, false, true));
            launch_joseph_pair_copy(d_volume, w.d_vol_zy.get(), grid.nx, grid.ny, grid.nz, false, true, s);
        }
        JosephSepFpLaunch p{};
        p.vol_zx = any_x ? w.d_vol_zx.get() : nullptr;
        p.vol_zy = any_y ? w.d_vol_zy.get() : nullptr;
        p.proj = d_proj;
        p.views = d_jviews.get() + first;
        p.vol = vol;
        p.nu = nu; p.nv = nv; p.count = count;
        p.accumulate = accumulate;
        launch_joseph_fp_slab(p, s);
    } else if (kind == YK_FP_JOSEPH) {
        bool need_t = false;
        std::vector<ProjViewDev> dummy;
        for (int i = first; i < first + count; ++i)
            if (main_axis(views[i], nu, nv) == 0) need_t = true;
        if (need_t) {
            w.d_vol_t.resize(size_t(grid.nx) * grid.ny * grid.nz);
            launch_transpose_xy(d_volume, w.d_vol_t.get(), grid.nx, grid.ny, grid.nz, s);
        }
        JosephLaunch p{};
        p.volume = d_volume;
        p.volume_t = need_t ? w.d_vol_t.get() : nullptr;
        p.proj = d_proj;
        p.views = d_jviews.get() + first;
        p.vol = vol;
        p.nu = nu; p.nv = nv; p.count = count;
        p.accumulate = accumulate;
        launch_joseph_fp(p, s);
    } else if (kind == YK_FP_SIDDON) {
        // every view set walks the [z][y][x] volume with siddon_fp_kernel:
        // its view-grouped tiles measured faster than the separable
        // wave-per-column walk over a Z-fastest copy on untilted scans too
        // (27.5 against 33.3 ms at 256^3 x 360), with bit-identical results
        SiddonLaunch p{};
        p.volume = d_volume;
        p.proj = d_proj;
        p.views = d_sviews.get() + first;
        p.vol = vol;
        p.nu = nu; p.nv = nv; p.count = count;
        p.accumulate = accumulate;
        launch_siddon_fp(p, s);
    } else {
        fail(YK_ERR_UNSUPPORTED, "unknown forward projector");
    }
}

void yk_projector::backproject(int kind, const float* d_proj, float* d_volume, int first, int count,
    int accumulate)
{
    if (first < 0 || count <= 0 || first + count > n_views) fail(YK_ERR_INVALID_CONFIG, "view range outside the plan");
    if (!d_volume || !d_proj) fail(YK_ERR_INVALID_BUFFER, "null device buffer");
    hipStream_t s = ctx->stream;
    yk_projector& w = scratch_owner ? *scratch_owner : *this;  // scratch arrays
    bool all_sep = true;
    for (int i = first; i < first + count; ++i) all_sep = all_sep && separable[i];
    if (kind == YK_BP_JOSEPH_ADJOINT && all_sep) {
        // column-major ray table of a chunk of views at a time (at most 1 GiB)
        const size_t per_view = size_t(nu) * nv;
        const size_t table = size_t(nu) * (nv + kSepPad);
        const int chunk = int(std::max<size_t>(1, std::min<size_t>(size_t(count),
            (size_t(1) << 30) / (table * (sizeof(float2) + sizeof(float)) + size_t(nu) * sizeof(float2)))));
        w.d_sep_ab1.resize(std::max(w.d_sep_ab1.size(), size_t(nu) * chunk));
        w.d_sep_ab2.resize(std::max(w.d_sep_ab2.size(), table * chunk));
        w.d_sep_pl.resize(std::max(w.d_sep_pl.size(), table * chunk));
        for (int c0 = 0; c0 < count; c0 += chunk) {
            JosephSepLaunch p{};
            p.proj = d_proj + per_view * c0;
            p.volume = d_volume;
            p.views = d_jviews.get() + first + c0;
            p.coeff32 = d_c32.get() + size_t(first + c0) * 12;
            p.ab1 = w.d_sep_ab1.get();
            p.ab2 = w.d_sep_ab2.get();
            p.pl = w.d_sep_pl.get();
            p.vol = vol;
            p.nu = nu; p.nv = nv; p.count = std::min(chunk, count - c0);
            p.accumulate = c0 > 0 ? 1 : accumulate;
            launch_joseph_adjoint_sep(p, s);
        }
    } else if (kind == YK_BP_JOSEPH_ADJOINT) {
        // the ray table holds a chunk of views at a time (at most 1 GiB)
        const size_t per_view = size_t(nu) * nv;
        const int chunk = int(std::max<size_t>(1, std::min<size_t>(size_t(count),
            (size_t(1) << 30) / (per_view * (sizeof(float4) + sizeof(float))))));
        w.d_ray_ab.resize(std::max(w.d_ray_ab.size(), per_view * chunk));
        w.d_ray_len.resize(std::max(w.d_ray_len.size(), per_view * chunk));
        for (int c0 = 0; c0 < count; c0 += chunk) {
            JosephAdjointLaunch p{};
            p.proj = d_proj + per_view * c0;
            p.volume = d_volume;
            p.views = d_jviews.get() + first + c0;
            p.coeff32 = d_c32.get() + size_t(first + c0) * 12;
            p.ray_ab = w.d_ray_ab.get();
            p.ray_len = w.d_ray_len.get();
            p.vol = vol;
            p.nu = nu; p.nv = nv; p.count = std::min(chunk, count - c0);
            p.accumulate = c0 > 0 ? 1 : accumulate;
            launch_joseph_adjoint(p, s);
        }
    } else if (kind == YK_BP_SIDDON_ADJOINT && std::all_of(siddon_separable.begin() + first,
                                                            siddon_separable.begin() + first + count,
                                                            [](char c) { return c != 0; })) {
        // column-major ray table of a chunk of views at a time (at most 1 GiB)
        const size_t per_view = size_t(nu) * nv;
        const size_t table = size_t(nu) * (nv + kSepPad);
        const int chunk = int(std::max<size_t>(1, std::min<size_t>(size_t(count),
            (size_t(1) << 30) / ((table + nu) * sizeof(float2)))));
        w.d_sid_col.resize(std::max(w.d_sid_col.size(), size_t(nu) * chunk));
        w.d_sid_ray.resize(std::max(w.d_sid_ray.size(), table * chunk));
        for (int c0 = 0; c0 < count; c0 += chunk) {
            SiddonSepLaunch p{};
            p.proj = d_proj + per_view * c0;
            p.volume = d_volume;
            p.views = d_sviews.get() + first + c0;
            p.coeff32 = d_c32.get() + size_t(first + c0) * 12;
            p.col_xy = w.d_sid_col.get();
            p.ray_wz = w.d_sid_ray.get();
            p.vol = vol;
            p.nu = nu; p.nv = nv; p.count = std::min(chunk, count - c0);
            p.accumulate = c0 > 0 ? 1 : accumulate;
            launch_siddon_adjoint_sep(p, s);
        }
    } else if (kind == YK_BP_SIDDON_ADJOINT) {
        SiddonAdjointLaunch p{};
        p.proj = d_proj;
        p.volume = d_volume;
        p.views = d_sviews.get() + first;
        p.coeff32 = d_c32.get() + size_t(first) * 12;
        p.vol = vol;
        p.nu = nu; p.nv = nv; p.count = count;
        p.accumulate = accumulate;
        launch_siddon_adjoint(p, s);
    } else if (kind == YK_BP_FDK || kind == YK_BP_FDK_PSEUDO_MATCHED || kind == YK_BP_JOSEPH_VOXEL) {
        OpBpLaunch p{};
        p.proj = d_proj;
        p.volume = d_volume;
        p.views = d_opviews.get() + first;
        p.vol = vol;
        p.nu = nu; p.nv = nv; p.count = count;
        p.scale = float(op_scale[first]);
        p.mode = kind == YK_BP_FDK_PSEUDO_MATCHED ? 1 : kind == YK_BP_JOSEPH_VOXEL ? 2 : 0;
        p.accumulate = accumulate;
        launch_fdk_operator_bp(p, s);
    } else {
        fail(YK_ERR_UNSUPPORTED, "unknown back projector");
    }
}

double yk_projector::dot(const float* a, const float* b, size_t n)
{
    hipStream_t s = ctx->stream;
    const int parts = int(std::min<size_t>(1024, std::max<size_t>(1, (n + 255) / 256)));
    launch_dot(a, b, n, d_partials.get(), parts, s);
    std::vector<double> h(parts);
    YK_HIP(hipMemcpyAsync(h.data(), d_partials.get(), parts * sizeof(double), hipMemcpyDeviceToHost, s));
    YK_HIP(hipStreamSynchronize(s));
    double acc = 0.0;
    for (double v : h) acc += v;
    return acc;
}

// Unified algebraic reconstruction (SIRT / SART / OS-SART) over the
// projector's forward model A and back model B:
//   x <- P[x + lambda C_s^-1 B_s R_s^-1 (b_s - A_s x)]
// with the reference's two weight models and schedules
// (AlgebraicReconstructorEx, src/Reconstruction/Iterative/Flat/
// YkAlgebraicReconstructorEx.hpp:15-180; AlgebraicDetailedWeightBackend,
// YkAlgebraicBackends.hpp:597-800; AlgebraicTigreBackend, :32-310;
// SirtBatchBackend, YkSirtBatchBackend.hpp:11-80; smoothed-TV plugin,
// src/Reconstruction/Iterative/Regularization/RegularizerPlugins.cpp:33-80).
// Subsets that are not a contiguous view range get a projector of their own
// views, so every operator call is a plain [first, first + count) range.
#include <algorithm>
#include <cmath>
#include <limits>
#include <memory>
#include <numeric>
#include <vector>

#include "fdk.hpp"

using namespace yk;

namespace {

struct Unit {
    std::vector<int> views;               // indices into the whole sequence
    yk_projector* op = nullptr;           // operators of this unit
    int first = 0;                        // first view of the unit within `op`
    std::unique_ptr<yk_projector> own;    // for non-contiguous subsets
    size_t row_offset = 0;                // of its row weights in the row cache
};

bool contiguous(const std::vector<int>& v)
{
    for (size_t i = 1; i < v.size(); ++i)
        if (v[i] != v[i - 1] + 1) return false;
    return true;
}

// golden-ratio visiting order of n subsets: step round(n / phi) made coprime
// with n (AlgebraicDetailedWeightBackend::init, YkAlgebraicBackends.hpp:669-684)
std::vector<int> golden_order(int n)
{
    std::vector<int> order(n);
    int step = std::max(1, int(std::lround(double(n) / 1.6180339887498949)));
    while (std::gcd(step, n) != 1) {
        ++step;
        if (step >= n) {
            step = 1;
            break;
        }
    }
    int cur = 0;
    for (int i = 0; i < n; ++i) {
        order[i] = cur;
        cur = (cur + step) % n;
    }
    return order;
}

}  // namespace

void run_algebraic(yk_projector* p, const yk_algebraic_desc& d, const float* d_b, float* d_x, yk_iterative_stats* out)
{
    const int na = p->n_views;
    if (d.method < YK_ALG_SIRT || d.method > YK_ALG_OSSART) fail(YK_ERR_INVALID_CONFIG, "unknown algebraic method");
    if (d.weight_model != YK_WEIGHTS_DETAILED && d.weight_model != YK_WEIGHTS_TIGRE)
        fail(YK_ERR_INVALID_CONFIG, "unknown weight model");
    if (d.subset_order != YK_ORDER_SEQUENTIAL && d.subset_order != YK_ORDER_GOLDEN)
        fail(YK_ERR_INVALID_CONFIG, "unknown subset order");
    // AlgebraicReconstructorEx::validate_ (YkAlgebraicReconstructorEx.hpp:365-381)
    if (d.iterations <= 0 || !std::isfinite(d.relaxation) || !(d.relaxation > 0.f) ||
        !std::isfinite(d.relaxation_reduction) || !(d.relaxation_reduction > 0.f) || !std::isfinite(d.epsilon) ||
        !(d.epsilon > 0.f))
        fail(YK_ERR_INVALID_CONFIG, "invalid iterations, relaxation or epsilon");
    if (d.method == YK_ALG_OSSART && (d.subsets < 1 || d.subsets > na))
        fail(YK_ERR_INVALID_CONFIG, "OS-SART subset count must lie in [1, views]");
    if (d.use_min && d.use_max && d.min_value > d.max_value)
        fail(YK_ERR_INVALID_CONFIG, "min_value exceeds max_value");
    if (d.tv_iterations < 0 || !std::isfinite(d.tv_strength) || d.tv_strength < 0.f ||
        (d.tv_iterations > 0 && d.tv_strength > 0.f && !(d.tv_epsilon > 0.f)) ||
        !std::isfinite(d.tv_strength_reduction) || !(d.tv_strength_reduction > 0.f))
        fail(YK_ERR_INVALID_CONFIG, "invalid TV strength, iterations, epsilon or reduction");
    if (d.tv_dimensions != 0 && d.tv_dimensions != 2 && d.tv_dimensions != 3)
        fail(YK_ERR_INVALID_CONFIG, "TV dimensionality must be 2 or 3");
    if (d.forward_kind != YK_FP_JOSEPH && d.forward_kind != YK_FP_SIDDON)
        fail(YK_ERR_UNSUPPORTED, "unknown forward projector");
    // validConvergenceConfig (src/Reconstruction/Iterative/Common/YkIterativeConvergence.hpp:39-46)
    if (!(d.relative_residual_tolerance >= 0.f) || !(d.relative_update_tolerance >= 0.f) ||
        !(d.relative_improvement_tolerance >= 0.f) || d.minimum_iterations < 0 || d.check_interval <= 0 ||
        d.patience <= 0)
        fail(YK_ERR_INVALID_CONFIG, "invalid convergence settings");
    if (out && out->struct_size < sizeof(yk_iterative_stats)) fail(YK_ERR_INVALID_CONFIG, "yk_iterative_stats is truncated");

    p->ctx->activate();
    hipStream_t s = p->ctx->stream;
    const size_t view_n = size_t(p->nu) * p->nv;
    const size_t nx = size_t(p->grid.nx) * p->grid.ny * p->grid.nz;
    const bool tigre = d.weight_model == YK_WEIGHTS_TIGRE;
    const int nsub = d.method == YK_ALG_SIRT ? 1 : d.method == YK_ALG_SART ? na : d.subsets;

    // update units and their visiting order
    std::vector<Unit> units;
    if (tigre) {
        const int bs = (na + nsub - 1) / nsub;
        for (int st = 0; st < na; st += bs) {
            Unit u;
            for (int i = st; i < std::min(na, st + bs); ++i) u.views.push_back(i);
            units.push_back(std::move(u));
        }
    } else {
        for (int k = 0; k < nsub; ++k) {
            Unit u;
            for (int i = k; i < na; i += nsub) u.views.push_back(i);
            units.push_back(std::move(u));
        }
    }
    const int nu_units = int(units.size());
    std::vector<int> order(nu_units);
    std::iota(order.begin(), order.end(), 0);
    if (!tigre && d.subset_order == YK_ORDER_GOLDEN && nu_units > 1) order = golden_order(nu_units);

    size_t max_k = 0, row_total = 0;
    for (Unit& u : units) {
        if (contiguous(u.views)) {
            u.op = p;
            u.first = u.views.front();
        } else {
            std::vector<yk_view> sub;
            for (int i : u.views) sub.push_back(p->views[i]);
            u.own.reset(new yk_projector(p->ctx, sub.data(), int(sub.size()), p->nu, p->nv, p->grid,
                float(p->scan_range)));
            u.own->scratch_owner = p;
            u.op = u.own.get();
            u.first = 0;
        }
        u.row_offset = row_total;
        row_total += u.views.size() * view_n;
        max_k = std::max(max_k, u.views.size());
    }

    // TIGRE row weights come from a 2 x 2 x 2 support volume
    std::unique_ptr<yk_projector> coarse;
    if (tigre) {
        const yk_view& v0 = p->views[0];
        const double dv = std::sqrt(double(v0.detV[0]) * v0.detV[0] + double(v0.detV[1]) * v0.detV[1] +
                                    double(v0.detV[2]) * v0.detV[2]);
        yk_volume_grid cg = p->grid;
        cg.nx = cg.ny = cg.nz = 2;
        cg.voxel_x_mm = float(p->grid.nx * double(p->grid.voxel_x_mm) * 1.1 / 2.0);
        cg.voxel_y_mm = float(p->grid.ny * double(p->grid.voxel_y_mm) * 1.1 / 2.0);
        cg.voxel_z_mm = float(std::max(p->nv * dv, p->grid.nz * double(p->grid.voxel_z_mm)) / 2.0);
        coarse.reset(new yk_projector(p->ctx, p->views.data(), na, p->nu, p->nv, cg, float(p->scan_range)));
    }
    const float min_voxel = std::min({ p->grid.voxel_x_mm, p->grid.voxel_y_mm, p->grid.voxel_z_mm });

    // weights: every unit's row weights; column weights cached while all of
    // them fit in 4 GiB, else recomputed at each update
    const bool cache_cols = double(nu_units) * double(nx) * 4.0 <= 4.0 * 1073741824.0;
    DeviceArray<float> rows(row_total), cols(cache_cols ? nx * nu_units : nx);
    DeviceArray<float> fwd(max_k * view_n), bsub(max_k * view_n), bp(nx), ones_vol(tigre ? 8 : nx),
        ones_proj(max_k * view_n), grad;
    launch_fill(ones_vol.get(), 1.f, tigre ? 8 : nx, s);
    launch_fill(ones_proj.get(), 1.f, max_k * view_n, s);
    // the SIRT batch backend and the TIGRE backend mask voxels no ray reaches
    const bool mask_cols = tigre || nu_units == 1;
    auto column_weights = [&](const Unit& u, float* out) {
        u.op->backproject(d.back_kind, ones_proj.get(), out, u.first, int(u.views.size()), 0);
        if (mask_cols) weight_prepare(out, 0.f, 0, nx, s);
    };
    for (int k = 0; k < nu_units; ++k) {
        const Unit& u = units[k];
        const int cnt = int(u.views.size());
        float* r = rows.get() + u.row_offset;
        if (tigre) {
            coarse->project(d.forward_kind, ones_vol.get(), r, u.views.front(), cnt, 0);
            weight_prepare(r, 0.5f * min_voxel, 1, size_t(cnt) * view_n, s);
        } else {
            u.op->project(d.forward_kind, ones_vol.get(), r, u.first, cnt, 0);
        }
        if (cache_cols) column_weights(u, cols.get() + size_t(k) * nx);
    }
    if (d.tv_iterations > 0 && d.tv_strength > 0.f) grad.resize(nx);

    // convergence monitor state (AlgebraicReconstructorEx::prepareConvergence_, :236-256)
    const bool watch_residual = d.relative_residual_tolerance > 0.f || d.relative_improvement_tolerance > 0.f;
    const bool watch_update = d.relative_update_tolerance > 0.f;
    DeviceArray<float> resid(watch_residual ? size_t(na) * view_n : 0), prev(watch_update ? nx : 0);
    if (watch_update) YK_HIP(hipMemcpyAsync(prev.get(), d_x, nx * sizeof(float), hipMemcpyDeviceToDevice, s));
    const double measured = watch_residual ? std::sqrt(std::max(0.0, p->dot(d_b, d_b, size_t(na) * view_n))) : 0.0;
    const double eps = d.epsilon;
    yk_iterative_stats st{};
    const double nan = std::numeric_limits<double>::quiet_NaN();
    st.projection_residual_l2 = st.relative_projection_residual = nan;
    st.relative_volume_update = st.relative_residual_improvement = nan;
    double previous_residual = nan;
    int patience = 0;

    float lambda = d.relaxation;
    const float per_update = tigre ? 1.f : std::pow(d.relaxation_reduction, 1.f / float(nu_units));
    float tv_strength = d.tv_strength;
    const VolumeGeom g = volume_geom(p->grid);
    for (int outer = 0; outer < d.iterations; ++outer) {
        for (int k = 0; k < nu_units; ++k) {
            const Unit& u = units[order[k]];
            const int cnt = int(u.views.size());
            const size_t n = size_t(cnt) * view_n;
            const float* b = d_b + size_t(u.views.front()) * view_n;
            if (!contiguous(u.views)) {
                for (int j = 0; j < cnt; ++j)
                    YK_HIP(hipMemcpyAsync(bsub.get() + size_t(j) * view_n, d_b + size_t(u.views[j]) * view_n,
                        view_n * sizeof(float), hipMemcpyDeviceToDevice, s));
                b = bsub.get();
            }
            const float* c = cols.get() + (cache_cols ? size_t(order[k]) * nx : 0);
            if (!cache_cols) column_weights(u, cols.get());
            u.op->project(d.forward_kind, d_x, fwd.get(), u.first, cnt, 0);
            algebraic_residual(fwd.get(), b, rows.get() + u.row_offset, d.epsilon, tigre ? 1 : 0, n, s);
            u.op->backproject(d.back_kind, fwd.get(), bp.get(), u.first, cnt, 0);
            algebraic_update(d_x, bp.get(), c, lambda, d.epsilon, d.use_min, d.min_value, d.use_max, d.max_value,
                nx, s);
            lambda *= per_update;
        }
        if (tigre) lambda *= d.relaxation_reduction;
        if (d.tv_iterations > 0 && tv_strength > 0.f) {
            p->ctx->reduce_scratch.resize(1024);
            for (int j = 0; j < d.tv_iterations; ++j) {
                launch_tv_gradient(d_x, grad.get(), p->ctx->reduce_scratch.get(), 1024, g, d.tv_epsilon, s,
                    d.tv_dimensions == 2 ? 2 : 3);
                launch_axpy(d_x, grad.get(), -tv_strength, nx, s);
            }
            tv_strength *= d.tv_strength_reduction;
        }
        if (d.use_min || d.use_max) launch_bounds_clamp(d_x, d.use_min, d.min_value, d.use_max, d.max_value, nx, s);
        st.completed_iterations = outer + 1;
        report_progress(d.progress_callback, d.progress_user_data, outer + 1, d.iterations, s);
        // shouldCheckConvergence_ / checkConvergence_ (YkAlgebraicReconstructorEx.hpp:258-330)
        if ((!watch_residual && !watch_update) ||
            ((outer + 1) % d.check_interval != 0 && outer + 1 != d.iterations))
            continue;
        ++st.convergence_checks;
        bool residual_ok = false, update_ok = false, stagnation_ok = false;
        if (watch_residual) {
            p->project(d.forward_kind, d_x, resid.get(), 0, na, 0);
            launch_xpby(resid.get(), d_b, -1.f, size_t(na) * view_n, s);  // b - A x
            const double r = std::sqrt(std::max(0.0, p->dot(resid.get(), resid.get(), size_t(na) * view_n)));
            st.projection_residual_l2 = r;
            st.relative_projection_residual = r / std::max(measured, eps);
            residual_ok = d.relative_residual_tolerance > 0.f &&
                          st.relative_projection_residual <= d.relative_residual_tolerance;
            if (std::isfinite(previous_residual)) {
                st.relative_residual_improvement = (previous_residual - r) / std::max(previous_residual, eps);
                stagnation_ok = d.relative_improvement_tolerance > 0.f && st.relative_residual_improvement >= 0.0 &&
                                st.relative_residual_improvement <= d.relative_improvement_tolerance;
            }
            previous_residual = r;
        }
        if (watch_update) {
            const double pn = std::sqrt(std::max(0.0, p->dot(prev.get(), prev.get(), nx)));
            launch_xpby(prev.get(), d_x, -1.f, nx, s);  // x - x_prev
            st.relative_volume_update = std::sqrt(std::max(0.0, p->dot(prev.get(), prev.get(), nx))) / std::max(pn, eps);
            update_ok = st.relative_volume_update <= d.relative_update_tolerance;
            YK_HIP(hipMemcpyAsync(prev.get(), d_x, nx * sizeof(float), hipMemcpyDeviceToDevice, s));
        }
        const bool satisfied =
            st.completed_iterations >= d.minimum_iterations && (residual_ok || update_ok || stagnation_ok);
        patience = satisfied ? patience + 1 : 0;
        if (patience < d.patience) continue;
        st.stopped_by_relative_residual = residual_ok;
        st.stopped_by_relative_update = update_ok;
        st.stopped_by_stagnation = stagnation_ok;
        break;
    }
    // the sub-projectors and workspaces are released when this returns: wait
    // for the queued work first
    YK_HIP(hipStreamSynchronize(s));
    if (out) {
        st.struct_size = sizeof(yk_iterative_stats);
        *out = st;
    }
}

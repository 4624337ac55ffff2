// CGLS with the reference's two strategies over the projector's forward model
// A and back model B (CglsReconstructorEx,
// src/Reconstruction/Iterative/Flat/YkCglsReconstructorEx.hpp:10-100;
// CglsAstraBackend, YkCglsBackends.hpp:22-190; CglsRobustBackend, :222-560;
// IterativeConvergenceConfig, src/Reconstruction/Iterative/Common/
// YkIterativeConvergence.hpp:9-47).  Dot products are reduced in FP64 on the
// device; the vectors stay FP32 in HBM.
#include <algorithm>
#include <cmath>
#include <limits>

#include "fdk.hpp"

using namespace yk;

namespace {

struct Vectors {
    DeviceArray<float> r, q, p, s, x_prev, ax;
};

bool convergence_enabled(const yk_cgls_desc& d)
{
    return d.relative_residual_tolerance > 0.f || d.relative_update_tolerance > 0.f ||
           d.relative_improvement_tolerance > 0.f;
}

}  // namespace

void run_cgls(yk_projector* p, const yk_cgls_desc& d, const float* d_b, float* d_x, yk_iterative_stats* out,
    double* residual_norms)
{
    // CglsReconstructorEx::prepare validation (YkCglsReconstructorEx.hpp:45-51)
    if (d.strategy != YK_CGLS_ROBUST_RESTART && d.strategy != YK_CGLS_ASTRA)
        fail(YK_ERR_INVALID_CONFIG, "unknown CGLS strategy");
    if (d.iterations <= 0 || !std::isfinite(d.epsilon) || !(d.epsilon > 0.f))
        fail(YK_ERR_INVALID_CONFIG, "CGLS needs iterations >= 1 and epsilon > 0");
    if (d.use_min && d.use_max && d.min_value > d.max_value)
        fail(YK_ERR_INVALID_CONFIG, "min_value exceeds max_value");
    if (!(d.relative_residual_tolerance >= 0.f) || !(d.relative_update_tolerance >= 0.f) ||
        !(d.relative_improvement_tolerance >= 0.f) || d.minimum_iterations < 0 || d.check_interval <= 0 ||
        d.patience <= 0)
        fail(YK_ERR_INVALID_CONFIG, "invalid convergence settings");
    if (d.strategy == YK_CGLS_ASTRA && convergence_enabled(d))
        fail(YK_ERR_INVALID_CONFIG, "the ASTRA strategy has no convergence monitor");
    if (d.forward_kind != YK_FP_JOSEPH && d.forward_kind != YK_FP_SIDDON)
        fail(YK_ERR_UNSUPPORTED, "unknown forward projector");
    if (out && out->struct_size < sizeof(yk_iterative_stats)) fail(YK_ERR_INVALID_CONFIG, "yk_iterative_stats is truncated");

    p->ctx->activate();
    hipStream_t s = p->ctx->stream;
    const int n = p->n_views;
    const size_t np = size_t(n) * p->nu * p->nv;
    const size_t nx = size_t(p->grid.nx) * p->grid.ny * p->grid.nz;
    const double eps = d.epsilon;
    yk_iterative_stats st{};
    st.projection_residual_l2 = std::numeric_limits<double>::quiet_NaN();
    st.relative_projection_residual = std::numeric_limits<double>::quiet_NaN();
    st.relative_volume_update = std::numeric_limits<double>::quiet_NaN();
    st.relative_residual_improvement = std::numeric_limits<double>::quiet_NaN();

    Vectors v;
    v.r.resize(np);
    v.q.resize(np);
    v.p.resize(nx);
    v.s.resize(nx);
    auto bounds = [&] {
        if (d.use_min || d.use_max) launch_bounds_clamp(d_x, d.use_min, d.min_value, d.use_max, d.max_value, nx, s);
    };
    // r = b - A x, p = B r, gamma = <p, p>
    auto initialize = [&]() {
        p->project(d.forward_kind, d_x, v.q.get(), 0, n, 0);
        launch_scale_copy(v.r.get(), d_b, 1.f, np, s);
        launch_axpy(v.r.get(), v.q.get(), -1.f, np, s);
        p->backproject(d.back_kind, v.r.get(), v.p.get(), 0, n, 0);
        return p->dot(v.p.get(), v.p.get(), nx);
    };

    if (d.strategy == YK_CGLS_ASTRA) {
        // CglsAstraBackend::run (YkCglsBackends.hpp:84-156)
        double gamma = initialize();
        if (residual_norms) residual_norms[0] = std::sqrt(p->dot(v.r.get(), v.r.get(), np));
        int k = 0;
        for (; k < d.iterations; ++k) {
            p->project(d.forward_kind, v.p.get(), v.q.get(), 0, n, 0);
            const double ww = p->dot(v.q.get(), v.q.get(), np);
            if (ww < eps) {
                st.stopped_by_breakdown = 1;
                break;
            }
            const double alpha = gamma / ww;
            launch_axpy(d_x, v.p.get(), float(alpha), nx, s);
            launch_axpy(v.r.get(), v.q.get(), float(-alpha), np, s);
            p->backproject(d.back_kind, v.r.get(), v.s.get(), 0, n, 0);
            const double gamma1 = p->dot(v.s.get(), v.s.get(), nx);
            const double beta = gamma1 / (gamma + eps);
            gamma = gamma1;
            launch_xpby(v.p.get(), v.s.get(), float(beta), nx, s);
            bounds();
            st.completed_iterations = k + 1;
            if (residual_norms) residual_norms[k + 1] = std::sqrt(p->dot(v.r.get(), v.r.get(), np));
            report_progress(d.progress_callback, d.progress_user_data, k + 1, d.iterations, s);
        }
        for (int j = k; j < d.iterations && residual_norms; ++j) residual_norms[j + 1] = residual_norms[j];
    } else {
        // CglsRobustBackend::run (YkCglsBackends.hpp:281-400): the true
        // residual after each step; a step that raises it is rolled back, and
        // the recursion restarts once from there (twice in a row stops it).
        v.x_prev.resize(nx);
        v.ax.resize(np);
        const double measured_norm = std::sqrt(std::max(0.0, p->dot(d_b, d_b, np)));
        double gamma = initialize();
        if (residual_norms) residual_norms[0] = std::sqrt(p->dot(v.r.get(), v.r.get(), np));
        int re_init_at = -1, patience = 0;
        double l2_prev = std::numeric_limits<double>::max();
        double previous_checked = std::numeric_limits<double>::quiet_NaN();
        int k = 0;
        for (; k < d.iterations; ++k) {
            YK_HIP(hipMemcpyAsync(v.x_prev.get(), d_x, nx * sizeof(float), hipMemcpyDeviceToDevice, s));
            p->project(d.forward_kind, v.p.get(), v.q.get(), 0, n, 0);
            const double qq = p->dot(v.q.get(), v.q.get(), np);
            if (qq < eps) {
                st.stopped_by_breakdown = 1;
                break;
            }
            const double alpha = gamma / qq;
            launch_axpy(d_x, v.p.get(), float(alpha), nx, s);
            p->project(d.forward_kind, d_x, v.ax.get(), 0, n, 0);
            launch_xpby(v.ax.get(), d_b, -1.f, np, s);  // b - A x
            const double l2 = std::sqrt(p->dot(v.ax.get(), v.ax.get(), np));
            if (k > 0 && l2 > l2_prev) {
                YK_HIP(hipMemcpyAsync(d_x, v.x_prev.get(), nx * sizeof(float), hipMemcpyDeviceToDevice, s));
                if (residual_norms) residual_norms[k + 1] = residual_norms[k];  // x unchanged
                if (re_init_at + 1 == k || !d.restart_on_divergence) {
                    st.stopped_by_divergence = 1;
                    break;
                }
                re_init_at = k;
                gamma = initialize();
                l2_prev = std::numeric_limits<double>::max();
                continue;
            }
            l2_prev = l2;
            launch_axpy(v.r.get(), v.q.get(), float(-alpha), np, s);
            p->backproject(d.back_kind, v.r.get(), v.s.get(), 0, n, 0);
            const double gamma1 = p->dot(v.s.get(), v.s.get(), nx);
            const double beta = gamma1 / (gamma + eps);
            gamma = gamma1;
            launch_xpby(v.p.get(), v.s.get(), float(beta), nx, s);
            bounds();
            st.completed_iterations = k + 1;
            if (residual_norms) residual_norms[k + 1] = l2;
            report_progress(d.progress_callback, d.progress_user_data, k + 1, d.iterations, s);
            // checkConvergence_ (YkCglsBackends.hpp:454-500)
            if (!convergence_enabled(d) ||
                (st.completed_iterations % d.check_interval != 0 && st.completed_iterations != d.iterations))
                continue;
            ++st.convergence_checks;
            st.projection_residual_l2 = l2;
            st.relative_projection_residual = l2 / std::max(measured_norm, eps);
            const bool residual_ok =
                d.relative_residual_tolerance > 0.f && st.relative_projection_residual <= d.relative_residual_tolerance;
            bool stagnation_ok = false;
            if (std::isfinite(previous_checked)) {
                st.relative_residual_improvement = (previous_checked - l2) / std::max(previous_checked, eps);
                stagnation_ok = d.relative_improvement_tolerance > 0.f && st.relative_residual_improvement >= 0.0 &&
                                st.relative_residual_improvement <= d.relative_improvement_tolerance;
            }
            previous_checked = l2;
            bool update_ok = false;
            if (d.relative_update_tolerance > 0.f) {
                const double prev_norm = std::sqrt(std::max(0.0, p->dot(v.x_prev.get(), v.x_prev.get(), nx)));
                launch_xpby(v.x_prev.get(), d_x, -1.f, nx, s);  // x - x_prev
                st.relative_volume_update =
                    std::sqrt(std::max(0.0, p->dot(v.x_prev.get(), v.x_prev.get(), nx))) / std::max(prev_norm, eps);
                update_ok = st.relative_volume_update <= d.relative_update_tolerance;
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
        // entries 0..k are written, and k + 1 too unless the step broke down
        const int written = st.stopped_by_breakdown ? k : k + 1;
        for (int j = std::min(written, d.iterations); j < d.iterations && residual_norms; ++j)
            residual_norms[j + 1] = residual_norms[j];
    }
    YK_HIP(hipStreamSynchronize(s));
    if (out) {
        st.struct_size = sizeof(yk_iterative_stats);
        *out = st;
    }
}

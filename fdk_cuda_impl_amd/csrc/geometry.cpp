// FP64 host geometry: regular flat circular scan builder, FDK derived
// geometry, affine projection coefficients and ramp-filter weights.
// All math is double precision; only the final records are rounded to float.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>

#include "yk_internal.hpp"

namespace yk {
namespace {

constexpr double kPi = 3.14159265358979323846;

inline double dot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void cross(const double* a, const double* b, double* o)
{
    const double r0 = a[1] * b[2] - a[2] * b[1];
    const double r1 = a[2] * b[0] - a[0] * b[2];
    const double r2 = a[0] * b[1] - a[1] * b[0];
    o[0] = r0; o[1] = r1; o[2] = r2;
}
inline void normalize(double* v)
{
    const double l2 = dot(v, v);
    if (!(l2 > 0.0)) { v[0] = v[1] = v[2] = 0.0; return; }
    const double s = 1.0 / std::sqrt(l2);
    v[0] *= s; v[1] *= s; v[2] *= s;
}
// Rodrigues rotation (geometry/YkDetectorGeometryBuilders.hpp:130-137)
inline void rotate_axis(double* value, const double* axis, double angle)
{
    double unit[3] = { axis[0], axis[1], axis[2] };
    normalize(unit);
    const double c = std::cos(angle), s = std::sin(angle);
    double cr[3];
    cross(unit, value, cr);
    const double d = dot(unit, value) * (1.0 - c);
    for (int k = 0; k < 3; ++k) value[k] = value[k] * c + cr[k] * s + unit[k] * d;
}

bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

}  // namespace

ViewD widen(const yk_view& v)
{
    ViewD d{};
    for (int k = 0; k < 3; ++k) {
        d.src[k] = v.src[k];
        d.detS[k] = v.detS[k];
        d.detU[k] = v.detU[k];
        d.detV[k] = v.detV[k];
    }
    d.angle = v.angle[0];
    return d;
}

// Regular flat circular scan (validateRegularScan / buildRegularAngles /
// buildScannerViewFramesImpl / detectorAxes / buildFlatProjectionGeometry,
// geometry/YkSystemGeometry.hpp:36-73, YkTrajectoryGeometryBuilder.hpp:29-67,
// YkDetectorGeometryBuilders.hpp:139-200).
void build_views(const yk_system& sys, std::vector<yk_view>& views)
{
    const yk_circular_scan& sc = sys.scan;
    const yk_flat_detector& dt = sys.detector;
    if (sc.total_views <= 0 || sc.views_per_turn < 2 ||
        (sc.rotation_direction != 1 && sc.rotation_direction != -1) ||
        !(sc.sid_mm > 0.f) || !(sc.sdd_mm > sc.sid_mm) ||
        !std::isfinite(sc.start_angle_rad) || !finite3(sc.source_offset_mm) ||
        !std::isfinite(sc.z_mm))
        fail(YK_ERR_INVALID_GEOMETRY, "invalid circular scan");
    if (dt.channels <= 0 || dt.rows <= 0 || !(dt.channel_size_mm > 0.f) ||
        !(dt.row_size_mm > 0.f) || !finite3(dt.pose.offset_unv_mm) ||
        !std::isfinite(dt.pose.tilt_u_rad) || !std::isfinite(dt.pose.tilt_v_rad) ||
        !std::isfinite(dt.pose.tilt_n_rad))
        fail(YK_ERR_INVALID_GEOMETRY, "invalid flat detector");
    views.resize(sc.total_views);
    const double step = sc.rotation_direction * 2.0 * kPi / sc.views_per_turn;
    const double det_dist = double(sc.sdd_mm) - double(sc.sid_mm);
    const double cu = 0.5 * (dt.channels - 1), cv = 0.5 * (dt.rows - 1);
    for (int i = 0; i < sc.total_views; ++i) {
        const double a = double(sc.start_angle_rad) + i * step;
        const double c = std::cos(a), s = std::sin(a);
        const double sl[3] = { sc.source_offset_mm[0], -double(sc.sid_mm) + sc.source_offset_mm[1],
                               sc.source_offset_mm[2] };
        const double src[3] = { c * sl[0] - s * sl[1], s * sl[0] + c * sl[1], sl[2] + sc.z_mm };
        const double prin[3] = { -s * det_dist, c * det_dist, double(sc.z_mm) };
        double u[3] = { c, s, 0.0 }, v[3] = { 0.0, 0.0, 1.0 }, n[3] = { s, -c, 0.0 };
        rotate_axis(v, u, dt.pose.tilt_u_rad);
        rotate_axis(n, u, dt.pose.tilt_u_rad);
        rotate_axis(u, v, dt.pose.tilt_v_rad);
        rotate_axis(n, v, dt.pose.tilt_v_rad);
        rotate_axis(u, n, dt.pose.tilt_n_rad);
        rotate_axis(v, n, dt.pose.tilt_n_rad);
        normalize(u);
        normalize(v);
        cross(u, v, n);
        normalize(n);
        const double d2s[3] = { src[0] - prin[0], src[1] - prin[1], src[2] - prin[2] };
        if (dot(n, d2s) < 0.0) { n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2]; }
        yk_view& o = views[i];
        std::memset(&o, 0, sizeof(o));
        const float* off = dt.pose.offset_unv_mm;
        for (int k = 0; k < 3; ++k) {
            const double center = prin[k] + u[k] * off[0] + n[k] * off[1] + v[k] * off[2];
            const double du = u[k] * dt.channel_size_mm, dv = v[k] * dt.row_size_mm;
            o.src[k] = float(src[k]);
            o.detS[k] = float(center - du * cu - dv * cv);
            o.detU[k] = float(du);
            o.detV[k] = float(dv);
        }
        o.angle[0] = float(a);
    }
}

// GeoDerivedManagerVec::build_geo_params
// (src/Reconstruction/Analytic/Circular/Flat/FDK/YkFDKVecGeoDerived.hpp:58-304)
// with deriveProjectionFrame (geometry/YkProjectionGeometry.hpp:147-223).
void derive_geometry(const std::vector<yk_view>& views, int nu, int nv, double scan_range,
    std::vector<DerivedD>& out)
{
    const int n = int(views.size());
    if (nu <= 0 || nv <= 0 || n <= 0) fail(YK_ERR_INVALID_GEOMETRY, "empty geometry");
    if (!(scan_range > 0.0) || !std::isfinite(scan_range))
        fail(YK_ERR_INVALID_CONFIG, "scan range must be positive");
    std::vector<double> theta(n);
    for (int i = 0; i < n; ++i) theta[i] = views[i].angle[0];
    for (int i = 1; i < n; ++i) {
        double t = theta[i];
        const double p = theta[i - 1];
        while (t - p > kPi) t -= 2.0 * kPi;
        while (t - p < -kPi) t += 2.0 * kPi;
        theta[i] = t;
    }
    out.assign(n, DerivedD{});
    for (int i = 0; i < n; ++i) {
        const ViewD g = widen(views[i]);
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(g.src[k]) || !std::isfinite(g.detS[k]) ||
                !std::isfinite(g.detU[k]) || !std::isfinite(g.detV[k]))
                fail(YK_ERR_INVALID_GEOMETRY, "non-finite view vector");
        DerivedD& d = out[i];
        d.theta = theta[i];
        double dt;
        if (n <= 1) dt = 1e-8;
        else if (i == 0) dt = theta[1] - theta[0];
        else if (i == n - 1) dt = theta[n - 1] - theta[n - 2];
        else dt = 0.5 * (theta[i + 1] - theta[i - 1]);
        dt = std::fabs(dt);
        if (dt < 1e-8) dt = 1e-8;
        d.dtheta = dt;
        d.fscale = 2.0 * kPi / scan_range;
        d.du = std::sqrt(dot(g.detU, g.detU));
        d.dv = std::sqrt(dot(g.detV, g.detV));
        const double cu = 0.5 * (nu - 1), cv = 0.5 * (nv - 1);
        double cdir[3], normal[3];
        for (int k = 0; k < 3; ++k) cdir[k] = g.detS[k] + cu * g.detU[k] + cv * g.detV[k] - g.src[k];
        cross(g.detV, g.detU, normal);
        if (!(dot(normal, normal) > 1e-12)) fail(YK_ERR_INVALID_GEOMETRY, "degenerate detector axes");
        normalize(normal);
        if (dot(normal, cdir) < 0.0) for (double& c : normal) c = -c;
        const double ds[3] = { g.detS[0] - g.src[0], g.detS[1] - g.src[1], g.detS[2] - g.src[2] };
        const double plane = dot(ds, normal);
        if (!(plane > 1e-6)) fail(YK_ERR_INVALID_GEOMETRY, "detector plane behind the source");
        double ray[3] = { -g.src[0], -g.src[1], 0.0 };
        const double sod = std::sqrt(dot(ray, ray));
        if (!(sod > 1e-10)) fail(YK_ERR_INVALID_GEOMETRY, "source on the rotation axis");
        ray[0] /= sod;
        ray[1] /= sod;
        const double den = dot(ray, normal);
        if (std::fabs(den) < 1e-12) fail(YK_ERR_INVALID_GEOMETRY, "radial ray parallel to detector");
        const double t = plane / den;
        if (t <= 0.0) fail(YK_ERR_INVALID_GEOMETRY, "detector behind the source");
        double r0[3], D[3];
        for (int k = 0; k < 3; ++k) {
            const double pp = g.src[k] + ray[k] * t;
            r0[k] = pp - g.src[k];
            D[k] = pp - g.detS[k];
        }
        const double UU = dot(g.detU, g.detU), VV = dot(g.detV, g.detV), UV = dot(g.detU, g.detV);
        const double gram = UU * VV - UV * UV;
        if (std::fabs(gram) < 1e-20) fail(YK_ERR_INVALID_GEOMETRY, "singular detector basis");
        const double DU = dot(D, g.detU), DV = dot(D, g.detV);
        d.sod = sod;
        d.sdd_radial = std::sqrt(dot(r0, r0));
        d.sdd_plane = plane;
        d.off_u = (DU * VV - DV * UV) / gram - cu;
        d.off_v = (-DU * UV + DV * UU) / gram - cv;
        for (int k = 0; k < 3; ++k) { d.radial[k] = ray[k]; d.normal[k] = normal[k]; }
        d.uu = UU; d.vv = VV; d.uv = UV;
    }
    // repeated adjacent angles make dtheta undefined
    // (validateApproximateFdkGeometry_, YkFdkPipeline.hpp:845-852)
    for (int i = 1; i < n; ++i)
        if (std::fabs(double(views[i].angle[0]) - double(views[i - 1].angle[0])) < 1e-5)
            fail(YK_ERR_INVALID_GEOMETRY, "repeated adjacent view angles");
}

// fdk_precompute_coeffs_kernel in FP64
// (.../FDK/kernels/YkFDKBpPrecompute.cu:74-160)
void fdk_coefficients(const std::vector<yk_view>& views, const std::vector<DerivedD>& derived,
    std::vector<CoeffD>& out)
{
    out.assign(views.size(), CoeffD{});
    for (size_t i = 0; i < views.size(); ++i) {
        const ViewD g = widen(views[i]);
        const DerivedD& d = derived[i];
        CoeffD& c = out[i];
        const double SDD = d.sdd_plane;
        const double s_n = dot(g.src, d.normal), s_u = dot(g.src, g.detU), s_v = dot(g.src, g.detV);
        double off_u = 0.0, off_v = 0.0;
        for (int k = 0; k < 3; ++k) {
            off_u += (g.detS[k] - g.src[k]) * g.detU[k];
            off_v += (g.detS[k] - g.src[k]) * g.detV[k];
        }
        double Cu[4], Cv[4];
        for (int k = 0; k < 3; ++k) {
            Cu[k] = SDD * g.detU[k] - off_u * d.normal[k];
            Cv[k] = SDD * g.detV[k] - off_v * d.normal[k];
        }
        Cu[3] = -SDD * s_u + off_u * s_n;
        Cv[3] = -SDD * s_v + off_v * s_n;
        const double gram = d.uu * d.vv - d.uv * d.uv;
        const double inv = std::fabs(gram) < 1e-20 ? 0.0 : 1.0 / gram;
        for (int k = 0; k < 4; ++k) {
            c.cu[k] = (Cu[k] * d.vv - Cv[k] * d.uv) * inv;
            c.cv[k] = (-Cu[k] * d.uv + Cv[k] * d.uu) * inv;
        }
        for (int k = 0; k < 3; ++k) { c.cd[k] = d.normal[k]; c.cr[k] = d.radial[k]; }
        c.cd[3] = -s_n;
        c.cr[3] = -dot(g.src, d.radial);
        c.wbase = d.sod * d.sod * d.dtheta * d.fscale;
    }
}

int padded_n(int nu)
{
    int n = 1;
    while (n < 2 * nu) n <<= 1;
    return n;
}

int start_u(int nu, int n, double offset_u_pix)
{
    const double axis = (nu - 1) * 0.5 + offset_u_pix;
    return int(std::lrint(n * 0.5 - axis));
}

// mainAxis of the world centre ray (src/common/cuda/operators/YkOperatorKernelTypes.cuh:133-141)
int main_axis(const yk_view& v, int nu, int nv)
{
    const ViewD g = widen(v);
    const double cu = 0.5 * (nu - 1), cv = 0.5 * (nv - 1);
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = std::fabs(g.detS[k] + cu * g.detU[k] + cv * g.detV[k] - g.src[k]);
    if (d[0] >= d[1] && d[0] >= d[2]) return 0;
    if (d[1] >= d[0] && d[1] >= d[2]) return 1;
    return 2;
}

// validAnalyticFilter (src/Filter/YkAnalyticFilterSettings.hpp:6-13)
void validate_filter(const yk_filter& p)
{
    if (p.type < YK_FILTER_NONE || p.type > YK_FILTER_CUSTOM || !std::isfinite(p.cutoff) ||
        !(p.cutoff > 0.f) || !(p.cutoff <= 0.5f) || !std::isfinite(p.gain) || !(p.gain > 0.f) ||
        !std::isfinite(p.butterworth_order) || !(p.butterworth_order > 0.f) ||
        !std::isfinite(p.kaiser_beta) || !(p.kaiser_beta >= 0.f) ||
        !std::isfinite(p.tukey_alpha) || !(p.tukey_alpha >= 0.f) || !(p.tukey_alpha <= 1.f))
        fail(YK_ERR_INVALID_CONFIG, "invalid analytic filter parameters");
    if (p.source < YK_FILTER_SOURCE_ANALYTIC_FREQ || p.source > YK_FILTER_SOURCE_SPATIAL_RAMP_FFT)
        fail(YK_ERR_INVALID_CONFIG, "unknown filter weight source");
    if (p.extract_mode != YK_RAMP_EXTRACT_REAL && p.extract_mode != YK_RAMP_EXTRACT_MAGNITUDE)
        fail(YK_ERR_INVALID_CONFIG, "unknown ramp extract mode");
    if (p.source == YK_FILTER_SOURCE_SPATIAL_RAMP_FFT) {
        if (!p.spatial_ramp || p.spatial_ramp_size <= 0 || (p.spatial_ramp_size & 1) == 0)
            fail(YK_ERR_INVALID_CONFIG, "spatial ramp needs an odd number of taps");
    } else if (p.type == YK_FILTER_CUSTOM) {
        if (!p.custom_weights || p.custom_weights_size <= 0)
            fail(YK_ERR_INVALID_CONFIG, "custom filter without weights");
    }
}

namespace {

// Abramowitz & Stegun 9.8.1 / 9.8.2 (the form of src/Filter/YkCreateFilterKernelHelpers.cuh:21-36)
double bessel_i0(double x)
{
    const double ax = std::fabs(x);
    if (ax < 3.75) {
        double y = x / 3.75;
        y *= y;
        return 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 +
            y * (0.2659732 + y * (0.0360768 + y * 0.0045813)))));
    }
    const double y = 3.75 / ax;
    return (std::exp(ax) / std::sqrt(ax)) * (0.39894228 + y * (0.01328592 +
        y * (0.00225319 + y * (-0.00157565 + y * (0.00916281 +
        y * (-0.02057706 + y * (0.02635537 + y * (-0.01647633 + y * 0.00392377))))))));
}

// window_shape_desc (src/Filter/YkCreateFilterKernelHelpers.cuh:38-103)
// window_shape_desc (parametric: Butterworth, Kaiser and Tukey use their
// parameters) or window_shape (those three give 1), the forms the analytic and
// the DFT-built sources use (src/Filter/YkCreateFilterKernelHelpers.cuh:38-103)
double window(const yk_filter& f, double x, bool parametric = true)
{
    x = std::min(std::max(x, 0.0), 1.0);
    if (!parametric && (f.type == YK_FILTER_BUTTERWORTH || f.type == YK_FILTER_KAISER || f.type == YK_FILTER_TUKEY))
        return 1.0;
    switch (f.type) {
    case YK_FILTER_RAMLAK: return 1.0;
    case YK_FILTER_SHEPPLOGAN: {
        const double t = kPi * 0.5 * x;
        return std::fabs(t) < 1e-8 ? 1.0 : std::sin(t) / t;
    }
    case YK_FILTER_COSINE: return std::cos(0.5 * kPi * x);
    case YK_FILTER_HANN: return 0.5 * (1.0 + std::cos(kPi * x));
    case YK_FILTER_HAMMING: return 0.54 + 0.46 * std::cos(kPi * x);
    case YK_FILTER_BLACKMAN: return 0.42 + 0.5 * std::cos(kPi * x) + 0.08 * std::cos(2.0 * kPi * x);
    case YK_FILTER_BUTTERWORTH: {
        const double o = std::max(double(f.butterworth_order), 0.01);
        return 1.0 / (1.0 + std::pow(x, 2.0 * o));
    }
    case YK_FILTER_KAISER: {
        const double b = std::max(double(f.kaiser_beta), 0.0);
        const double t = 2.0 * x - 1.0;
        return bessel_i0(b * std::sqrt(std::max(0.0, 1.0 - t * t))) / bessel_i0(b);
    }
    case YK_FILTER_TUKEY: {
        const double a = std::min(std::max(double(f.tukey_alpha), 0.0), 1.0);
        if (a <= 0.0) return 1.0;
        if (x < a * 0.5) return 0.5 * (1.0 + std::cos(kPi * (2.0 * x / a - 1.0)));
        if (x <= 1.0 - a * 0.5) return 1.0;
        return 0.5 * (1.0 + std::cos(kPi * (2.0 * x / a - 2.0 / a + 1.0)));
    }
    default: return 1.0;
    }
}

// Bins 0..n/2 of the forward DFT of a real n-periodic sequence given by its
// nonzero taps (index, value), as the weight the extract mode takes: the real
// part or the magnitude (kernel_extract_weights_from_fft,
// src/Filter/YkCreateFilterKernelKernels.cu:299-320), in FP64.
void dft_extract(const std::vector<std::pair<int, double>>& taps, int n, int mode, std::vector<double>& w)
{
    std::vector<double> c(n), s(n);
    for (int j = 0; j < n; ++j) {
        c[j] = std::cos(2.0 * kPi * j / n);
        s[j] = std::sin(2.0 * kPi * j / n);
    }
    const int nc = n / 2 + 1;
    w.assign(nc, 0.0);
    for (int k = 0; k < nc; ++k) {
        double re = 0.0, im = 0.0;
        for (const auto& t : taps) {
            const int j = int((long long)k * t.first % n);
            re += t.second * c[j];
            im -= t.second * s[j];
        }
        w[k] = mode == YK_RAMP_EXTRACT_MAGNITUDE ? std::hypot(re, im) : re;
    }
}

}  // namespace

// Frequency weights with 1/N baked in, restating build_weights and
// build_custom_weights (src/Filter/YkCreateFilterKernelFFT.hpp:114-258) and
// the kernels they launch (src/Filter/YkCreateFilterKernelKernels.cu:135-345),
// in FP64:
//   SpatialRampFFT (any kind): the taps centred on index 0 of an N-periodic
//     row, / N, DFT, extract; unless Ram-Lak, zero above the cutoff and x
//     gain window_shape(f / cutoff); DC := 0; x gain / du (so a windowed
//     spatial ramp carries the gain twice, as the reference's does).
//   Custom: the caller's N/2 + 1 weights x gain / N / du.
//   None: gain / N.
//   AnalyticFreq: gain f window_shape_desc(f / cutoff) / N / du, f <= cutoff.
//   DiscreteRLFFT: the du = 1 sampled Ram-Lak kernel (1/4 at 0, -1/(pi^2 n^2)
//     at odd n, wrapped at N/2) / N, DFT, extract, DC := 0, zero above the
//     cutoff and x gain window_shape(f / cutoff), / du.
void filter_weights(const yk_filter& f, int n, double du, std::vector<float>& out)
{
    validate_filter(f);
    if (n < 2 || (n & (n - 1))) fail(YK_ERR_INVALID_CONFIG, "padded length must be a power of two");
    if (!(du > 0.0)) fail(YK_ERR_INVALID_GEOMETRY, "detector pixel size must be positive");
    const int nc = n / 2 + 1;
    out.assign(nc, 0.f);
    const double gain = f.gain, cc = f.cutoff > 0.f ? f.cutoff : 0.5;
    if (f.source == YK_FILTER_SOURCE_SPATIAL_RAMP_FFT) {
        const int m = f.spatial_ramp_size;
        if (m > n) fail(YK_ERR_INVALID_CONFIG, "spatial ramp longer than the padded row");
        for (int i = 0; i < m; ++i)
            if (!std::isfinite(f.spatial_ramp[i])) fail(YK_ERR_INVALID_CONFIG, "spatial ramp tap not finite");
        for (int i = 0; i < m / 2; ++i)
            if (!(std::fabs(f.spatial_ramp[i] - f.spatial_ramp[m - 1 - i]) < 1e-4f))
                fail(YK_ERR_INVALID_CONFIG, "spatial ramp not symmetric");
        std::vector<std::pair<int, double>> taps;
        for (int j = 0; j < m; ++j) {
            int idx = j - m / 2;
            if (idx < 0) idx += n;
            taps.emplace_back(idx, double(f.spatial_ramp[j]) / n);
        }
        std::vector<double> w;
        dft_extract(taps, n, f.extract_mode, w);
        if (f.type != YK_FILTER_RAMLAK) {
            for (int k = 0; k < nc; ++k) {
                const double fr = double(k) / n;
                w[k] = fr > cc ? 0.0 : w[k] * gain * window(f, fr / cc, false);
            }
        }
        w[0] = 0.0;
        for (int k = 0; k < nc; ++k) out[k] = float(w[k] * gain / du);
        return;
    }
    if (f.type == YK_FILTER_CUSTOM) {
        if (f.custom_weights_size != nc) fail(YK_ERR_INVALID_CONFIG, "custom filter needs padded N / 2 + 1 weights");
        for (int k = 0; k < nc; ++k) {
            if (!std::isfinite(f.custom_weights[k])) fail(YK_ERR_INVALID_CONFIG, "custom filter weight not finite");
            out[k] = float(double(f.custom_weights[k]) * gain / n / du);
        }
        return;
    }
    if (f.type == YK_FILTER_NONE) {
        for (int k = 0; k < nc; ++k) out[k] = float(gain / n);
        return;
    }
    if (f.source == YK_FILTER_SOURCE_ANALYTIC_FREQ) {
        for (int k = 0; k < nc; ++k) {
            const double fr = double(k) / n;
            const double w = fr > cc ? 0.0 : gain * fr * window(f, fr / cc) / n;
            out[k] = float(w / du);
        }
        return;
    }
    std::vector<std::pair<int, double>> taps;
    taps.emplace_back(0, 0.25 / n);
    for (int u = 1; u < n; ++u) {
        const int nn = u <= n / 2 ? u : u - n;
        if (nn & 1) taps.emplace_back(u, -1.0 / (kPi * kPi * double(nn) * double(nn)) / n);
    }
    std::vector<double> w;
    dft_extract(taps, n, f.extract_mode, w);
    w[0] = 0.0;
    for (int k = 0; k < nc; ++k) {
        const double fr = double(k) / n;
        out[k] = float((fr > cc ? 0.0 : w[k] * gain * window(f, fr / cc, false)) / du);
    }
}


double filter_response_l1(const std::vector<float>& weights, int n)
{
    if (n < 2 || int(weights.size()) != n / 2 + 1) fail(YK_ERR_INVALID_CONFIG, "filter weights of another length");
    // cos(2 pi j / n) for j in [0, n): the argument k m reduces modulo n
    std::vector<double> c(n);
    for (int j = 0; j < n; ++j) c[j] = std::cos(2.0 * M_PI * j / n);
    double l1 = 0.0;
    for (int m = 0; m <= n / 2; ++m) {
        // h[m] = w[0] + (-1)^m w[n/2] + 2 sum_{0<k<n/2} w[k] cos(2 pi k m / n)
        double h = double(weights[0]) + ((m & 1) ? -1.0 : 1.0) * double(weights[n / 2]);
        size_t idx = 0;
        for (int k = 1; k < n / 2; ++k) {
            idx += size_t(m);
            if (idx >= size_t(n)) idx -= size_t(n);
            h += 2.0 * double(weights[k]) * c[idx];
        }
        // h is even: h[n - m] = h[m]
        l1 += (m == 0 || m == n / 2) ? std::fabs(h) : 2.0 * std::fabs(h);
    }
    return l1;
}

VolumeGeom volume_geom(const yk_volume_grid& g)
{
    VolumeGeom v{};
    v.nx = g.nx; v.ny = g.ny; v.nz = g.nz;
    v.vx = g.voxel_x_mm; v.vy = g.voxel_y_mm; v.vz = g.voxel_z_mm;
    // SVolGeom::origin (geometry/YkVolumeGeometry.hpp:48-54), evaluated in FP64
    v.ox = float(double(g.center_mm[0]) - (g.nx - 1) * 0.5 * double(g.voxel_x_mm));
    v.oy = float(double(g.center_mm[1]) - (g.ny - 1) * 0.5 * double(g.voxel_y_mm));
    v.oz = float(double(g.center_mm[2]) - (g.nz - 1) * 0.5 * double(g.voxel_z_mm));
    return v;
}

// buildVolumeGeometry (geometry/YkSystemGeometry.hpp:99-112)
void validate_volume(const yk_volume_grid& g)
{
    if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0 || !(g.voxel_x_mm > 0.f) || !(g.voxel_y_mm > 0.f) ||
        !(g.voxel_z_mm > 0.f) || !finite3(g.center_mm))
        fail(YK_ERR_INVALID_GEOMETRY, "invalid volume grid");
    if (g.nx > 32767 || g.ny > 32767 || g.nz > 32767)
        fail(YK_ERR_UNSUPPORTED, "volume dimension above 32767");
}

}  // namespace yk

namespace yk {

// The projective part of fdk_precompute_coeffs_kernel
// (.../FDK/kernels/YkFDKBpPrecompute.cu:96-147) from the view alone: detector
// normal oriented away from the source, plane distance, Gram-inverse fold.
void projection_map(const yk_view& view, double c[12])
{
    const ViewD g = widen(view);
    double n[3];
    cross(g.detV, g.detU, n);
    if (!(dot(n, n) > 1e-12)) fail(YK_ERR_INVALID_GEOMETRY, "degenerate detector axes");
    normalize(n);
    const double ds[3] = { g.detS[0] - g.src[0], g.detS[1] - g.src[1], g.detS[2] - g.src[2] };
    double plane = dot(ds, n);
    if (plane < 0.0) { for (double& k : n) k = -k; plane = -plane; }
    if (!(plane > 1e-6)) fail(YK_ERR_INVALID_GEOMETRY, "detector plane through the source");
    const double s_n = dot(g.src, n), s_u = dot(g.src, g.detU), s_v = dot(g.src, g.detV);
    const double off_u = dot(ds, g.detU), off_v = dot(ds, g.detV);
    double Cu[4], Cv[4];
    for (int k = 0; k < 3; ++k) {
        Cu[k] = plane * g.detU[k] - off_u * n[k];
        Cv[k] = plane * g.detV[k] - off_v * n[k];
    }
    Cu[3] = -plane * s_u + off_u * s_n;
    Cv[3] = -plane * s_v + off_v * s_n;
    const double uu = dot(g.detU, g.detU), vv = dot(g.detV, g.detV), uv = dot(g.detU, g.detV);
    const double gram = uu * vv - uv * uv;
    if (std::fabs(gram) < 1e-20) fail(YK_ERR_INVALID_GEOMETRY, "singular detector basis");
    const double inv = 1.0 / gram;
    for (int k = 0; k < 4; ++k) {
        c[k] = (Cu[k] * vv - Cv[k] * uv) * inv;
        c[4 + k] = (-Cu[k] * uv + Cv[k] * uu) * inv;
    }
    c[8] = n[0]; c[9] = n[1]; c[10] = n[2];
    c[11] = -s_n;
}

}  // namespace yk

// Twiddles of the filter kernel's mixed radix-8/4/2 Stockham FFT, pass by pass:
// for span ns and radix R the block holds exp(-2 pi i r k / (ns R)) at
// [(r - 1) ns + k], r = 1..R-1, k = 0..ns-1.  n - 1 entries in all (n >= 2),
// padded to n.
void yk::fft_twiddles(int n, std::vector<float2>& out)
{
    out.clear();
    for (int ns = 1; ns < n;) {
        const int R = ns * 8 <= n ? 8 : (ns * 4 <= n ? 4 : 2);
        for (int r = 1; r < R; ++r)
            for (int k = 0; k < ns; ++k) {
                const double a = -2.0 * 3.14159265358979323846 * double(r) * double(k) / double(ns * R);
                out.push_back(make_float2(float(std::cos(a)), float(std::sin(a))));
            }
        ns *= R;
    }
    out.resize(size_t(std::max(n, 1)), make_float2(0.f, 0.f));
}

/*
 * yk_oracle.c -- FP64 CPU restatement of the YKCBCT flat circular FDK path.
 *
 * TEST INFRASTRUCTURE ONLY (see yk_oracle.h).  Not part of the product.
 * Parity of every function here is pinned in tests/test_oracle_*.py.
 */
#include "yk_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int yko_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* small vector helpers                                                      */
/* ------------------------------------------------------------------------ */
static double dot3(const double* a, const double* b)
{ return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void cross3(const double* a, const double* b, double* out)
{
    double r[3] = { a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                    a[0] * b[1] - a[1] * b[0] };
    out[0] = r[0]; out[1] = r[1]; out[2] = r[2];
}

static void normalize3(double* v)
{
    const double l2 = dot3(v, v);
    if (!(l2 > 0.0)) { v[0] = v[1] = v[2] = 0.0; return; }
    const double s = 1.0 / sqrt(l2);
    v[0] *= s; v[1] *= s; v[2] *= s;
}

/* Rodrigues rotation of value about axis, as
 * include/YKCBCT/geometry/YkDetectorGeometryBuilders.hpp:130-137. */
static void rotate_axis(const double* value, const double* axis, double angle,
    double* out)
{
    double unit[3] = { axis[0], axis[1], axis[2] };
    normalize3(unit);
    const double c = cos(angle), s = sin(angle);
    double cr[3];
    cross3(unit, value, cr);
    const double d = dot3(unit, value) * (1.0 - c);
    double r[3];
    for (int k = 0; k < 3; ++k) r[k] = value[k] * c + cr[k] * s + unit[k] * d;
    out[0] = r[0]; out[1] = r[1]; out[2] = r[2];
}

/* ------------------------------------------------------------------------ */
/* geometry builder                                                          */
/* ------------------------------------------------------------------------ */
/*
 * Regular flat circular scan -> per-view cone vectors.
 *   angles: include/YKCBCT/geometry/YkSystemGeometry.hpp:49-57
 *   frames: include/YKCBCT/geometry/YkTrajectoryGeometryBuilder.hpp:29-67
 *   axes:   include/YKCBCT/geometry/YkDetectorGeometryBuilders.hpp:139-157
 *   views:  include/YKCBCT/geometry/YkDetectorGeometryBuilders.hpp:173-200
 */
int yko_build_circular_flat_views(
    int total_views, int views_per_turn, double start_angle_rad, int direction,
    double sid_mm, double sdd_mm, double z_mm, const double so[3],
    int channels, int rows, double channel_mm, double row_mm,
    const double off[3], double tilt_u, double tilt_v, double tilt_n,
    double* views_out)
{
    if (total_views <= 0 || views_per_turn < 2 ||
        (direction != 1 && direction != -1) || !(sid_mm > 0.0) ||
        !(sdd_mm > sid_mm) || channels <= 0 || rows <= 0 ||
        !(channel_mm > 0.0) || !(row_mm > 0.0))
        return -1;
    const double step = (double)direction * 2.0 * M_PI / (double)views_per_turn;
    const double det_dist = sdd_mm - sid_mm;
    const double cu = 0.5 * (channels - 1), cv = 0.5 * (rows - 1);
    for (int i = 0; i < total_views; ++i) {
        const double a = start_angle_rad + i * step;
        const double c = cos(a), s = sin(a);
        /* source_local = (ox, -sid + oy, oz); detector_local = (0, sdd - sid, 0) */
        const double sl[3] = { so[0], -sid_mm + so[1], so[2] };
        double src[3] = { c * sl[0] - s * sl[1], s * sl[0] + c * sl[1], sl[2] + z_mm };
        double prin[3] = { -s * det_dist, c * det_dist, z_mm };
        double u[3] = { c, s, 0.0 };
        double v[3] = { 0.0, 0.0, 1.0 };
        double n[3] = { s, -c, 0.0 };  /* Rz(0,-1,0) */
        rotate_axis(v, u, tilt_u, v);
        rotate_axis(n, u, tilt_u, n);
        rotate_axis(u, v, tilt_v, u);
        rotate_axis(n, v, tilt_v, n);
        rotate_axis(u, n, tilt_n, u);
        rotate_axis(v, n, tilt_n, v);
        normalize3(u);
        normalize3(v);
        cross3(u, v, n);
        normalize3(n);
        const double d2s[3] = { src[0] - prin[0], src[1] - prin[1], src[2] - prin[2] };
        if (dot3(n, d2s) < 0.0) { n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2]; }
        double* o = views_out + (size_t)i * YKO_VIEW_STRIDE;
        for (int k = 0; k < 3; ++k) {
            const double center = prin[k] + u[k] * off[0] + n[k] * off[1] + v[k] * off[2];
            const double du = u[k] * channel_mm, dv = v[k] * row_mm;
            o[0 + k] = src[k];
            o[3 + k] = center - du * cu - dv * cv;
            o[6 + k] = du;
            o[9 + k] = dv;
        }
        o[12] = a;
    }
    return 0;
}

/*
 * FDK derived geometry (restates GeoDerivedManagerVec::build_geo_params,
 * src/Reconstruction/Analytic/Circular/Flat/FDK/YkFDKVecGeoDerived.hpp:58-304,
 * and deriveProjectionFrame, include/YKCBCT/geometry/YkProjectionGeometry.hpp:147-223).
 */
int yko_derive(const double* views, int n_views, int nu, int nv,
    double scan_range_rad, double* out)
{
    if (nu <= 0 || nv <= 0 || n_views <= 0) return -1;
    double* theta = (double*)malloc(sizeof(double) * (size_t)n_views);
    for (int i = 0; i < n_views; ++i) theta[i] = views[(size_t)i * YKO_VIEW_STRIDE + 12];
    /* unwrap (YkFDKVecGeoDerived.hpp:158-167) */
    for (int i = 1; i < n_views; ++i) {
        double t = theta[i];
        const double p = theta[i - 1];
        while (t - p > M_PI) t -= 2.0 * M_PI;
        while (t - p < -M_PI) t += 2.0 * M_PI;
        theta[i] = t;
    }
    int status = 0;
    for (int i = 0; i < n_views && status == 0; ++i) {
        const double* g = views + (size_t)i * YKO_VIEW_STRIDE;
        double* d = out + (size_t)i * YKO_DERIVED_STRIDE;
        memset(d, 0, sizeof(double) * YKO_DERIVED_STRIDE);
        d[YKO_D_THETA] = theta[i];
        /* dtheta (YkFDKVecGeoDerived.hpp:172-184) */
        double dt;
        if (n_views <= 1) dt = 1e-8;
        else if (i == 0) dt = theta[1] - theta[0];
        else if (i == n_views - 1) dt = theta[n_views - 1] - theta[n_views - 2];
        else dt = 0.5 * (theta[i + 1] - theta[i - 1]);
        if (dt < 0.0) dt = -dt;
        if (dt < 1e-8) dt = 1e-8;
        d[YKO_D_DTHETA] = dt;
        d[YKO_D_FSCALE] = 2.0 * M_PI / scan_range_rad;
        const double* src = g;
        const double* detS = g + 3;
        const double* detU = g + 6;
        const double* detV = g + 9;
        d[YKO_D_DU] = sqrt(dot3(detU, detU));
        d[YKO_D_DV] = sqrt(dot3(detV, detV));
        /* projection frame */
        const double cu = 0.5 * (nu - 1), cv = 0.5 * (nv - 1);
        double center[3], cdir[3], normal[3];
        for (int k = 0; k < 3; ++k) {
            center[k] = detS[k] + cu * detU[k] + cv * detV[k];
            cdir[k] = center[k] - src[k];
        }
        cross3(detV, detU, normal);
        if (!(dot3(normal, normal) > 1e-12)) { status = -2; break; }
        normalize3(normal);
        if (dot3(normal, cdir) < 0.0) { normal[0] = -normal[0]; normal[1] = -normal[1]; normal[2] = -normal[2]; }
        const double ds[3] = { detS[0] - src[0], detS[1] - src[1], detS[2] - src[2] };
        const double plane = dot3(ds, normal);
        if (!(plane > 1e-6)) { status = -2; break; }
        /* radial ray: source -> closest point of the z axis through the origin
         * (YkFDKVecGeoDerived.hpp:194-209) */
        double ray[3] = { -src[0], -src[1], 0.0 };
        const double sod = sqrt(dot3(ray, ray));
        if (!(sod > 1e-10)) { status = -2; break; }
        ray[0] /= sod; ray[1] /= sod;
        /* principal point, SDD and pixel offsets (YkFDKVecGeoDerived.hpp:257-304) */
        const double den = dot3(ray, normal);
        if (fabs(den) < 1e-12) { status = -2; break; }
        const double t = plane / den;
        if (t <= 0.0) { status = -2; break; }
        double pp[3], r0[3], D[3];
        for (int k = 0; k < 3; ++k) {
            pp[k] = src[k] + ray[k] * t;
            r0[k] = pp[k] - src[k];
            D[k] = pp[k] - detS[k];
        }
        const double UU = dot3(detU, detU), VV = dot3(detV, detV), UV = dot3(detU, detV);
        const double DU = dot3(D, detU), DV = dot3(D, detV);
        const double gram = UU * VV - UV * UV;
        if (fabs(gram) < 1e-20) { status = -2; break; }
        const double up = (DU * VV - DV * UV) / gram;
        const double vp = (-DU * UV + DV * UU) / gram;
        d[YKO_D_SOD] = sod;
        d[YKO_D_SDD_RADIAL] = sqrt(dot3(r0, r0));
        d[YKO_D_SDD_PLANE] = plane;
        d[YKO_D_OFFU] = up - cu;
        d[YKO_D_OFFV] = vp - cv;
        for (int k = 0; k < 3; ++k) {
            d[YKO_D_RADIAL + k] = ray[k];
            d[YKO_D_NORMAL + k] = normal[k];
        }
        d[YKO_D_UU] = UU; d[YKO_D_VV] = VV; d[YKO_D_UV] = UV;
    }
    free(theta);
    return status;
}

/*
 * Affine projection coefficients, all in FP64
 * (restates fdk_precompute_coeffs_kernel,
 * src/Reconstruction/Analytic/Circular/Flat/FDK/kernels/YkFDKBpPrecompute.cu:74-160).
 *   u = Cu.[X,1] / Cd.[X,1],  v = Cv.[X,1] / Cd.[X,1],  depth = Cr.[X,1]
 *   wbase = SOD^2 * dtheta * (2 pi / range)
 */
int yko_fdk_coeffs(const double* views, const double* derived, int n_views,
    double* out)
{
    for (int i = 0; i < n_views; ++i) {
        const double* g = views + (size_t)i * YKO_VIEW_STRIDE;
        const double* d = derived + (size_t)i * YKO_DERIVED_STRIDE;
        double* c = out + (size_t)i * YKO_COEFF_STRIDE;
        const double* s = g;
        const double* ds = g + 3;
        const double* u = g + 6;
        const double* v = g + 9;
        const double* n = d + YKO_D_NORMAL;
        const double* rr = d + YKO_D_RADIAL;
        const double SDD = d[YKO_D_SDD_PLANE];
        const double s_n = dot3(s, n), s_u = dot3(s, u), s_v = dot3(s, v);
        const double off_u = (ds[0] - s[0]) * u[0] + (ds[1] - s[1]) * u[1] + (ds[2] - s[2]) * u[2];
        const double off_v = (ds[0] - s[0]) * v[0] + (ds[1] - s[1]) * v[1] + (ds[2] - s[2]) * v[2];
        double Cu[4], Cv[4];
        for (int k = 0; k < 3; ++k) {
            Cu[k] = SDD * u[k] - off_u * n[k];
            Cv[k] = SDD * v[k] - off_v * n[k];
        }
        Cu[3] = -SDD * s_u + off_u * s_n;
        Cv[3] = -SDD * s_v + off_v * s_n;
        const double UU = d[YKO_D_UU], VV = d[YKO_D_VV], UV = d[YKO_D_UV];
        const double gram = UU * VV - UV * UV;
        const double inv = fabs(gram) < 1e-20 ? 0.0 : 1.0 / gram;
        /* component order x, y, z, w */
        const int order[4] = { 0, 1, 2, 3 };
        for (int k = 0; k < 4; ++k) {
            const int q = order[k];
            c[YKO_C_CU + k] = (Cu[q] * VV - Cv[q] * UV) * inv;
            c[YKO_C_CV + k] = (-Cu[q] * UV + Cv[q] * UU) * inv;
        }
        c[YKO_C_CD + 0] = n[0]; c[YKO_C_CD + 1] = n[1]; c[YKO_C_CD + 2] = n[2];
        c[YKO_C_CD + 3] = -s_n;
        c[YKO_C_CR + 0] = rr[0]; c[YKO_C_CR + 1] = rr[1]; c[YKO_C_CR + 2] = rr[2];
        c[YKO_C_CR + 3] = -dot3(s, rr);
        c[YKO_C_WBASE] = d[YKO_D_SOD] * d[YKO_D_SOD] * d[YKO_D_DTHETA] * d[YKO_D_FSCALE];
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* preweight and Parker                                                      */
/* ------------------------------------------------------------------------ */
/*
 * q(u,v) = p(u,v) * SDD_r / |detS + u detU + v detV - src| * SDD_r / (2 SOD)
 * (restates preweight_vec_chunk_rowwarp_shfl_kernel_v2,
 * src/Reconstruction/Analytic/Circular/Flat/FDK/kernels/YkFDKPreWeightKernels.cu:289-405)
 */
void yko_preweight(const double* views, const double* derived, int n_views,
    int nu, int nv, const double* in, double* out)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < n_views; ++i) {
        for (int v = 0; v < nv; ++v) {
            const double* g = views + (size_t)i * YKO_VIEW_STRIDE;
            const double* d = derived + (size_t)i * YKO_DERIVED_STRIDE;
            const double sddr = d[YKO_D_SDD_RADIAL], sod = d[YKO_D_SOD];
            const double wgeom = sddr / sod / 2.0;
            const size_t base = ((size_t)i * nv + v) * nu;
            for (int u = 0; u < nu; ++u) {
                double q[3];
                for (int k = 0; k < 3; ++k)
                    q[k] = g[3 + k] + g[9 + k] * v + g[6 + k] * u - g[k];
                double r2 = dot3(q, q);
                if (r2 < 1e-20) r2 = 1e-20;
                const double w = sddr / sqrt(r2);
                out[base + u] = in[base + u] * (w * wgeom);
            }
        }
    }
}

/*
 * Parker short-scan weights times range/pi, in place
 * (restates parker_weight_kernel and pk_uploadAngles,
 * src/Reconstruction/Analytic/Circular/Flat/FDK/kernels/YkFDKParkerKernels.cu:46-145;
 * scale from YkFDKParkerWeightProcessor.hpp:37).
 */
void yko_parker(const double* views, const double* derived, int n_views,
    int nu, int nv, double start_angle_rad, int direction,
    double scan_range_rad, double* data)
{
    const double fscale = scan_range_rad / M_PI;
    const double quad_pi = 0.25 * M_PI;
    const double eps = 1e-6;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n_views; ++i) {
        const double* g = views + (size_t)i * YKO_VIEW_STRIDE;
        const double* d = derived + (size_t)i * YKO_DERIVED_STRIDE;
        double beta = direction * (g[12] - start_angle_rad);
        while (beta < 0.0) beta += 2.0 * M_PI;
        while (beta >= 2.0 * M_PI) beta -= 2.0 * M_PI;
        const double rx = d[YKO_D_RADIAL], ry = d[YKO_D_RADIAL + 1];
        const double cvr = 0.5 * (nv - 1);
#define YKO_FAN(du_) (atan2(((g[3] + (du_) * g[6] + cvr * g[9] - g[0]) * ry - \
                             (g[4] + (du_) * g[7] + cvr * g[10] - g[1]) * rx), \
                            ((g[3] + (du_) * g[6] + cvr * g[9] - g[0]) * rx + \
                             (g[4] + (du_) * g[7] + cvr * g[10] - g[1]) * ry)))
        const double f0 = fabs(YKO_FAN(0.0)), f1 = fabs(YKO_FAN((double)(nu - 1)));
        const double Gamma = f0 > f1 ? f0 : f1;
        for (int u = 0; u < nu; ++u) {
            const double gamma = direction * YKO_FAN((double)u);
            const double t1 = 2.0 * (Gamma + gamma);
            const double t2 = M_PI + 2.0 * gamma;
            const double t3 = M_PI + 2.0 * Gamma;
            double w;
            if (beta <= 0.0) w = 0.0;
            else if (beta < t1) {
                const double dd = Gamma + gamma;
                const double s = sin(quad_pi * beta / (fabs(dd) > eps ? dd : eps));
                w = s * s;
            } else if (beta <= t2) w = 1.0;
            else if (beta < t3) {
                const double dd = Gamma - gamma;
                const double s = sin(quad_pi * (t3 - beta) / (fabs(dd) > eps ? dd : eps));
                w = s * s;
            } else w = 0.0;
            for (int v = 0; v < nv; ++v)
                data[((size_t)i * nv + v) * nu + u] *= w * fscale;
        }
#undef YKO_FAN
    }
}

/* ------------------------------------------------------------------------ */
/* filter                                                                    */
/* ------------------------------------------------------------------------ */
/* smallest power of two >= 2 Nu
 * (.../FDK/kernels/YkFDKFilterHelpers.cuh:29-35) */
int yko_padded_n(int nu)
{
    int n = 1;
    while (n < 2 * nu) n <<= 1;
    return n;
}

/* rint(N/2 - ((Nu-1)/2 + offsetU)) (.../FDK/kernels/YkFDKFilterHelpers.cuh:38-42) */
int yko_start_u(int nu, int padded_n, double offset_u_pix)
{
    const double axis = (nu - 1) * 0.5 + offset_u_pix;
    return (int)lrint(padded_n * 0.5 - axis);
}

/* Abramowitz & Stegun 9.8.1 / 9.8.2 polynomial for I0, the form the reference
 * uses (src/Filter/YkCreateFilterKernelHelpers.cuh:21-36). */
double yko_bessel_i0_as(double x)
{
    const double ax = fabs(x);
    if (ax < 3.75) {
        double y = x / 3.75;
        y *= y;
        return 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 +
            y * (0.2659732 + y * (0.0360768 + y * 0.0045813)))));
    }
    const double y = 3.75 / ax;
    return (exp(ax) / sqrt(ax)) * (0.39894228 + y * (0.01328592 +
        y * (0.00225319 + y * (-0.00157565 + y * (0.00916281 +
        y * (-0.02057706 + y * (0.02635537 + y * (-0.01647633 +
        y * 0.00392377))))))));
}

/* power series sum_k ((x/2)^k / k!)^2, used only to pin the A&S form */
double yko_bessel_i0_series(double x)
{
    const double q = 0.25 * x * x;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 500; ++k) {
        term *= q / ((double)k * (double)k);
        sum += term;
        if (term < 1e-18 * sum) break;
    }
    return sum;
}

/*
 * Window shapes W(x), x = f / cutoff clamped to [0,1]
 * (restates window_shape / window_shape_desc,
 * src/Filter/YkCreateFilterKernelHelpers.cuh:38-103).
 * kind: 0 None, 1 RamLak, 2 SheppLogan, 3 Cosine, 4 Hann, 5 Hamming,
 *       6 Blackman, 7 Butterworth, 8 Kaiser, 9 Tukey
 * parametric = 1 uses window_shape_desc (Butterworth/Kaiser/Tukey active),
 * parametric = 0 uses window_shape (those three return 1).
 */
double yko_window(int kind, double x, double order, double beta, double alpha,
    int parametric)
{
    if (x < 0.0) x = 0.0;
    if (x > 1.0) x = 1.0;
    if (parametric) {
        if (kind == 7) {
            const double o = order > 0.01 ? order : 0.01;
            return 1.0 / (1.0 + pow(x, 2.0 * o));
        }
        if (kind == 8) {
            const double b = beta > 0.0 ? beta : 0.0;
            const double t = 2.0 * x - 1.0;
            const double r = 1.0 - t * t;
            return yko_bessel_i0_as(b * sqrt(r > 0.0 ? r : 0.0)) / yko_bessel_i0_as(b);
        }
        if (kind == 9) {
            double a = alpha < 0.0 ? 0.0 : (alpha > 1.0 ? 1.0 : alpha);
            if (a <= 0.0) return 1.0;
            if (x < a * 0.5) return 0.5 * (1.0 + cos(M_PI * (2.0 * x / a - 1.0)));
            if (x <= 1.0 - a * 0.5) return 1.0;
            return 0.5 * (1.0 + cos(M_PI * (2.0 * x / a - 2.0 / a + 1.0)));
        }
    }
    switch (kind) {
    case 1: return 1.0;
    case 2: {
        const double t = M_PI * 0.5 * x;
        return fabs(t) < 1e-8 ? 1.0 : sin(t) / t;
    }
    case 3: return cos(0.5 * M_PI * x);
    case 4: return 0.5 * (1.0 + cos(M_PI * x));
    case 5: return 0.54 + 0.46 * cos(M_PI * x);
    case 6: return 0.42 + 0.5 * cos(M_PI * x) + 0.08 * cos(2.0 * M_PI * x);
    default: return 1.0;
    }
}

/* Bins 0..N/2 of the forward DFT of the N-periodic real sequence h, as its
 * real part (mode 0) or magnitude (mode 1)
 * (kernel_extract_weights_from_fft, src/Filter/YkCreateFilterKernelKernels.cu:299-320). */
static void yko_dft_extract(const double* h, int N, int mode, double* w)
{
    for (int k = 0; k <= N / 2; ++k) {
        double re = 0.0, im = 0.0;
        for (int u = 0; u < N; ++u) {
            if (h[u] == 0.0) continue;
            const double a = 2.0 * M_PI * (double)(((long long)k * u) % N) / (double)N;
            re += h[u] * cos(a);
            im -= h[u] * sin(a);
        }
        w[k] = mode == 1 ? sqrt(re * re + im * im) : re;
    }
}

/*
 * Frequency weights w[k], k = 0..N/2, with 1/N baked in
 * (restates CreateFilterKernelFromFFT::build_weights and build_custom_weights,
 * src/Filter/YkCreateFilterKernelFFT.hpp:114-258, and the kernels in
 * src/Filter/YkCreateFilterKernelKernels.cu:135-345).
 *   source 2 (SpatialRampFFT), any kind: taps ramp[0..m) centred on index 0
 *                  (ramp[j] at (j - m/2) mod N) / N, DFT, extract; unless
 *                  Ram-Lak, 0 above the cutoff and x gain W(f/c) (window_shape);
 *                  DC := 0; x gain / du.
 *   kind 10 (Custom): custom[k] x gain / N / du.
 *   kind 0 (None): gain / N.
 *   AnalyticFreq:  w = gain f W(f/c) / N / du for f = k/N <= c, else 0.
 *   DiscreteRLFFT: extract of the DFT of the du=1 spatial Ram-Lak kernel / N,
 *                  DC := 0, times gain W(f/c) (f <= c, else 0), then / du.
 */
int yko_filter_weights_ext(int kind, int source, int extract, double cutoff, double gain,
    double order, double beta, double tukey_alpha, int N, double du_mm,
    const double* ramp, int m, const double* custom, int nc_custom, double* w)
{
    const int nc = N / 2 + 1;
    const double cc = cutoff > 0.0 ? cutoff : 0.5;
    if (source == 2) {
        if (!ramp || m <= 0 || !(m & 1) || m > N) return -1;
        double* h = (double*)calloc((size_t)N, sizeof(double));
        if (!h) return -2;
        for (int j = 0; j < m; ++j) {
            int idx = j - m / 2;
            if (idx < 0) idx += N;
            h[idx] = ramp[j] / (double)N;
        }
        yko_dft_extract(h, N, extract, w);
        free(h);
        if (kind != 1) {
            for (int k = 0; k < nc; ++k) {
                const double f = (double)k / (double)N;
                w[k] = f > cc ? 0.0 : w[k] * gain * yko_window(kind, f / cc, order, beta, tukey_alpha, 0);
            }
        }
        w[0] = 0.0;
        for (int k = 0; k < nc; ++k) w[k] *= gain / du_mm;
        return 0;
    }
    if (kind == 10) {
        if (!custom || nc_custom != nc) return -1;
        for (int k = 0; k < nc; ++k) w[k] = custom[k] * gain / (double)N / du_mm;
        return 0;
    }
    if (kind == 0) {
        for (int k = 0; k < nc; ++k) w[k] = gain / (double)N;
        return 0;
    }
    if (source == 0) {
        for (int k = 0; k < nc; ++k) {
            const double f = (double)k / (double)N;
            w[k] = f > cc ? 0.0
                : gain * f * yko_window(kind, f / cc, order, beta, tukey_alpha, 1) / (double)N;
            w[k] /= du_mm;
        }
        return 0;
    }
    /* DiscreteRLFFT: h[0] = 1/4, h[odd n] = -1/(pi^2 n^2), wrapped for n > N/2 */
    double* h = (double*)calloc((size_t)N, sizeof(double));
    if (!h) return -2;
    h[0] = 0.25 / (double)N;
    for (int u = 1; u < N; ++u) {
        const int n = u <= N / 2 ? u : u - N;
        const int an = n < 0 ? -n : n;
        if (an & 1) h[u] = -1.0 / (M_PI * M_PI * (double)n * (double)n) / (double)N;
    }
    yko_dft_extract(h, N, extract, w);
    free(h);
    w[0] = 0.0;
    for (int k = 0; k < nc; ++k) {
        const double f = (double)k / (double)N;
        if (f > cc) { w[k] = 0.0; continue; }
        w[k] *= gain * yko_window(kind, f / cc, order, beta, tukey_alpha, 0);
    }
    for (int k = 0; k < nc; ++k) w[k] /= du_mm;
    return 0;
}

int yko_filter_weights(int kind, int source, double cutoff, double gain,
    double order, double beta, double tukey_alpha, int N, double du_mm,
    double* w)
{
    return yko_filter_weights_ext(kind, source, 0, cutoff, gain, order, beta, tukey_alpha, N, du_mm,
        NULL, 0, NULL, 0, w);
}

/* h[n] = sum_{k=0}^{N-1} W[k] cos(2 pi k n / N), W[N-k] = w[k]: the impulse
 * response of R2C -> (x w) -> unnormalized C2R. */
void yko_filter_kernel_from_weights(const double* w, int N, double* h)
{
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; ++n) {
        double acc = w[0];
        for (int k = 1; k < N / 2; ++k)
            acc += 2.0 * w[k] * cos(2.0 * M_PI * (double)(((long long)k * n) % N) / (double)N);
        acc += w[N / 2] * ((n & 1) ? -1.0 : 1.0);
        h[n] = acc;
    }
}

/* iterative radix-2 complex FFT, forward = exp(-i...), inverse unnormalized */
void yko_fft(double* re, double* im, int n, int inverse)
{
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const double ang = (inverse ? 2.0 : -2.0) * M_PI / (double)len;
        for (int i = 0; i < n; i += len) {
            for (int j = 0; j < len / 2; ++j) {
                const double wr = cos(ang * j), wi = sin(ang * j);
                const int a = i + j, b = i + j + len / 2;
                const double xr = re[b] * wr - im[b] * wi;
                const double xi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - xr; im[b] = im[a] - xi;
                re[a] += xr; im[a] += xi;
            }
        }
    }
}

/*
 * Row filtering: place each row at start_u in a zero row of length N
 * (dropping samples outside [0,N)), circularly convolve, read back at
 * start_u (zero outside [0,N)).  Restates the pad / R2C / multiply / C2R /
 * crop chain (.../FDK/kernels/YkFDKFilterKernels.cu:32-140,
 * src/Filter/YkConvKernels.cu:109-220).  start_u is per view.
 */
/* radix-2 FFT with a precomputed twiddle table c[k] + i s[k] = exp(-2 pi i k / N) */
static void fft_table(double* re, double* im, int n, const double* c, const double* s, int inverse)
{
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int j = 0; j < len / 2; ++j) {
                const double wr = c[j * step], wi = inverse ? -s[j * step] : s[j * step];
                const int a = i + j, b = i + j + len / 2;
                const double xr = re[b] * wr - im[b] * wi;
                const double xi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - xr; im[b] = im[a] - xi;
                re[a] += xr; im[a] += xi;
            }
        }
    }
}

void yko_filter_rows(const double* in, int n_views, int nv, int nu,
    int N, const double* w, const int* start_u, int method, double* out)
{
    double* h = NULL;
    double* tc = NULL;
    double* ts = NULL;
    if (method == 0) {
        h = (double*)malloc(sizeof(double) * (size_t)N);
        yko_filter_kernel_from_weights(w, N, h);
    } else {
        tc = (double*)malloc(sizeof(double) * (size_t)N);
        ts = (double*)malloc(sizeof(double) * (size_t)N);
        for (int k = 0; k < N; ++k) {
            tc[k] = cos(-2.0 * M_PI * k / N);
            ts[k] = sin(-2.0 * M_PI * k / N);
        }
    }
    const long long rows = (long long)n_views * nv;
#pragma omp parallel
    {
        double* re = (double*)malloc(sizeof(double) * (size_t)N);
        double* im = (double*)malloc(sizeof(double) * (size_t)N);
        if (method == 0) {
#pragma omp for schedule(dynamic, 16)
            for (long long r = 0; r < rows; ++r) {
                const int su = start_u[r / nv];
                const double* x = in + (size_t)r * nu;
                double* y = out + (size_t)r * nu;
                for (int a = 0; a < nu; ++a) {
                    const int j = su + a;
                    if (j < 0 || j >= N) { y[a] = 0.0; continue; }
                    double acc = 0.0;
                    for (int b = 0; b < nu; ++b) {
                        const int m = su + b;
                        if (m < 0 || m >= N) continue;
                        int idx = (a - b) % N;
                        if (idx < 0) idx += N;
                        acc += x[b] * h[idx];
                    }
                    y[a] = acc;
                }
            }
        } else {
            /* two real rows per complex transform: the response is real and
             * even, so Re and Im filter independently */
#pragma omp for schedule(dynamic, 8)
            for (long long p = 0; p < (rows + 1) / 2; ++p) {
                const long long ra = 2 * p, rb = 2 * p + 1;
                const int sa = start_u[ra / nv];
                const int sb = rb < rows ? start_u[rb / nv] : 0;
                for (int k = 0; k < N; ++k) { re[k] = 0.0; im[k] = 0.0; }
                for (int b = 0; b < nu; ++b) {
                    const int ma = sa + b;
                    if (ma >= 0 && ma < N) re[ma] = in[(size_t)ra * nu + b];
                    if (rb < rows) {
                        const int mb = sb + b;
                        if (mb >= 0 && mb < N) im[mb] = in[(size_t)rb * nu + b];
                    }
                }
                fft_table(re, im, N, tc, ts, 0);
                for (int k = 0; k < N; ++k) {
                    const double wk = w[k <= N / 2 ? k : N - k];
                    re[k] *= wk; im[k] *= wk;
                }
                fft_table(re, im, N, tc, ts, 1);
                for (int a = 0; a < nu; ++a) {
                    const int ja = sa + a;
                    out[(size_t)ra * nu + a] = (ja < 0 || ja >= N) ? 0.0 : re[ja];
                    if (rb < rows) {
                        const int jb = sb + a;
                        out[(size_t)rb * nu + a] = (jb < 0 || jb >= N) ? 0.0 : im[jb];
                    }
                }
            }
        }
        free(re);
        free(im);
    }
    free(h);
    free(tc);
    free(ts);
}

/* ------------------------------------------------------------------------ */
/* FDK backprojection                                                        */
/* ------------------------------------------------------------------------ */
/* exact bilinear, texel (iu, iv) centre at integer coordinates, zero outside
 * (the unnormalized Linear/Border tex2D(u + .5, v + .5) of
 * .../FDK/kernels/YkFDKBpKernels.cu:170 without weight quantization) */
double yko_bilinear(const double* p, int nu, int nv, double u, double v)
{
    const double fu0 = floor(u), fv0 = floor(v);
    if (fu0 < -2.0 || fv0 < -2.0 || fu0 > nu + 1.0 || fv0 > nv + 1.0) return 0.0;
    const int iu = (int)fu0, iv = (int)fv0;
    const double fu = u - fu0, fv = v - fv0;
    double t[2][2] = { { 0.0, 0.0 }, { 0.0, 0.0 } };
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            const int r = iv + a, c = iu + b;
            if (r >= 0 && r < nv && c >= 0 && c < nu) t[a][b] = p[(size_t)r * nu + c];
        }
    const double r0 = t[0][0] + fu * (t[0][1] - t[0][0]);
    const double r1 = t[1][0] + fu * (t[1][1] - t[1][0]);
    return r0 + fv * (r1 - r0);
}

/*
 * vol(X) += sum_i bilinear(q_i; u, v) * SOD^2 dtheta fscale / ((X - src).r)^2
 * (restates fdk_bp_kernel<ZSIZE> precomputed form,
 * src/Reconstruction/Analytic/Circular/Flat/FDK/kernels/YkFDKBpKernels.cu:112-209).
 * voxel_index NULL: all voxels, out is [nz][ny][nx]; otherwise out[k] for the
 * listed linear voxel indices.  out is overwritten.
 */
void yko_bp_fdk(const double* coeffs, int n_views, const double* filtered,
    int nu, int nv, const yko_volume* vol, const int64_t* voxel_index,
    int64_t n_voxels, double* out)
{
    const int64_t total = voxel_index ? n_voxels
        : (int64_t)vol->nx * vol->ny * vol->nz;
    const double ox = vol->cx - (vol->nx - 1) * 0.5 * vol->vx;
    const double oy = vol->cy - (vol->ny - 1) * 0.5 * vol->vy;
    const double oz = vol->cz - (vol->nz - 1) * 0.5 * vol->vz;
    const size_t plane = (size_t)nu * nv;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t k = 0; k < total; ++k) {
        const int64_t lin = voxel_index ? voxel_index[k] : k;
        const int64_t x = lin % vol->nx;
        const int64_t y = (lin / vol->nx) % vol->ny;
        const int64_t z = lin / ((int64_t)vol->nx * vol->ny);
        const double X = ox + x * vol->vx, Y = oy + y * vol->vy, Z = oz + z * vol->vz;
        double acc = 0.0;
        for (int i = 0; i < n_views; ++i) {
            const double* c = coeffs + (size_t)i * YKO_COEFF_STRIDE;
            const double den = c[YKO_C_CD + 3] + X * c[YKO_C_CD] + Y * c[YKO_C_CD + 1] + Z * c[YKO_C_CD + 2];
            const double un = c[YKO_C_CU + 3] + X * c[YKO_C_CU] + Y * c[YKO_C_CU + 1] + Z * c[YKO_C_CU + 2];
            const double vn = c[YKO_C_CV + 3] + X * c[YKO_C_CV] + Y * c[YKO_C_CV + 1] + Z * c[YKO_C_CV + 2];
            const double dep = c[YKO_C_CR + 3] + X * c[YKO_C_CR] + Y * c[YKO_C_CR + 1] + Z * c[YKO_C_CR + 2];
            const double s = yko_bilinear(filtered + plane * i, nu, nv, un / den, vn / den);
            acc += s * c[YKO_C_WBASE] / (dep * dep);
        }
        out[k] = acc;
    }
}

/*
 * Operator-form FDK backprojections used as the iterative B
 * (fdk_bp_kernel_fdk_weight_v2 / fdk_bp_kernel_matched_weight,
 * src/FlatFpBp/BP/kernels/YkBpFdkKernel.cu:228-466), on unfiltered views:
 *   FDK:     w = SOD^2 dtheta fscale / (SOD - src.X / SOD)^2
 *   matched: w = L^3 / (|X - src|^2 |SDD_plane|) * vx vy vz / (du0 dv0),
 *            L = |detS + u detU + v detV - src|, du0/dv0 of the first view
 * A sample is skipped when den < 1e-8, |SOD - src.X/SOD| < 1e-8, L^2 < 1e-8
 * or |X - src|^2 < 1e-8.  matched = 2 selects the voxel-driven Joseph model
 * (joseph_bp_v3_kernel, src/FlatFpBp/BP/kernels/YkBPJosephKernel.cu:624-735;
 * the same model as joseph_bp_v2_kernel, :430-560): w = the physical length
 * of the ray X - src through one voxel along its main axis, m = argmax
 * |d_k / vox_k| (ties X, then Y), w = vox_m |d| / |d_m|, skipped only when
 * den < 1e-8.  out ([nz][ny][nx]) is overwritten.
 */
void yko_bp_fdk_operator(const double* views, const double* derived, const double* coeffs,
    int n_views, const double* proj, int nu, int nv, const yko_volume* vol, int matched,
    double* out)
{
    const double ox = vol->cx - (vol->nx - 1) * 0.5 * vol->vx;
    const double oy = vol->cy - (vol->ny - 1) * 0.5 * vol->vy;
    const double oz = vol->cz - (vol->nz - 1) * 0.5 * vol->vz;
    const size_t plane = (size_t)nu * nv;
    const double scale = vol->vx * vol->vy * vol->vz /
        (derived[YKO_D_DU] * derived[YKO_D_DV]);
    const int64_t total = (int64_t)vol->nx * vol->ny * vol->nz;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t k = 0; k < total; ++k) {
        const int64_t x = k % vol->nx;
        const int64_t y = (k / vol->nx) % vol->ny;
        const int64_t z = k / ((int64_t)vol->nx * vol->ny);
        const double X[3] = { ox + x * vol->vx, oy + y * vol->vy, oz + z * vol->vz };
        double acc = 0.0;
        for (int i = 0; i < n_views; ++i) {
            const double* g = views + (size_t)i * YKO_VIEW_STRIDE;
            const double* d = derived + (size_t)i * YKO_DERIVED_STRIDE;
            const double* c = coeffs + (size_t)i * YKO_COEFF_STRIDE;
            const double den = c[YKO_C_CD + 3] + X[0] * c[YKO_C_CD] + X[1] * c[YKO_C_CD + 1] + X[2] * c[YKO_C_CD + 2];
            if (!(den >= 1e-8)) continue;
            const double u = (c[YKO_C_CU + 3] + X[0] * c[YKO_C_CU] + X[1] * c[YKO_C_CU + 1] + X[2] * c[YKO_C_CU + 2]) / den;
            const double v = (c[YKO_C_CV + 3] + X[0] * c[YKO_C_CV] + X[1] * c[YKO_C_CV + 1] + X[2] * c[YKO_C_CV + 2]) / den;
            double w;
            if (matched == 2) {
                const double d3[3] = { X[0] - g[0], X[1] - g[1], X[2] - g[2] };
                const double vox3[3] = { vol->vx, vol->vy, vol->vz };
                const double ax = fabs(d3[0] / vox3[0]), ay = fabs(d3[1] / vox3[1]), az = fabs(d3[2] / vox3[2]);
                const int m = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
                const double len = sqrt(d3[0] * d3[0] + d3[1] * d3[1] + d3[2] * d3[2]);
                w = vox3[m] * len / fabs(d3[m]);
            } else if (matched) {
                double P[3], L2 = 0.0, lsq = 0.0;
                for (int a = 0; a < 3; ++a) {
                    P[a] = g[3 + a] + u * g[6 + a] + v * g[9 + a] - g[a];
                    L2 += P[a] * P[a];
                    lsq += (X[a] - g[a]) * (X[a] - g[a]);
                }
                if (!(L2 >= 1e-8) || !(lsq >= 1e-8)) continue;
                const double sdd = fabs(d[YKO_D_SDD_PLANE]);
                w = L2 * sqrt(L2) / lsq * (sdd > 1e-12 ? 1.0 / sdd : 0.0) * scale;
            } else {
                const double sod = d[YKO_D_SOD];
                const double cn = sod - (g[0] * X[0] + g[1] * X[1] + g[2] * X[2]) / sod;
                if (!(fabs(cn) >= 1e-8)) continue;
                w = sod * sod * d[YKO_D_DTHETA] * d[YKO_D_FSCALE] / (cn * cn);
            }
            acc += yko_bilinear(proj + plane * i, nu, nv, u, v) * w;
        }
        out[k] = acc;
    }
}

/* ------------------------------------------------------------------------ */
/* forward projectors                                                        */
/* ------------------------------------------------------------------------ */
/* main axis of the world centre ray (mainAxis / forEachAxisRun,
 * src/common/cuda/operators/YkOperatorKernelTypes.cuh:133-210):
 * 0 = X, 1 = Y, 2 = Z */
int yko_main_axis(const double* g, int nu, int nv)
{
    const double cu = 0.5 * (nu - 1), cv = 0.5 * (nv - 1);
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = g[3 + k] + cu * g[6 + k] + cv * g[9 + k] - g[k];
    const double ax = fabs(d[0]), ay = fabs(d[1]), az = fabs(d[2]);
    if (ax >= ay && ax >= az) return 0;
    if (ay >= ax && ay >= az) return 1;
    return 2;
}

/* axis permutation: c0 = perm[0], c1 = perm[1], c2 = perm[2]
 * (AxisX/AxisY/AxisZ, YkOperatorKernelTypes.cuh:214-275) */
static void axis_perm(int axis, int perm[3])
{
    if (axis == 0) { perm[0] = 0; perm[1] = 1; perm[2] = 2; }
    else if (axis == 1) { perm[0] = 1; perm[1] = 0; perm[2] = 2; }
    else { perm[0] = 2; perm[1] = 0; perm[2] = 1; }
}

typedef struct {
    int perm[3];
    int n0, n1, n2;
    double v0, v1, v2;
    double src[3], detS[3], detU[3], detV[3];   /* voxel-normalized, volume-centred */
} joseph_view;

static void joseph_prepare(const double* g, int nu, int nv, const yko_volume* vol,
    joseph_view* jv)
{
    axis_perm(yko_main_axis(g, nu, nv), jv->perm);
    const int dims[3] = { vol->nx, vol->ny, vol->nz };
    const double vox[3] = { vol->vx, vol->vy, vol->vz };
    const double cen[3] = { vol->cx, vol->cy, vol->cz };
    jv->n0 = dims[jv->perm[0]]; jv->n1 = dims[jv->perm[1]]; jv->n2 = dims[jv->perm[2]];
    jv->v0 = vox[jv->perm[0]]; jv->v1 = vox[jv->perm[1]]; jv->v2 = vox[jv->perm[2]];
    /* normalizeToVoxel (YkOperatorKernelTypes.cuh:148-168) */
    for (int k = 0; k < 3; ++k) {
        jv->src[k] = (g[k] - cen[k]) / vox[k];
        jv->detS[k] = (g[3 + k] - cen[k]) / vox[k];
        jv->detU[k] = g[6 + k] / vox[k];
        jv->detV[k] = g[9 + k] / vox[k];
    }
}

/* Per-ray Joseph line: c1 = a1 c0 + b1, c2 = a2 c0 + b2 in volume-centred
 * voxel units, weight L = sqrt(v0^2 + (a1 v1)^2 + (a2 v2)^2)
 * (fp_joseph_kernel, src/FlatFpBp/FP/kernels/YkFPKernels.cu:33-197). */
static int joseph_ray(const joseph_view* jv, int u, int v, double* a1, double* b1,
    double* a2, double* b2, double* L)
{
    double det[3];
    for (int k = 0; k < 3; ++k) det[k] = jv->detS[k] + u * jv->detU[k] + v * jv->detV[k];
    const double s0 = jv->src[jv->perm[0]], s1 = jv->src[jv->perm[1]], s2 = jv->src[jv->perm[2]];
    const double d0 = det[jv->perm[0]], d1 = det[jv->perm[1]], d2 = det[jv->perm[2]];
    const double den = s0 - d0;
    if (den == 0.0) return 0;
    *a1 = (s1 - d1) / den;
    *a2 = (s2 - d2) / den;
    *b1 = s1 - *a1 * s0;
    *b2 = s2 - *a2 * s0;
    const double t1 = *a1 * jv->v1, t2 = *a2 * jv->v2;
    *L = sqrt(jv->v0 * jv->v0 + t1 * t1 + t2 * t2);
    return 1;
}

static size_t vol_index(const int perm[3], int s, int i1, int i2, const yko_volume* vol)
{
    int idx[3];
    idx[perm[0]] = s; idx[perm[1]] = i1; idx[perm[2]] = i2;
    return ((size_t)idx[2] * vol->ny + idx[1]) * vol->nx + idx[0];
}

void yko_fp_joseph(const double* views, int n_views, int nu, int nv,
    const yko_volume* vol, const float* volume, double* proj_out)
{
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int i = 0; i < n_views; ++i) {
        for (int v = 0; v < nv; ++v) {
            joseph_view jv;
            joseph_prepare(views + (size_t)i * YKO_VIEW_STRIDE, nu, nv, vol, &jv);
            for (int u = 0; u < nu; ++u) {
                double a1, b1, a2, b2, L, acc = 0.0;
                double* o = proj_out + ((size_t)i * nv + v) * nu + u;
                if (!joseph_ray(&jv, u, v, &a1, &b1, &a2, &b2, &L)) { *o = 0.0; continue; }
                for (int s = 0; s < jv.n0; ++s) {
                    const double c0 = s - 0.5 * (jv.n0 - 1);
                    const double f1 = a1 * c0 + b1 + 0.5 * (jv.n1 - 1);
                    const double f2 = a2 * c0 + b2 + 0.5 * (jv.n2 - 1);
                    const double g1 = floor(f1), g2 = floor(f2);
                    if (g1 < -1.0 || g2 < -1.0 || g1 > jv.n1 - 1.0 || g2 > jv.n2 - 1.0) continue;
                    const int i1 = (int)g1, i2 = (int)g2;
                    const double w1 = f1 - g1, w2 = f2 - g2;
                    for (int a = 0; a < 2; ++a)
                        for (int b = 0; b < 2; ++b) {
                            const int j1 = i1 + b, j2 = i2 + a;
                            if (j1 < 0 || j1 >= jv.n1 || j2 < 0 || j2 >= jv.n2) continue;
                            const double wt = (b ? w1 : 1.0 - w1) * (a ? w2 : 1.0 - w2);
                            acc += wt * volume[vol_index(jv.perm, s, j1, j2, vol)];
                        }
                }
                *o = acc * L;
            }
        }
    }
}

/* exact transpose of yko_fp_joseph (scatter, serial per view for determinism) */
void yko_bp_joseph_adjoint(const double* views, int n_views, int nu, int nv,
    const yko_volume* vol, const double* proj, double* out)
{
    const size_t nvox = (size_t)vol->nx * vol->ny * vol->nz;
    memset(out, 0, sizeof(double) * nvox);
    for (int i = 0; i < n_views; ++i) {
        joseph_view jv;
        joseph_prepare(views + (size_t)i * YKO_VIEW_STRIDE, nu, nv, vol, &jv);
        for (int v = 0; v < nv; ++v)
            for (int u = 0; u < nu; ++u) {
                double a1, b1, a2, b2, L;
                if (!joseph_ray(&jv, u, v, &a1, &b1, &a2, &b2, &L)) continue;
                const double y = proj[((size_t)i * nv + v) * nu + u] * L;
                if (y == 0.0) continue;
                for (int s = 0; s < jv.n0; ++s) {
                    const double c0 = s - 0.5 * (jv.n0 - 1);
                    const double f1 = a1 * c0 + b1 + 0.5 * (jv.n1 - 1);
                    const double f2 = a2 * c0 + b2 + 0.5 * (jv.n2 - 1);
                    const double g1 = floor(f1), g2 = floor(f2);
                    if (g1 < -1.0 || g2 < -1.0 || g1 > jv.n1 - 1.0 || g2 > jv.n2 - 1.0) continue;
                    const int i1 = (int)g1, i2 = (int)g2;
                    const double w1 = f1 - g1, w2 = f2 - g2;
                    for (int a = 0; a < 2; ++a)
                        for (int b = 0; b < 2; ++b) {
                            const int j1 = i1 + b, j2 = i2 + a;
                            if (j1 < 0 || j1 >= jv.n1 || j2 < 0 || j2 >= jv.n2) continue;
                            const double wt = (b ? w1 : 1.0 - w1) * (a ? w2 : 1.0 - w2);
                            out[vol_index(jv.perm, s, j1, j2, vol)] += wt * y;
                        }
                }
            }
    }
}

/*
 * Exact Siddon line integral of the voxel grid along src -> dst, with the
 * segment clipped to t in [0, 1] and to the volume box
 * [origin - vox/2, origin + (N - 1/2) vox]
 * (siddon_fp_kernel, src/FlatFpBp/FP/kernels/YkFPSiddonKernel.cu:19-124):
 * all plane crossings are merged and each interval contributes
 * mu(voxel at the interval midpoint) * dt * |ray|.
 */
typedef void (*siddon_visit)(void* ctx, size_t voxel, double length);

static void siddon_walk(const double src[3], const double dst[3],
    const yko_volume* vol, siddon_visit visit, void* ctx)
{
    const int dims[3] = { vol->nx, vol->ny, vol->nz };
    const double vox[3] = { vol->vx, vol->vy, vol->vz };
    const double cen[3] = { vol->cx, vol->cy, vol->cz };
    double origin[3], lo[3], hi[3], ray[3];
    for (int k = 0; k < 3; ++k) {
        origin[k] = cen[k] - (dims[k] - 1) * 0.5 * vox[k];
        lo[k] = origin[k] - 0.5 * vox[k];
        hi[k] = origin[k] + (dims[k] - 0.5) * vox[k];
        ray[k] = dst[k] - src[k];
    }
    double tmin = 0.0, tmax = 1.0;
    for (int k = 0; k < 3; ++k) {
        if (fabs(ray[k]) > 1e-12) {
            double ta = (lo[k] - src[k]) / ray[k], tb = (hi[k] - src[k]) / ray[k];
            if (ta > tb) { const double t = ta; ta = tb; tb = t; }
            if (ta > tmin) tmin = ta;
            if (tb < tmax) tmax = tb;
        } else if (src[k] < lo[k] || src[k] > hi[k]) return;
    }
    if (!(tmin < tmax)) return;
    const double len = sqrt(dot3(ray, ray));
    /* collect crossings */
    int cap = dims[0] + dims[1] + dims[2] + 4;
    double* ts = (double*)malloc(sizeof(double) * (size_t)cap);
    int nt = 0;
    ts[nt++] = tmin;
    for (int k = 0; k < 3; ++k) {
        if (fabs(ray[k]) <= 1e-12) continue;
        for (int p = 0; p <= dims[k]; ++p) {
            const double plane = lo[k] + p * vox[k];
            const double t = (plane - src[k]) / ray[k];
            if (t > tmin && t < tmax) ts[nt++] = t;
        }
    }
    ts[nt++] = tmax;
    /* insertion sort is fine: crossings per axis are already monotone */
    for (int a = 1; a < nt; ++a) {
        const double t = ts[a];
        int b = a - 1;
        while (b >= 0 && ts[b] > t) { ts[b + 1] = ts[b]; --b; }
        ts[b + 1] = t;
    }
    for (int a = 0; a + 1 < nt; ++a) {
        const double dt = ts[a + 1] - ts[a];
        if (dt <= 0.0) continue;
        const double tm = 0.5 * (ts[a] + ts[a + 1]);
        int idx[3], ok = 1;
        for (int k = 0; k < 3; ++k) {
            const double p = src[k] + tm * ray[k];
            idx[k] = (int)floor((p - lo[k]) / vox[k]);
            if (idx[k] < 0 || idx[k] >= dims[k]) ok = 0;
        }
        if (!ok) continue;
        visit(ctx, ((size_t)idx[2] * vol->ny + idx[1]) * vol->nx + idx[0], dt * len);
    }
    free(ts);
}

typedef struct { const float* volume; double acc; } siddon_sum_ctx;
static void siddon_sum(void* c, size_t voxel, double length)
{
    siddon_sum_ctx* s = (siddon_sum_ctx*)c;
    s->acc += s->volume[voxel] * length;
}

double yko_siddon_ray(const double src[3], const double dst[3],
    const yko_volume* vol, const float* volume)
{
    siddon_sum_ctx c = { volume, 0.0 };
    siddon_walk(src, dst, vol, siddon_sum, &c);
    return c.acc;
}

void yko_fp_siddon(const double* views, int n_views, int nu, int nv,
    const yko_volume* vol, const float* volume, double* proj_out)
{
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int i = 0; i < n_views; ++i)
        for (int v = 0; v < nv; ++v) {
            const double* g = views + (size_t)i * YKO_VIEW_STRIDE;
            for (int u = 0; u < nu; ++u) {
                double dst[3];
                for (int k = 0; k < 3; ++k) dst[k] = g[3 + k] + u * g[6 + k] + v * g[9 + k];
                proj_out[((size_t)i * nv + v) * nu + u] = yko_siddon_ray(g, dst, vol, volume);
            }
        }
}

typedef struct { double* out; double y; } siddon_scatter_ctx;
static void siddon_scatter(void* c, size_t voxel, double length)
{
    siddon_scatter_ctx* s = (siddon_scatter_ctx*)c;
    s->out[voxel] += s->y * length;
}

void yko_bp_siddon_adjoint(const double* views, int n_views, int nu, int nv,
    const yko_volume* vol, const double* proj, double* out)
{
    const size_t nvox = (size_t)vol->nx * vol->ny * vol->nz;
    memset(out, 0, sizeof(double) * nvox);
    for (int i = 0; i < n_views; ++i) {
        const double* g = views + (size_t)i * YKO_VIEW_STRIDE;
        for (int v = 0; v < nv; ++v)
            for (int u = 0; u < nu; ++u) {
                const double y = proj[((size_t)i * nv + v) * nu + u];
                if (y == 0.0) continue;
                double dst[3];
                for (int k = 0; k < 3; ++k) dst[k] = g[3 + k] + u * g[6 + k] + v * g[9 + k];
                siddon_scatter_ctx c = { out, y };
                siddon_walk(g, dst, vol, siddon_scatter, &c);
            }
    }
}

/* ------------------------------------------------------------------------ */
/* phantoms                                                                  */
/* ------------------------------------------------------------------------ */
static int inside_ellipsoid(const double* e, double x, double y, double z)
{
    const double d[3] = { x - e[0], y - e[1], z - e[2] };
    const double* R = e + 6;
    double q = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double r = R[3 * k] * d[0] + R[3 * k + 1] * d[1] + R[3 * k + 2] * d[2];
        q += (r / e[3 + k]) * (r / e[3 + k]);
    }
    return q <= 1.0;
}

void yko_phantom_ellipsoids(const double* ell, int n, const yko_volume* vol,
    int clamp_nonnegative, float* out)
{
    const double ox = vol->cx - (vol->nx - 1) * 0.5 * vol->vx;
    const double oy = vol->cy - (vol->ny - 1) * 0.5 * vol->vy;
    const double oz = vol->cz - (vol->nz - 1) * 0.5 * vol->vz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < vol->nz; ++z)
        for (int y = 0; y < vol->ny; ++y)
            for (int x = 0; x < vol->nx; ++x) {
                const double X = ox + x * vol->vx, Y = oy + y * vol->vy, Z = oz + z * vol->vz;
                double val = 0.0;
                for (int k = 0; k < n; ++k) {
                    const double* e = ell + (size_t)k * YKO_ELLIPSOID_STRIDE;
                    if (inside_ellipsoid(e, X, Y, Z)) val += e[15];
                }
                if (clamp_nonnegative && val < 0.0) val = 0.0;
                out[((size_t)z * vol->ny + y) * vol->nx + x] = (float)val;
            }
}

/* chord length of src->dst (the whole line segment, t in [0,1]) through each
 * ellipsoid, times its value */
void yko_ellipsoid_line_integrals(const double* ell, int n,
    const double* views, int n_views, int nu, int nv, double* proj_out)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < n_views; ++i)
        for (int v = 0; v < nv; ++v) {
            const double* g = views + (size_t)i * YKO_VIEW_STRIDE;
            for (int u = 0; u < nu; ++u) {
                double dst[3], dir[3];
                for (int k = 0; k < 3; ++k) {
                    dst[k] = g[3 + k] + u * g[6 + k] + v * g[9 + k];
                    dir[k] = dst[k] - g[k];
                }
                const double len = sqrt(dot3(dir, dir));
                double acc = 0.0;
                for (int e = 0; e < n; ++e) {
                    const double* E = ell + (size_t)e * YKO_ELLIPSOID_STRIDE;
                    const double* R = E + 6;
                    double p[3], q[3];
                    for (int k = 0; k < 3; ++k) {
                        const double s = g[0] - E[0], t = g[1] - E[1], w = g[2] - E[2];
                        p[k] = (R[3 * k] * s + R[3 * k + 1] * t + R[3 * k + 2] * w) / E[3 + k];
                        q[k] = (R[3 * k] * dir[0] + R[3 * k + 1] * dir[1] + R[3 * k + 2] * dir[2]) / E[3 + k];
                    }
                    const double A = dot3(q, q), B = 2.0 * dot3(p, q), C = dot3(p, p) - 1.0;
                    const double disc = B * B - 4.0 * A * C;
                    if (!(disc > 0.0) || !(A > 0.0)) continue;
                    const double sq = sqrt(disc);
                    double t0 = (-B - sq) / (2.0 * A), t1 = (-B + sq) / (2.0 * A);
                    if (t0 < 0.0) t0 = 0.0;
                    if (t1 > 1.0) t1 = 1.0;
                    if (t1 > t0) acc += E[15] * (t1 - t0) * len;
                }
                proj_out[((size_t)i * nv + v) * nu + u] = acc;
            }
        }
}

/* ------------------------------------------------------------------------ */
/* smoothed TV                                                               */
/* ------------------------------------------------------------------------ */
void yko_tv_value_grad(const double* x, const yko_volume* vol, double eps,
    double* value_out, double* grad_out)
{
    const int nx = vol->nx, ny = vol->ny, nz = vol->nz;
    const double ix = 1.0 / vol->vx, iy = 1.0 / vol->vy, iz = 1.0 / vol->vz;
    const size_t n = (size_t)nx * ny * nz;
    double* px = (double*)malloc(sizeof(double) * n);
    double* py = (double*)malloc(sizeof(double) * n);
    double* pz = (double*)malloc(sizeof(double) * n);
    double value = 0.0;
#pragma omp parallel for collapse(2) reduction(+:value) schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int xx = 0; xx < nx; ++xx) {
                const size_t c = ((size_t)z * ny + y) * nx + xx;
                const double gx = xx + 1 < nx ? (x[c + 1] - x[c]) * ix : 0.0;
                const double gy = y + 1 < ny ? (x[c + nx] - x[c]) * iy : 0.0;
                const double gz = z + 1 < nz ? (x[c + (size_t)nx * ny] - x[c]) * iz : 0.0;
                const double den = sqrt(gx * gx + gy * gy + gz * gz + eps * eps);
                value += den;
                px[c] = gx / den; py[c] = gy / den; pz[c] = gz / den;
            }
    if (grad_out) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int xx = 0; xx < nx; ++xx) {
                    const size_t c = ((size_t)z * ny + y) * nx + xx;
                    const double div = (px[c] - (xx > 0 ? px[c - 1] : 0.0)) * ix +
                        (py[c] - (y > 0 ? py[c - nx] : 0.0)) * iy +
                        (pz[c] - (z > 0 ? pz[c - (size_t)nx * ny] : 0.0)) * iz;
                    grad_out[c] = -div;
                }
    }
    free(px); free(py); free(pz);
    if (value_out) *value_out = value;
}

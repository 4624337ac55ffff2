/*
 * yk_oracle.h -- FP64 CPU restatement of the YKCBCT flat-panel circular-orbit
 * reconstruction path (geometry builder, FDK derived geometry and affine
 * coefficients, cosine preweight, Parker weights, ramp-filter weights and row
 * filtering, voxel-driven FDK backprojection, Joseph and Siddon forward
 * projection and their exact transposes).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  The product library never links or calls it.
 *
 * Every function cites the reference file:line whose formula it restates.
 * All arithmetic is double precision; inputs that the product receives as
 * float (projections, volumes, view vectors) are taken as float and widened.
 *
 * Layouts (identical to the reference, doc/geometry-and-data.md:105-106):
 *   projection [view][v][u], volume [z][y][x], row-major, contiguous.
 * View vectors are 13 doubles: src[3], detS[3], detU[3], detV[3], angle.
 *   detS is the centre of pixel (0,0); detU/detV are one-pixel steps
 *   (include/YKCBCT/geometry/YkProjectionGeometry.hpp:10-18).
 *
 * Parity pinning: see tests/test_oracle_*.py.  Each function below is pinned
 * by closed forms, invariants or the reference's own published numbers.
 */
#ifndef YK_ORACLE_H
#define YK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YKO_VIEW_STRIDE 13
#define YKO_DERIVED_STRIDE 20
#define YKO_COEFF_STRIDE 17

/* Derived-geometry record offsets (restates SFDKGeoParamPerView,
 * src/global/YkFdkKernelTypes.hpp:7-32). */
enum {
    YKO_D_THETA = 0, YKO_D_DTHETA = 1, YKO_D_FSCALE = 2, YKO_D_SOD = 3,
    YKO_D_SDD_RADIAL = 4, YKO_D_SDD_PLANE = 5, YKO_D_OFFU = 6, YKO_D_OFFV = 7,
    YKO_D_DU = 8, YKO_D_DV = 9, YKO_D_RADIAL = 10 /* 3 */, YKO_D_NORMAL = 13 /* 3 */,
    YKO_D_UU = 16, YKO_D_VV = 17, YKO_D_UV = 18
};

/* Coefficient record offsets (restates FdkAffineCoeff,
 * src/global/YkFdkKernelTypes.hpp:35-55). */
enum {
    YKO_C_CU = 0 /* 4 */, YKO_C_CV = 4 /* 4 */, YKO_C_CD = 8 /* 4 */,
    YKO_C_CR = 12 /* 4 */, YKO_C_WBASE = 16
};

typedef struct {
    int nx, ny, nz;
    double vx, vy, vz;
    double cx, cy, cz;
} yko_volume;

/* OpenMP threads the oracle uses */
int yko_num_threads(void);

/* ---- geometry ---------------------------------------------------------- */
int yko_build_circular_flat_views(
    int total_views, int views_per_turn, double start_angle_rad, int direction,
    double sid_mm, double sdd_mm, double z_mm, const double source_offset_mm[3],
    int channels, int rows, double channel_mm, double row_mm,
    const double offset_unv_mm[3], double tilt_u, double tilt_v, double tilt_n,
    double* views_out);

int yko_derive(const double* views, int n_views, int nu, int nv,
    double scan_range_rad, double* derived_out);

int yko_fdk_coeffs(const double* views, const double* derived, int n_views,
    double* coeffs_out);

/* ---- projection-domain weights ---------------------------------------- */
void yko_preweight(const double* views, const double* derived, int n_views,
    int nu, int nv, const double* in, double* out);

void yko_parker(const double* views, const double* derived, int n_views,
    int nu, int nv, double start_angle_rad, int direction,
    double scan_range_rad, double* data);

/* ---- filter ------------------------------------------------------------ */
int yko_padded_n(int nu);
int yko_start_u(int nu, int padded_n, double offset_u_pix);
double yko_bessel_i0_as(double x);
double yko_bessel_i0_series(double x);
double yko_window(int kind, double x, double order, double beta, double tukey_alpha,
    int parametric);
/* source: 0 = AnalyticFreq, 1 = DiscreteRLFFT */
int yko_filter_weights(int kind, int source, double cutoff, double gain,
    double order, double beta, double tukey_alpha, int padded_n, double du_mm,
    double* weights_out);
int yko_filter_weights_ext(int kind, int source, int extract, double cutoff, double gain,
    double order, double beta, double tukey_alpha, int N, double du_mm,
    const double* ramp, int m, const double* custom, int nc_custom, double* w);
void yko_filter_kernel_from_weights(const double* weights, int padded_n,
    double* h_out);
/* method: 0 = direct circular convolution with h = IDFT(w), 1 = FP64 FFT */
void yko_filter_rows(const double* in, int n_views, int nv, int nu,
    int padded_n, const double* weights, const int* start_u, int method,
    double* out);
void yko_fft(double* re, double* im, int n, int inverse);

/* ---- FDK backprojection ----------------------------------------------- */
double yko_bilinear(const double* plane, int nu, int nv, double u, double v);
void yko_bp_fdk(const double* coeffs, int n_views, const double* filtered,
    int nu, int nv, const yko_volume* vol, const int64_t* voxel_index,
    int64_t n_voxels, double* out);
/* operator-form voxel-driven backprojections of unfiltered views: matched 0
 * FDK weight, 1 pseudo-matched (src/FlatFpBp/BP/kernels/YkBpFdkKernel.cu:228-466),
 * 2 voxel-driven Joseph (YkBPJosephKernel.cu:624-735) */
void yko_bp_fdk_operator(const double* views, const double* derived, const double* coeffs,
    int n_views, const double* proj, int nu, int nv, const yko_volume* vol, int matched,
    double* out);

/* ---- forward projectors and transposes -------------------------------- */
int yko_main_axis(const double* view, int nu, int nv);
void yko_fp_joseph(const double* views, int n_views, int nu, int nv,
    const yko_volume* vol, const float* volume, double* proj_out);
void yko_bp_joseph_adjoint(const double* views, int n_views, int nu, int nv,
    const yko_volume* vol, const double* proj, double* volume_out);
void yko_fp_siddon(const double* views, int n_views, int nu, int nv,
    const yko_volume* vol, const float* volume, double* proj_out);
void yko_bp_siddon_adjoint(const double* views, int n_views, int nu, int nv,
    const yko_volume* vol, const double* proj, double* volume_out);
double yko_siddon_ray(const double src[3], const double dst[3],
    const yko_volume* vol, const float* volume);

/* ---- phantoms, regularizer --------------------------------------------- */
/* Ellipsoid record (16 doubles, world mm): centre[3], semi_axes[3],
 * R[9] (row-major; body coordinates r = R (p - centre)), value.
 * A point is inside when sum_k (r_k / a_k)^2 <= 1. */
#define YKO_ELLIPSOID_STRIDE 16
void yko_phantom_ellipsoids(const double* ellipsoids, int n, const yko_volume* vol,
    int clamp_nonnegative, float* out);
void yko_ellipsoid_line_integrals(const double* ellipsoids, int n,
    const double* views, int n_views, int nu, int nv, double* proj_out);
/* Smoothed TV with forward differences (zero past the last voxel), divided by
 * the voxel spacing; value and exact gradient.  Restates
 * src/Reconstruction/Iterative/Common/kernels/YkTvRegularizationKernels.cu:17-73
 * and doc/algorithms.md:257-261. */
void yko_tv_value_grad(const double* x, const yko_volume* vol, double eps,
    double* value_out, double* grad_out);

#ifdef __cplusplus
}
#endif
#endif

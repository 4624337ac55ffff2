// Vector helpers for the iterative callers (dot, axpy, fills), smoothed TV
// value/gradient and the ellipsoid phantom, for gfx950.
#include <hip/hip_runtime.h>

#include "yk_internal.hpp"

namespace yk {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ double block_sum(double v, double* red)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kThreads / 64; ++w) s += red[w];
    return s;
}

// per-block FP64 partial sums of a.b (FP32 products, FP64 accumulation);
// the host adds the partials (the reference finishes dots on the host too,
// src/Reconstruction/Iterative/Common/kernels/YkIterKernels.cu:399-432)
__global__ void __launch_bounds__(kThreads)
dot_kernel(const float* __restrict__ a, const float* __restrict__ b, size_t n,
    double* __restrict__ partials)
{
    __shared__ double red[kThreads / 64];
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kThreads)
        acc += double(a[i]) * double(b[i]);
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void axpy_kernel(float* __restrict__ y, const float* __restrict__ x, float a, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = fmaf(a, x[i], y[i]);
}

__global__ void xpby_kernel(float* __restrict__ y, const float* __restrict__ x, float b, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = fmaf(b, y[i], x[i]);
}

__global__ void scale_copy_kernel(float* __restrict__ y, const float* __restrict__ x, float a, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = a * x[i];
}

__global__ void fill_kernel(float* __restrict__ y, float v, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = v;
}

// Algebraic (SIRT / SART / OS-SART) element steps
// (src/Reconstruction/Iterative/Common/kernels/YkIterKernels.cu:202-256, 299-318):
//   mode 0: res = row > eps ? (b - res) / row : 0     (residual, divide_launch)
//   mode 1: res = (b - res) * row                     (residual, multiply_launch)
__global__ void algebraic_residual_kernel(float* __restrict__ res, const float* __restrict__ b,
    const float* __restrict__ row, float eps, int mode, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float r = b[i] - res[i];
        const float w = row[i];
        res[i] = mode == 0 ? (w > eps ? r / w : 0.f) : r * w;
    }
}

// x += lambda bp / (col + eps), then the optional bounds (update_launch,
// clamp_min_launch, clamp_max_launch)
__global__ void algebraic_update_kernel(float* __restrict__ x, const float* __restrict__ bp,
    const float* __restrict__ col, float lambda, float eps, int use_min, float lo, int use_max, float hi,
    size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float v = x[i] + lambda * bp[i] / (col[i] + eps);
        if (use_min && v < lo) v = lo;
        if (use_max && v > hi) v = hi;
        x[i] = v;
    }
}

__global__ void bounds_kernel(float* __restrict__ x, int use_min, float lo, int use_max, float hi, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float v = x[i];
        if (use_min && v < lo) v = lo;
        if (use_max && v > hi) v = hi;
        x[i] = v;
    }
}

// mode 0: x <= t -> 1e30 (threshold_inf_launch); mode 1: additionally x = 1 / x
__global__ void weight_prepare_kernel(float* __restrict__ x, float t, int reciprocal, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float v = x[i];
        if (v <= t) v = 1e30f;
        x[i] = reciprocal ? 1.f / v : v;
    }
}

// forward differences with zero past the last voxel, divided by the spacing
// (src/Reconstruction/Iterative/Common/kernels/YkTvRegularizationKernels.cu:17-39)
__device__ __forceinline__ void tv_field(const float* __restrict__ x, int i, int j, int k,
    const VolumeGeom& g, float ix, float iy, float iz, float eps, float& px, float& py, float& pz,
    float& den)
{
    const size_t c = ((size_t)k * g.ny + j) * g.nx + i;
    const float v = x[c];
    const float gx = i + 1 < g.nx ? (x[c + 1] - v) * ix : 0.f;
    const float gy = j + 1 < g.ny ? (x[c + g.nx] - v) * iy : 0.f;
    const float gz = k + 1 < g.nz ? (x[c + (size_t)g.nx * g.ny] - v) * iz : 0.f;
    den = sqrtf(gx * gx + gy * gy + gz * gz + eps * eps);
    px = gx / den;
    py = gy / den;
    pz = gz / den;
}

// gradient = -div(grad x / sqrt(|grad x|^2 + eps^2)), value = sum sqrt(...)
// (YkTvRegularizationKernels.cu:41-73; doc/algorithms.md:257-261)
// (iz = 0: the slice-wise 2-D form, no z differences; TvDimensionality::Slice2D)
__global__ void __launch_bounds__(kThreads)
tv_kernel(const float* __restrict__ x, float* __restrict__ grad, double* __restrict__ partials,
    VolumeGeom g, float eps, float iz)
{
    __shared__ double red[kThreads / 64];
    const float ix = 1.f / g.vx, iy = 1.f / g.vy;
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    double val = 0.0;
    for (size_t c = (size_t)blockIdx.x * kThreads + threadIdx.x; c < n; c += (size_t)gridDim.x * kThreads) {
        const int i = int(c % g.nx);
        const int j = int((c / g.nx) % g.ny);
        const int k = int(c / ((size_t)g.nx * g.ny));
        float px, py, pz, den, qx, qy, qz, d2;
        tv_field(x, i, j, k, g, ix, iy, iz, eps, px, py, pz, den);
        val += den;
        float div = 0.f;
        float pxm = 0.f, pym = 0.f, pzm = 0.f;
        if (i > 0) { tv_field(x, i - 1, j, k, g, ix, iy, iz, eps, qx, qy, qz, d2); pxm = qx; }
        if (j > 0) { tv_field(x, i, j - 1, k, g, ix, iy, iz, eps, qx, qy, qz, d2); pym = qy; }
        if (k > 0) { tv_field(x, i, j, k - 1, g, ix, iy, iz, eps, qx, qy, qz, d2); pzm = qz; }
        div = (px - pxm) * ix + (py - pym) * iy + (pz - pzm) * iz;
        if (grad) grad[c] = -div;
    }
    const double s = block_sum(val, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// point-sampled ellipsoid phantom; record = centre[3], semi_axes[3], R[9], value
__global__ void __launch_bounds__(kThreads)
phantom_kernel(const float* __restrict__ ell, int n, VolumeGeom g, int clamp, float* __restrict__ out)
{
    const size_t total = (size_t)g.nx * g.ny * g.nz;
    for (size_t c = (size_t)blockIdx.x * kThreads + threadIdx.x; c < total; c += (size_t)gridDim.x * kThreads) {
        const int i = int(c % g.nx);
        const int j = int((c / g.nx) % g.ny);
        const int k = int(c / ((size_t)g.nx * g.ny));
        const double X = double(g.ox) + double(i) * g.vx;
        const double Y = double(g.oy) + double(j) * g.vy;
        const double Z = double(g.oz) + double(k) * g.vz;
        double val = 0.0;
        for (int e = 0; e < n; ++e) {
            const float* E = ell + (size_t)e * 16;
            const double d0 = X - E[0], d1 = Y - E[1], d2 = Z - E[2];
            double q = 0.0;
            for (int r = 0; r < 3; ++r) {
                const double b = (E[6 + 3 * r] * d0 + E[7 + 3 * r] * d1 + E[8 + 3 * r] * d2) / E[3 + r];
                q += b * b;
            }
            if (q <= 1.0) val += E[15];
        }
        if (clamp && val < 0.0) val = 0.0;
        out[c] = float(val);
    }
}

int grid_for(size_t n)
{
    const size_t b = (n + 255) / 256;
    return int(b < 4096 ? (b == 0 ? 1 : b) : 4096);
}

}  // namespace

void launch_dot(const float* a, const float* b, size_t n, double* partials, int n_partials, hipStream_t s)
{
    hipLaunchKernelGGL(dot_kernel, dim3(n_partials), dim3(kThreads), 0, s, a, b, n, partials);
    YK_HIP(hipGetLastError());
}

void launch_axpy(float* y, const float* x, float a, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(n)), dim3(256), 0, s, y, x, a, n);
    YK_HIP(hipGetLastError());
}

void launch_xpby(float* y, const float* x, float b, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(xpby_kernel, dim3(grid_for(n)), dim3(256), 0, s, y, x, b, n);
    YK_HIP(hipGetLastError());
}

void launch_scale_copy(float* y, const float* x, float a, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(scale_copy_kernel, dim3(grid_for(n)), dim3(256), 0, s, y, x, a, n);
    YK_HIP(hipGetLastError());
}

void launch_fill(float* y, float v, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, s, y, v, n);
    YK_HIP(hipGetLastError());
}

void algebraic_residual(float* res, const float* b, const float* row, float eps, int mode, size_t n,
    hipStream_t s)
{
    hipLaunchKernelGGL(algebraic_residual_kernel, dim3(grid_for(n)), dim3(256), 0, s, res, b, row, eps, mode, n);
    YK_HIP(hipGetLastError());
}

void algebraic_update(float* x, const float* bp, const float* col, float lambda, float eps, int use_min,
    float lo, int use_max, float hi, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(algebraic_update_kernel, dim3(grid_for(n)), dim3(256), 0, s, x, bp, col, lambda, eps,
        use_min, lo, use_max, hi, n);
    YK_HIP(hipGetLastError());
}

void launch_bounds_clamp(float* x, int use_min, float lo, int use_max, float hi, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(bounds_kernel, dim3(grid_for(n)), dim3(256), 0, s, x, use_min, lo, use_max, hi, n);
    YK_HIP(hipGetLastError());
}

void weight_prepare(float* x, float threshold, int reciprocal, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(weight_prepare_kernel, dim3(grid_for(n)), dim3(256), 0, s, x, threshold, reciprocal, n);
    YK_HIP(hipGetLastError());
}

void launch_tv_gradient(const float* x, float* grad, double* partials, int n_partials,
    const VolumeGeom& g, float eps, hipStream_t s, int dims)
{
    hipLaunchKernelGGL(tv_kernel, dim3(n_partials), dim3(kThreads), 0, s, x, grad, partials, g, eps,
        dims == 2 ? 0.f : 1.f / g.vz);
    YK_HIP(hipGetLastError());
}

void launch_phantom(const float* ellipsoids, int n, const VolumeGeom& g, int clamp, float* out,
    hipStream_t s)
{
    const size_t total = (size_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(phantom_kernel, dim3(grid_for(total)), dim3(kThreads), 0, s, ellipsoids, n, g,
        clamp, out);
    YK_HIP(hipGetLastError());
}

}  // namespace yk

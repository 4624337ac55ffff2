// FDK projection-domain stage for gfx950: cosine preweight, optional Parker
// weight, zero padding, ramp filter and crop fused into one LDS-resident
// kernel per pair of detector rows.
//
// Two real rows are packed as the real and imaginary parts of one complex
// sequence.  The ramp response is real and even (w[N-k] = w[k]), so
//   IFFT(FFT(a + i b) * W) = (a * h) + i (b * h)
// and both rows come out of one forward and one inverse transform.  The
// transform is a mixed radix-4/2 Stockham FFT over two LDS ping-pong buffers;
// twiddles come from an FP64-computed table.  Equivalent to the reference's
// pad -> cuFFT R2C -> pointwise * w -> C2R -> crop chain
// (src/Reconstruction/Analytic/Circular/Flat/FDK/kernels/YkFDKFilterKernels.cu:32-140,
// src/Filter/YkConvKernels.cu:109-220) with the preweight of
// .../FDK/kernels/YkFDKPreWeightKernels.cu:289-405 applied on the way in.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include "yk_internal.hpp"

namespace yk {
namespace {

constexpr int kFilterThreads = 256;

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// one Stockham pass of radix R over span Ns (src -> dst)
template <int R>
__device__ __forceinline__ void stockham_pass(const float2* __restrict__ src,
    float2* __restrict__ dst, const float2* __restrict__ tw, int n, int ns)
{
    const int quarter = n / R;
    const int stride = n / (ns * R);
    for (int j = threadIdx.x; j < quarter; j += kFilterThreads) {
        const int k = j & (ns - 1);
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = src[j + r * quarter];
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[(r * k * stride) & (n - 1)]);
        float2 o[R];
        if constexpr (R == 4) {
            const float2 s02 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
            const float2 d02 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
            const float2 s13 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
            const float2 d13 = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
            o[0] = make_float2(s02.x + s13.x, s02.y + s13.y);
            o[2] = make_float2(s02.x - s13.x, s02.y - s13.y);
            // -i * d13 = (d13.y, -d13.x)
            o[1] = make_float2(d02.x + d13.y, d02.y - d13.x);
            o[3] = make_float2(d02.x - d13.y, d02.y + d13.x);
        } else {
            o[0] = make_float2(v[0].x + v[1].x, v[0].y + v[1].y);
            o[1] = make_float2(v[0].x - v[1].x, v[0].y - v[1].y);
        }
        const int idx = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[idx + r * ns] = o[r];
    }
}

// forward FFT of `a` in place; `b` is scratch.  Returns the buffer holding the result.
__device__ float2* fft_forward(float2* a, float2* b, const float2* __restrict__ tw, int n)
{
    float2* src = a;
    float2* dst = b;
    for (int ns = 1; ns < n;) {
        if (ns * 4 <= n) {
            stockham_pass<4>(src, dst, tw, n, ns);
            ns *= 4;
        } else {
            stockham_pass<2>(src, dst, tw, n, ns);
            ns *= 2;
        }
        __syncthreads();
        float2* t = src;
        src = dst;
        dst = t;
    }
    return src;
}

__device__ __forceinline__ float row_value(const float* __restrict__ proj, const FilterView& fv,
    const float* __restrict__ parker_row, int v, int u, int nu)
{
    const float x = proj[(size_t)v * nu + u];
    // q = detS + v detV + u detU - src
    const float qx = (fv.detS[0] + fv.detV[0] * v - fv.src[0]) + fv.detU[0] * u;
    const float qy = (fv.detS[1] + fv.detV[1] * v - fv.src[1]) + fv.detU[1] * u;
    const float qz = (fv.detS[2] + fv.detV[2] * v - fv.src[2]) + fv.detU[2] * u;
    const float r2 = fmaxf(qx * qx + qy * qy + qz * qz, 1e-20f);
    float w = fv.sdd_radial * rsqrtf(r2) * fv.wgeom;
    if (parker_row) w *= parker_row[u];
    return x * w;
}

__device__ __forceinline__ float block_max(float v, float* red)
{
    // wave64 reduction through LDS-free shuffles, then across waves
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[wave] = v;
    __syncthreads();
    float m = red[0];
    for (int w = 1; w < kFilterThreads / 64; ++w) m = fmaxf(m, red[w]);
    __syncthreads();
    return m;
}

// power-of-two scale s with max|row| / s in (2^13, 2^14]; 1 for an all-zero row
__device__ __forceinline__ float row_scale(float maxabs)
{
    if (!(maxabs > 0.f) || !isfinite(maxabs)) return 1.f;
    int e;
    const float m = frexpf(maxabs, &e);  // maxabs = m 2^e, m in [0.5, 1)
    (void)m;
    return ldexpf(1.f, e - 14);
}

template <int STORAGE>
__device__ __forceinline__ void store_row(void* out, size_t base, int a, float val, float inv_scale)
{
    if constexpr (STORAGE == YK_F32) {
        static_cast<float*>(out)[base + a] = val;
    } else if constexpr (STORAGE == YK_F16) {
        static_cast<__half*>(out)[base + a] = __float2half_rn(val * inv_scale);
    } else {
        static_cast<__hip_bfloat16*>(out)[base + a] = __float2bfloat16(val * inv_scale);
    }
}

template <int STORAGE>
__global__ void __launch_bounds__(kFilterThreads)
fdk_filter_kernel(const float* __restrict__ proj, void* __restrict__ out,
    float* __restrict__ row_scales, const FilterView* __restrict__ views,
    const float* __restrict__ parker, const float* __restrict__ weights,
    const float2* __restrict__ tw, int nu, int nv, int rows, int n)
{
    extern __shared__ float4 smem_raw[];
    float2* bufa = reinterpret_cast<float2*>(smem_raw);
    float2* bufb = bufa + n;
    __shared__ float red[kFilterThreads / 64];

    const int ra = 2 * blockIdx.x;
    const int rb = ra + 1;
    const bool has_b = rb < rows;
    const int ia = ra / nv, va = ra - ia * nv;
    const int ib = has_b ? rb / nv : ia, vb = has_b ? rb - ib * nv : va;
    const FilterView fa = views[ia];
    const FilterView fb = views[ib];
    const float* pa = proj + (size_t)ia * nv * nu;
    const float* pb = proj + (size_t)ib * nv * nu;
    const float* park_a = parker ? parker + (size_t)ia * nu : nullptr;
    const float* park_b = parker ? parker + (size_t)ib * nu : nullptr;

    for (int m = threadIdx.x; m < n; m += kFilterThreads) {
        const int ua = m - fa.start_u;
        const int ub = m - fb.start_u;
        float xa = 0.f, xb = 0.f;
        if (ua >= 0 && ua < nu) xa = row_value(pa, fa, park_a, va, ua, nu);
        if (has_b && ub >= 0 && ub < nu) xb = row_value(pb, fb, park_b, vb, ub, nu);
        bufa[m] = make_float2(xa, xb);
    }
    __syncthreads();

    float2* spec = fft_forward(bufa, bufb, tw, n);
    // Y = conj(X W): the second forward transform then yields conj(IFFT(X W))
    for (int k = threadIdx.x; k < n; k += kFilterThreads) {
        const float w = weights[k <= n / 2 ? k : n - k];
        const float2 x = spec[k];
        spec[k] = make_float2(x.x * w, -x.y * w);
    }
    __syncthreads();
    float2* other = spec == bufa ? bufb : bufa;
    float2* res = fft_forward(spec, other, tw, n);

    // crop: row a = Re(res), row b = -Im(res)
    float inv_a = 1.f, inv_b = 1.f;
    if constexpr (STORAGE != YK_F32) {
        float ma = 0.f, mb = 0.f;
        for (int a = threadIdx.x; a < nu; a += kFilterThreads) {
            const int ja = fa.start_u + a, jb = fb.start_u + a;
            if (ja >= 0 && ja < n) ma = fmaxf(ma, fabsf(res[ja].x));
            if (jb >= 0 && jb < n) mb = fmaxf(mb, fabsf(res[jb].y));
        }
        ma = block_max(ma, red);
        mb = block_max(mb, red);
        const float sa = row_scale(ma), sb = row_scale(mb);
        inv_a = 1.f / sa;
        inv_b = 1.f / sb;
        if (threadIdx.x == 0) {
            row_scales[ra] = sa;
            if (has_b) row_scales[rb] = sb;
        }
    }
    const size_t base_a = (size_t)ra * nu, base_b = (size_t)rb * nu;
    for (int a = threadIdx.x; a < nu; a += kFilterThreads) {
        const int ja = fa.start_u + a;
        store_row<STORAGE>(out, base_a, a, (ja >= 0 && ja < n) ? res[ja].x : 0.f, inv_a);
        if (has_b) {
            const int jb = fb.start_u + a;
            store_row<STORAGE>(out, base_b, a, (jb >= 0 && jb < n) ? -res[jb].y : 0.f, inv_b);
        }
    }
}

}  // namespace

void launch_fdk_filter(const FilterLaunch& p, hipStream_t s)
{
    const int rows = p.count * p.nv;
    const int pairs = (rows + 1) / 2;
    const size_t lds = size_t(2) * p.n * sizeof(float2);
    if (lds > 160 * 1024) fail(YK_ERR_UNSUPPORTED, "padded row too long for the LDS filter");
    dim3 grid(pairs), block(kFilterThreads);
    const void* fns[3] = { reinterpret_cast<const void*>(fdk_filter_kernel<YK_F32>),
                           reinterpret_cast<const void*>(fdk_filter_kernel<YK_F16>),
                           reinterpret_cast<const void*>(fdk_filter_kernel<YK_BF16>) };
    if (p.storage >= 0 && p.storage < 3)
        YK_HIP(hipFuncSetAttribute(fns[p.storage], hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    switch (p.storage) {
    case YK_F32:
        hipLaunchKernelGGL(fdk_filter_kernel<YK_F32>, grid, block, lds, s, p.proj, p.out,
            p.row_scales, p.views, p.parker, p.weights, p.twiddles, p.nu, p.nv, rows, p.n);
        break;
    case YK_F16:
        hipLaunchKernelGGL(fdk_filter_kernel<YK_F16>, grid, block, lds, s, p.proj, p.out,
            p.row_scales, p.views, p.parker, p.weights, p.twiddles, p.nu, p.nv, rows, p.n);
        break;
    case YK_BF16:
        hipLaunchKernelGGL(fdk_filter_kernel<YK_BF16>, grid, block, lds, s, p.proj, p.out,
            p.row_scales, p.views, p.parker, p.weights, p.twiddles, p.nu, p.nv, rows, p.n);
        break;
    default: fail(YK_ERR_INVALID_CONFIG, "unknown storage type");
    }
    YK_HIP(hipGetLastError());
}

}  // namespace yk

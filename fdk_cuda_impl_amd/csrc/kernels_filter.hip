// FDK projection-domain stage for gfx950: cosine preweight, optional Parker
// weight, zero padding, ramp filter and crop fused into one LDS-resident
// kernel per pair of detector rows.
//
// Two real rows are packed as the real and imaginary parts of one complex
// sequence.  The ramp response is real and even (w[N-k] = w[k]), so
//   IFFT(FFT(a + i b) * W) = (a * h) + i (b * h)
// and both rows come out of one forward and one inverse transform.  The
// transform is a mixed radix-8/4/2 Stockham FFT over two padded LDS ping-pong
// buffers; twiddles come from an FP64-computed table, ordered pass by pass
// (fft_twiddles), read from global memory.  Equivalent to the reference's
// pad -> cuFFT R2C -> pointwise * w -> C2R -> crop chain
// (src/Reconstruction/Analytic/Circular/Flat/FDK/kernels/YkFDKFilterKernels.cu:32-140,
// src/Filter/YkConvKernels.cu:109-220) with the preweight of
// .../FDK/kernels/YkFDKPreWeightKernels.cu:289-405 applied on the way in.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <algorithm>

#include "yk_internal.hpp"

namespace yk {
namespace {

constexpr int kFilterThreads = 256;

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// LDS index with one pad slot per 8 complex values: the radix-8 passes write
// runs of 8 consecutive values per lane, which the padding spreads over banks
__device__ __forceinline__ int pidx(int i) { return i + (i >> 3); }

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// -i a
__device__ __forceinline__ float2 cmi(float2 a) { return make_float2(a.y, -a.x); }

// forward DFT-4 in place
__device__ __forceinline__ void dft4(float2& a, float2& b, float2& c, float2& d)
{
    const float2 s02 = cadd(a, c), d02 = csub(a, c), s13 = cadd(b, d), d13 = csub(b, d);
    a = cadd(s02, s13);
    c = csub(s02, s13);
    b = cadd(d02, cmi(d13));
    d = csub(d02, cmi(d13));
}

// forward DFT-R of v (R = 8, 4 or 2), result in natural order
template <int R>
__device__ __forceinline__ void dft(float2 (&v)[R])
{
    if constexpr (R == 2) {
        const float2 a = v[0];
        v[0] = cadd(a, v[1]);
        v[1] = csub(a, v[1]);
    } else if constexpr (R == 4) {
        dft4(v[0], v[1], v[2], v[3]);
    } else {
        // even / odd DFT-4s, then X[m] = E[m] + W8^m O[m], X[m+4] = E[m] - W8^m O[m]
        float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
        float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
        dft4(e0, e1, e2, e3);
        dft4(o0, o1, o2, o3);
        const float h = 0.70710678118654752f;
        const float2 t1 = make_float2((o1.x + o1.y) * h, (o1.y - o1.x) * h);   // W8^1 o1
        const float2 t2 = cmi(o2);                                           // W8^2 o2
        const float2 t3 = make_float2((o3.y - o3.x) * h, -(o3.x + o3.y) * h);  // W8^3 o3
        v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
        v[1] = cadd(e1, t1); v[5] = csub(e1, t1);
        v[2] = cadd(e2, t2); v[6] = csub(e2, t2);
        v[3] = cadd(e3, t3); v[7] = csub(e3, t3);
    }
}

// one Stockham pass of radix R over span ns (src -> dst).  tw is this pass's
// twiddle block: tw[(r - 1) ns + k] = exp(-2 pi i r k / (ns R)), so the lanes of
// a wave read consecutive entries.
template <int R>
__device__ __forceinline__ void stockham_pass(const float2* src, float2* dst, const float2* tw, int n, int ns)
{
    const int m = n / R;
    for (int j = threadIdx.x; j < m; j += kFilterThreads) {
        const int k = j & (ns - 1);
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = src[pidx(j + r * m)];
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[(r - 1) * ns + k]);
        dft<R>(v);
        const int idx = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[pidx(idx + r * ns)] = v[r];
    }
}

// forward FFT of `a` (padded layout); `b` is scratch.  Returns the buffer holding the result.
__device__ float2* fft_forward(float2* a, float2* b, const float2* tw, int n)
{
    float2* src = a;
    float2* dst = b;
    for (int ns = 1; ns < n;) {
        if (ns * 8 <= n) {
            stockham_pass<8>(src, dst, tw, n, ns);
            tw += 7 * ns;
            ns *= 8;
        } else if (ns * 4 <= n) {
            stockham_pass<4>(src, dst, tw, n, ns);
            tw += 3 * ns;
            ns *= 4;
        } else {
            stockham_pass<2>(src, dst, tw, n, ns);
            tw += ns;
            ns *= 2;
        }
        __syncthreads();
        float2* t = src;
        src = dst;
        dst = t;
    }
    return src;
}

// The same transform for a length known at compile time (n = NPT x the
// workgroup size, every length of 256 and above): the pass sequence, spans
// and lane loops are constants, so the index arithmetic folds to shifts and
// masks, and the first pass (span 1, all twiddles 1) multiplies by nothing.
// The operations on the data are those of fft_forward.
template <int R, int N, int NS, bool SPECTRAL = false>
__device__ __forceinline__ void stockham_pass_ct(const float2* src, float2* dst, const float2* tw,
    const float* __restrict__ weights = nullptr)
{
    constexpr int m = N / R;
#pragma unroll 1
    for (int j0 = 0; j0 < m; j0 += kFilterThreads) {
        const int j = j0 + int(threadIdx.x);
        if (m % kFilterThreads != 0 && j >= m) break;
        const int k = j & (NS - 1);
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = src[pidx(j + r * m)];
        if constexpr (NS > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[(r - 1) * NS + k]);
        }
        dft<R>(v);
        const int idx = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if constexpr (SPECTRAL) {
                // last pass: bin idx + r NS in natural order; Y = conj(X W)
                const int b = idx + r * NS;
                const float w = weights[b <= N / 2 ? b : N - b];
                v[r] = make_float2(v[r].x * w, -v[r].y * w);
            }
            dst[pidx(idx + r * NS)] = v[r];
        }
    }
}

// SPECTRAL: the last pass also applies the filter weights and conjugates
template <int N, int NS, bool SPECTRAL>
__device__ __forceinline__ float2* fft_forward_ct(float2* src, float2* dst, const float2* tw,
    const float* __restrict__ weights)
{
    if constexpr (NS >= N) {
        (void)dst; (void)tw; (void)weights;
        return src;
    } else {
        constexpr int R = NS * 8 <= N ? 8 : (NS * 4 <= N ? 4 : 2);
        stockham_pass_ct<R, N, NS, SPECTRAL && NS * R >= N>(src, dst, tw, weights);
        __syncthreads();
        return fft_forward_ct<N, NS * R, SPECTRAL>(dst, src, tw + (R - 1) * NS, weights);
    }
}

// NPT > 1 means n = NPT x the workgroup size exactly (the launcher picks
// NPT = n / 256 for power-of-two n >= 512); NPT = 1 also serves n < 256
template <int NPT>
__device__ __forceinline__ float2* fft_rows(float2* a, float2* b, const float2* tw, int n)
{
    if constexpr (NPT > 1) {
        (void)n;
        return fft_forward_ct<NPT * kFilterThreads, 1, false>(a, b, tw, nullptr);
    } else {
        return fft_forward(a, b, tw, n);
    }
}

// cosine weight SDD_r / |q| * SDD_r / (2 SOD), q = detS + u detU + v detV - src
// (.../FDK/kernels/YkFDKPreWeightKernels.cu:374,401-402), times Parker
__device__ __forceinline__ float row_weight(const FilterView& fv, const float* __restrict__ parker_row,
    int v, int u)
{
    const float qx = (fv.detS[0] + fv.detV[0] * v - fv.src[0]) + fv.detU[0] * u;
    const float qy = (fv.detS[1] + fv.detV[1] * v - fv.src[1]) + fv.detU[1] * u;
    const float qz = (fv.detS[2] + fv.detV[2] * v - fv.src[2]) + fv.detU[2] * u;
    const float r2 = fmaxf(qx * qx + qy * qy + qz * qz, 1e-20f);
    float w = fv.sdd_radial * rsqrtf(r2) * fv.wgeom;
    if (parker_row) w *= parker_row[u];
    return w;
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

// Detector rows of one padded pair: row r of the flattened [view][v] list.
struct RowRef {
    int view, v;
};

__device__ __forceinline__ RowRef row_ref(int r, int nv)
{
    RowRef x;
    x.view = r / nv;
    x.v = r - x.view * nv;
    return x;
}

// Persistent workgroups: each filters row pairs p = blockIdx.x, + gridDim.x,
// ... .  The raw texels of the next pair (NPT padded positions per lane and
// row) are loaded into registers while the current pair is transformed.
template <int STORAGE, int NPT>
__global__ void __launch_bounds__(kFilterThreads)
fdk_filter_kernel(const float* __restrict__ proj, void* __restrict__ out,
    float* __restrict__ row_scales, const FilterView* __restrict__ views,
    const float* __restrict__ parker, const float* __restrict__ weights,
    const float2* __restrict__ tw, int nu, int nv, int rows, int n)
{
    extern __shared__ float4 smem_raw[];
    const int padded = pidx(n);
    float2* bufa = reinterpret_cast<float2*>(smem_raw);
    float2* bufb = bufa + padded;
    const float2* tws = tw;  // pass-ordered table, read through the L1
    __shared__ float red[kFilterThreads / 64];

    const int pairs = (rows + 1) / 2;
    float xa[NPT], xb[NPT];
    auto prefetch = [&](int p) {
        const int ra = 2 * p, rb = ra + 1;
        const RowRef A = row_ref(ra, nv), B = row_ref(rb < rows ? rb : ra, nv);
        const int sa = views[A.view].start_u, sb = views[B.view].start_u;
        const float* pa = proj + ((size_t)A.view * nv + A.v) * nu;
        const float* pb = proj + ((size_t)B.view * nv + B.v) * nu;
#pragma unroll
        for (int e = 0; e < NPT; ++e) {
            const int m = int(threadIdx.x) + e * kFilterThreads;
            const int ua = m - sa, ub = m - sb;
            xa[e] = (m < n && ua >= 0 && ua < nu) ? pa[ua] : 0.f;
            xb[e] = (m < n && rb < rows && ub >= 0 && ub < nu) ? pb[ub] : 0.f;
        }
    };
    if (int(blockIdx.x) < pairs) prefetch(blockIdx.x);

    for (int p = blockIdx.x; p < pairs; p += gridDim.x) {
        const int ra = 2 * p, rb = ra + 1;
        const bool has_b = rb < rows;
        const RowRef A = row_ref(ra, nv), B = row_ref(has_b ? rb : ra, nv);
        const FilterView fa = views[A.view];
        const FilterView fb = views[B.view];
        const float* park_a = parker ? parker + (size_t)A.view * nu : nullptr;
        const float* park_b = parker ? parker + (size_t)B.view * nu : nullptr;
        // Y = conj(X W): the second forward transform then yields conj(IFFT(X W))
        float2* spec;
        if constexpr (NPT >= 8) {
            // The first radix-8 pass (span 1, no twiddles) reads elements
            // j + r N/8, which for N/8 a multiple of the workgroup size are
            // this lane's own prefetched texels: weight them (cosine and
            // Parker) in registers and transform them straight into LDS.
            constexpr int N = NPT * kFilterThreads, M = N / 8, S = M / kFilterThreads;
            float2 wv[NPT];
#pragma unroll
            for (int e = 0; e < NPT; ++e) {
                const int m = int(threadIdx.x) + e * kFilterThreads;
                const int ua = m - fa.start_u, ub = m - fb.start_u;
                const float wa = (ua >= 0 && ua < nu) ? row_weight(fa, park_a, A.v, ua) : 0.f;
                const float wb = (has_b && ub >= 0 && ub < nu) ? row_weight(fb, park_b, B.v, ub) : 0.f;
                wv[e] = make_float2(xa[e] * wa, xb[e] * wb);
            }
#pragma unroll
            for (int t = 0; t < S; ++t) {
                float2 v[8];
#pragma unroll
                for (int r = 0; r < 8; ++r) v[r] = wv[t + r * S];
                dft<8>(v);
                const int j = t * kFilterThreads + int(threadIdx.x);
#pragma unroll
                for (int r = 0; r < 8; ++r) bufb[pidx(j * 8 + r)] = v[r];
            }
            __syncthreads();
            if (p + int(gridDim.x) < pairs) prefetch(p + gridDim.x);
            spec = fft_forward_ct<N, 8, true>(bufb, bufa, tws + 7, weights);
        } else {
            // cosine (and Parker) weights on the way into LDS
#pragma unroll
            for (int e = 0; e < NPT; ++e) {
                const int m = int(threadIdx.x) + e * kFilterThreads;
                if (m < n) {
                    const int ua = m - fa.start_u, ub = m - fb.start_u;
                    const float wa = (ua >= 0 && ua < nu) ? row_weight(fa, park_a, A.v, ua) : 0.f;
                    const float wb = (has_b && ub >= 0 && ub < nu) ? row_weight(fb, park_b, B.v, ub) : 0.f;
                    bufa[pidx(m)] = make_float2(xa[e] * wa, xb[e] * wb);
                }
            }
            __syncthreads();
            if (p + int(gridDim.x) < pairs) prefetch(p + gridDim.x);
            if constexpr (NPT > 1) {
                spec = fft_forward_ct<NPT * kFilterThreads, 1, true>(bufa, bufb, tws, weights);
            } else {
                spec = fft_forward(bufa, bufb, tws, n);
                for (int k = threadIdx.x; k < n; k += kFilterThreads) {
                    const float w = weights[k <= n / 2 ? k : n - k];
                    const float2 x = spec[pidx(k)];
                    spec[pidx(k)] = make_float2(x.x * w, -x.y * w);
                }
                __syncthreads();
            }
        }
        float2* other = spec == bufa ? bufb : bufa;
        float2* res = fft_rows<NPT>(spec, other, tws, n);

        // crop: row a = Re(res), row b = -Im(res)
        float inv_a = 1.f, inv_b = 1.f;
        if constexpr (STORAGE != YK_F32) {
            float ma = 0.f, mb = 0.f;
            for (int a = threadIdx.x; a < nu; a += kFilterThreads) {
                const int ja = fa.start_u + a, jb = fb.start_u + a;
                if (ja >= 0 && ja < n) ma = fmaxf(ma, fabsf(res[pidx(ja)].x));
                if (jb >= 0 && jb < n) mb = fmaxf(mb, fabsf(res[pidx(jb)].y));
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
            store_row<STORAGE>(out, base_a, a, (ja >= 0 && ja < n) ? res[pidx(ja)].x : 0.f, inv_a);
            if (has_b) {
                const int jb = fb.start_u + a;
                store_row<STORAGE>(out, base_b, a, (jb >= 0 && jb < n) ? -res[pidx(jb)].y : 0.f, inv_b);
            }
        }
        // the next pair overwrites bufa, which the crop (res) may be
        __syncthreads();
    }
}

template <int STORAGE, int NPT>
void launch_npt(const FilterLaunch& p, int rows, size_t lds, hipStream_t s)
{
    auto k = fdk_filter_kernel<STORAGE, NPT>;
    YK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
        int(lds)));
    int dev = 0, cus = 256;
    YK_HIP(hipGetDevice(&dev));
    YK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int per_cu = std::max(1, int((160 * 1024) / lds));
    const int pairs = (rows + 1) / 2;
    const int grid = std::max(1, std::min(pairs, cus * std::min(per_cu, 8)));
    hipLaunchKernelGGL(k, dim3(grid), dim3(kFilterThreads), lds, s, p.proj, p.out, p.row_scales, p.views,
        p.parker, p.weights, p.twiddles, p.nu, p.nv, rows, p.n);
}

template <int STORAGE>
void launch_storage(const FilterLaunch& p, int rows, size_t lds, hipStream_t s)
{
    const int npt = std::max(1, p.n / kFilterThreads);
    switch (npt) {
    case 1: launch_npt<STORAGE, 1>(p, rows, lds, s); break;
    case 2: launch_npt<STORAGE, 2>(p, rows, lds, s); break;
    case 4: launch_npt<STORAGE, 4>(p, rows, lds, s); break;
    case 8: launch_npt<STORAGE, 8>(p, rows, lds, s); break;
    case 16: launch_npt<STORAGE, 16>(p, rows, lds, s); break;
    default: fail(YK_ERR_UNSUPPORTED, "padded row length outside the filter kernel's range");
    }
}

// weighted projections alone (the stage the filter consumes): out = p * w
__global__ void __launch_bounds__(256)
preweight_kernel(const float* __restrict__ proj, float* __restrict__ out, const FilterView* __restrict__ views,
    const float* __restrict__ parker, int nu, int nv)
{
    const int u = blockIdx.x * 256 + int(threadIdx.x);
    const int v = blockIdx.y, i = blockIdx.z;
    if (u >= nu) return;
    const size_t at = ((size_t)i * nv + v) * nu + u;
    out[at] = proj[at] * row_weight(views[i], parker ? parker + (size_t)i * nu : nullptr, v, u);
}

}  // namespace

void launch_preweight(const FilterLaunch& p, float* out, hipStream_t s)
{
    if (p.count <= 0) return;
    dim3 grid((p.nu + 255) / 256, p.nv, p.count);
    hipLaunchKernelGGL(preweight_kernel, grid, dim3(256), 0, s, p.proj, out, p.views, p.parker, p.nu, p.nv);
    YK_HIP(hipGetLastError());
}

void launch_fdk_filter(const FilterLaunch& p, hipStream_t s)
{
    const int rows = p.count * p.nv;
    // two padded ping-pong buffers (the twiddles stay in global memory: an
    // LDS copy would cost a workgroup per CU)
    const size_t lds = (size_t(2) * (p.n + p.n / 8)) * sizeof(float2);
    if (lds > 160 * 1024) fail(YK_ERR_UNSUPPORTED, "padded row too long for the LDS filter");
    switch (p.storage) {
    case YK_F32: launch_storage<YK_F32>(p, rows, lds, s); break;
    case YK_F16: launch_storage<YK_F16>(p, rows, lds, s); break;
    case YK_BF16: launch_storage<YK_BF16>(p, rows, lds, s); break;
    default: fail(YK_ERR_INVALID_CONFIG, "unknown storage type");
    }
    YK_HIP(hipGetLastError());
}

}  // namespace yk

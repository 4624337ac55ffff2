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
// (fft_twiddles): each lane keeps those of its butterflies in registers
// (spans of 64 and above) or reads them from a small LDS copy (span 8).
// Equivalent to the reference's pad -> cuFFT R2C -> pointwise * w -> C2R ->
// crop chain
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

// LDS index of the runtime-length transform: one pad slot per 8 complex
// values (the radix-8 passes write runs of 8 consecutive values per lane,
// which the padding spreads over banks)
__device__ __forceinline__ int pidx(int i) { return i + (i >> 3); }

// LDS index of the compile-time transforms: C pad slots per 32 complex
// values.  A half-wave's 32 lanes then read any 32 consecutive values of an
// aligned block (every pass reads so) from distinct banks, and each pass
// writes its output in the layout that spreads its own write pattern: C = 4
// after the span-8 pass (runs of 8 values at a 64-value stride), C = 1 after
// every other pass (runs of 8 single values at an 8-value stride, or
// contiguous runs).  Both layouts fit in n + n/8 slots.
template <int C>
__device__ __forceinline__ int slot(int i) { return i + C * (i >> 5); }

// pad slots of the layout a pass of span ns writes, and of the one it reads
constexpr int fft_out_pad(int ns) { return ns == 8 ? 4 : 1; }
constexpr int fft_in_pad(int ns) { return ns == 64 ? 4 : 1; }

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
// workgroup size, every length of 512 and above): the pass sequence, spans
// and lane loops are constants, so the index arithmetic folds to shifts and
// masks, and the first pass (span 1, all twiddles 1) multiplies by nothing.
// Every lane keeps the twiddles of its butterflies (and the filter weights
// of its last-pass bins) in registers for all the row pairs it filters: the
// same entries of the pass-ordered table, so the operations on the data are
// those of fft_forward.
constexpr int fft_radix(int n, int ns) { return ns * 8 <= n ? 8 : (ns * 4 <= n ? 4 : 2); }

// butterflies per lane of a pass of radix r
constexpr int fft_iters(int n, int r) { return (n / r + kFilterThreads - 1) / kFilterThreads; }

// The span-8 pass has 8 distinct twiddle sets (k = lane mod 8): it reads
// them from a 56-entry LDS copy (8 distinct addresses per read: no bank
// conflicts), which spares 14 registers.  Passes of span 64 and above keep
// theirs in registers.
constexpr int kTwLdsSpan = 8;
constexpr bool fft_tw_in_regs(int ns) { return ns > kTwLdsSpan; }

// registers of a pass's twiddles
constexpr int fft_pass_tw_regs(int n, int ns)
{
    return fft_tw_in_regs(ns) ? fft_iters(n, fft_radix(n, ns)) * (fft_radix(n, ns) - 1) : 0;
}

// twiddle registers of the passes from span ns on
constexpr int fft_tw_regs(int n, int ns)
{
    int total = 0;
    while (ns < n) {
        total += fft_pass_tw_regs(n, ns);
        ns *= fft_radix(n, ns);
    }
    return total;
}

// span of the last pass
constexpr int fft_last_ns(int n)
{
    int ns = 1;
    while (ns * fft_radix(n, ns) < n) ns *= fft_radix(n, ns);
    return ns;
}

// this lane's twiddles of the passes from span NS on (tw: that pass's block)
template <int N, int NS>
__device__ __forceinline__ void load_twiddles(float2* reg, const float2* __restrict__ tw)
{
    if constexpr (NS < N) {
        constexpr int R = fft_radix(N, NS), M = N / R, IT = fft_iters(N, R);
        if constexpr (fft_tw_in_regs(NS)) {
#pragma unroll
            for (int it = 0; it < IT; ++it) {
                const int j = it * kFilterThreads + int(threadIdx.x);
                const int k = j & (NS - 1);
                const bool live = M % kFilterThreads == 0 || j < M;
#pragma unroll
                for (int r = 1; r < R; ++r)
                    reg[it * (R - 1) + r - 1] = live ? tw[(r - 1) * NS + k] : make_float2(0.f, 0.f);
            }
        }
        load_twiddles<N, NS * R>(reg + fft_pass_tw_regs(N, NS), tw + (R - 1) * NS);
    }
}

// this lane's filter weights of the last pass: bin idx + r NS of each of its
// butterflies, w[b] for b <= N/2 and w[N - b] above (the response is even)
template <int N>
__device__ __forceinline__ void load_spectral_weights(float* reg, const float* __restrict__ weights)
{
    constexpr int NS = fft_last_ns(N), R = fft_radix(N, NS), M = N / R, IT = fft_iters(N, R);
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int j = it * kFilterThreads + int(threadIdx.x);
        const int k = j & (NS - 1);
        const int idx = (j - k) * R + k;
        const bool live = M % kFilterThreads == 0 || j < M;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int b = idx + r * NS;
            reg[it * R + r] = live ? weights[b <= N / 2 ? b : N - b] : 0.f;
        }
    }
}

template <int R, int N, int NS, bool SPECTRAL = false>
__device__ __forceinline__ void stockham_pass_ct(const float2* src, float2* dst, const float2* twr,
    const float* wr, const float2* tw_lds)
{
    constexpr int m = N / R;
#pragma unroll
    for (int it = 0; it < fft_iters(N, R); ++it) {
        const int j = it * kFilterThreads + int(threadIdx.x);
        if (m % kFilterThreads != 0 && j >= m) break;
        const int k = j & (NS - 1);
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = src[slot<fft_in_pad(NS)>(j + r * m)];
        if constexpr (fft_tw_in_regs(NS)) {
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twr[it * (R - 1) + r - 1]);
        } else if constexpr (NS > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw_lds[(r - 1) * NS + k]);
        }
        dft<R>(v);
        const int idx = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if constexpr (SPECTRAL) {
                // last pass: bin idx + r NS in natural order; Y = conj(X W)
                const float w = wr[it * R + r];
                v[r] = make_float2(v[r].x * w, -v[r].y * w);
            }
            dst[slot<fft_out_pad(NS)>(idx + r * NS)] = v[r];
        }
    }
}

// SPECTRAL: the last pass also applies the filter weights and conjugates.
// twr: the twiddle registers of the passes of span NS and after.
template <int N, int NS, bool SPECTRAL>
__device__ __forceinline__ float2* fft_forward_ct(float2* src, float2* dst, const float2* twr, const float* wr,
    const float2* tw_lds)
{
    if constexpr (NS >= N) {
        (void)dst; (void)twr; (void)wr; (void)tw_lds;
        return src;
    } else {
        constexpr int R = fft_radix(N, NS);
        stockham_pass_ct<R, N, NS, SPECTRAL && NS * R >= N>(src, dst, twr, wr, tw_lds);
        __syncthreads();
        return fft_forward_ct<N, NS * R, SPECTRAL>(dst, src, twr + fft_pass_tw_regs(N, NS), wr, tw_lds);
    }
}

// Cosine weight SDD_r / |q| * SDD_r / (2 SOD) of one detector row, q =
// detS + u detU + v detV - src (.../FDK/kernels/YkFDKPreWeightKernels.cu:374,401-402),
// times Parker: |q|^2 = |b|^2 + u (2 b.detU + u |detU|^2) with b the
// row's u = 0 offset, a quadratic in u set up once per row.
struct RowWeight {
    float a, b, c, k;
    const float* parker;
};

__device__ __forceinline__ RowWeight row_weight_setup(const FilterView& fv, const float* __restrict__ parker_row,
    int v)
{
    const float bx = fv.detS[0] + fv.detV[0] * v - fv.src[0];
    const float by = fv.detS[1] + fv.detV[1] * v - fv.src[1];
    const float bz = fv.detS[2] + fv.detV[2] * v - fv.src[2];
    RowWeight w;
    w.a = bx * bx + by * by + bz * bz;
    w.b = 2.f * (bx * fv.detU[0] + by * fv.detU[1] + bz * fv.detU[2]);
    w.c = fv.detU[0] * fv.detU[0] + fv.detU[1] * fv.detU[1] + fv.detU[2] * fv.detU[2];
    w.k = fv.sdd_radial * fv.wgeom;
    w.parker = parker_row;
    return w;
}

__device__ __forceinline__ float row_weight(const RowWeight& w, int u)
{
    const float uf = float(u);
    const float r2 = fmaxf(fmaf(uf, fmaf(uf, w.c, w.b), w.a), 1e-20f);
    float x = w.k * rsqrtf(r2);
    if (w.parker) x *= w.parker[u];
    return x;
}

// wave maxima of a and b (both rows of a pair at once), stored to red[wave]
// for the reads after the next workgroup barrier (wave_max2_read)
__device__ __forceinline__ void wave_max2_store(float a, float b, float2* red)
{
    for (int off = 32; off > 0; off >>= 1) {
        a = fmaxf(a, __shfl_xor(a, off, 64));
        b = fmaxf(b, __shfl_xor(b, off, 64));
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = make_float2(a, b);
}

__device__ __forceinline__ float2 wave_max2_read(const float2* red)
{
    float2 m = red[0];
    for (int w = 1; w < kFilterThreads / 64; ++w) m = make_float2(fmaxf(m.x, red[w].x), fmaxf(m.y, red[w].y));
    return m;
}

// power-of-two scale s = 2^e with bound / s in (2^13, 2^14], where bound
// >= max|row| (the largest weighted input texel times the filter's response
// L1 norm); e = 0 for an all-zero (or non-finite) row.  Returns e: s and 1 / s
// are exact powers of two.  Half and bfloat16 keep their relative precision
// at any such scale, so a loose bound only moves the row's exponent range.
__device__ __forceinline__ int row_scale_exp(float maxabs)
{
    if (!(maxabs > 0.f) || !isfinite(maxabs)) return 0;
    int e;
    const float m = frexpf(maxabs, &e);  // maxabs = m 2^e, m in [0.5, 1)
    (void)m;
    return e - 14;
}

// LDS index of the rows entering and leaving the transforms: the layout the
// first pass reads and the last pass writes (pad 1 per 32 for compile-time
// lengths; every pass sequence of n >= 512 ends on a span other than 8)
template <int NPT>
__device__ __forceinline__ int io_slot(int i) { return NPT > 1 ? slot<1>(i) : pidx(i); }

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
// row) are loaded into registers while the current pair is stored.  Lengths
// of 512 and above run at most 128 registers per lane (4 workgroups per CU,
// the LDS limit at n = 2048); 4096 at 2 workgroups per CU.
template <int STORAGE, int NPT>
__global__ void __launch_bounds__(kFilterThreads, NPT >= 16 ? 2 : 4)
fdk_filter_kernel(const float* __restrict__ proj, void* __restrict__ out,
    float* __restrict__ row_scales, const FilterView* __restrict__ views,
    const float* __restrict__ parker, const float* __restrict__ weights,
    const float2* __restrict__ tw, int nu, int nv, int rows, int n, float response_l1)
{
    extern __shared__ float4 smem_raw[];
    const int padded = pidx(n);
    float2* bufa = reinterpret_cast<float2*>(smem_raw);
    float2* bufb = bufa + padded;
    __shared__ float2 red[kFilterThreads / 64];
    // compile-time lengths: this lane's twiddles and last-pass filter
    // weights, loaded once (the pass-ordered table and the weights)
    constexpr int NC = NPT > 1 ? NPT * kFilterThreads : 2;
    float2 twr[NPT > 1 ? fft_tw_regs(NC, 1) : 1];
    float wr[NPT > 1 ? fft_iters(NC, fft_radix(NC, fft_last_ns(NC))) * fft_radix(NC, fft_last_ns(NC)) : 1];
    __shared__ float2 tw_lds[7 * kTwLdsSpan];
    if constexpr (NPT > 1) {
        load_twiddles<NC, 1>(twr, tw);
        load_spectral_weights<NC>(wr, weights);
        // the span-8 block of the pass-ordered table (after the 7 entries of span 1)
        if (threadIdx.x < 7 * kTwLdsSpan) tw_lds[threadIdx.x] = tw[7 + threadIdx.x];
        __syncthreads();
    }

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
        const RowWeight wta = row_weight_setup(fa, parker ? parker + (size_t)A.view * nu : nullptr, A.v);
        const RowWeight wtb = row_weight_setup(fb, parker ? parker + (size_t)B.view * nu : nullptr, B.v);
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
                const float wa = (ua >= 0 && ua < nu) ? row_weight(wta, ua) : 0.f;
                const float wb = (has_b && ub >= 0 && ub < nu) ? row_weight(wtb, ub) : 0.f;
                wv[e] = make_float2(xa[e] * wa, xb[e] * wb);
            }
            if constexpr (STORAGE != YK_F32) {
                float ma = 0.f, mb = 0.f;
#pragma unroll
                for (int e = 0; e < NPT; ++e) {
                    ma = fmaxf(ma, fabsf(wv[e].x));
                    mb = fmaxf(mb, fabsf(wv[e].y));
                }
                wave_max2_store(ma, mb, red);
            }
#pragma unroll
            for (int t = 0; t < S; ++t) {
                float2 v[8];
#pragma unroll
                for (int r = 0; r < 8; ++r) v[r] = wv[t + r * S];
                dft<8>(v);
                const int j = t * kFilterThreads + int(threadIdx.x);
#pragma unroll
                for (int r = 0; r < 8; ++r) bufb[slot<fft_out_pad(1)>(j * 8 + r)] = v[r];
            }
            __syncthreads();
            __builtin_amdgcn_s_setprio(0);
            spec = fft_forward_ct<N, 8, true>(bufb, bufa, twr, wr, tw_lds);
        } else {
            // cosine (and Parker) weights on the way into LDS
            float ma = 0.f, mb = 0.f;
#pragma unroll
            for (int e = 0; e < NPT; ++e) {
                const int m = int(threadIdx.x) + e * kFilterThreads;
                if (m < n) {
                    const int ua = m - fa.start_u, ub = m - fb.start_u;
                    const float wa = (ua >= 0 && ua < nu) ? row_weight(wta, ua) : 0.f;
                    const float wb = (has_b && ub >= 0 && ub < nu) ? row_weight(wtb, ub) : 0.f;
                    const float2 x = make_float2(xa[e] * wa, xb[e] * wb);
                    ma = fmaxf(ma, fabsf(x.x));
                    mb = fmaxf(mb, fabsf(x.y));
                    bufa[io_slot<NPT>(m)] = x;
                }
            }
            if constexpr (STORAGE != YK_F32) wave_max2_store(ma, mb, red);
            __syncthreads();
            __builtin_amdgcn_s_setprio(0);
            if constexpr (NPT > 1) {
                spec = fft_forward_ct<NC, 1, true>(bufa, bufb, twr, wr, tw_lds);
            } else {
                spec = fft_forward(bufa, bufb, tw, n);
                for (int k = threadIdx.x; k < n; k += kFilterThreads) {
                    const float w = weights[k <= n / 2 ? k : n - k];
                    const float2 x = spec[pidx(k)];
                    spec[pidx(k)] = make_float2(x.x * w, -x.y * w);
                }
                __syncthreads();
            }
        }
        float2* other = spec == bufa ? bufb : bufa;
        float2* res;
        if constexpr (NPT > 1) res = fft_forward_ct<NC, 1, false>(spec, other, twr, nullptr, tw_lds);
        else res = fft_forward(spec, other, tw, n);

        // crop: row a = Re(res), row b = -Im(res)
        float inv_a = 1.f, inv_b = 1.f;
        if constexpr (STORAGE != YK_F32) {
            // red holds the wave maxima of the weighted inputs, stored before
            // the first barrier of this pair's transforms
            const float2 mx = wave_max2_read(red);
            const int ea = row_scale_exp(mx.x * response_l1), eb = row_scale_exp(mx.y * response_l1);
            const float sa = ldexpf(1.f, ea), sb = ldexpf(1.f, eb);
            inv_a = ldexpf(1.f, -ea);
            inv_b = ldexpf(1.f, -eb);
            if (threadIdx.x == 0) {
                row_scales[ra] = sa;
                if (has_b) row_scales[rb] = sb;
            }
        }
        // the next pair's texels load while this pair is stored (issued
        // here, not before the transforms, so that they hold no registers
        // through the passes and the row maxima)
        // From here (the next pair's loads, this pair's crop and store, the
        // barrier) through the next pair's weighting and first pass, the
        // waves run at a raised priority; the transforms' LDS passes run at
        // the normal one.  The scheduler then favours the workgroups between
        // transforms over those inside one: -6 % for the FP16 output at
        // n = 2048, -7 % for the FP32 output at n = 4096.
        __builtin_amdgcn_s_setprio(3);
        if (p + int(gridDim.x) < pairs) prefetch(p + gridDim.x);
        const size_t base_a = (size_t)ra * nu, base_b = (size_t)rb * nu;
        for (int a = threadIdx.x; a < nu; a += kFilterThreads) {
            const int ja = fa.start_u + a;
            store_row<STORAGE>(out, base_a, a, (ja >= 0 && ja < n) ? res[io_slot<NPT>(ja)].x : 0.f, inv_a);
            if (has_b) {
                const int jb = fb.start_u + a;
                store_row<STORAGE>(out, base_b, a, (jb >= 0 && jb < n) ? -res[io_slot<NPT>(jb)].y : 0.f, inv_b);
            }
        }
        // the next pair overwrites bufa, which the crop (res) may be
        __syncthreads();
    }
    __builtin_amdgcn_s_setprio(0);
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
        p.parker, p.weights, p.twiddles, p.nu, p.nv, rows, p.n, p.response_l1);
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
    out[at] = proj[at] * row_weight(row_weight_setup(views[i], parker ? parker + (size_t)i * nu : nullptr, v), u);
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
    // two padded ping-pong buffers (the twiddles live in registers: a full
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

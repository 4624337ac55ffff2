// Flat-panel forward projectors and their exact transposes for gfx950.
//
// Joseph: per ray, sum over every slice of the per-view main axis of the
// bilinear (minor-axis) sample of the volume at the ray's crossing, times
// sqrt(v0^2 + (a1 v1)^2 + (a2 v2)^2) (fp_joseph_kernel,
// src/FlatFpBp/FP/kernels/YkFPKernels.cu:10-202; axis conventions
// src/common/cuda/operators/YkOperatorKernelTypes.cuh:133-275).  The volume is
// read with software bilinear in place of a 3-D texture; for X-main-axis
// views a [z][x][y] copy keeps the minor-axis gathers contiguous.
//
// Siddon: exact voxel walk from the source to the pixel centre clipped to
// t in [0, 1] and the volume box (siddon_fp_kernel,
// src/FlatFpBp/FP/kernels/YkFPSiddonKernel.cu:19-124); voxel boundaries are
// recomputed from integer indices instead of accumulated.
//
// Transposes are deterministic gathers (no atomics): for each voxel and view
// the candidate rays are the detector pixels inside the projection of the
// voxel's support, and each candidate ray is re-evaluated with exactly the
// arithmetic of the forward kernel (shared functions below, FP contraction
// off), so <A x, y> = <x, A^T y> up to summation order.
#include <hip/hip_runtime.h>

#include "yk_internal.hpp"

namespace yk {
namespace {

struct AxisInfo {
    int p0, p1, p2;     // component index of c0, c1, c2
    int n0, n1, n2;
    float v0, v1, v2;
};

__device__ __forceinline__ AxisInfo axis_info(int axis, const VolumeGeom& g)
{
    AxisInfo a;
    if (axis == 0) { a.p0 = 0; a.p1 = 1; a.p2 = 2; a.n0 = g.nx; a.n1 = g.ny; a.n2 = g.nz; a.v0 = g.vx; a.v1 = g.vy; a.v2 = g.vz; }
    else if (axis == 1) { a.p0 = 1; a.p1 = 0; a.p2 = 2; a.n0 = g.ny; a.n1 = g.nx; a.n2 = g.nz; a.v0 = g.vy; a.v1 = g.vx; a.v2 = g.vz; }
    else { a.p0 = 2; a.p1 = 0; a.p2 = 1; a.n0 = g.nz; a.n1 = g.nx; a.n2 = g.ny; a.v0 = g.vz; a.v1 = g.vx; a.v2 = g.vy; }
    return a;
}

template <typename T>
__device__ __forceinline__ T pick(int k, T a, T b, T c) { return k == 0 ? a : (k == 1 ? b : c); }

struct JRay {
    float a1, b1, a2, b2, L;
};

// Per-ray Joseph line in volume-centred voxel units.  Shared by the forward
// kernel and the transpose's ray table, with contraction disabled so both
// evaluate the same IEEE operations and obtain bit-identical parameters.
__device__ __forceinline__ bool joseph_ray(const ProjViewDev& pv, const AxisInfo& ax,
    int u, int v, JRay& r)
{
#pragma clang fp contract(off)
    const float dx = (pv.detS[0] + float(u) * pv.detU[0]) + float(v) * pv.detV[0];
    const float dy = (pv.detS[1] + float(u) * pv.detU[1]) + float(v) * pv.detV[1];
    const float dz = (pv.detS[2] + float(u) * pv.detU[2]) + float(v) * pv.detV[2];
    const float sx = pv.src[0], sy = pv.src[1], sz = pv.src[2];
    const float s0 = pick(ax.p0, sx, sy, sz), s1 = pick(ax.p1, sx, sy, sz), s2 = pick(ax.p2, sx, sy, sz);
    const float d0 = pick(ax.p0, dx, dy, dz), d1 = pick(ax.p1, dx, dy, dz), d2 = pick(ax.p2, dx, dy, dz);
    const float den = s0 - d0;
    if (den == 0.f) return false;
    r.a1 = (s1 - d1) / den;
    r.a2 = (s2 - d2) / den;
    r.b1 = s1 - r.a1 * s0;
    r.b2 = s2 - r.a2 * s0;
    const float t1 = r.a1 * ax.v1, t2 = r.a2 * ax.v2;
    r.L = sqrtf(ax.v0 * ax.v0 + t1 * t1 + t2 * t2);
    return true;
}

// centred main-axis coordinate of slice s
__device__ __forceinline__ float joseph_c0(int s, int n0)
{
#pragma clang fp contract(off)
    return float(s) - 0.5f * float(n0 - 1);
}

// one minor-axis sample position (index space, texel k at k) of the line
// c = a c0 + b
__device__ __forceinline__ float joseph_minor(float a, float b, float c0, int n)
{
#pragma clang fp contract(off)
    return (a * c0 + b) + 0.5f * float(n - 1);
}

// minor-axis sample position of slice s
__device__ __forceinline__ void joseph_sample(const JRay& r, const AxisInfo& ax, int s, float& f1, float& f2)
{
    const float c0 = joseph_c0(s, ax.n0);
    f1 = joseph_minor(r.a1, r.b1, c0, ax.n1);
    f2 = joseph_minor(r.a2, r.b2, c0, ax.n2);
}

// slices whose sample can touch the volume: -1 < f < n on both minor axes
__device__ __forceinline__ void joseph_slice_range(const JRay& r, const AxisInfo& ax, int& lo, int& hi)
{
    float slo = 0.f, shi = float(ax.n0 - 1);
    const float h0 = 0.5f * float(ax.n0 - 1);
    const float a[2] = { r.a1, r.a2 };
    const float b[2] = { r.b1 + 0.5f * float(ax.n1 - 1), r.b2 + 0.5f * float(ax.n2 - 1) };
    const float n[2] = { float(ax.n1), float(ax.n2) };
    for (int k = 0; k < 2; ++k) {
        if (fabsf(a[k]) > 1e-12f) {
            float t0 = (-1.f - b[k]) / a[k] + h0;
            float t1 = (n[k] - b[k]) / a[k] + h0;
            if (t0 > t1) { const float t = t0; t0 = t1; t1 = t; }
            slo = fmaxf(slo, t0);
            shi = fminf(shi, t1);
        } else if (b[k] <= -1.f || b[k] >= n[k]) {
            lo = 1; hi = 0;
            return;
        }
    }
    // one slice of margin on each side; out-of-range samples weigh zero anyway
    lo = max(0, int(floorf(slo)) - 1);
    hi = min(ax.n0 - 1, int(ceilf(shi)) + 1);
}

// One ray per lane, 64 x 4 rays per workgroup.  The lanes of a wave march
// the same slice index (the union of their slice ranges, lanes outside their
// own range weigh zero), so a sample costs four unconditional loads at
// clamped in-volume positions, with the off-volume taps zeroed in the weights;
// four samples are issued before any is consumed.
__global__ void __launch_bounds__(256)
joseph_fp_kernel(const float* __restrict__ vol, const float* __restrict__ vol_t,
    float* __restrict__ proj, const ProjViewDev* __restrict__ views, VolumeGeom g,
    int nu, int nv, int accumulate)
{
    const int u = blockIdx.x * 64 + threadIdx.x;
    const int v = blockIdx.y * 4 + threadIdx.y;
    const int i = blockIdx.z;
    const bool pix = u < nu && v < nv;
    const ProjViewDev pv = views[i];
    const AxisInfo ax = axis_info(pv.axis, g);
    // memory strides of (s, i1, i2)
    const float* base;
    size_t st0, st2;
    if (pv.axis == 0) { base = vol_t; st0 = (size_t)g.ny; st2 = (size_t)g.nx * g.ny; }
    else if (pv.axis == 1) { base = vol; st0 = (size_t)g.nx; st2 = (size_t)g.nx * g.ny; }
    else { base = vol; st0 = (size_t)g.nx * g.ny; st2 = (size_t)g.nx; }
    JRay r{};
    int lo = 1, hi = 0;
    if (pix && joseph_ray(pv, ax, u, v, r)) joseph_slice_range(r, ax, lo, hi);
    // wave-wide slice range
    int wlo = lo <= hi ? lo : 0x7fffffff, whi = lo <= hi ? hi : -1;
    for (int off = 32; off > 0; off >>= 1) {
        wlo = min(wlo, __shfl_xor(wlo, off, 64));
        whi = max(whi, __shfl_xor(whi, off, 64));
    }
    const int n1 = ax.n1, n2 = ax.n2;
    float acc = 0.f;
    constexpr int kUnroll = 4;
    for (int s0 = wlo; s0 <= whi; s0 += kUnroll) {
        float t[kUnroll][4], wa[kUnroll], wb[kUnroll], wc[kUnroll], wd[kUnroll];
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) {
            const int s = min(s0 + k, whi);
            float f1, f2;
            joseph_sample(r, ax, s, f1, f2);
            const float g1 = floorf(f1), g2 = floorf(f2);
            const float w1 = f1 - g1, w2 = f2 - g2;
            const bool live = s0 + k <= whi && s >= lo && s <= hi && g1 >= -1.f && g2 >= -1.f &&
                g1 <= float(n1 - 1) && g2 <= float(n2 - 1);
            const int j1 = live ? int(g1) : 0, j2 = live ? int(g2) : 0;
            // taps (j1 + a, j2 + b) inside the volume, the others weigh zero
            const bool a0 = j1 >= 0, a1 = j1 + 1 < n1, b0 = j2 >= 0, b1 = j2 + 1 < n2;
            const float x0 = live && a0 ? 1.f - w1 : 0.f, x1 = live && a1 ? w1 : 0.f;
            const float y0 = b0 ? 1.f - w2 : 0.f, y1 = b1 ? w2 : 0.f;
            wa[k] = x0 * y0; wb[k] = x1 * y0; wc[k] = x0 * y1; wd[k] = x1 * y1;
            const int c1 = max(j1, 0), c1b = min(j1 + 1, n1 - 1);
            const int c2 = max(j2, 0), c2b = min(j2 + 1, n2 - 1);
            const float* sl = base + (size_t)s * st0;
            t[k][0] = sl[(size_t)c2 * st2 + c1];
            t[k][1] = sl[(size_t)c2 * st2 + c1b];
            t[k][2] = sl[(size_t)c2b * st2 + c1];
            t[k][3] = sl[(size_t)c2b * st2 + c1b];
        }
#pragma unroll
        for (int k = 0; k < kUnroll; ++k)
            acc += (wa[k] * t[k][0] + wb[k] * t[k][1]) + (wc[k] * t[k][2] + wd[k] * t[k][3]);
    }
    if (!pix) return;
    const float out = lo <= hi ? acc * r.L : 0.f;
    float* dst = proj + ((size_t)i * nv + v) * nu + u;
    *dst = accumulate ? *dst + out : out;
}

// final store of one Joseph ray: the sum times the ray's step length, or
// zero for a ray that misses the volume
__device__ __forceinline__ void finish_joseph(float* proj, int i, int nu, int nv, int u, int v, bool pix,
    bool hit, float acc, float L, int accumulate)
{
    if (!pix) return;
    const float out = hit ? acc * L : 0.f;
    float* dst = proj + ((size_t)i * nv + v) * nu + u;
    *dst = accumulate ? *dst + out : out;
}

// Paired guarded volume copies of the Joseph projections.  A guarded copy
// has kJosephGuard zero planes on both sides of both minor axes; its paired
// form stores, at entry (slice c, row r, column e), the two guarded values
// of rows r and r + 1 in column e.  One 16-byte load at (c, r, e) then
// returns the four taps (r, e), (r + 1, e), (r, e + 1), (r + 1, e + 1) of a
// bilinear sample: a wave64 load of 8 or 16 bytes per lane takes 16 cycles
// of a CU's vector memory pipe on gfx950, so the one wider load halves the
// cost of two 8-byte row loads.  Rows run 0 .. n_r + 2 guards - 2 (the last
// row pairs with the trailing guard plane).
constexpr int kJosephGuard = 2;
typedef float float4u __attribute__((ext_vector_type(4), aligned(8)));

// Joseph projection for separable views (main axis X or Y, detV along Z
// only).  One wave per detector column (64 consecutive rays in v): a1, b1 and
// hence f1 and its taps are the same for every lane, and the lanes' Z taps are
// neighbouring entries of a paired Z-fastest copy (rows along y, columns
// along z for X-main views; rows along x for Y-main views), so the gathers of
// a sample touch a few cache lines per wave.  With the sample position
// clamped to [-2, n] every tap lies in the copy, and a position outside
// (-1, n) reads only zeros, so the kernel needs no per-sample range test (a
// slice outside a lane's own range is outside (-1, n) on some minor axis).
// A sample is one 16-byte load and three lerps; the first-minor weight and
// tap row come from a per-wave LDS table built 64 slices at a time.  Sample
// positions are those of joseph_fp_kernel (the same IEEE operations) and the
// ray sum keeps its slice order.
template <int CW>
__global__ void __launch_bounds__(512)
joseph_fp_sep_kernel(const float* __restrict__ vzx, const float* __restrict__ vzy,
    float* __restrict__ proj, const ProjViewDev* __restrict__ views, VolumeGeom g,
    int nu, int nv, int count, int accumulate)
{
    // per wave, 64 slices at a time: the first-minor weight and the byte
    // offset of its tap row (the same in every lane: a1, b1 are exact
    // functions of the column u), built once per slice by one lane each and
    // read back per sample by a broadcast LDS read
    __shared__ uint2 tab[8][64];
    const int wave = int(threadIdx.x >> 6), lane = int(threadIdx.x & 63);
    // The eight waves take CW neighbouring detector columns of 8 / CW
    // consecutive views: their rays nearly coincide, so the repeats of a tap
    // row hit the CU's L1 (eight views measured 8 % faster than four at
    // 512^3 x 720, sixteen 7 % slower than eight).  Column groups are remapped so that each XCD
    // (workgroups go round-robin to the 8 XCDs; gridDim.x is a multiple of
    // 8) sweeps a contiguous range of columns, whose neighbouring rays share
    // tap rows in its L2.
    const int group = int(blockIdx.x & 7u) * int(gridDim.x >> 3) + int(blockIdx.x >> 3);
    const int u = group * CW + wave % CW;
    const int v = blockIdx.y * 64 + lane;
    const int i = int(blockIdx.z) * (8 / CW) + wave / CW;
    if (u >= nu || i >= count) return;  // whole wave
    const bool pix = v < nv;
    const ProjViewDev pv = views[i];
    const AxisInfo ax = axis_info(pv.axis, g);
    const float* base = pv.axis == 0 ? vzx : vzy;
    JRay r{};
    int lo = 1, hi = 0;
    if (pix && joseph_ray(pv, ax, u, v, r)) joseph_slice_range(r, ax, lo, hi);
    int wlo = lo <= hi ? lo : 0x7fffffff, whi = lo <= hi ? hi : -1;
    for (int off = 32; off > 0; off >>= 1) {
        wlo = min(wlo, __shfl_xor(wlo, off, 64));
        whi = max(whi, __shfl_xor(whi, off, 64));
    }
    wlo = __builtin_amdgcn_readfirstlane(wlo);
    whi = __builtin_amdgcn_readfirstlane(whi);
    // lane 0 (v = 64 blockIdx.y < nv) holds the column's a1, b1
    const float a1 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(r.a1)));
    const float b1 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(r.b1)));
    const int p2 = ax.n2 + 2 * kJosephGuard;
    const unsigned row_bytes = unsigned(p2) * 8u;
    const size_t slice_bytes = size_t(ax.n1 + 2 * kJosephGuard - 1) * row_bytes;
    const float f1max = float(ax.n1), f2max = float(ax.n2);
    float acc = 0.f;
    uint2* t = tab[wave];
    for (int c = wlo; c <= whi; c += 64) {
        {
            const int st = min(c + lane, whi);
            float f1 = joseph_minor(a1, b1, joseph_c0(st, ax.n0), ax.n1);
            f1 = __builtin_amdgcn_fmed3f(f1, -2.f, f1max);
            const int j1 = int(floorf(f1)) + kJosephGuard;
            t[lane] = make_uint2(__float_as_uint(__builtin_amdgcn_fractf(f1)), unsigned(j1) * row_bytes);
        }
        __builtin_amdgcn_wave_barrier();
        const int cnt = min(64, whi - c + 1);
        // the slice coordinate steps by exactly 1 (half-integers below 2^22);
        // the slice base is wave-uniform
        float c0 = joseph_c0(c, ax.n0);
        const char* sl = reinterpret_cast<const char*>(base) + size_t(c) * slice_bytes;
        int k = 0;
        constexpr int kUnroll = 4;
        for (; k + kUnroll <= cnt; k += kUnroll) {
            float4u q[kUnroll];
            float w1[kUnroll], w2[kUnroll];
#pragma unroll
            for (int j = 0; j < kUnroll; ++j) {
                const uint2 e = t[k + j];
                float f2 = joseph_minor(r.a2, r.b2, c0, ax.n2);
                c0 += 1.f;
                f2 = __builtin_amdgcn_fmed3f(f2, -2.f, f2max);
                w1[j] = __uint_as_float(e.x);
                w2[j] = __builtin_amdgcn_fractf(f2);
                const unsigned o = e.y + unsigned(int(floorf(f2)) + kJosephGuard) * 8u;
                q[j] = *reinterpret_cast<const float4u*>(sl + o);
                sl += slice_bytes;
            }
#pragma unroll
            for (int j = 0; j < kUnroll; ++j) {
                const float a = fmaf(w2[j], q[j].z - q[j].x, q[j].x);
                const float b = fmaf(w2[j], q[j].w - q[j].y, q[j].y);
                acc += fmaf(w1[j], b - a, a);
            }
        }
        for (; k < cnt; ++k) {
            const uint2 e = t[k];
            float f2 = joseph_minor(r.a2, r.b2, c0, ax.n2);
            c0 += 1.f;
            f2 = __builtin_amdgcn_fmed3f(f2, -2.f, f2max);
            const float w2 = __builtin_amdgcn_fractf(f2);
            const unsigned o = e.y + unsigned(int(floorf(f2)) + kJosephGuard) * 8u;
            const float4u q = *reinterpret_cast<const float4u*>(sl + o);
            sl += slice_bytes;
            const float a = fmaf(w2, q.z - q.x, q.x);
            const float b = fmaf(w2, q.w - q.y, q.y);
            acc += fmaf(__uint_as_float(e.x), b - a, a);
        }
        __builtin_amdgcn_wave_barrier();
    }
    finish_joseph(proj, i, nu, nv, u, v, pix, lo <= hi, acc, r.L, accumulate);
}

// Joseph projection for any X- or Y-main view set (tilted detectors, lifted
// sources), one lane per ray as joseph_fp_kernel, over paired slab copies:
// rows along z and columns along the first minor axis (y for X-main, x for
// Y-main views).  A slice of the main axis is one contiguous slab (scalar
// base, 32-bit lane offsets); the lanes of a wave (consecutive u) read
// neighbouring columns.  As in joseph_fp_sep_kernel the sample position is
// clamped to [-2, n], positions outside (-1, n) read only zeros, and a
// sample is one 16-byte load and three lerps.
__global__ void __launch_bounds__(512)
joseph_fp_slab_kernel(const float* __restrict__ gx, const float* __restrict__ gy,
    float* __restrict__ proj, const ProjViewDev* __restrict__ views, VolumeGeom g,
    int nu, int nv, int count, int accumulate)
{
    // the eight waves take one 64-ray row segment of eight consecutive
    // views (nearly coincident rays: the repeats of a sample's cache lines
    // hit the CU's L1; eight measured 7 % faster than four at 512^3 x 720)
    const int u = blockIdx.x * 64 + threadIdx.x;
    const int v = blockIdx.y;
    const int i = blockIdx.z * 8 + threadIdx.y;
    if (i >= count) return;  // whole wave
    const bool pix = u < nu && v < nv;
    const ProjViewDev pv = views[i];
    const AxisInfo ax = axis_info(pv.axis, g);
    const float* base = pv.axis == 0 ? gx : gy;
    JRay r{};
    int lo = 1, hi = 0;
    if (pix && joseph_ray(pv, ax, u, v, r)) joseph_slice_range(r, ax, lo, hi);
    int wlo = lo <= hi ? lo : 0x7fffffff, whi = lo <= hi ? hi : -1;
    for (int off = 32; off > 0; off >>= 1) {
        wlo = min(wlo, __shfl_xor(wlo, off, 64));
        whi = max(whi, __shfl_xor(whi, off, 64));
    }
    wlo = __builtin_amdgcn_readfirstlane(wlo);
    whi = __builtin_amdgcn_readfirstlane(whi);
    const int p1 = ax.n1 + 2 * kJosephGuard;
    const size_t slab_bytes = size_t(ax.n2 + 2 * kJosephGuard - 1) * size_t(p1) * 8u;
    const float f1max = float(ax.n1), f2max = float(ax.n2);
    float acc = 0.f;
    if (wlo > whi) return finish_joseph(proj, i, nu, nv, u, v, pix, lo <= hi, 0.f, r.L, accumulate);
    // the slice coordinate steps by exactly 1 (half-integers below 2^22)
    float c0 = joseph_c0(wlo, ax.n0);
    const char* sl = reinterpret_cast<const char*>(base) + size_t(wlo) * slab_bytes;
    constexpr int kUnroll = 4;
    int s = wlo;
    for (; s + kUnroll - 1 <= whi; s += kUnroll) {
        float4u q[kUnroll];
        float w1[kUnroll], w2[kUnroll];
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) {
            float f1 = joseph_minor(r.a1, r.b1, c0, ax.n1);
            float f2 = joseph_minor(r.a2, r.b2, c0, ax.n2);
            c0 += 1.f;
            f1 = __builtin_amdgcn_fmed3f(f1, -2.f, f1max);
            f2 = __builtin_amdgcn_fmed3f(f2, -2.f, f2max);
            w1[k] = __builtin_amdgcn_fractf(f1);
            w2[k] = __builtin_amdgcn_fractf(f2);
            const int j1 = int(floorf(f1)) + kJosephGuard, j2 = int(floorf(f2)) + kJosephGuard;
            const unsigned o = unsigned(__mul24(j2, p1) + j1) * 8u;
            q[k] = *reinterpret_cast<const float4u*>(sl + o);
            sl += slab_bytes;
        }
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) {
            const float a = fmaf(w1[k], q[k].z - q[k].x, q[k].x);
            const float b = fmaf(w1[k], q[k].w - q[k].y, q[k].y);
            acc += fmaf(w2[k], b - a, a);
        }
    }
    for (; s <= whi; ++s) {
        float f1 = joseph_minor(r.a1, r.b1, c0, ax.n1);
        float f2 = joseph_minor(r.a2, r.b2, c0, ax.n2);
        c0 += 1.f;
        f1 = __builtin_amdgcn_fmed3f(f1, -2.f, f1max);
        f2 = __builtin_amdgcn_fmed3f(f2, -2.f, f2max);
        const float w1 = __builtin_amdgcn_fractf(f1), w2 = __builtin_amdgcn_fractf(f2);
        const int j1 = int(floorf(f1)) + kJosephGuard, j2 = int(floorf(f2)) + kJosephGuard;
        const unsigned o = unsigned(__mul24(j2, p1) + j1) * 8u;
        const float4u q = *reinterpret_cast<const float4u*>(sl + o);
        sl += slab_bytes;
        const float a = fmaf(w1, q.z - q.x, q.x);
        const float b = fmaf(w1, q.w - q.y, q.y);
        acc += fmaf(w2, b - a, a);
    }
    finish_joseph(proj, i, nu, nv, u, v, pix, lo <= hi, acc, r.L, accumulate);
}

// paired separable copy, X-main: [x][row y + guard][column z + guard];
// blockIdx.z is the row, 32 x 32 (x, z) tiles of the two source rows
// through LDS
__global__ void pair_zfast_x_kernel(const float* __restrict__ in, float2* __restrict__ out,
    int nx, int ny, int nz)
{
    __shared__ float tile[2][32][33];
    const int r = blockIdx.z;
    const int x0 = blockIdx.x * 32, z0 = blockIdx.y * 32;
    const size_t rows = size_t(ny) + 2 * kJosephGuard - 1, cols = size_t(nz) + 2 * kJosephGuard;
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int z = z0 + k, x = x0 + threadIdx.x;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int y = r + h - kJosephGuard;
            tile[h][k][threadIdx.x] = x < nx && z < nz && y >= 0 && y < ny ? in[(size_t(z) * ny + y) * nx + x] : 0.f;
        }
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int x = x0 + k, z = z0 + threadIdx.x;
        if (x < nx && z < nz)
            out[(size_t(x) * rows + r) * cols + z + kJosephGuard] = make_float2(tile[0][threadIdx.x][k], tile[1][threadIdx.x][k]);
    }
}

// paired separable copy, Y-main: [y][row x + guard][column z + guard];
// blockIdx.z is y, 32 x 32 (row, z) tiles (33 source columns of x) through
// LDS
__global__ void pair_zfast_y_kernel(const float* __restrict__ in, float2* __restrict__ out,
    int nx, int ny, int nz)
{
    __shared__ float tile[32][34];
    const int y = blockIdx.z;
    const int r0 = blockIdx.x * 32, z0 = blockIdx.y * 32;
    const int rows = nx + 2 * kJosephGuard - 1;
    const size_t cols = size_t(nz) + 2 * kJosephGuard;
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int z = z0 + k;
        for (int t = threadIdx.x; t < 33; t += 32) {
            const int x = r0 + t - kJosephGuard;
            tile[k][t] = z < nz && x >= 0 && x < nx ? in[(size_t(z) * ny + y) * nx + x] : 0.f;
        }
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int r = r0 + k, z = z0 + threadIdx.x;
        if (r < rows && z < nz)
            out[(size_t(y) * rows + r) * cols + z + kJosephGuard] = make_float2(tile[threadIdx.x][k], tile[threadIdx.x][k + 1]);
    }
}

// paired slab copy, X-main: [x][row z + guard][column y + guard];
// blockIdx.z is the row, 32 x 32 (x, y) tiles of the two source planes
// through LDS
__global__ void pair_slab_x_kernel(const float* __restrict__ in, float2* __restrict__ out,
    int nx, int ny, int nz)
{
    __shared__ float tile[2][32][33];
    const int r = blockIdx.z;
    const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
    const size_t rows = size_t(nz) + 2 * kJosephGuard - 1, cols = size_t(ny) + 2 * kJosephGuard;
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int y = y0 + k, x = x0 + threadIdx.x;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int z = r + h - kJosephGuard;
            tile[h][k][threadIdx.x] = x < nx && y < ny && z >= 0 && z < nz ? in[(size_t(z) * ny + y) * nx + x] : 0.f;
        }
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int x = x0 + k, y = y0 + threadIdx.x;
        if (x < nx && y < ny)
            out[(size_t(x) * rows + r) * cols + y + kJosephGuard] = make_float2(tile[0][threadIdx.x][k], tile[1][threadIdx.x][k]);
    }
}

// paired slab copy, Y-main: [y][row z + guard][column x + guard] (x stays
// fastest: row copies)
__global__ void pair_slab_y_kernel(const float* __restrict__ in, float2* __restrict__ out,
    int nx, int ny, int nz)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y, r = blockIdx.z;
    if (x >= nx) return;
    const size_t rows = size_t(nz) + 2 * kJosephGuard - 1, cols = size_t(nx) + 2 * kJosephGuard;
    const int z = r - kJosephGuard;
    const float a = z >= 0 && z < nz ? in[(size_t(z) * ny + y) * nx + x] : 0.f;
    const float b = z + 1 >= 0 && z + 1 < nz ? in[(size_t(z + 1) * ny + y) * nx + x] : 0.f;
    out[(size_t(y) * rows + r) * cols + x + kJosephGuard] = make_float2(a, b);
}

// [z][y][x] -> [z][x][y]
__global__ void transpose_xy_kernel(const float* __restrict__ in, float* __restrict__ out,
    int nx, int ny)
{
    __shared__ float tile[32][33];
    const int z = blockIdx.z;
    const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
    const size_t plane = (size_t)nx * ny;
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int y = y0 + k, x = x0 + threadIdx.x;
        if (x < nx && y < ny) tile[k][threadIdx.x] = in[z * plane + (size_t)y * nx + x];
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int x = x0 + k, y = y0 + threadIdx.x;
        if (x < nx && y < ny) out[z * plane + (size_t)x * ny + y] = tile[threadIdx.x][k];
    }
}

// Ray table of the Joseph transpose: the parameters of every ray of `count`
// views, from the same joseph_ray routine the forward kernel calls (so the
// transpose reuses bit-identical lines); L = 0 marks a ray the forward skips.
__global__ void __launch_bounds__(256)
joseph_ray_table_kernel(const ProjViewDev* __restrict__ views, VolumeGeom g, int nu, int nv,
    float4* __restrict__ ab, float* __restrict__ len)
{
    const int u = blockIdx.x * 64 + threadIdx.x;
    const int v = blockIdx.y * 4 + threadIdx.y;
    const int i = blockIdx.z;
    if (u >= nu || v >= nv) return;
    const ProjViewDev pv = views[i];
    const AxisInfo ax = axis_info(pv.axis, g);
    JRay r;
    const size_t k = ((size_t)i * nv + v) * nu + u;
    if (joseph_ray(pv, ax, u, v, r)) {
        ab[k] = make_float4(r.a1, r.b1, r.a2, r.b2);
        len[k] = r.L;
    } else {
        ab[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        len[k] = 0.f;
    }
}

// Joseph transpose: one thread per voxel, gather over views and candidate
// rays.  The candidates of a voxel are the pixels inside the projection of its
// bilinear support on its main-axis slice (FP32 projective map, 0.05 px of
// margin against rounding); each is then tested with the ray's own sample
// position, exactly as the forward kernel computes it.
__global__ void __launch_bounds__(256)
joseph_adjoint_kernel(const float* __restrict__ proj, float* __restrict__ volume,
    const ProjViewDev* __restrict__ views, const float* __restrict__ c32,
    const float4* __restrict__ ray_ab, const float* __restrict__ ray_len, VolumeGeom g,
    int nu, int nv, int count, int accumulate)
{
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny) return;
    const float margin = 0.05f;
    float acc = 0.f;
    for (int i = 0; i < count; ++i) {
        const ProjViewDev pv = views[i];
        const AxisInfo ax = axis_info(pv.axis, g);
        const int s = pick(ax.p0, x, y, z), i1 = pick(ax.p1, x, y, z), i2 = pick(ax.p2, x, y, z);
        const float* c = c32 + (size_t)i * 12;
        // world coordinates of the main-axis plane and of the minor neighbours
        const float m0 = pick(ax.p0, g.ox, g.oy, g.oz) + float(s) * pick(ax.p0, g.vx, g.vy, g.vz);
        const float o1 = pick(ax.p1, g.ox, g.oy, g.oz), w1 = pick(ax.p1, g.vx, g.vy, g.vz);
        const float o2 = pick(ax.p2, g.ox, g.oy, g.oz), w2 = pick(ax.p2, g.vx, g.vy, g.vz);
        float umin = 1e30f, umax = -1e30f, vmin = 1e30f, vmax = -1e30f;
        bool ok = true;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float m1 = o1 + float(i1 + ((k & 1) ? 1 : -1)) * w1;
            const float m2 = o2 + float(i2 + ((k & 2) ? 1 : -1)) * w2;
            float p[3];
            p[0] = ax.p0 == 0 ? m0 : (ax.p1 == 0 ? m1 : m2);
            p[1] = ax.p0 == 1 ? m0 : (ax.p1 == 1 ? m1 : m2);
            p[2] = ax.p0 == 2 ? m0 : (ax.p1 == 2 ? m1 : m2);
            const float den = c[8] * p[0] + c[9] * p[1] + c[10] * p[2] + c[11];
            if (!(den > 1e-6f)) { ok = false; break; }
            const float rd = 1.f / den;
            const float uu = (c[0] * p[0] + c[1] * p[1] + c[2] * p[2] + c[3]) * rd;
            const float vv = (c[4] * p[0] + c[5] * p[1] + c[6] * p[2] + c[7]) * rd;
            umin = fminf(umin, uu); umax = fmaxf(umax, uu);
            vmin = fminf(vmin, vv); vmax = fmaxf(vmax, vv);
        }
        if (!ok) continue;
        const int ulo = max(0, int(ceilf(umin - margin)));
        const int uhi = min(nu - 1, int(floorf(umax + margin)));
        const int vlo = max(0, int(ceilf(vmin - margin)));
        const int vhi = min(nv - 1, int(floorf(vmax + margin)));
        const size_t vb = (size_t)i * nu * nv;
        const float* pi = proj + vb;
        const float4* abi = ray_ab + vb;
        const float* li = ray_len + vb;
        for (int vv = vlo; vv <= vhi; ++vv)
            for (int uu = ulo; uu <= uhi; ++uu) {
                const size_t k = (size_t)vv * nu + uu;
                // The candidate test is branch-free, so the three loads issue
                // at once (a skip after the first load made them wait in
                // turn); a rejected candidate leaves the sum as it is.
                const float L = li[k];
                const float4 a = abi[k];
                const float py = pi[k];
                JRay r;
                r.a1 = a.x; r.b1 = a.y; r.a2 = a.z; r.b2 = a.w; r.L = L;
                float f1, f2;
                joseph_sample(r, ax, s, f1, f2);
                const float g1 = floorf(f1), g2 = floorf(f2);
                const float w1 = f1 - g1, w2 = f2 - g2;
                // the voxel is tap j or j + 1 of the ray's sample on each axis
                const bool a0 = float(i1) == g1, a1 = float(i1) == g1 + 1.f;
                const bool b0 = float(i2) == g2, b1 = float(i2) == g2 + 1.f;
                const float wa = a0 ? 1.f - w1 : w1, wb = b0 ? 1.f - w2 : w2;
                const float t = fmaf(py * L, wa * wb, acc);
                acc = (L != 0.f && (a0 || a1) && (b0 || b1)) ? t : acc;
            }
    }
    float* dst = volume + ((size_t)z * g.ny + y) * g.nx + x;
    *dst = accumulate ? *dst + acc : acc;
}

// ---------------------------------------------------------------------------
// Joseph transpose for separable views: main axis X or Y and detector columns
// parallel to Z (detV has no component on the main or first minor axis, as
// for an untilted circular scan).  Then d0 and d1 of a ray do not depend on v,
// so a1, b1 - and the first-minor sample f1 of every slice - are those of the
// whole detector column u, while f2 is affine in v.
// ---------------------------------------------------------------------------

// Column-major ray table of the separable transpose: ab1[view][u] = (a1, b1)
// of column u; ab2[view][u][v] = (a2, b2) and pl[view][u][v] = p * L of ray
// (u, v) (0 for a ray the forward projector skips), each column padded to
// nvp = nv + kSepPad entries with zeros.  Same joseph_ray as the
// forward kernel; the 32 x 32 tile is transposed through LDS so both the
// projection reads and the table writes are contiguous.
__global__ void __launch_bounds__(256)
joseph_sep_table_kernel(const float* __restrict__ proj, const ProjViewDev* __restrict__ views,
    VolumeGeom g, int nu, int nv, int nvp, float2* __restrict__ ab1, float2* __restrict__ ab2,
    float* __restrict__ pl)
{
    __shared__ float2 s_ab[32][33];
    __shared__ float s_pl[32][33];
    const int i = blockIdx.z;
    const int u0 = blockIdx.x * 32, v0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const ProjViewDev pv = views[i];
    const AxisInfo ax = axis_info(pv.axis, g);
    const float* pi = proj + (size_t)i * nu * nv;
    for (int k = ty; k < 32; k += 8) {
        const int u = u0 + tx, v = v0 + k;
        float2 ab = make_float2(0.f, 0.f);
        float w = 0.f;
        if (u < nu && v < nv) {
            JRay r;
            const bool ok = joseph_ray(pv, ax, u, v, r);
            if (ok) {
                ab = make_float2(r.a2, r.b2);
                w = pi[(size_t)v * nu + u] * r.L;
            }
            if (v == 0) ab1[(size_t)i * nu + u] = ok ? make_float2(r.a1, r.b1) : make_float2(0.f, 0.f);
        }
        s_ab[k][tx] = ab;
        s_pl[k][tx] = w;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int u = u0 + k, v = v0 + tx;
        if (u < nu && v < nvp) {
            const size_t o = ((size_t)i * nu + u) * nvp + v;
            ab2[o] = s_ab[tx][k];
            pl[o] = s_pl[tx][k];
        }
    }
}

// world projective map (u or v numerator row r, denominator row d) of a point
__device__ __forceinline__ float proj_coord(const float* c, int row, float x, float y, float z, float rd)
{
    return (c[row] * x + c[row + 1] * y + c[row + 2] * z + c[row + 3]) * rd;
}

// One lane per (x, y) column of kSepZ voxels, 16 x 16 columns per workgroup.
// Per view the column sits on one main-axis slice s and one first-minor index
// i1, so its candidate rays are whole detector columns u with
// floor(f1(u)) in {i1 - 1, i1}, and in each such column the rays whose f2
// lies in [z0 - 1, z0 + kSepZ).  Those are walked in v order; every ray adds
// p L wa (1 - w2) and p L wa w2 to the two voxels its sample straddles.  The
// sums live in LDS slots private to the lane (slot-major, lane-minor: no
// bank conflicts, no sharing: plain read-modify-write, no atomics), so the
// order of additions per voxel is fixed and the result deterministic.  Candidate ranges come from the FP32
// projective map with 0.05 px of margin; membership is decided only by the
// forward kernel's own f1, f2.
constexpr int kSepZ = 32;
constexpr int kSepU = kSepPad;  // rays of a column loaded before any is consumed

__global__ void __launch_bounds__(256)
joseph_adjoint_sep_kernel(const float2* __restrict__ ab1, const float2* __restrict__ ab2,
    const float* __restrict__ pl, const ProjViewDev* __restrict__ views, const float* __restrict__ c32,
    VolumeGeom g, int nu, int nv, int nvp, int count, int accumulate, float* __restrict__ volume)
{
    __shared__ float slot[(kSepZ + 2) * 256];
    const int lane = threadIdx.x;
    const int x = blockIdx.x * 16 + (lane & 15);
    const int y = blockIdx.y * 16 + (lane >> 4);
    const int z0 = blockIdx.z * kSepZ;
#pragma unroll
    for (int k = 0; k < kSepZ + 2; ++k) slot[k * 256 + lane] = 0.f;
    if (x >= g.nx || y >= g.ny) return;
    const float margin = 0.05f;
    const float zlo = g.oz + float(z0 - 1) * g.vz;
    const float zhi = g.oz + float(z0 + kSepZ) * g.vz;
    const float zmid = g.oz + (float(z0) + 0.5f * float(kSepZ - 1)) * g.vz;
    for (int i = 0; i < count; ++i) {
        const ProjViewDev pv = views[i];
        const AxisInfo ax = axis_info(pv.axis, g);
        const bool xmain = pv.axis == 0;
        const int s = xmain ? x : y, i1 = xmain ? y : x;
        const float c0 = joseph_c0(s, ax.n0);
        const float* c = c32 + (size_t)i * 12;
        const float m0 = (xmain ? g.ox : g.oy) + float(s) * ax.v0;
        const float o1 = xmain ? g.oy : g.ox;
        // detector columns whose rays cross slice s within (i1 - 1, i1 + 1)
        float ua, ub;
        {
            const float ma = o1 + float(i1 - 1) * ax.v1, mb = o1 + float(i1 + 1) * ax.v1;
            const float xa = xmain ? m0 : ma, ya = xmain ? ma : m0;
            const float xb = xmain ? m0 : mb, yb = xmain ? mb : m0;
            const float da = c[8] * xa + c[9] * ya + c[10] * zmid + c[11];
            const float db = c[8] * xb + c[9] * yb + c[10] * zmid + c[11];
            if (!(da > 1e-6f) || !(db > 1e-6f)) continue;
            ua = proj_coord(c, 0, xa, ya, zmid, 1.f / da);
            ub = proj_coord(c, 0, xb, yb, zmid, 1.f / db);
        }
        const int ulo = max(0, int(ceilf(fminf(ua, ub) - margin)));
        const int uhi = min(nu - 1, int(floorf(fmaxf(ua, ub) + margin)));
        for (int u = ulo; u <= uhi; ++u) {
            const float2 q1 = ab1[(size_t)i * nu + u];
            const float f1 = joseph_minor(q1.x, q1.y, c0, ax.n1);
            const float g1 = floorf(f1);
            if (g1 < -1.f || g1 > float(ax.n1 - 1)) continue;
            const int j1 = int(g1);
            const float w1 = f1 - g1;
            float wa;
            if (j1 == i1) wa = 1.f - w1;
            else if (j1 + 1 == i1) wa = w1;
            else continue;
            // rays of column u whose f2 falls in [z0 - 1, z0 + kSepZ)
            const float m1 = o1 + f1 * ax.v1;
            const float xc = xmain ? m0 : m1, yc = xmain ? m1 : m0;
            const float dl = c[8] * xc + c[9] * yc + c[10] * zlo + c[11];
            const float dh = c[8] * xc + c[9] * yc + c[10] * zhi + c[11];
            if (!(dl > 1e-6f) || !(dh > 1e-6f)) continue;
            const float va = proj_coord(c, 4, xc, yc, zlo, 1.f / dl);
            const float vb = proj_coord(c, 4, xc, yc, zhi, 1.f / dh);
            const int vlo = max(0, int(ceilf(fminf(va, vb) - margin)));
            const int vhi = min(nv - 1, int(floorf(fmaxf(va, vb) + margin)));
            // kSepU rays per step, loads first; the column carries kSepU
            // zero-weight pad entries, so a step may run past vhi (rays
            // beyond the candidate range fall outside the slots by f2)
            const float2* q2 = ab2 + ((size_t)i * nu + u) * nvp + vlo;
            const float* w = pl + ((size_t)i * nu + u) * nvp + vlo;
            const int nr = vhi - vlo + 1;
            for (int r0 = 0; r0 < nr; r0 += kSepU) {
                float2 q[kSepU];
                float pw[kSepU];
#pragma unroll
                for (int j = 0; j < kSepU; ++j) {
                    q[j] = q2[r0 + j];
                    pw[j] = w[r0 + j];
                }
#pragma unroll
                for (int j = 0; j < kSepU; ++j) {
                    const float f2 = joseph_minor(q[j].x, q[j].y, c0, ax.n2);
                    const float g2 = floorf(f2);
                    // slot of voxel floor(f2): 0 .. kSepZ covers f2 in
                    // [z0 - 1, z0 + kSepZ); slots of voxels outside the
                    // volume are never stored, which stands in for the
                    // forward kernel's g2 range test
                    const int k = int(g2) - (z0 - 1);
                    const bool live = unsigned(k) <= unsigned(kSepZ);
                    const float w2 = f2 - g2;
                    const float cw = live ? pw[j] * wa : 0.f;
                    float* sk = slot + ((live ? k : 0) * 256 + lane);
                    sk[0] += cw * (1.f - w2);
                    sk[256] += cw * w2;
                }
            }
        }
    }
    const size_t plane = (size_t)g.nx * g.ny;
    const size_t colx = (size_t)y * g.nx + x;
#pragma unroll
    for (int k = 1; k <= kSepZ; ++k) {
        const int z = z0 + k - 1;
        if (z >= g.nz) break;
        const float val = slot[k * 256 + lane];
        float* dst = volume + z * plane + colx;
        *dst = accumulate ? *dst + val : val;
    }
}

// ---------------------------------------------------------------------------
// Siddon
// ---------------------------------------------------------------------------
struct Box {
    float lo[3], hi[3];
};

__device__ __forceinline__ Box volume_box(const VolumeGeom& g)
{
    Box b;
    b.lo[0] = g.ox - 0.5f * g.vx; b.hi[0] = g.ox + (float(g.nx) - 0.5f) * g.vx;
    b.lo[1] = g.oy - 0.5f * g.vy; b.hi[1] = g.oy + (float(g.ny) - 0.5f) * g.vy;
    b.lo[2] = g.oz - 0.5f * g.vz; b.hi[2] = g.oz + (float(g.nz) - 0.5f) * g.vz;
    return b;
}

// one slab of the box clip: narrows [t0, t1] to lo <= s + t r <= hi; a ray
// parallel to the slab and outside it misses
__device__ __forceinline__ bool slab_clip(float s, float r, float lo, float hi, float& t0, float& t1)
{
    if (fabsf(r) > 1e-8f) {
        float ta = (lo - s) / r, tb = (hi - s) / r;
        if (ta > tb) { const float t = ta; ta = tb; tb = t; }
        t0 = fmaxf(t0, ta);
        t1 = fminf(t1, tb);
        return true;
    }
    return !(s < lo || s > hi);
}

// parametric segment [t0, t1] of src + t ray inside the box, clipped to [0, 1]
__device__ __forceinline__ bool clip_ray(const float* src, const float* ray, const float* lo,
    const float* hi, float& t0, float& t1)
{
    t0 = 0.f;
    t1 = 1.f;
    for (int k = 0; k < 3; ++k)
        if (!slab_clip(src[k], ray[k], lo[k], hi[k], t0, t1)) return false;
    return t0 < t1;
}

__device__ __forceinline__ void siddon_ray_geom(const SiddonViewDev& sv, int u, int v, float* src, float* ray)
{
#pragma clang fp contract(off)
    for (int k = 0; k < 3; ++k) {
        src[k] = sv.src[k];
        ray[k] = (sv.detS[k] + float(u) * sv.detU[k] + float(v) * sv.detV[k]) - sv.src[k];
    }
}

// Siddon projection of any view set: one lane per ray, an exact Amanatides-
// Woo walk through the [z][y][x] volume, clipped to t in [0, 1] and to the
// volume box.  Plane crossings are recomputed from integer indices as
// (plane - src) * (1 / ray), the values the transposes rebuild, and the ray
// sum adds the segments in walk order.
__global__ void __launch_bounds__(256)
siddon_fp_kernel(const float* __restrict__ vol, float* __restrict__ proj,
    const SiddonViewDev* __restrict__ views, VolumeGeom g, int nu, int nv, int count, int accumulate)
{
    // a workgroup is an 8 x 4 pixel tile of eight consecutive views: their
    // rays nearly coincide, so the scattered voxel reads of the walk share
    // cache lines in the CU's L1 (the walk is bound by those reads): 822 ->
    // 389 ms at 512^3 x 720 against a 16 x 16 tile of one view; a 4 x 4 tile
    // of sixteen views measured 409 ms
    const int u = blockIdx.x * 8 + (threadIdx.x & 7);
    const int v = blockIdx.y * 4 + ((threadIdx.x >> 3) & 3);
    const int i = int(blockIdx.z) * 8 + int(threadIdx.x >> 5);
    if (u >= nu || v >= nv || i >= count) return;
    const SiddonViewDev sv = views[i];
    float src[3], ray[3];
    siddon_ray_geom(sv, u, v, src, ray);
    const Box b = volume_box(g);
    const int dims[3] = { g.nx, g.ny, g.nz };
    const float vox[3] = { g.vx, g.vy, g.vz };
    float t0, t1, acc = 0.f;
    if (clip_ray(src, ray, b.lo, b.hi, t0, t1)) {
        const float len = sqrtf(ray[0] * ray[0] + ray[1] * ray[1] + ray[2] * ray[2]);
        int idx[3], step[3];
        float tn[3], rcp[3];
        bool valid[3];
        for (int k = 0; k < 3; ++k) {
            const float e = src[k] + t0 * ray[k];
            idx[k] = min(dims[k] - 1, max(0, int(floorf((e - b.lo[k]) / vox[k]))));
            step[k] = ray[k] >= 0.f ? 1 : -1;
            valid[k] = fabsf(ray[k]) > 1e-8f;
            rcp[k] = valid[k] ? 1.f / ray[k] : 0.f;
            const float plane = b.lo[k] + float(idx[k] + (step[k] > 0 ? 1 : 0)) * vox[k];
            tn[k] = valid[k] ? (plane - src[k]) * rcp[k] : 3.0e38f;
        }
        float t = t0;
        const int max_steps = g.nx + g.ny + g.nz + 3;
        for (int n = 0; n < max_steps && t < t1; ++n) {
            const float tnext = fminf(fminf(tn[0], tn[1]), fminf(tn[2], t1));
            acc += vol[((size_t)idx[2] * g.ny + idx[1]) * g.nx + idx[0]] * (tnext - t);
            t = tnext;
            int k;
            if (tn[0] <= tn[1] && tn[0] <= tn[2]) k = 0;
            else if (tn[1] <= tn[2]) k = 1;
            else k = 2;
            idx[k] += step[k];
            if (idx[k] < 0 || idx[k] >= dims[k]) break;
            const float plane = b.lo[k] + float(idx[k] + (step[k] > 0 ? 1 : 0)) * vox[k];
            tn[k] = (plane - src[k]) * rcp[k];
        }
        acc *= len;
    }
    float* dst = proj + ((size_t)i * nv + v) * nu + u;
    *dst = accumulate ? *dst + acc : acc;
}

// t-interval of the walk inside cell c of one axis: the forward walk's own
// plane crossings (plane - src) * (1 / ray); a ray parallel to the axis stays
// in the cell its source coordinate lies in
__device__ __forceinline__ void cell_interval(float s, float r, float lo, float vox, int c, int n,
    float& ta, float& tb)
{
    if (fabsf(r) > 1e-8f) {
        const float rc = 1.f / r;
        const float t0 = (lo + float(c) * vox - s) * rc, t1 = (lo + float(c + 1) * vox - s) * rc;
        ta = fmaxf(ta, fminf(t0, t1));
        tb = fminf(tb, fmaxf(t0, t1));
    } else if (min(n - 1, max(0, int(floorf((s - lo) / vox)))) != c) {
        tb = -3.0e38f;
    }
}

// Siddon transpose: one thread per voxel; for each view the candidate rays
// are the pixels inside the projection of the voxel cube, and each candidate
// contributes y * |segment inside the cube, clipped like the forward walk|.
__global__ void __launch_bounds__(256)
siddon_adjoint_kernel(const float* __restrict__ proj, float* __restrict__ volume,
    const SiddonViewDev* __restrict__ views, const float* __restrict__ c32, VolumeGeom g,
    int nu, int nv, int count, int accumulate)
{
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny) return;
    const Box b = volume_box(g);
    float clo[3], chi[3];  // the voxel's cube
    clo[0] = b.lo[0] + float(x) * g.vx; chi[0] = b.lo[0] + float(x + 1) * g.vx;
    clo[1] = b.lo[1] + float(y) * g.vy; chi[1] = b.lo[1] + float(y + 1) * g.vy;
    clo[2] = b.lo[2] + float(z) * g.vz; chi[2] = b.lo[2] + float(z + 1) * g.vz;
    const float margin = 0.05f;
    const bool surface = x == 0 || x == g.nx - 1 || y == 0 || y == g.ny - 1 || z == 0 || z == g.nz - 1;
    float acc = 0.f;
    for (int i = 0; i < count; ++i) {
        // candidate pixels: the projected box of the voxel's 8 corners (FP32
        // map; the 0.05 px margin covers its rounding, and every candidate
        // is then tested with its own segment)
        const float* c = c32 + (size_t)i * 12;
        float umin = 3.0e38f, umax = -3.0e38f, vmin = 3.0e38f, vmax = -3.0e38f;
        bool ok = true;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float X = (k & 1) ? chi[0] : clo[0];
            const float Y = (k & 2) ? chi[1] : clo[1];
            const float Z = (k & 4) ? chi[2] : clo[2];
            const float den = c[8] * X + c[9] * Y + c[10] * Z + c[11];
            ok = ok && den > 1e-6f;
            const float rd = __builtin_amdgcn_rcpf(den);
            const float uu = proj_coord(c, 0, X, Y, Z, rd), vv = proj_coord(c, 4, X, Y, Z, rd);
            umin = fminf(umin, uu); umax = fmaxf(umax, uu);
            vmin = fminf(vmin, vv); vmax = fmaxf(vmax, vv);
        }
        if (!ok) continue;
        const int ulo = max(0, int(ceilf(umin - margin)));
        const int uhi = min(nu - 1, int(floorf(umax + margin)));
        const int vlo = max(0, int(ceilf(vmin - margin)));
        const int vhi = min(nv - 1, int(floorf(vmax + margin)));
        if (ulo > uhi || vlo > vhi) continue;
        const SiddonViewDev sv = views[i];
        const float* pi = proj + (size_t)i * nu * nv;
        for (int vv = vlo; vv <= vhi; ++vv)
            for (int uu = ulo; uu <= uhi; ++uu) {
                float src[3], ray[3];
                siddon_ray_geom(sv, uu, vv, src, ray);
                // the voxel's interval from the forward walk's own plane
                // crossings; the volume clip (whose faces are a voxel or more
                // away from an interior voxel's) binds only on the surface
                float a = -3.0e38f, e = 3.0e38f;
                cell_interval(src[0], ray[0], b.lo[0], g.vx, x, g.nx, a, e);
                cell_interval(src[1], ray[1], b.lo[1], g.vy, y, g.ny, a, e);
                cell_interval(src[2], ray[2], b.lo[2], g.vz, z, g.nz, a, e);
                if (!(e > a)) continue;
                if (surface) {
                    float t0, t1;
                    if (!clip_ray(src, ray, b.lo, b.hi, t0, t1)) continue;
                    a = fmaxf(a, t0);
                    e = fminf(e, t1);
                } else {
                    a = fmaxf(a, 0.f);
                    e = fminf(e, 1.f);
                }
                if (!(e > a)) continue;
                const float len = sqrtf(ray[0] * ray[0] + ray[1] * ray[1] + ray[2] * ray[2]);
                acc += pi[(size_t)vv * nu + uu] * ((e - a) * len);
            }
    }
    float* dst = volume + ((size_t)z * g.ny + y) * g.nx + x;
    *dst = accumulate ? *dst + acc : acc;
}

// ---------------------------------------------------------------------------
// Siddon transpose for separable views (detV along Z only).  The rays of a
// detector column u share ray_x and ray_y (independent of v), hence their XY
// path: the t-interval of an (x, y) cell and the XY part of the volume clip
// are column constants, and each ray only resolves the Z voxels it crosses
// inside that interval.  Lane / slot layout of the separable Joseph transpose,
// with kSidSepZ slices per lane column.
// ---------------------------------------------------------------------------

// col_xy[view][u] = (ray_x, ray_y) of column u; ray_wz[view][u][v] =
// (p L, ray_z) of ray (u, v), L = |ray|, each column padded to nvp entries of
// zeros (p L = 0 marks a ray that adds nothing)
__global__ void __launch_bounds__(256)
siddon_sep_table_kernel(const float* __restrict__ proj, const SiddonViewDev* __restrict__ views,
    int nu, int nv, int nvp, float2* __restrict__ col_xy, float2* __restrict__ ray_wz)
{
    __shared__ float2 s_t[32][33];
    const int i = blockIdx.z;
    const int u0 = blockIdx.x * 32, v0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const SiddonViewDev sv = views[i];
    const float* pi = proj + (size_t)i * nu * nv;
    for (int k = ty; k < 32; k += 8) {
        const int u = u0 + tx, v = v0 + k;
        float2 e = make_float2(0.f, 0.f);
        if (u < nu && v < nv) {
            float src[3], ray[3];
            siddon_ray_geom(sv, u, v, src, ray);
            const float len = sqrtf(ray[0] * ray[0] + ray[1] * ray[1] + ray[2] * ray[2]);
            e = make_float2(pi[(size_t)v * nu + u] * len, ray[2]);
            if (v == 0) col_xy[(size_t)i * nu + u] = make_float2(ray[0], ray[1]);
        }
        s_t[k][tx] = e;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int u = u0 + k, v = v0 + tx;
        if (u < nu && v < nvp) ray_wz[((size_t)i * nu + u) * nvp + v] = s_t[tx][k];
    }
}

// z-slices per lane column of the Siddon transpose: 16 (18 LDS slots a lane,
// 8 workgroups per CU) measured 18 % faster than 32 at 256^3 x 360; the
// Joseph transpose keeps kSepZ = 32, where 16 measured 12 % slower
constexpr int kSidSepZ = 16;

__global__ void __launch_bounds__(256)
siddon_adjoint_sep_kernel(const float2* __restrict__ col_xy, const float2* __restrict__ ray_wz,
    const SiddonViewDev* __restrict__ views, const float* __restrict__ c32, VolumeGeom g, int nu, int nv,
    int nvp, int count, int accumulate, float* __restrict__ volume)
{
    __shared__ float slot[(kSidSepZ + 2) * 256];
    const int lane = threadIdx.x;
    const int x = blockIdx.x * 16 + (lane & 15);
    const int y = blockIdx.y * 16 + (lane >> 4);
    const int z0 = blockIdx.z * kSidSepZ;
#pragma unroll
    for (int k = 0; k < kSidSepZ + 2; ++k) slot[k * 256 + lane] = 0.f;
    if (x >= g.nx || y >= g.ny) return;
    const Box b = volume_box(g);
    const float margin = 0.05f;
    const int zend = min(z0 + kSidSepZ, g.nz);  // one past the block's last slice
    const float xl = b.lo[0] + float(x) * g.vx, xh = b.lo[0] + float(x + 1) * g.vx;
    const float yl = b.lo[1] + float(y) * g.vy, yh = b.lo[1] + float(y + 1) * g.vy;
    const float zl = b.lo[2] + float(z0) * g.vz, zh = b.lo[2] + float(zend) * g.vz;
    const float inv_vz = 1.f / g.vz;
    // only a block holding the volume's first or last slice can see the Z
    // faces of the volume clip bind inside its voxels
    const bool zface = z0 == 0 || zend == g.nz;
    for (int i = 0; i < count; ++i) {
        const float* c = c32 + (size_t)i * 12;
        // detector range of the cell x block (8 corners, FP32 map, margin)
        float umin = 3.0e38f, umax = -3.0e38f, vmin = 3.0e38f, vmax = -3.0e38f;
        bool ok = true;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float X = (k & 1) ? xh : xl, Y = (k & 2) ? yh : yl, Z = (k & 4) ? zh : zl;
            const float den = c[8] * X + c[9] * Y + c[10] * Z + c[11];
            ok = ok && den > 1e-6f;
            const float rd = 1.f / den;
            const float uu = proj_coord(c, 0, X, Y, Z, rd), vv = proj_coord(c, 4, X, Y, Z, rd);
            umin = fminf(umin, uu); umax = fmaxf(umax, uu);
            vmin = fminf(vmin, vv); vmax = fmaxf(vmax, vv);
        }
        if (!ok) continue;
        const int ulo = max(0, int(ceilf(umin - margin)));
        const int uhi = min(nu - 1, int(floorf(umax + margin)));
        const int vlo = max(0, int(ceilf(vmin - margin)));
        const int vhi = min(nv - 1, int(floorf(vmax + margin)));
        const SiddonViewDev sv = views[i];
        const float sx = sv.src[0], sy = sv.src[1], sz = sv.src[2];
        for (int u = ulo; u <= uhi; ++u) {
            const float2 r = col_xy[(size_t)i * nu + u];
            // XY part of the volume clip and of the cell, shared by the column
            float a0 = 0.f, e0 = 1.f;
            if (!slab_clip(sx, r.x, b.lo[0], b.hi[0], a0, e0)) continue;
            if (!slab_clip(sy, r.y, b.lo[1], b.hi[1], a0, e0)) continue;
            float ca = -3.0e38f, cb = 3.0e38f;
            cell_interval(sx, r.x, b.lo[0], g.vx, x, g.nx, ca, cb);
            cell_interval(sy, r.y, b.lo[1], g.vy, y, g.ny, ca, cb);
            a0 = fmaxf(a0, ca);
            e0 = fminf(e0, cb);
            if (!(e0 > a0)) continue;
            const float2* q = ray_wz + ((size_t)i * nu + u) * nvp + vlo;
            const int nr = vhi - vlo + 1;
            for (int r0 = 0; r0 < nr; r0 += kSepU) {
                float2 e[kSepU];
#pragma unroll
                for (int j = 0; j < kSepU; ++j) e[j] = q[r0 + j];
#pragma unroll
                for (int j = 0; j < kSepU; ++j) {
                    const float pw = e[j].x, rz = e[j].y;
                    float a = a0, t1 = e0;
                    bool live = pw != 0.f;
                    if (zface) live = live && slab_clip(sz, rz, b.lo[2], b.hi[2], a, t1);
                    live = live && t1 > a;
                    if (!live) continue;
                    // Z voxels of the block the ray crosses inside [a, t1]
                    const bool zvalid = fabsf(rz) > 1e-8f;
                    const float rc = zvalid ? 1.f / rz : 0.f;
                    const float za = sz + a * rz, zb = sz + t1 * rz;
                    int ka = int(floorf((fminf(za, zb) - b.lo[2]) * inv_vz));
                    int kb = int(floorf((fmaxf(za, zb) - b.lo[2]) * inv_vz));
                    ka = max(ka, z0);
                    kb = min(kb, zend - 1);
                    for (int k = ka; k <= kb; ++k) {
                        float lo = a, hi = t1;
                        if (zvalid) {
                            const float tl = (b.lo[2] + float(k) * g.vz - sz) * rc;
                            const float th = (b.lo[2] + float(k + 1) * g.vz - sz) * rc;
                            lo = fmaxf(lo, fminf(tl, th));
                            hi = fminf(hi, fmaxf(tl, th));
                        }
                        const float seg = hi - lo;
                        if (seg > 0.f) slot[(k - z0 + 1) * 256 + lane] += pw * seg;
                    }
                }
            }
        }
    }
    const size_t plane = (size_t)g.nx * g.ny;
    const size_t colx = (size_t)y * g.nx + x;
#pragma unroll
    for (int k = 1; k <= kSidSepZ; ++k) {
        const int z = z0 + k - 1;
        if (z >= g.nz) break;
        const float val = slot[k * 256 + lane];
        float* dst = volume + z * plane + colx;
        *dst = accumulate ? *dst + val : val;
    }
}

// Operator-form FDK backprojections, the iterative B of the reference's
// automatic flat selection (fdk_bp_kernel_fdk_weight_v2 and
// fdk_bp_kernel_matched_weight, src/FlatFpBp/BP/kernels/YkBpFdkKernel.cu:228-466):
// unfiltered views, exact bilinear sample at the projected voxel centre (zero
// outside the detector), per-sample weight
//   FDK:     SOD^2 dtheta fscale / (Cc . [X, 1])^2,  Cc = (-src / SOD, SOD)
//   matched: L^3 / (|X - src|^2 SDD_plane) * vx vy vz / (du dv),
//            L^2 = |detS + u detU + v detV - src|^2;
//   voxel-driven Joseph (joseph_bp_v3_kernel,
//   src/FlatFpBp/BP/kernels/YkBPJosephKernel.cu:624-735): the physical length
//   of X - src through one voxel along its main axis m = argmax |d_k / v_k|
//   (ties X, then Y): v_m |d| / |d_m|.
constexpr int kOpBpZ = 8;
enum OpBpMode { kOpFdk = 0, kOpMatched = 1, kOpJosephVoxel = 2 };

__device__ __forceinline__ float bilinear_view(const float* __restrict__ p, int nu, int nv, float u, float v)
{
    const float fu = floorf(u), fv = floorf(v);
    const int iu = int(fu), iv = int(fv);
    const float a = u - fu, b = v - fv;
    const bool u0 = iu >= 0 && iu < nu, u1 = iu + 1 >= 0 && iu + 1 < nu;
    const bool v0 = iv >= 0 && iv < nv, v1 = iv + 1 >= 0 && iv + 1 < nv;
    const float t00 = (u0 && v0) ? p[(size_t)iv * nu + iu] : 0.f;
    const float t01 = (u1 && v0) ? p[(size_t)iv * nu + iu + 1] : 0.f;
    const float t10 = (u0 && v1) ? p[(size_t)(iv + 1) * nu + iu] : 0.f;
    const float t11 = (u1 && v1) ? p[(size_t)(iv + 1) * nu + iu + 1] : 0.f;
    return (1.f - b) * ((1.f - a) * t00 + a * t01) + b * ((1.f - a) * t10 + a * t11);
}

template <int MODE>
__global__ void __launch_bounds__(256)
fdk_operator_bp_kernel(const float* __restrict__ proj, float* __restrict__ volume,
    const OpBpViewDev* __restrict__ views, VolumeGeom g, int nu, int nv, int count, float scale,
    int accumulate)
{
    const int x = blockIdx.x * 16 + int(threadIdx.x % 16);
    const int y = blockIdx.y * 16 + int(threadIdx.x / 16);
    const int z0 = blockIdx.z * kOpBpZ;
    if (x >= g.nx || y >= g.ny) return;
    const float X = g.ox + x * g.vx, Y = g.oy + y * g.vy;
    float acc[kOpBpZ];
#pragma unroll
    for (int k = 0; k < kOpBpZ; ++k) acc[k] = 0.f;
    // reciprocal voxel sizes for the voxel-driven Joseph main-axis test
    const float ivx = 1.f / g.vx, ivy = 1.f / g.vy, ivz = 1.f / g.vz;
    for (int i = 0; i < count; ++i) {
        const OpBpViewDev c = views[i];
        const float* p = proj + (size_t)i * nu * nv;
        // the view's map at this voxel column; each slice adds its Z term
        const float d0 = fmaf(c.cd[1], Y, fmaf(c.cd[0], X, c.cd[3]));
        const float u0 = fmaf(c.cu[1], Y, fmaf(c.cu[0], X, c.cu[3]));
        const float v0 = fmaf(c.cv[1], Y, fmaf(c.cv[0], X, c.cv[3]));
#pragma unroll
        for (int k = 0; k < kOpBpZ; ++k) {
            if (z0 + k >= g.nz) break;
            const float Z = g.oz + (z0 + k) * g.vz;
            const float den = fmaf(c.cd[2], Z, d0);
            if (!(den >= 1e-8f)) continue;
            // one Newton step on the hardware reciprocal: u and v within an
            // ulp or so of the quotients (the sample position is the
            // sensitive input), for two FMAs instead of two divisions
            float rd = __builtin_amdgcn_rcpf(den);
            rd = fmaf(rd, fmaf(-den, rd, 1.f), rd);
            const float u = fmaf(c.cu[2], Z, u0) * rd;
            const float v = fmaf(c.cv[2], Z, v0) * rd;
            float w;
            if constexpr (MODE == kOpJosephVoxel) {
                const float dx = X - c.src[0], dy = Y - c.src[1], dz = Z - c.src[2];
                const float ax = fabsf(dx * ivx), ay = fabsf(dy * ivy), az = fabsf(dz * ivz);
                const bool mx = ax >= ay && ax >= az, my = !mx && ay >= az;
                const float vm = mx ? g.vx : my ? g.vy : g.vz;
                const float dm = mx ? dx : my ? dy : dz;
                w = vm * sqrtf(dx * dx + dy * dy + dz * dz) * __builtin_amdgcn_rcpf(fabsf(dm));
            } else if constexpr (MODE == kOpMatched) {
                const float L2 = c.l2[0] + u * (c.l2[1] + u * c.l2[3]) + v * (c.l2[2] + u * c.l2[4] + v * c.l2[5]);
                if (!(L2 >= 1e-8f)) continue;
                const float dx = X - c.src[0], dy = Y - c.src[1], dz = Z - c.src[2];
                const float lsq = dx * dx + dy * dy + dz * dz;
                if (!(lsq >= 1e-8f)) continue;
                w = L2 * sqrtf(L2) * __builtin_amdgcn_rcpf(lsq) * c.inv_sdd_plane * scale;
            } else {
                const float cn = c.cc[0] * X + c.cc[1] * Y + c.cc[2] * Z + c.cc[3];
                if (!(fabsf(cn) >= 1e-8f)) continue;
                const float rc = __builtin_amdgcn_rcpf(cn);
                w = c.wbase * rc * rc;
            }
            acc[k] += bilinear_view(p, nu, nv, u, v) * w;
        }
    }
#pragma unroll
    for (int k = 0; k < kOpBpZ; ++k) {
        if (z0 + k >= g.nz) break;
        float* dst = volume + ((size_t)(z0 + k) * g.ny + y) * g.nx + x;
        *dst = accumulate ? *dst + acc[k] : acc[k];
    }
}

}  // namespace

void launch_joseph_fp(const JosephLaunch& p, hipStream_t s)
{
    dim3 block(64, 4), grid((p.nu + 63) / 64, (p.nv + 3) / 4, p.count);
    hipLaunchKernelGGL(joseph_fp_kernel, grid, block, 0, s, p.volume, p.volume_t, p.proj, p.views,
        p.vol, p.nu, p.nv, p.accumulate);
    YK_HIP(hipGetLastError());
}

void launch_joseph_fp_sep(const JosephSepFpLaunch& p, hipStream_t s)
{
    constexpr int kCols = 1;  // columns per workgroup (the other waves: further views)
    const int groups = (p.nu + kCols - 1) / kCols;
    dim3 block(512), grid((groups + 7) / 8 * 8, (p.nv + 63) / 64, (p.count + 8 / kCols - 1) / (8 / kCols));
    hipLaunchKernelGGL(joseph_fp_sep_kernel<kCols>, grid, block, 0, s, p.vol_zx, p.vol_zy, p.proj, p.views,
        p.vol, p.nu, p.nv, p.count, p.accumulate);
    YK_HIP(hipGetLastError());
}

size_t joseph_pair_size(int nx, int ny, int nz, bool x_main, bool slab)
{
    const size_t n0 = size_t(x_main ? nx : ny), n1 = size_t(x_main ? ny : nx);
    const size_t g = 2 * kJosephGuard;
    // rows x columns of float2 per slice: (n_r + 2g - 1) x (n_e + 2g)
    const size_t rows = (slab ? size_t(nz) : n1) + g - 1, cols = (slab ? n1 : size_t(nz)) + g;
    return n0 * rows * cols * 2;
}

void launch_joseph_pair_copy(const float* in, float* out, int nx, int ny, int nz, bool x_main, bool slab,
    hipStream_t s)
{
    // the guard columns stay zero from the fill; every row is written
    launch_fill(out, 0.f, joseph_pair_size(nx, ny, nz, x_main, slab), s);
    float2* o = reinterpret_cast<float2*>(out);
    const int g = 2 * kJosephGuard;
    if (!slab && x_main) {
        dim3 block(32, 8), grid((nx + 31) / 32, (nz + 31) / 32, ny + g - 1);
        hipLaunchKernelGGL(pair_zfast_x_kernel, grid, block, 0, s, in, o, nx, ny, nz);
    } else if (!slab) {
        dim3 block(32, 8), grid((nx + g - 1 + 31) / 32, (nz + 31) / 32, ny);
        hipLaunchKernelGGL(pair_zfast_y_kernel, grid, block, 0, s, in, o, nx, ny, nz);
    } else if (x_main) {
        dim3 block(32, 8), grid((nx + 31) / 32, (ny + 31) / 32, nz + g - 1);
        hipLaunchKernelGGL(pair_slab_x_kernel, grid, block, 0, s, in, o, nx, ny, nz);
    } else {
        dim3 block(256), grid((nx + 255) / 256, ny, nz + g - 1);
        hipLaunchKernelGGL(pair_slab_y_kernel, grid, block, 0, s, in, o, nx, ny, nz);
    }
    YK_HIP(hipGetLastError());
}

void launch_joseph_fp_slab(const JosephSepFpLaunch& p, hipStream_t s)
{
    dim3 block(64, 8), grid((p.nu + 63) / 64, p.nv, (p.count + 7) / 8);
    hipLaunchKernelGGL(joseph_fp_slab_kernel, grid, block, 0, s, p.vol_zx, p.vol_zy, p.proj, p.views, p.vol,
        p.nu, p.nv, p.count, p.accumulate);
    YK_HIP(hipGetLastError());
}

void launch_transpose_xy(const float* in, float* out, int nx, int ny, int nz, hipStream_t s)
{
    dim3 block(32, 8), grid((nx + 31) / 32, (ny + 31) / 32, nz);
    hipLaunchKernelGGL(transpose_xy_kernel, grid, block, 0, s, in, out, nx, ny);
    YK_HIP(hipGetLastError());
}

void launch_joseph_adjoint(const JosephAdjointLaunch& p, hipStream_t s)
{
    dim3 tb(64, 4), tg((p.nu + 63) / 64, (p.nv + 3) / 4, p.count);
    hipLaunchKernelGGL(joseph_ray_table_kernel, tg, tb, 0, s, p.views, p.vol, p.nu, p.nv, p.ray_ab, p.ray_len);
    dim3 block(256), grid((p.vol.nx + 15) / 16, (p.vol.ny + 15) / 16, p.vol.nz);
    hipLaunchKernelGGL(joseph_adjoint_kernel, grid, block, 0, s, p.proj, p.volume, p.views, p.coeff32,
        p.ray_ab, p.ray_len, p.vol, p.nu, p.nv, p.count, p.accumulate);
    YK_HIP(hipGetLastError());
}

void launch_joseph_adjoint_sep(const JosephSepLaunch& p, hipStream_t s)
{
    const int nvp = p.nv + kSepPad;
    dim3 tb(256), tg((p.nu + 31) / 32, (nvp + 31) / 32, p.count);
    hipLaunchKernelGGL(joseph_sep_table_kernel, tg, tb, 0, s, p.proj, p.views, p.vol, p.nu, p.nv, nvp, p.ab1,
        p.ab2, p.pl);
    dim3 block(256), grid((p.vol.nx + 15) / 16, (p.vol.ny + 15) / 16, (p.vol.nz + kSepZ - 1) / kSepZ);
    hipLaunchKernelGGL(joseph_adjoint_sep_kernel, grid, block, 0, s, p.ab1, p.ab2, p.pl, p.views, p.coeff32,
        p.vol, p.nu, p.nv, nvp, p.count, p.accumulate, p.volume);
    YK_HIP(hipGetLastError());
}

void launch_siddon_fp(const SiddonLaunch& p, hipStream_t s)
{
    dim3 block(256), grid((p.nu + 7) / 8, (p.nv + 3) / 4, (p.count + 7) / 8);
    hipLaunchKernelGGL(siddon_fp_kernel, grid, block, 0, s, p.volume, p.proj, p.views, p.vol, p.nu,
        p.nv, p.count, p.accumulate);
    YK_HIP(hipGetLastError());
}

void launch_siddon_adjoint_sep(const SiddonSepLaunch& p, hipStream_t s)
{
    const int nvp = p.nv + kSepPad;
    dim3 tb(256), tg((p.nu + 31) / 32, (nvp + 31) / 32, p.count);
    hipLaunchKernelGGL(siddon_sep_table_kernel, tg, tb, 0, s, p.proj, p.views, p.nu, p.nv, nvp, p.col_xy,
        p.ray_wz);
    dim3 block(256), grid((p.vol.nx + 15) / 16, (p.vol.ny + 15) / 16, (p.vol.nz + kSidSepZ - 1) / kSidSepZ);
    hipLaunchKernelGGL(siddon_adjoint_sep_kernel, grid, block, 0, s, p.col_xy, p.ray_wz, p.views, p.coeff32,
        p.vol, p.nu, p.nv, nvp, p.count, p.accumulate, p.volume);
    YK_HIP(hipGetLastError());
}

void launch_siddon_adjoint(const SiddonAdjointLaunch& p, hipStream_t s)
{
    dim3 block(256), grid((p.vol.nx + 15) / 16, (p.vol.ny + 15) / 16, p.vol.nz);
    hipLaunchKernelGGL(siddon_adjoint_kernel, grid, block, 0, s, p.proj, p.volume, p.views, p.coeff32,
        p.vol, p.nu, p.nv, p.count, p.accumulate);
    YK_HIP(hipGetLastError());
}

void launch_fdk_operator_bp(const OpBpLaunch& p, hipStream_t s)
{
    dim3 block(256), grid((p.vol.nx + 15) / 16, (p.vol.ny + 15) / 16, (p.vol.nz + kOpBpZ - 1) / kOpBpZ);
    if (p.mode == kOpMatched)
        hipLaunchKernelGGL(fdk_operator_bp_kernel<kOpMatched>, grid, block, 0, s, p.proj, p.volume, p.views,
            p.vol, p.nu, p.nv, p.count, p.scale, p.accumulate);
    else if (p.mode == kOpJosephVoxel)
        hipLaunchKernelGGL(fdk_operator_bp_kernel<kOpJosephVoxel>, grid, block, 0, s, p.proj, p.volume, p.views,
            p.vol, p.nu, p.nv, p.count, p.scale, p.accumulate);
    else
        hipLaunchKernelGGL(fdk_operator_bp_kernel<kOpFdk>, grid, block, 0, s, p.proj, p.volume, p.views, p.vol,
            p.nu, p.nv, p.count, p.scale, p.accumulate);
    YK_HIP(hipGetLastError());
}

}  // namespace yk

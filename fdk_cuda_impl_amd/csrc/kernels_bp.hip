// Voxel-driven FDK backprojection for gfx950.
//
// Work decomposition: a workgroup owns a 16 x 16 column tile of voxels and a
// run of BZ consecutive z slices; each lane owns one (x, y) column and keeps
// BZ accumulators in registers for the whole view loop, so the volume is read
// and written once per launch.  Per view the workgroup stages the detector
// footprint of its voxel block into LDS as "quads": for texel (r, c) the four
// floats (q[r][c], q[r][c+1], q[r+1][c]-q[r][c], q[r+1][c+1]-q[r][c+1]), so
// one 16-byte LDS read feeds a whole bilinear sample.  The footprint of every
// (block, view) pair is computed once per plan in FP64 (bp_tiles_kernel).
//
// For a detector whose rows stay parallel to the rotation axis (the regular
// circular geometry; Cu_z = Cd_z = 0 exactly) u, the u-weights and the
// distance weight are constant along a column, v is linear in z and a sample
// costs two FMAs for the v-lerp and two for the weighted u-lerp.  Any other
// geometry takes the general path with a per-sample perspective divide.
//
// Storage variants: float; half or bfloat16 rows with power-of-two row
// scales, staged into LDS as half quads renormalized by the global maximum
// row scale (exact power-of-two rescale), lerped with packed FP16 FMAs and
// accumulated in FP32 by v_dot2c_f32_f16.  Optional Kahan compensation across
// blocks of views.
//
// Math restates fdk_bp_kernel<ZSIZE> (precomputed form),
// src/Reconstruction/Analytic/Circular/Flat/FDK/kernels/YkFDKBpKernels.cu:112-209:
//   vol(X) += sum_i bilinear(q_i; u, v) * SOD^2 dtheta fscale / ((X - src).r)^2
// with exact software bilinear (texel (c, r) centred at integer (u, v), zero
// outside the detector) in place of the CUDA Linear/Border texture.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <algorithm>

#include "yk_internal.hpp"

namespace yk {
namespace {

constexpr int kBX = 16;
constexpr int kBY = 16;
constexpr int kThreads = kBX * kBY;
constexpr int kKahanViews = 16;

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// ---------------------------------------------------------------------------
// footprint table: per (block, view) the quad range [u0, u0+w) x [v0, v0+h)
// that contains floor(u), floor(v) of every voxel of the block, with one quad
// of margin on each side; w = h = 0 when the block misses the detector.
// ---------------------------------------------------------------------------
__global__ void bp_tiles_kernel(const double* __restrict__ c64, int n_views, VolumeGeom vol,
    int bx, int by, int bz, int nbx, int nby, int nbz, int nu, int nv,
    short4* __restrict__ tiles, int* __restrict__ maxes)
{
    const long long total = (long long)nbx * nby * nbz * n_views;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int view = int(idx % n_views);
    const long long block = idx / n_views;
    const int ib = int(block % nbx);
    const int jb = int((block / nbx) % nby);
    const int kb = int(block / ((long long)nbx * nby));
    const double* c = c64 + (size_t)view * 12;
    const double xs[2] = { double(vol.ox) + double(ib * bx) * vol.vx,
                           double(vol.ox) + double(ib * bx + bx - 1) * vol.vx };
    const double ys[2] = { double(vol.oy) + double(jb * by) * vol.vy,
                           double(vol.oy) + double(jb * by + by - 1) * vol.vy };
    const double zs[2] = { double(vol.oz) + double(kb * bz) * vol.vz,
                           double(vol.oz) + double(kb * bz + bz - 1) * vol.vz };
    double umin = 1e300, umax = -1e300, vmin = 1e300, vmax = -1e300;
    bool bad = false;
    for (int a = 0; a < 8; ++a) {
        const double X = xs[a & 1], Y = ys[(a >> 1) & 1], Z = zs[a >> 2];
        const double den = c[8] * X + c[9] * Y + c[10] * Z + c[11];
        if (!(den > 1e-6)) { bad = true; break; }
        const double u = (c[0] * X + c[1] * Y + c[2] * Z + c[3]) / den;
        const double v = (c[4] * X + c[5] * Y + c[6] * Z + c[7]) / den;
        umin = fmin(umin, u); umax = fmax(umax, u);
        vmin = fmin(vmin, v); vmax = fmax(vmax, v);
    }
    short4 t = make_short4(0, 0, 0, 0);
    if (!bad) {
        const double fu0 = floor(umin) - 1.0, fu1 = floor(umax) + 1.0;
        const double fv0 = floor(vmin) - 1.0, fv1 = floor(vmax) + 1.0;
        // quads with index in [-1, n-1] can touch a texel
        const bool miss = fu1 < -1.0 || fu0 > nu - 1.0 || fv1 < -1.0 || fv0 > nv - 1.0;
        if (!miss) {
            if (fu0 < -30000.0 || fv0 < -30000.0 || fu1 > 30000.0 || fv1 > 30000.0 ||
                fu1 - fu0 + 1.0 > 4096.0 || fv1 - fv0 + 1.0 > 4096.0) {
                bad = true;
            } else {
                t.x = short(fu0);
                t.y = short(fv0);
                t.z = short(fu1 - fu0 + 1.0);
                t.w = short(fv1 - fv0 + 1.0);
                atomicMax(&maxes[0], int(t.z));
                atomicMax(&maxes[1], int(t.w));
                atomicMax(&maxes[2], int(t.z) * int(t.w));
            }
        }
    }
    if (bad) atomicMax(&maxes[3], 1);
    tiles[idx] = t;
}

// small fill used instead of hipMemsetAsync so that callers may capture the
// launch sequence in a HIP graph
__global__ void set_ints_kernel(int* __restrict__ p, int n, int value)
{
    if (int(threadIdx.x) < n) p[threadIdx.x] = value;
}

// global maximum of the row scales (power-of-two values)
__global__ void max_scale_kernel(const float* __restrict__ s, int n, int* __restrict__ out_bits)
{
    float m = 0.f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        m = fmaxf(m, s[i]);
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out_bits, __float_as_int(m));  // m >= 0
}

// ---------------------------------------------------------------------------
// staging: the quads of tile (u0, v0, qw, qh) of one view.  Lanes take
// consecutive quad columns; each thread owns a run of quad rows in one column
// and slides a two-row window down it.  For tiles up to kMaxRows rows per
// thread the texels of the NEXT view are fetched into registers before the
// current view is accumulated, so global latency hides behind the LDS work.
// ---------------------------------------------------------------------------
constexpr int kMaxRows = 16;  // texel rows a thread can hold (quad rows + 1)

struct StageMap {
    int col;      // quad column
    int r0;       // first quad row
    int rows;     // quad rows written by this thread (0: idle)
    int rows_per; // quad rows per thread (uniform)
    bool fast;    // fits the register prefetch
};

__device__ __forceinline__ StageMap stage_map(short4 t)
{
    StageMap m;
    const int qw = t.z, qh = t.w;
    const int span = min(qw, kThreads);
    const int chunks = kThreads / span;
    const int rows_per = (qh + chunks - 1) / chunks;
    const int rc = int(threadIdx.x) / span;
    m.col = int(threadIdx.x) % span;
    m.r0 = rc * rows_per;
    m.rows = rc < chunks ? max(0, min(rows_per, qh - m.r0)) : 0;
    m.rows_per = rows_per;
    m.fast = qw <= kThreads && rows_per + 1 <= kMaxRows;
    return m;
}

__device__ __forceinline__ bool tile_inside(short4 t, int nu, int nv)
{
    return t.x >= 0 && t.y >= 0 && t.x + t.z + 1 <= nu && t.y + t.w + 1 <= nv;
}

// raw texel pair (col, col+1) of one detector row: float2 for float storage,
// two 16-bit patterns packed lo/hi for half / bfloat16; zero outside
template <int STORAGE>
struct Pair;
template <>
struct Pair<YK_F32> { using T = float2; };
template <>
struct Pair<YK_F16> { using T = unsigned; };
template <>
struct Pair<YK_BF16> { using T = unsigned; };

template <int STORAGE>
__device__ __forceinline__ typename Pair<STORAGE>::T fetch_pair(const void* __restrict__ base, int nu,
    int nv, int row, int col, bool inside)
{
    if constexpr (STORAGE == YK_F32) {
        const float* p = static_cast<const float*>(base);
        if (inside) {
            const float* r = p + (size_t)row * nu + col;
            return make_float2(r[0], r[1]);
        }
        float2 v = make_float2(0.f, 0.f);
        if (row >= 0 && row < nv) {
            const float* r = p + (size_t)row * nu;
            if (col >= 0 && col < nu) v.x = r[col];
            if (col + 1 >= 0 && col + 1 < nu) v.y = r[col + 1];
        }
        return v;
    } else {
        const unsigned short* p = static_cast<const unsigned short*>(base);
        unsigned lo = 0, hi = 0;
        if (inside) {
            const unsigned short* r = p + (size_t)row * nu + col;
            lo = r[0];
            hi = r[1];
        } else if (row >= 0 && row < nv) {
            const unsigned short* r = p + (size_t)row * nu;
            if (col >= 0 && col < nu) lo = r[col];
            if (col + 1 >= 0 && col + 1 < nu) hi = r[col + 1];
        }
        return lo | (hi << 16);
    }
}

// row normalizer: stored * scale / global scale (an exact power of two)
__device__ __forceinline__ float row_factor(const float* __restrict__ scales, int nv, int row, float inv_g)
{
    return (row >= 0 && row < nv) ? scales[row] * inv_g : 0.f;
}

// normalized half pair of a raw pair
template <int STORAGE>
__device__ __forceinline__ half2_t to_half2(unsigned raw, float f)
{
    if constexpr (STORAGE == YK_F16) {
        const half2_t h = __builtin_bit_cast(half2_t, raw);
        const _Float16 fh = (_Float16)f;  // power of two: exact when representable
        const half2_t ff = { fh, fh };
        return h * ff;
    } else {
        const float a = __uint_as_float(raw << 16) * f;
        const float b = __uint_as_float(raw & 0xffff0000u) * f;
        const half2_t r = { (_Float16)a, (_Float16)b };
        return r;
    }
}

template <int STORAGE>
__device__ __forceinline__ void write_quad(void* lds, int index, const typename Pair<STORAGE>::T& top,
    const typename Pair<STORAGE>::T& bot, float ftop, float fbot)
{
    if constexpr (STORAGE == YK_F32) {
        static_cast<float4*>(lds)[index] = make_float4(top.x, top.y, bot.x - top.x, bot.y - top.y);
    } else {
        const half2_t a = to_half2<STORAGE>(top, ftop);
        const half2_t b = to_half2<STORAGE>(bot, fbot);
        const half2_t d = b - a;
        uint2 q;
        q.x = __builtin_bit_cast(unsigned, a);
        q.y = __builtin_bit_cast(unsigned, d);
        static_cast<uint2*>(lds)[index] = q;
    }
}

// direct staging (no register prefetch), used for very tall or wide tiles
template <int STORAGE>
__device__ void stage_direct(void* lds, const void* __restrict__ base, const float* __restrict__ scales,
    int nu, int nv, short4 t, float inv_g)
{
    const int qw = t.z, qh = t.w;
    const bool inside = tile_inside(t, nu, nv);
    const int span = min(qw, kThreads);
    const int chunks = kThreads / span;
    const int rows_per = (qh + chunks - 1) / chunks;
    const int rc = int(threadIdx.x) / span;
    if (rc >= chunks) return;
    const int r0 = rc * rows_per;
    const int r1 = min(r0 + rows_per, qh);
    for (int c = int(threadIdx.x) % span; c < qw; c += span) {
        if (r0 >= r1) break;
        auto top = fetch_pair<STORAGE>(base, nu, nv, t.y + r0, t.x + c, inside);
        float ft = 1.f;
        if constexpr (STORAGE != YK_F32) ft = row_factor(scales, nv, t.y + r0, inv_g);
        for (int r = r0; r < r1; ++r) {
            auto bot = fetch_pair<STORAGE>(base, nu, nv, t.y + r + 1, t.x + c, inside);
            float fb = 1.f;
            if constexpr (STORAGE != YK_F32) fb = row_factor(scales, nv, t.y + r + 1, inv_g);
            write_quad<STORAGE>(lds, r * qw + c, top, bot, ft, fb);
            top = bot;
            ft = fb;
        }
    }
}

// Register prefetch: loads only, no use of the loaded values, so the compiler
// can leave every load in flight until commit.  Rows and columns are clamped
// into the detector so each load is in bounds; commit zeroes what lies
// outside.  Requires nu >= 2.
template <int STORAGE>
struct Prefetch {
    typename Pair<STORAGE>::T v[kMaxRows];
    float f[STORAGE == YK_F32 ? 1 : kMaxRows];
};

template <int STORAGE>
__device__ __forceinline__ typename Pair<STORAGE>::T raw_pair(const void* __restrict__ base, int nu,
    int row, int col)
{
    if constexpr (STORAGE == YK_F32) {
        const float* r = static_cast<const float*>(base) + (size_t)row * nu + col;
        return make_float2(r[0], r[1]);
    } else {
        // two adjacent 16-bit texels as one (possibly 2-byte aligned) dword
        const unsigned short* r = static_cast<const unsigned short*>(base) + (size_t)row * nu + col;
        return unsigned(r[0]) | (unsigned(r[1]) << 16);
    }
}

template <int STORAGE>
__device__ __forceinline__ void prefetch_tile(Prefetch<STORAGE>& pf, const StageMap& m,
    const void* __restrict__ base, const float* __restrict__ scales, int nu, int nv, short4 t)
{
    const int cl = min(max(t.x + m.col, 0), nu - 2);
    const int rbase = t.y + m.r0;
    const int rows_per = m.rows_per;  // uniform across the workgroup
#pragma unroll
    for (int j = 0; j < kMaxRows; ++j) {
        if (j <= rows_per) {
            const int row = min(max(rbase + j, 0), nv - 1);
            pf.v[j] = raw_pair<STORAGE>(base, nu, row, cl);
            if constexpr (STORAGE != YK_F32) pf.f[j] = scales[row];
        }
    }
}

// shift = requested first column - loaded first column (0 inside the detector)
template <int STORAGE>
__device__ __forceinline__ typename Pair<STORAGE>::T fix_pair(typename Pair<STORAGE>::T v, int shift, bool row_ok)
{
    if constexpr (STORAGE == YK_F32) {
        float2 r;
        r.x = shift == 0 ? v.x : (shift == 1 ? v.y : 0.f);
        r.y = shift == 0 ? v.y : (shift == -1 ? v.x : 0.f);
        return row_ok ? r : make_float2(0.f, 0.f);
    } else {
        unsigned r = shift == 0 ? v : (shift == 1 ? (v >> 16) : (shift == -1 ? (v << 16) : 0u));
        return row_ok ? r : 0u;
    }
}

template <int STORAGE>
__device__ __forceinline__ void commit_tile(void* lds, const Prefetch<STORAGE>& pf, const StageMap& m,
    short4 t, int nu, int nv, float inv_g)
{
    if (m.rows <= 0) return;
    const int col0 = t.x + m.col;
    const int shift = col0 - min(max(col0, 0), nu - 2);
    const int rbase = t.y + m.r0;
    if (tile_inside(t, nu, nv)) {
#pragma unroll
        for (int j = 0; j + 1 < kMaxRows; ++j) {
            if (j < m.rows) {
                float ft = 1.f, fb = 1.f;
                if constexpr (STORAGE != YK_F32) { ft = pf.f[j] * inv_g; fb = pf.f[j + 1] * inv_g; }
                write_quad<STORAGE>(lds, (m.r0 + j) * t.z + m.col, pf.v[j], pf.v[j + 1], ft, fb);
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j + 1 < kMaxRows; ++j) {
            if (j < m.rows) {
                const int rt = rbase + j, rb = rt + 1;
                const bool okt = rt >= 0 && rt < nv, okb = rb >= 0 && rb < nv;
                float ft = 1.f, fb = 1.f;
                if constexpr (STORAGE != YK_F32) { ft = pf.f[j] * inv_g; fb = pf.f[j + 1] * inv_g; }
                write_quad<STORAGE>(lds, (m.r0 + j) * t.z + m.col, fix_pair<STORAGE>(pf.v[j], shift, okt),
                    fix_pair<STORAGE>(pf.v[j + 1], shift, okb), ft, fb);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// per-view accumulation into acc[BZ]
// ---------------------------------------------------------------------------
constexpr int kBatch = 8;  // LDS reads issued before their results are consumed

template <int STORAGE, int BZ, bool GENERAL>
__device__ __forceinline__ void accumulate_view(const void* lds, const BpCoeff& c, short4 t,
    float X, float Y, float Z0, float vz, float (&acc)[BZ])
{
    const float ul0 = float(t.x), vl0 = float(t.y);
    const int rowq = t.z;
    // distance weight: the radial ray has no z component, so it is z-invariant
    const float depth = c.cr[3] + c.cr[0] * X + c.cr[1] * Y;
    const float rd = __builtin_amdgcn_rcpf(depth);
    const float w = c.wbase * rd * rd;
    if constexpr (!GENERAL) {
        const float den = c.cd[3] + c.cd[0] * X + c.cd[1] * Y;
        const float r = __builtin_amdgcn_rcpf(den);
        const float u = (c.cu[3] + c.cu[0] * X + c.cu[1] * Y) * r - ul0;
        const float v0 = (c.cv[3] + c.cv[0] * X + c.cv[1] * Y + c.cv[2] * Z0) * r - vl0;
        const float dv = c.cv[2] * vz * r;
        const int iu = int(u);
        const float fu = u - float(iu);
        const float gw = (1.f - fu) * w, fw = fu * w;
        if constexpr (STORAGE == YK_F32) {
            const char* base = static_cast<const char*>(lds) + iu * 16;
            const int rowb = rowq * 16;
#pragma unroll
            for (int k0 = 0; k0 < BZ; k0 += kBatch) {
                float4 q[kBatch];
                float fv[kBatch];
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const float v = fmaf(float(k0 + j), dv, v0);
                    fv[j] = __builtin_amdgcn_fractf(v);
                    q[j] = *reinterpret_cast<const float4*>(base + __mul24(int(v), rowb));
                }
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const float c0 = fmaf(fv[j], q[j].z, q[j].x);
                    const float c1 = fmaf(fv[j], q[j].w, q[j].y);
                    acc[k0 + j] = fmaf(c1, fw, fmaf(c0, gw, acc[k0 + j]));
                }
            }
        } else {
            const char* base = static_cast<const char*>(lds) + iu * 8;
            const int rowb = rowq * 8;
            const half2_t gf = { (_Float16)gw, (_Float16)fw };
#pragma unroll
            for (int k0 = 0; k0 < BZ; k0 += kBatch) {
                uint2 q[kBatch];
                half2_t fp[kBatch / 2];
#pragma unroll
                for (int j = 0; j < kBatch; j += 2) {
                    const float va = fmaf(float(k0 + j), dv, v0);
                    const float vb = fmaf(float(k0 + j + 1), dv, v0);
                    fp[j / 2] = __builtin_bit_cast(half2_t,
                        __builtin_amdgcn_cvt_pkrtz(__builtin_amdgcn_fractf(va), __builtin_amdgcn_fractf(vb)));
                    q[j] = *reinterpret_cast<const uint2*>(base + __mul24(int(va), rowb));
                    q[j + 1] = *reinterpret_cast<const uint2*>(base + __mul24(int(vb), rowb));
                }
#pragma unroll
                for (int j = 0; j < kBatch; j += 2) {
                    const half2_t fa = { fp[j / 2].x, fp[j / 2].x };
                    const half2_t fb = { fp[j / 2].y, fp[j / 2].y };
                    const half2_t ca = __builtin_bit_cast(half2_t, q[j].y) * fa + __builtin_bit_cast(half2_t, q[j].x);
                    const half2_t cb = __builtin_bit_cast(half2_t, q[j + 1].y) * fb +
                        __builtin_bit_cast(half2_t, q[j + 1].x);
                    acc[k0 + j] = __builtin_amdgcn_fdot2(ca, gf, acc[k0 + j], false);
                    acc[k0 + j + 1] = __builtin_amdgcn_fdot2(cb, gf, acc[k0 + j + 1], false);
                }
            }
        }
    } else {
        // general perspective path: u, v per sample
        float den = c.cd[3] + c.cd[0] * X + c.cd[1] * Y + c.cd[2] * Z0;
        float un = c.cu[3] + c.cu[0] * X + c.cu[1] * Y + c.cu[2] * Z0;
        float vn = c.cv[3] + c.cv[0] * X + c.cv[1] * Y + c.cv[2] * Z0;
        const float dd = c.cd[2] * vz, du = c.cu[2] * vz, dvn = c.cv[2] * vz;
#pragma unroll 4
        for (int k = 0; k < BZ; ++k) {
            const float r = __builtin_amdgcn_rcpf(den);
            const float u = un * r - ul0;
            const float v = vn * r - vl0;
            const int iu = int(u), iv = int(v);
            const float fu = u - float(iu), fv = v - float(iv);
            float c0, c1;
            if constexpr (STORAGE == YK_F32) {
                const float4 q = static_cast<const float4*>(lds)[iv * rowq + iu];
                c0 = fmaf(fv, q.z, q.x);
                c1 = fmaf(fv, q.w, q.y);
            } else {
                const uint2 q = static_cast<const uint2*>(lds)[iv * rowq + iu];
                const half2_t h = __builtin_bit_cast(half2_t, q.x);
                const half2_t d = __builtin_bit_cast(half2_t, q.y);
                c0 = fmaf(fv, float(d.x), float(h.x));
                c1 = fmaf(fv, float(d.y), float(h.y));
            }
            acc[k] = fmaf(w, fmaf(fu, c1 - c0, c0), acc[k]);
            den += dd;
            un += du;
            vn += dvn;
        }
    }
}

template <int STORAGE, int BZ, bool GENERAL, bool KAHAN>
__global__ void __launch_bounds__(kThreads)
fdk_bp_kernel(const void* __restrict__ proj, const float* __restrict__ row_scales,
    const int* __restrict__ gscale_bits, const BpCoeff* __restrict__ coeffs,
    const short4* __restrict__ tiles, int tile_stride, int tile_offset,
    float* __restrict__ volume, VolumeGeom vol, int nu, int nv, int count, int accumulate)
{
    extern __shared__ float4 lds_raw[];
    void* lds = lds_raw;
    const int bxi = blockIdx.x, byi = blockIdx.y, bzi = blockIdx.z;
    const int x = bxi * kBX + (threadIdx.x % kBX);
    const int y = byi * kBY + (threadIdx.x / kBX);
    const int z0 = bzi * BZ;
    const bool active = x < vol.nx && y < vol.ny;
    const float X = vol.ox + x * vol.vx;
    const float Y = vol.oy + y * vol.vy;
    const float Z0 = vol.oz + z0 * vol.vz;
    const size_t block = ((size_t)bzi * gridDim.y + byi) * gridDim.x + bxi;
    const short4* my_tiles = tiles + block * tile_stride + tile_offset;
    const size_t view_elems = (size_t)nu * nv;
    const size_t elem_bytes = STORAGE == YK_F32 ? 4 : 2;
    auto view_base = [&](int i) {
        return static_cast<const void*>(static_cast<const char*>(proj) + view_elems * elem_bytes * i);
    };
    auto view_scales = [&](int i) { return row_scales ? row_scales + (size_t)i * nv : nullptr; };

    float g = 1.f, inv_g = 1.f;
    if constexpr (STORAGE != YK_F32) {
        g = __int_as_float(*gscale_bits);
        if (!(g > 0.f)) g = 1.f;
        inv_g = 1.f / g;
    }

    float acc[BZ];
#pragma unroll
    for (int k = 0; k < BZ; ++k) acc[k] = 0.f;
    float sum[KAHAN ? BZ : 1], comp[KAHAN ? BZ : 1];
    if constexpr (KAHAN) {
#pragma unroll
        for (int k = 0; k < BZ; ++k) { sum[k] = 0.f; comp[k] = 0.f; }
    }

    // stage view 0
    Prefetch<STORAGE> pf;
    short4 t_cur = my_tiles[0];
    if (t_cur.z > 0) {
        const StageMap m = stage_map(t_cur);
        if (m.fast) {
            prefetch_tile<STORAGE>(pf, m, view_base(0), view_scales(0), nu, nv, t_cur);
            commit_tile<STORAGE>(lds, pf, m, t_cur, nu, nv, inv_g);
        } else {
            stage_direct<STORAGE>(lds, view_base(0), view_scales(0), nu, nv, t_cur, inv_g);
        }
    }
    __syncthreads();

#pragma unroll 1
    for (int i = 0; i < count; ++i) {
        const short4 t_next = (i + 1 < count) ? my_tiles[i + 1] : make_short4(0, 0, 0, 0);
        StageMap m_next;
        m_next.fast = false;
        m_next.rows = 0;
        if (t_next.z > 0) {
            m_next = stage_map(t_next);
            if (m_next.fast)
                prefetch_tile<STORAGE>(pf, m_next, view_base(i + 1), view_scales(i + 1), nu, nv, t_next);
        }
        if (t_cur.z > 0 && active)
            accumulate_view<STORAGE, BZ, GENERAL>(lds, coeffs[i], t_cur, X, Y, Z0, vol.vz, acc);
        if constexpr (KAHAN) {
            if ((i + 1) % kKahanViews == 0 || i + 1 == count) {
#pragma unroll
                for (int k = 0; k < BZ; ++k) {
                    const float yk = acc[k] - comp[k];
                    const float tk = sum[k] + yk;
                    comp[k] = (tk - sum[k]) - yk;
                    sum[k] = tk;
                    acc[k] = 0.f;
                }
            }
        }
        if (t_next.z > 0) {
            __syncthreads();
            if (m_next.fast) commit_tile<STORAGE>(lds, pf, m_next, t_next, nu, nv, inv_g);
            else stage_direct<STORAGE>(lds, view_base(i + 1), view_scales(i + 1), nu, nv, t_next, inv_g);
            __syncthreads();
        }
        t_cur = t_next;
    }
    if (!active) return;
    const size_t plane = (size_t)vol.nx * vol.ny;
    const size_t col = (size_t)y * vol.nx + x;
#pragma unroll
    for (int k = 0; k < BZ; ++k) {
        const int z = z0 + k;
        if (z >= vol.nz) break;
        float val;
        if constexpr (KAHAN) val = sum[k] - comp[k];
        else val = acc[k];
        val *= g;
        float* dst = volume + (size_t)z * plane + col;
        *dst = accumulate ? *dst + val : val;
    }
}

template <int STORAGE, int BZ, bool GENERAL, bool KAHAN>
void launch_bz(const BpLaunch& p, const int* gbits, hipStream_t s)
{
    dim3 grid(p.nbx, p.nby, p.nbz), block(kThreads);
    const size_t lds = size_t(p.lds_quads) * (STORAGE == YK_F32 ? 16 : 8);
    auto k = fdk_bp_kernel<STORAGE, BZ, GENERAL, KAHAN>;
    YK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    hipLaunchKernelGGL(k, grid, block, lds, s, p.proj, p.row_scales, gbits, p.coeffs, p.tiles,
        p.tile_view_stride, p.tile_view_offset, p.volume, p.vol, p.nu, p.nv, p.count, p.accumulate);
}

template <int STORAGE, bool GENERAL, bool KAHAN>
void launch_variant(const BpLaunch& p, const int* gbits, hipStream_t s, int bz)
{
    if (bz == 32) launch_bz<STORAGE, 32, GENERAL, KAHAN>(p, gbits, s);
    else if (bz == 16) launch_bz<STORAGE, 16, GENERAL, KAHAN>(p, gbits, s);
    else if (bz == 8) launch_bz<STORAGE, 8, GENERAL, KAHAN>(p, gbits, s);
    else fail(YK_ERR_UNSUPPORTED, "unsupported block depth");
}

template <int STORAGE>
void launch_storage(const BpLaunch& p, const int* gbits, hipStream_t s, int bz)
{
    if (p.general) {
        if (p.kahan) launch_variant<STORAGE, true, true>(p, gbits, s, bz);
        else launch_variant<STORAGE, true, false>(p, gbits, s, bz);
    } else {
        if (p.kahan) launch_variant<STORAGE, false, true>(p, gbits, s, bz);
        else launch_variant<STORAGE, false, false>(p, gbits, s, bz);
    }
}

}  // namespace

int bp_block_z(int general) { return general ? 16 : 32; }

void launch_bp_tiles(const TileLaunch& p, hipStream_t s)
{
    const long long total = (long long)p.nbx * p.nby * p.nbz * p.n_views;
    const int threads = 256;
    const long long blocks = (total + threads - 1) / threads;
    if (blocks > 0x7fffffffLL) fail(YK_ERR_UNSUPPORTED, "footprint table too large");
    hipLaunchKernelGGL(set_ints_kernel, dim3(1), dim3(64), 0, s, p.max_quads, 4, 0);
    hipLaunchKernelGGL(bp_tiles_kernel, dim3(unsigned(blocks)), dim3(threads), 0, s, p.coeff64,
        p.n_views, p.vol, p.bx, p.by, p.bz, p.nbx, p.nby, p.nbz, p.nu, p.nv, p.tiles, p.max_quads);
    YK_HIP(hipGetLastError());
}

void launch_max_scale(const float* scales, int n, int* out_bits, hipStream_t s)
{
    hipLaunchKernelGGL(set_ints_kernel, dim3(1), dim3(64), 0, s, out_bits, 1, 0);
    const int threads = 256;
    const int blocks = std::max(1, std::min(1024, (n + threads - 1) / threads));
    hipLaunchKernelGGL(max_scale_kernel, dim3(blocks), dim3(threads), 0, s, scales, n, out_bits);
    YK_HIP(hipGetLastError());
}

void launch_fdk_bp(const BpLaunch& p, hipStream_t s)
{
    const int bz = p.bz;
    const int* gbits = p.gscale_bits;
    switch (p.storage) {
    case YK_F32: launch_storage<YK_F32>(p, gbits, s, bz); break;
    case YK_F16: launch_storage<YK_F16>(p, gbits, s, bz); break;
    case YK_BF16: launch_storage<YK_BF16>(p, gbits, s, bz); break;
    default: fail(YK_ERR_INVALID_CONFIG, "unknown storage type");
    }
    YK_HIP(hipGetLastError());
}

}  // namespace yk

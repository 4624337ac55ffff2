// Voxel-driven FDK backprojection for gfx950.
//
// Work decomposition: a workgroup owns a 16 x 16 column tile of voxels and a
// run of BZ consecutive z slices; each lane owns one (x, y) column and keeps
// BZ accumulators in registers for the whole view loop, so the volume is read
// and written once per launch.  Per view the workgroup stages the detector
// footprint of its voxel block into LDS as "quads": for texel (r, c) the four
// values (q[r][c], q[r][c+1], q[r+1][c]-q[r][c], q[r+1][c+1]-q[r][c+1]), so
// one LDS read feeds a whole bilinear sample.  The footprint of every
// (block, view) pair is computed once per plan in FP64 (bp_tiles_kernel).
// The texels of view i+1 are loaded into registers while view i is
// accumulated, then written to LDS between two barriers.
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
#include <mutex>

#include "yk_internal.hpp"

// Fixed scratch of the allocation-free backprojection entry point
// (yk_fdk_backproject_noalloc): word 0 receives the global row-scale bits,
// words [kNoAllocHeader, kNoAllocHeader + kNoAllocRowFactors) the normalized
// row factors of the views being backprojected.  One copy per GPU, held until
// the process ends.  In host code `scratch` names a host-side stand-in, never
// the device memory: kernels get the address bp_noalloc_scratch() looked up.
static __device__ __align__(256) unsigned scratch[yk::kNoAllocHeader + yk::kNoAllocRowFactors];

namespace yk {

unsigned* bp_noalloc_scratch(int device)
{
    // looked up once per device, with that device current (the caller's
    // hipSetDevice); each GPU holds its own copy of the array
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static unsigned* addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) fail(YK_ERR_INVALID_CONFIG, "device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        int cur = -1;
        YK_HIP(hipGetDevice(&cur));
        if (cur != device) fail(YK_ERR_DEVICE, "scratch lookup with another device current");
        void* p = nullptr;
        YK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) fail(YK_ERR_DEVICE, "no device address for the backprojection scratch");
        addr[device] = static_cast<unsigned*>(p);
    }
    return addr[device];
}

namespace {

// internal storage tag of the kernel: half texels already multiplied by their
// normalized row factor (normalize_rows_kernel), so staging applies nothing
constexpr int kHalfNorm = 16;
constexpr bool has_factors(int storage) { return storage == YK_F16 || storage == YK_BF16; }

constexpr int kBX = 16;
constexpr int kBY = 16;
constexpr int kThreads = kBX * kBY;
constexpr int kKahanViews = 64;
constexpr int kMaxRows = 12;  // texel rows a thread prefetches (quad rows + 1)
constexpr double kFootMargin = 1.0 / 32.0;  // pixels of margin around a block's footprint

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// staging layout of a tile of qw x qh quads: lanes take `span` consecutive
// columns, `chunks` row ranges of `rows_per` quad rows each
struct StageShape {
    int span, chunks, rows_per;
};

__host__ __device__ __forceinline__ StageShape stage_shape(int qw, int qh)
{
    StageShape s;
    s.span = qw < kThreads ? qw : kThreads;
    s.chunks = kThreads / s.span;
    s.rows_per = (qh + s.chunks - 1) / s.chunks;
    return s;
}

// LDS row pitch in quads: odd, so that lanes reading the same column of
// neighbouring rows fall on different banks
__host__ __device__ __forceinline__ int lds_pitch(int qw) { return qw | 1; }

// quads of LDS a tile needs: the staging writes rows [0, qh) only (the last
// row range is moved back to end at qh, see stage_map)
__host__ __device__ __forceinline__ int stage_quads(int qw, int qh) { return lds_pitch(qw) * qh; }

// ---------------------------------------------------------------------------
// footprint table: per (block, view) the quad range [u0, u0+w) x [v0, v0+h)
// that contains floor(u), floor(v) of every voxel of the block, widened by
// kFootMargin pixels on each side; w = h = 0 when the block misses the
// detector.  The kernel's FP32 u, v stay within ~1e-3 pixel of the FP64
// values here (FP32 coefficients, one reciprocal, at most 32 incremental
// steps), so a 1/32-pixel margin keeps every sample inside its tile.
// ---------------------------------------------------------------------------
__global__ void bp_tiles_kernel(const double* __restrict__ c64, int n_views, VolumeGeom vol,
    int bx, int by, int bz, int nbx, int nby, int nbz, int nu, int nv,
    short4* __restrict__ tiles, uint2* __restrict__ stage, int* __restrict__ maxes)
{
    const long long total = (long long)nbx * nby * nbz * n_views;
    const long long idx_raw = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    // every lane stays to the wave reduction below; lanes past the end
    // evaluate entry 0 and store nothing
    const bool valid = idx_raw < total;
    const long long idx = valid ? idx_raw : 0;
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
    uint2 st = make_uint2(0u, 0u);
    int m_w = 0, m_h = 0, m_q = 0;
    if (!bad) {
        const double fu0 = floor(umin - kFootMargin), fu1 = floor(umax + kFootMargin);
        const double fv0 = floor(vmin - kFootMargin), fv1 = floor(vmax + kFootMargin);
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
                m_w = int(t.z);
                m_h = int(t.w);
                m_q = stage_quads(t.z, t.w);
                // staging shape, so the backprojection divides nothing per view:
                // x = 16-bit reciprocal of the span, y = chunks | rows_per << 16
                const StageShape sh = stage_shape(t.z, t.w);
                st.x = (65536u + unsigned(sh.span) - 1u) / unsigned(sh.span);
                st.y = unsigned(sh.chunks) | (unsigned(min(sh.rows_per, 65535)) << 16);
            }
        }
    }
    // one atomic per wave and counter (a global atomic per entry serializes
    // millions of updates on four addresses)
    int m_bad = valid && bad ? 1 : 0;
    if (!valid) m_w = m_h = m_q = 0;
    for (int off = 32; off > 0; off >>= 1) {
        m_w = max(m_w, __shfl_xor(m_w, off, 64));
        m_h = max(m_h, __shfl_xor(m_h, off, 64));
        m_q = max(m_q, __shfl_xor(m_q, off, 64));
        m_bad = max(m_bad, __shfl_xor(m_bad, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        if (m_w) atomicMax(&maxes[0], m_w);
        if (m_h) atomicMax(&maxes[1], m_h);
        if (m_q) atomicMax(&maxes[2], m_q);
        if (m_bad) atomicMax(&maxes[3], 1);
    }
    if (!valid) return;
    tiles[idx] = t;
    stage[idx] = st;
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

// normalized row factors: scale / (global max scale), a power of two <= 1.
// half storage: packed (f, f) as two halves (rows below 2^-24 of the largest
// row scale become 0; their texels lie below 2^-22 of the largest texel);
// bfloat16 storage: the float factor.
__global__ void row_factor_kernel(const float* __restrict__ s, int n, const int* __restrict__ gbits,
    int storage, unsigned* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float g = __int_as_float(*gbits);
    if (!(g > 0.f)) g = 1.f;
    const float f = s[i] / g;
    if (storage == YK_F16) {
        const _Float16 h = (_Float16)f;
        const half2_t p = { h, h };
        out[i] = __builtin_bit_cast(unsigned, p);
    } else {
        out[i] = __float_as_uint(f);
    }
}

// half / bfloat16 rows times their normalized row factor, as half texels:
// the values the staging would otherwise form once per (block, view) pair
// (exact for half: the factor is a power of two; rows below 2^-24 of the
// largest row scale become 0 as in row_factor_kernel).  One workgroup per row
// at a time, four texel pairs (16 bytes) per lane and step; nu % 8 == 0.
template <int STORAGE>
__device__ __forceinline__ unsigned normalize_pair(unsigned raw, unsigned f)
{
    half2_t h;
    if constexpr (STORAGE == YK_F16) {
        h = __builtin_bit_cast(half2_t, raw) * __builtin_bit_cast(half2_t, f);
    } else {
        const float ff = __uint_as_float(f);
        h = __builtin_bit_cast(half2_t, __builtin_amdgcn_cvt_pkrtz(
            __uint_as_float(raw << 16) * ff, __uint_as_float(raw & 0xffff0000u) * ff));
    }
    return __builtin_bit_cast(unsigned, h);
}

template <int STORAGE>
__global__ void __launch_bounds__(256)
normalize_rows_kernel(const uint4* __restrict__ in, const unsigned* __restrict__ factors,
    uint4* __restrict__ out, int quads_per_row, int rows)
{
    for (int r = blockIdx.x; r < rows; r += gridDim.x) {
        const unsigned f = factors[r];
        const size_t base = (size_t)r * quads_per_row;
        for (int c = threadIdx.x; c < quads_per_row; c += blockDim.x) {
            const uint4 v = in[base + c];
            out[base + c] = make_uint4(normalize_pair<STORAGE>(v.x, f), normalize_pair<STORAGE>(v.y, f),
                                       normalize_pair<STORAGE>(v.z, f), normalize_pair<STORAGE>(v.w, f));
        }
    }
}

// ---------------------------------------------------------------------------
// staging
// ---------------------------------------------------------------------------
constexpr int kRsrcWord3 = 0x00020000;  // raw buffer, 32-bit data format

// Buffer resource over `bytes` bytes: a load at an offset at or past the end
// (negative offsets wrap to large unsigned values) returns 0, which zeroes
// every texel of a row above or below the detector without a branch.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, int(bytes), kRsrcWord3);
}

struct StageMap {
    StageShape s;
    int col;      // quad column of this lane within the tile
    int cl;       // first loaded texel column, clamped into [0, nu - 2]
    int shift;    // requested - loaded first column: -1, 0 or 1 on the detector edges
    int r0;       // first quad row of this lane (rows [r0, r0 + rows_per) lie in [0, qh))
    bool lane_on; // lane has a row range
    bool fast;    // tile fits the register prefetch
};

__device__ __forceinline__ StageMap stage_map(short4 t, uint2 st, int nu)
{
    StageMap m;
    m.s.span = min(int(t.z), kThreads);
    m.s.chunks = int(st.y & 0xffffu);
    m.s.rows_per = __builtin_amdgcn_readfirstlane(int(st.y >> 16));
    // threadIdx.x / span by the 16-bit reciprocal st.x: exact for
    // threadIdx.x < 256 and span <= 256 (the rounding excess stays below 1/256)
    const int rc = int(__umul24(threadIdx.x, st.x) >> 16);
    m.col = int(threadIdx.x) - __mul24(rc, m.s.span);
    const int col = t.x + m.col;
    m.cl = min(max(col, 0), nu - 2);
    m.shift = col - m.cl;
    // the last row range ends at qh: it overlaps its neighbour (the rows
    // written twice get the same values) instead of spilling past the tile
    m.r0 = min(__mul24(rc, m.s.rows_per), int(t.w) - m.s.rows_per);
    m.lane_on = rc < m.s.chunks;
    m.fast = t.z <= kThreads && m.s.rows_per + 1 <= kMaxRows;
    return m;
}

template <int STORAGE>
struct Pair;
template <>
struct Pair<YK_F32> { using T = float2; };
template <>
struct Pair<YK_F16> { using T = unsigned; };
template <>
struct Pair<YK_BF16> { using T = unsigned; };
template <>
struct Pair<kHalfNorm> { using T = unsigned; };

// texel pair (col, col + 1) of one row at byte offset `off`: float2, or the
// two 16-bit texels as one dword (lo = col), read by a single dword load at a
// 2-byte aligned offset (the queues of ROCm run with unaligned buffer access
// enabled).  The pair never straddles two rows (the column is clamped), so it
// is either wholly in range or wholly zero.
template <int STORAGE>
__device__ __forceinline__ typename Pair<STORAGE>::T load_pair(__amdgpu_buffer_rsrc_t r, unsigned off)
{
    if constexpr (STORAGE == YK_F32) {
        return make_float2(__uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0)),
                           __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off + 4, 0, 0)));
    } else {
        return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
    }
}

// byte offset of a row factor; the empty asm keeps the compiler from moving
// the row step into the instruction offset, so a negative row stays a
// wrapped (out-of-range) offset rather than summing back into range
__device__ __forceinline__ unsigned factor_offset(int row)
{
    unsigned o = unsigned(row) * 4u;
    asm volatile("" : "+v"(o));
    return o;
}

template <int STORAGE>
struct Prefetch {
    typename Pair<STORAGE>::T v[kMaxRows];
    unsigned f[has_factors(STORAGE) ? kMaxRows : 1];
};

// Loads only: nothing consumes the loaded values before commit, so every load
// stays in flight while the current view is accumulated.
template <int STORAGE>
__device__ __forceinline__ void prefetch_tile(Prefetch<STORAGE>& pf, const StageMap& m,
    __amdgpu_buffer_rsrc_t data, __amdgpu_buffer_rsrc_t fac, int nu, short4 t)
{
    if (!m.lane_on) return;
    constexpr int eb = STORAGE == YK_F32 ? 4 : 2;
    const int row0 = t.y + m.r0;
    // byte offset of the first row; the empty asm keeps the row steps below
    // from being folded back into a per-row multiply
    int off0 = (__mul24(row0, nu) + m.cl) * eb;
    asm volatile("" : "+v"(off0));
    const int rowb = nu * eb;
    const int rows_per = m.s.rows_per;  // uniform across the workgroup
#pragma unroll
    for (int j = 0; j < kMaxRows; ++j) {
        if (j <= rows_per) {
            pf.v[j] = load_pair<STORAGE>(data, unsigned(off0 + j * rowb));
            if constexpr (has_factors(STORAGE))
                pf.f[j] = __builtin_amdgcn_raw_buffer_load_b32(fac, factor_offset(row0 + j), 0, 0);
        }
    }
}

// move the pair onto the requested columns, zeroing the off-detector texel
template <int STORAGE>
__device__ __forceinline__ typename Pair<STORAGE>::T shift_pair(typename Pair<STORAGE>::T v, int shift)
{
    if constexpr (STORAGE == YK_F32) {
        float2 r;
        r.x = shift == 0 ? v.x : (shift == 1 ? v.y : 0.f);
        r.y = shift == 0 ? v.y : (shift == -1 ? v.x : 0.f);
        return r;
    } else {
        return shift == 0 ? v : (shift == 1 ? (v >> 16) : (shift == -1 ? (v << 16) : 0u));
    }
}

// normalized half pair of a raw pair and its row factor
template <int STORAGE>
__device__ __forceinline__ half2_t to_half2(unsigned raw, unsigned f)
{
    if constexpr (STORAGE == kHalfNorm) {
        (void)f;
        return __builtin_bit_cast(half2_t, raw);
    } else if constexpr (STORAGE == YK_F16) {
        return __builtin_bit_cast(half2_t, raw) * __builtin_bit_cast(half2_t, f);
    } else {
        const float ff = __uint_as_float(f);
        return __builtin_bit_cast(half2_t, __builtin_amdgcn_cvt_pkrtz(
            __uint_as_float(raw << 16) * ff, __uint_as_float(raw & 0xffff0000u) * ff));
    }
}

// Write quad rows [r0, r0 + rows_per) of this lane's column.  Q32: float
// quads (always for float storage; half and bfloat16 texels are widened to
// FP32 here for the general-geometry path, whose per-sample arithmetic is
// FP32)
template <int STORAGE, bool Q32>
__device__ __forceinline__ void commit_tile(void* lds, Prefetch<STORAGE>& pf, const StageMap& m, short4 t)
{
    if (!m.lane_on) return;
    const int rows_per = m.s.rows_per;
    if (m.shift != 0) {  // lanes on the detector's first or last column only
#pragma unroll
        for (int j = 0; j < kMaxRows; ++j)
            if (j <= rows_per) pf.v[j] = shift_pair<STORAGE>(pf.v[j], m.shift);
    }
    constexpr int qb = Q32 ? 16 : 8;
    const int pitch = lds_pitch(t.z);
    char* dst = static_cast<char*>(lds) + unsigned(__mul24(m.r0, pitch) + m.col) * unsigned(qb);
    const int rowb = pitch * qb;
    if constexpr (STORAGE == YK_F32) {
#pragma unroll
        for (int j = 0; j + 1 < kMaxRows; ++j) {
            if (j < rows_per) {
                const float2 a = pf.v[j], b = pf.v[j + 1];
                *reinterpret_cast<float4*>(dst + j * rowb) = make_float4(a.x, a.y, b.x - a.x, b.y - a.y);
            }
        }
    } else if constexpr (Q32) {
        float2 x[kMaxRows];
#pragma unroll
        for (int j = 0; j < kMaxRows; ++j) {
            if (j <= rows_per) {
                const half2_t h = to_half2<STORAGE>(pf.v[j], pf.f[has_factors(STORAGE) ? j : 0]);
                x[j] = make_float2(float(h.x), float(h.y));
            }
        }
#pragma unroll
        for (int j = 0; j + 1 < kMaxRows; ++j) {
            if (j < rows_per) {
                const float2 a = x[j], b = x[j + 1];
                *reinterpret_cast<float4*>(dst + j * rowb) = make_float4(a.x, a.y, b.x - a.x, b.y - a.y);
            }
        }
    } else {
        half2_t x[kMaxRows];
#pragma unroll
        for (int j = 0; j < kMaxRows; ++j)
            if (j <= rows_per) x[j] = to_half2<STORAGE>(pf.v[j], pf.f[has_factors(STORAGE) ? j : 0]);
#pragma unroll
        for (int j = 0; j + 1 < kMaxRows; ++j) {
            if (j < rows_per) {
                uint2 q;
                q.x = __builtin_bit_cast(unsigned, x[j]);
                q.y = __builtin_bit_cast(unsigned, x[j + 1] - x[j]);
                *reinterpret_cast<uint2*>(dst + j * rowb) = q;
            }
        }
    }
}

// direct staging for tiles too tall or wide for the register prefetch (Q32
// as in commit_tile)
template <int STORAGE, bool Q32>
__device__ void stage_direct(void* lds, __amdgpu_buffer_rsrc_t data, __amdgpu_buffer_rsrc_t fac,
    int nu, short4 t)
{
    constexpr int eb = STORAGE == YK_F32 ? 4 : 2;
    const int qw = t.z, qh = t.w, pitch = lds_pitch(qw);
    for (int e = threadIdx.x; e < qw * qh; e += kThreads) {
        const int r = e / qw, c = e - r * qw;
        const int slot = r * pitch + c;
        const int col = t.x + c;
        const int cl = min(max(col, 0), nu - 2);
        const int rt = t.y + r;
        const auto a = shift_pair<STORAGE>(load_pair<STORAGE>(data, unsigned((rt * nu + cl) * eb)), col - cl);
        const auto b = shift_pair<STORAGE>(load_pair<STORAGE>(data, unsigned(((rt + 1) * nu + cl) * eb)), col - cl);
        if constexpr (STORAGE == YK_F32) {
            static_cast<float4*>(lds)[slot] = make_float4(a.x, a.y, b.x - a.x, b.y - a.y);
        } else {
            half2_t xa, xb;
            if constexpr (has_factors(STORAGE)) {
                xa = to_half2<STORAGE>(a, __builtin_amdgcn_raw_buffer_load_b32(fac, factor_offset(rt), 0, 0));
                xb = to_half2<STORAGE>(b, __builtin_amdgcn_raw_buffer_load_b32(fac, factor_offset(rt + 1), 0, 0));
            } else {
                xa = to_half2<STORAGE>(a, 0u);
                xb = to_half2<STORAGE>(b, 0u);
            }
            if constexpr (Q32) {
                const float a0 = float(xa.x), a1 = float(xa.y);
                static_cast<float4*>(lds)[slot] = make_float4(a0, a1, float(xb.x) - a0, float(xb.y) - a1);
            } else {
                uint2 q;
                q.x = __builtin_bit_cast(unsigned, xa);
                q.y = __builtin_bit_cast(unsigned, xb - xa);
                static_cast<uint2*>(lds)[slot] = q;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// per-view accumulation into acc[BZ]
// ---------------------------------------------------------------------------
constexpr int kBatch = 8;  // LDS reads issued before their results are consumed

// Geometry modes of the kernel: the regular circular detector; any other
// detector with one reciprocal per sample; any other detector whose depth
// denominator changes by at most 2e-4 (relative) per slice (fdk.cpp
// general_mode), where u and v are quadratic in the slice to far below FP32
// resolution of a pixel and are carried by forward differences (the mode's
// name is that of its first form, one Newton step per sample).
constexpr int kRegular = 0, kGeneralRcp = 1, kGeneralNewton = 2;

// float LDS quads: float storage, and every storage on the general path
constexpr bool bp_float_quads(int storage, int geo) { return storage == YK_F32 || geo != kRegular; }

// byte offset of an LDS pointer, and a 16-byte LDS load at a byte offset
typedef __attribute__((address_space(3))) const char lds_cchar;
__device__ __forceinline__ unsigned lds_offset(const void* p)
{
    return unsigned(reinterpret_cast<uintptr_t>((lds_cchar*)p));
}
typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 lds_load_f4(unsigned offset)
{
    const f32x4 v = *(__attribute__((address_space(3))) const f32x4*)(uintptr_t)offset;
    return make_float4(v.x, v.y, v.z, v.w);
}

template <int STORAGE, int BZ, int GEO>
__device__ __forceinline__ void accumulate_view(const void* lds, const BpCoeff& c, short4 t,
    float X, float Y, float Z0, float vz, float (&acc)[BZ])
{
    const float ul0 = float(t.x), vl0 = float(t.y);
    const int rowq = lds_pitch(t.z);
    // distance weight: the radial ray has no z component, so it is z-invariant
    const float depth = c.cr[3] + c.cr[0] * X + c.cr[1] * Y;
    const float rd = __builtin_amdgcn_rcpf(depth);
    const float w = c.wbase * rd * rd;
    if constexpr (GEO == kRegular) {
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
        // float quads for every storage (half texels are widened when staged):
        // this path is VALU-bound, and packed half lerps cost as many cycles
        // per FMA as FP32 ones on gfx950 plus a conversion per sample
        const int rowb = rowq * 16;
        if constexpr (GEO == kGeneralNewton) {
            // The depth denominator d0 + d1 k changes by at most 2e-4
            // (relative) per slice on this path, so u(k) = (n0 + n1 k) /
            // (d0 + d1 k) is quadratic in the slice k to far below 1e-4 pixel
            // over a block, and is carried by forward differences of the
            // tile-local coordinate: four additions per sample instead of a
            // reciprocal step and three products.  The differences are the
            // exact ones, not differences of rounded values (whose reciprocal
            // rounding, ~1e-4 pixel at |u| ~ 1000, the recurrence would
            // multiply by k and k^2):
            //   du0 = u(1) - u(0) = (n1 - u(0) d1) / (d0 + d1),
            //   d2u = du(1) - du(0) = -2 du0 d1 / (d0 + 2 d1).
            // Over BZ slices the position stays within ~1e-4 pixel of the
            // exact quotient, as close as the per-sample reciprocal.
            const float r0 = __builtin_amdgcn_rcpf(den);
            const float r1 = __builtin_amdgcn_rcpf(den + dd);
            const float r2 = __builtin_amdgcn_rcpf(fmaf(2.f, dd, den));
            const float ug = un * r0, vg = vn * r0;
            float u = ug - ul0, v = vg - vl0;
            float su = fmaf(-ug, dd, du) * r1, sv = fmaf(-vg, dd, dvn) * r1;
            const float m2 = -2.f * dd * r2;
            const float s2u = su * m2, s2v = sv * m2;
            const unsigned base = lds_offset(lds);
#pragma unroll
            for (int k = 0; k < BZ; ++k) {
                const float fu = __builtin_amdgcn_fractf(u), fv = __builtin_amdgcn_fractf(v);
                // one lshl_add and one 24-bit mad (u, v >= 0 in the tile); the
                // empty asm keeps the compiler from re-associating the sum
                // into a multiply, a shift and a three-way add
                unsigned ub = (unsigned(int(u)) << 4) + base;
                asm volatile("" : "+v"(ub));
                const unsigned qa = __umul24(unsigned(int(v)), unsigned(rowb)) + ub;
                const float4 q = lds_load_f4(qa);
                const float c0 = fmaf(fv, q.z, q.x);
                const float c1 = fmaf(fv, q.w, q.y);
                acc[k] = fmaf(w, fmaf(fu, c1 - c0, c0), acc[k]);
                u += su;
                v += sv;
                su += s2u;
                sv += s2v;
            }
            return;
        }
        // fully unrolled: a partial unroll indexes acc[] through the register
        // index mode (a move in and out per sample).  (Issuing the LDS reads of
        // eight samples before consuming them measured slower: the extra VGPRs
        // cost more than the exposed latency.)
        const unsigned base = lds_offset(lds);
#pragma unroll
        for (int k = 0; k < BZ; ++k) {
            const float r = __builtin_amdgcn_rcpf(den);
            // u, v >= 1 inside the footprint: truncation is floor
            const float u = un * r - ul0;
            const float v = vn * r - vl0;
            const float fu = __builtin_amdgcn_fractf(u), fv = __builtin_amdgcn_fractf(v);
            unsigned ub = (unsigned(int(u)) << 4) + base;
            asm volatile("" : "+v"(ub));
            const float4 q = lds_load_f4(__umul24(unsigned(int(v)), unsigned(rowb)) + ub);
            const float c0 = fmaf(fv, q.z, q.x);
            const float c1 = fmaf(fv, q.w, q.y);
            acc[k] = fmaf(w, fmaf(fu, c1 - c0, c0), acc[k]);
            den += dd;
            un += du;
            vn += dvn;
        }
    }
}

// workgroups per CU the register allocation must leave room for: four (128
// VGPRs) for the half / bfloat16 Kahan variant, whose state would otherwise
// just exceed that budget; the others fit it or need more on their own
constexpr int bp_min_blocks(int storage, int bz, bool kahan)
{
    return (storage != YK_F32 && kahan && bz == 32) ? 4 : 1;
}

template <int STORAGE, int BZ, int GEO, bool KAHAN>
__global__ void __launch_bounds__(kThreads, bp_min_blocks(STORAGE, BZ, KAHAN))
fdk_bp_kernel(const void* __restrict__ proj, const unsigned* __restrict__ row_factors,
    const int* __restrict__ gscale_bits, const BpCoeff* __restrict__ coeffs,
    const short4* __restrict__ tiles, const uint2* __restrict__ stage, int tile_stride, int tile_offset,
    float* __restrict__ volume, VolumeGeom vol, int nu, int nv, int count, int accumulate,
    unsigned buf_bytes, int double_buffer)
{
    // one LDS tile buffer, or two alternating ones (then a single barrier per
    // view orders the writes of view i+1 after every read of view i-1)
    constexpr bool Q32 = bp_float_quads(STORAGE, GEO);
    extern __shared__ float4 lds_raw[];
    char* const lds0 = reinterpret_cast<char*>(lds_raw);
    unsigned cur = 0, nxt = double_buffer ? buf_bytes : 0u;
    const int bxi = blockIdx.x, byi = blockIdx.y, bzi = blockIdx.z;
    // lane -> column: each wave covers 16 x 4 columns, and every 8 consecutive
    // lanes a 4 x 2 patch, so the lanes one LDS cycle serves read nearby quads
    // of distinct banks whatever the view angle
    const int l = int(threadIdx.x) & 63;
    const int x = bxi * kBX + ((l & 3) | (((l >> 3) & 3) << 2));
    const int y = byi * kBY + 4 * (int(threadIdx.x) >> 6) + (((l >> 2) & 1) | ((l >> 5) << 1));
    const int z0 = bzi * BZ;
    const bool active = x < vol.nx && y < vol.ny;
    const float X = vol.ox + x * vol.vx;
    const float Y = vol.oy + y * vol.vy;
    const float Z0 = vol.oz + z0 * vol.vz;
    const size_t block = ((size_t)bzi * gridDim.y + byi) * gridDim.x + bxi;
    const short4* my_tiles = tiles + block * tile_stride + tile_offset;
    const uint2* my_stage = stage + block * tile_stride + tile_offset;
    const unsigned view_bytes = unsigned(nu) * unsigned(nv) * (STORAGE == YK_F32 ? 4u : 2u);
    auto data_rsrc = [&](int i) {
        return make_rsrc(static_cast<const char*>(proj) + (size_t)view_bytes * i, view_bytes);
    };
    auto fac_rsrc = [&](int i) {
        return make_rsrc(row_factors + (has_factors(STORAGE) ? (size_t)nv * i : 0), unsigned(nv) * 4u);
    };

    float g = 1.f;
    if constexpr (STORAGE != YK_F32) {
        g = __int_as_float(*gscale_bits);
        if (!(g > 0.f)) g = 1.f;
    }

    float acc[BZ];
#pragma unroll
    for (int k = 0; k < BZ; ++k) acc[k] = 0.f;
    // Kahan summation over blocks of kKahanViews views: acc holds the block
    // partial, sum the running total; the rounding error of each sum += acc is
    // carried into the next block by starting acc at minus that error (the
    // compensation of Kahan's y = x - c, kept in acc instead of a third array)
    float sum[KAHAN ? BZ : 1];
    if constexpr (KAHAN) {
#pragma unroll
        for (int k = 0; k < BZ; ++k) sum[k] = 0.f;
    }

    // stage view 0
    Prefetch<STORAGE> pf;
    short4 t_cur = my_tiles[0];
    if (t_cur.z > 0) {
        const StageMap m = stage_map(t_cur, my_stage[0], nu);
        if (m.fast) {
            prefetch_tile<STORAGE>(pf, m, data_rsrc(0), fac_rsrc(0), nu, t_cur);
            commit_tile<STORAGE, Q32>(lds0, pf, m, t_cur);
        } else {
            stage_direct<STORAGE, Q32>(lds0, data_rsrc(0), fac_rsrc(0), nu, t_cur);
        }
    }
    __syncthreads();

#pragma unroll 1
    for (int i = 0; i < count; ++i) {
        const short4 t_next = (i + 1 < count) ? my_tiles[i + 1] : make_short4(0, 0, 0, 0);
        StageMap m_next{};
        // The staging steps (issuing the next view's loads; the barrier,
        // commit and barrier after this view) run at a raised wave priority,
        // so the scheduler issues them ahead of the other workgroups'
        // sampling and the workgroup spends less time held at its barriers:
        // FP16 -3 % at 512^3 x 720, FP32 unchanged.
        if (t_next.z > 0) {
            __builtin_amdgcn_s_setprio(3);
            m_next = stage_map(t_next, my_stage[i + 1], nu);
            if (m_next.fast)
                prefetch_tile<STORAGE>(pf, m_next, data_rsrc(i + 1), fac_rsrc(i + 1), nu, t_next);
            __builtin_amdgcn_s_setprio(0);
        }
        if (t_cur.z > 0 && active)
            accumulate_view<STORAGE, BZ, GEO>(lds0 + cur, coeffs[i], t_cur, X, Y, Z0, vol.vz, acc);
        if constexpr (KAHAN) {
            if ((i + 1) % kKahanViews == 0 || i + 1 == count) {
#pragma unroll
                for (int k = 0; k < BZ; ++k) {
                    const float t = sum[k] + acc[k];
                    acc[k] = acc[k] - (t - sum[k]);  // what t lost of acc
                    sum[k] = t;
                }
            }
        }
        if (t_next.z > 0) {
            if (!double_buffer) __syncthreads();
            __builtin_amdgcn_s_setprio(3);
            if (m_next.fast) commit_tile<STORAGE, Q32>(lds0 + nxt, pf, m_next, t_next);
            else stage_direct<STORAGE, Q32>(lds0 + nxt, data_rsrc(i + 1), fac_rsrc(i + 1), nu, t_next);
            __syncthreads();
            __builtin_amdgcn_s_setprio(0);
            const unsigned t = cur;
            cur = nxt;
            nxt = t;
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
        if constexpr (KAHAN) val = sum[k] + acc[k];
        else val = acc[k];
        val *= g;
        float* dst = volume + (size_t)z * plane + col;
        *dst = accumulate ? *dst + val : val;
    }
}

template <int STORAGE, int BZ, int GEO, bool KAHAN>
void launch_bz(const BpLaunch& p, const int* gbits, hipStream_t s)
{
    dim3 grid(p.nbx, p.nby, p.nbz), block(kThreads);
    const size_t buf = size_t(p.lds_quads) * (bp_float_quads(STORAGE, GEO) ? 16 : 8);
    auto k = fdk_bp_kernel<STORAGE, BZ, GEO, KAHAN>;
    // double-buffer the tile when that costs no occupancy: workgroups per CU
    // allowed by the VGPRs (one wave per SIMD each) vs by 160 KB of LDS
    hipFuncAttributes attr{};
    YK_HIP(hipFuncGetAttributes(&attr, reinterpret_cast<const void*>(k)));
    const int vgprs = std::max(8, (attr.numRegs + 7) / 8 * 8);
    const size_t lds_cu = 160 * 1024;
    const int wg_vgpr = std::min(8, 512 / vgprs);
    const int wg_db = int(lds_cu / std::max<size_t>(1, 2 * buf));
    int db = (wg_db >= wg_vgpr && 2 * buf <= 64 * 1024) ? 1 : 0;
    const size_t lds = db ? 2 * buf : buf;
    YK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    hipLaunchKernelGGL(k, grid, block, lds, s, p.proj, p.row_factors, gbits, p.coeffs, p.tiles, p.stage,
        p.tile_view_stride, p.tile_view_offset, p.volume, p.vol, p.nu, p.nv, p.count, p.accumulate,
        unsigned(buf), db);
}

template <int STORAGE, int GEO, bool KAHAN>
void launch_variant(const BpLaunch& p, const int* gbits, hipStream_t s, int bz)
{
    if (bz == 32) launch_bz<STORAGE, 32, GEO, KAHAN>(p, gbits, s);
    else if (bz == 16) launch_bz<STORAGE, 16, GEO, KAHAN>(p, gbits, s);
    else if (bz == 8) launch_bz<STORAGE, 8, GEO, KAHAN>(p, gbits, s);
    else fail(YK_ERR_UNSUPPORTED, "unsupported block depth");
}

template <int STORAGE, int GEO>
void launch_geo(const BpLaunch& p, const int* gbits, hipStream_t s, int bz)
{
    if (p.kahan) launch_variant<STORAGE, GEO, true>(p, gbits, s, bz);
    else launch_variant<STORAGE, GEO, false>(p, gbits, s, bz);
}

template <int STORAGE>
void launch_storage(const BpLaunch& p, const int* gbits, hipStream_t s, int bz)
{
    switch (p.general) {
    case kRegular: launch_geo<STORAGE, kRegular>(p, gbits, s, bz); break;
    case kGeneralRcp: launch_geo<STORAGE, kGeneralRcp>(p, gbits, s, bz); break;
    case kGeneralNewton: launch_geo<STORAGE, kGeneralNewton>(p, gbits, s, bz); break;
    default: fail(YK_ERR_INVALID_CONFIG, "unknown geometry mode");
    }
}

}  // namespace

int bp_block_z(int general, int storage, int kahan)
{
    // general path (float quads for every storage): 16 slices (32 measured
    // 12 % slower at 512^3 x 720 with a tilted detector); regular path: 32
    (void)kahan;
    (void)storage;
    return general ? 16 : 32;
}

size_t bp_quad_bytes(int storage, int general) { return storage == YK_F32 || general ? 16 : 8; }

void launch_bp_tiles(const TileLaunch& p, hipStream_t s)
{
    const long long total = (long long)p.nbx * p.nby * p.nbz * p.n_views;
    const int threads = 256;
    const long long blocks = (total + threads - 1) / threads;
    if (blocks > 0x7fffffffLL) fail(YK_ERR_UNSUPPORTED, "footprint table too large");
    hipLaunchKernelGGL(set_ints_kernel, dim3(1), dim3(64), 0, s, p.max_quads, 4, 0);
    hipLaunchKernelGGL(bp_tiles_kernel, dim3(unsigned(blocks)), dim3(threads), 0, s, p.coeff64,
        p.n_views, p.vol, p.bx, p.by, p.bz, p.nbx, p.nby, p.nbz, p.nu, p.nv, p.tiles, p.stage, p.max_quads);
    YK_HIP(hipGetLastError());
}

void launch_row_factors(const float* scales, int n, int storage, int* gbits, unsigned* factors,
    hipStream_t s)
{
    hipLaunchKernelGGL(set_ints_kernel, dim3(1), dim3(64), 0, s, gbits, 1, 0);
    const int threads = 256;
    const int blocks = std::max(1, std::min(1024, (n + threads - 1) / threads));
    hipLaunchKernelGGL(max_scale_kernel, dim3(blocks), dim3(threads), 0, s, scales, n, gbits);
    hipLaunchKernelGGL(row_factor_kernel, dim3((n + threads - 1) / threads), dim3(threads), 0, s, scales, n,
        gbits, storage, factors);
    YK_HIP(hipGetLastError());
}

void launch_normalize_rows(const void* in, const unsigned* factors, int storage, int rows, int nu, void* out,
    hipStream_t s)
{
    if (nu % 8) fail(YK_ERR_UNSUPPORTED, "detector width not a multiple of 8");
    const int quads = nu / 8;  // 16-byte groups of 8 texels
    const int threads = quads >= 256 ? 256 : (quads + 63) / 64 * 64;
    const unsigned blocks = unsigned(std::min(rows, 16384));
    if (storage == YK_F16)
        hipLaunchKernelGGL(normalize_rows_kernel<YK_F16>, dim3(blocks), dim3(threads), 0, s,
            static_cast<const uint4*>(in), factors, static_cast<uint4*>(out), quads, rows);
    else
        hipLaunchKernelGGL(normalize_rows_kernel<YK_BF16>, dim3(blocks), dim3(threads), 0, s,
            static_cast<const uint4*>(in), factors, static_cast<uint4*>(out), quads, rows);
    YK_HIP(hipGetLastError());
}

void launch_fdk_bp(const BpLaunch& p, hipStream_t s)
{
    // staging offsets are 32-bit byte offsets of rows in [-4400, nv + 4400)
    if ((size_t(p.nv) + 8800) * size_t(p.nu) * 4 >= (size_t(1) << 31))
        fail(YK_ERR_UNSUPPORTED, "detector too large for 32-bit staging offsets");
    if ((p.storage == YK_F16 || p.storage == YK_BF16) && !p.row_factors)
        fail(YK_ERR_INVALID_BUFFER, "row factors missing");
    const int bz = p.bz;
    const int* gbits = p.gscale_bits;
    switch (p.storage) {
    case YK_F32: launch_storage<YK_F32>(p, gbits, s, bz); break;
    case YK_F16: launch_storage<YK_F16>(p, gbits, s, bz); break;
    case YK_BF16: launch_storage<YK_BF16>(p, gbits, s, bz); break;
    case kBpHalfNormalized: launch_storage<kHalfNorm>(p, gbits, s, bz); break;
    default: fail(YK_ERR_INVALID_CONFIG, "unknown storage type");
    }
    YK_HIP(hipGetLastError());
}

}  // namespace yk

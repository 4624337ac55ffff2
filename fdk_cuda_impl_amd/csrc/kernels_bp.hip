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
constexpr int kKahanViews = 8;

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
// staging: fill LDS quads for tile (u0, v0, qw, qh) of one view
// ---------------------------------------------------------------------------
template <int STORAGE>
__device__ __forceinline__ float texel(const void* __restrict__ proj, size_t off)
{
    if constexpr (STORAGE == YK_F32) return static_cast<const float*>(proj)[off];
    else if constexpr (STORAGE == YK_F16) return __half2float(static_cast<const __half*>(proj)[off]);
    else return __bfloat162float(static_cast<const __hip_bfloat16*>(proj)[off]);
}

template <int STORAGE>
__device__ __forceinline__ void load_pair(const void* __restrict__ view_base,
    const float* __restrict__ view_scales_row, int nu, int nv, int row, int col, bool inside,
    float inv_g, float& a, float& b)
{
    if (inside || (row >= 0 && row < nv)) {
        const size_t off = (size_t)row * nu;
        float s = 1.f;
        if constexpr (STORAGE != YK_F32) s = view_scales_row[row] * inv_g;
        if (inside) {
            a = texel<STORAGE>(view_base, off + col) * s;
            b = texel<STORAGE>(view_base, off + col + 1) * s;
        } else {
            a = (col >= 0 && col < nu) ? texel<STORAGE>(view_base, off + col) * s : 0.f;
            b = (col + 1 >= 0 && col + 1 < nu) ? texel<STORAGE>(view_base, off + col + 1) * s : 0.f;
        }
    } else {
        a = 0.f;
        b = 0.f;
    }
}

template <int STORAGE>
__device__ __forceinline__ void store_quad(void* lds, int index, float t0, float t1, float b0, float b1)
{
    if constexpr (STORAGE == YK_F32) {
        static_cast<float4*>(lds)[index] = make_float4(t0, t1, b0 - t0, b1 - t1);
    } else {
        const half2_t h = { (_Float16)t0, (_Float16)t1 };
        const half2_t d = { (_Float16)(b0 - t0), (_Float16)(b1 - t1) };
        uint2 q;
        q.x = __builtin_bit_cast(unsigned, h);
        q.y = __builtin_bit_cast(unsigned, d);
        static_cast<uint2*>(lds)[index] = q;
    }
}

template <int STORAGE>
__device__ __forceinline__ void stage_tile(void* lds, const void* __restrict__ view_base,
    const float* __restrict__ scales_row, int nu, int nv, short4 t, float inv_g)
{
    const int qw = t.z, qh = t.w;
    const bool inside = t.x >= 0 && t.y >= 0 && t.x + qw + 1 <= nu && t.y + qh + 1 <= nv;
    // lanes walk down columns; a pass covers `span` columns with `chunks`
    // row ranges each, so every thread slides a 2-row window over its range
    const int span = min(qw, kThreads);
    const int chunks = kThreads / span;
    const int rows_per = (qh + chunks - 1) / chunks;
    const int tid = threadIdx.x;
    const int rc = tid / span;
    if (rc >= chunks) return;
    const int r0 = rc * rows_per;
    const int r1 = min(r0 + rows_per, qh);
    if (r0 >= r1) return;
    for (int c = tid % span; c < qw; c += span) {
        const int col = t.x + c;
        float a0, a1;
        load_pair<STORAGE>(view_base, scales_row, nu, nv, t.y + r0, col, inside, inv_g, a0, a1);
        for (int r = r0; r < r1; ++r) {
            float b0, b1;
            load_pair<STORAGE>(view_base, scales_row, nu, nv, t.y + r + 1, col, inside, inv_g, b0, b1);
            store_quad<STORAGE>(lds, r * qw + c, a0, a1, b0, b1);
            a0 = b0;
            a1 = b1;
        }
    }
}

// ---------------------------------------------------------------------------
// per-view accumulation into acc[BZ]
// ---------------------------------------------------------------------------
template <int STORAGE, int BZ, bool GENERAL>
__device__ __forceinline__ void accumulate_view(const void* lds, const BpCoeff& c, short4 t,
    float X, float Y, float Z0, float vz, float wscale, float (&acc)[BZ])
{
    const float ul0 = float(t.x), vl0 = float(t.y);
    const int rowq = t.z;
    // distance weight: the radial ray has no z component, so it is z-invariant
    const float depth = c.cr[3] + c.cr[0] * X + c.cr[1] * Y;
    const float w = c.wbase / (depth * depth) * wscale;
    if constexpr (!GENERAL) {
        const float den = c.cd[3] + c.cd[0] * X + c.cd[1] * Y;
        const float r = 1.f / den;
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
            for (int k = 0; k < BZ; ++k) {
                const float v = fmaf(float(k), dv, v0);
                const int iv = int(v);
                const float fv = __builtin_amdgcn_fractf(v);
                const float4 q = *reinterpret_cast<const float4*>(base + __mul24(iv, rowb));
                const float c0 = fmaf(fv, q.z, q.x);
                const float c1 = fmaf(fv, q.w, q.y);
                acc[k] = fmaf(c1, fw, fmaf(c0, gw, acc[k]));
            }
        } else {
            const char* base = static_cast<const char*>(lds) + iu * 8;
            const int rowb = rowq * 8;
            const half2_t gf = { (_Float16)gw, (_Float16)fw };
#pragma unroll
            for (int k = 0; k < BZ; k += 2) {
                const float va = fmaf(float(k), dv, v0);
                const float vb = fmaf(float(k + 1), dv, v0);
                const int ia = int(va), ib = int(vb);
                const half2_t fp = __builtin_bit_cast(half2_t,
                    __builtin_amdgcn_cvt_pkrtz(__builtin_amdgcn_fractf(va), __builtin_amdgcn_fractf(vb)));
                const uint2 qa = *reinterpret_cast<const uint2*>(base + __mul24(ia, rowb));
                const uint2 qb = *reinterpret_cast<const uint2*>(base + __mul24(ib, rowb));
                const half2_t fa = { fp.x, fp.x };
                const half2_t fb = { fp.y, fp.y };
                const half2_t ca = __builtin_bit_cast(half2_t, qa.y) * fa + __builtin_bit_cast(half2_t, qa.x);
                const half2_t cb = __builtin_bit_cast(half2_t, qb.y) * fb + __builtin_bit_cast(half2_t, qb.x);
                acc[k] = __builtin_amdgcn_fdot2(ca, gf, acc[k], false);
                acc[k + 1] = __builtin_amdgcn_fdot2(cb, gf, acc[k + 1], false);
            }
        }
    } else {
        // general perspective path: u, v per sample
        float den = c.cd[3] + c.cd[0] * X + c.cd[1] * Y + c.cd[2] * Z0;
        float un = c.cu[3] + c.cu[0] * X + c.cu[1] * Y + c.cu[2] * Z0;
        float vn = c.cv[3] + c.cv[0] * X + c.cv[1] * Y + c.cv[2] * Z0;
        const float dd = c.cd[2] * vz, du = c.cu[2] * vz, dvn = c.cv[2] * vz;
        const int rowe = rowq;
#pragma unroll 4
        for (int k = 0; k < BZ; ++k) {
            const float r = 1.f / den;
            const float u = un * r - ul0;
            const float v = vn * r - vl0;
            const int iu = int(u), iv = int(v);
            const float fu = u - float(iu), fv = v - float(iv);
            float c0, c1;
            if constexpr (STORAGE == YK_F32) {
                const float4 q = static_cast<const float4*>(lds)[iv * rowe + iu];
                c0 = fmaf(fv, q.z, q.x);
                c1 = fmaf(fv, q.w, q.y);
            } else {
                const uint2 q = static_cast<const uint2*>(lds)[iv * rowe + iu];
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

    const size_t view_elems = (size_t)nu * nv;
    for (int i = 0; i < count; ++i) {
        const short4 t = my_tiles[i];
        if (t.z > 0) {
            __syncthreads();
            const void* view_base;
            if constexpr (STORAGE == YK_F32) view_base = static_cast<const float*>(proj) + view_elems * i;
            else view_base = static_cast<const unsigned short*>(proj) + view_elems * i;
            stage_tile<STORAGE>(lds, view_base, row_scales ? row_scales + (size_t)i * nv : nullptr,
                nu, nv, t, inv_g);
            __syncthreads();
            if (active)
                accumulate_view<STORAGE, BZ, GENERAL>(lds, coeffs[i], t, X, Y, Z0, vol.vz, 1.f, acc);
        }
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

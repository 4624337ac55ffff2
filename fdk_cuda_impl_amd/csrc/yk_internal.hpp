// Internal host-side declarations shared by the C ABI, the plans and the
// kernel launchers of libyk_hip.so.  Nothing here is part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/yk_hip.h"

namespace yk {

// ---------------------------------------------------------------------------
// errors: every failure is a yk::Error carrying a yk_status; the C ABI layer
// converts it to a return code and a thread-local message (no aborts).
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
    yk_status status;
    Error(yk_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
};

[[noreturn]] void fail(yk_status status, const std::string& message);
void check_hip(hipError_t err, const char* what);

#define YK_HIP(call) ::yk::check_hip((call), #call)

// ---------------------------------------------------------------------------
// device memory RAII (hipMalloc'd, freed on destruction)
// ---------------------------------------------------------------------------
template <typename T>
class DeviceArray {
public:
    DeviceArray() = default;
    explicit DeviceArray(size_t n) { resize(n); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    DeviceArray(DeviceArray&& o) noexcept : ptr_(o.ptr_), n_(o.n_) { o.ptr_ = nullptr; o.n_ = 0; }
    DeviceArray& operator=(DeviceArray&& o) noexcept
    {
        if (this != &o) { reset(); ptr_ = o.ptr_; n_ = o.n_; o.ptr_ = nullptr; o.n_ = 0; }
        return *this;
    }
    ~DeviceArray() { reset(); }
    void resize(size_t n)
    {
        if (n == n_) return;
        reset();
        if (n) {
            void* p = nullptr;
            YK_HIP(hipMalloc(&p, n * sizeof(T)));
            ptr_ = static_cast<T*>(p);
            n_ = n;
        }
    }
    void reset()
    {
        if (ptr_) (void)hipFree(ptr_);
        ptr_ = nullptr;
        n_ = 0;
    }
    T* get() const { return ptr_; }
    size_t size() const { return n_; }
    void upload(const T* host, size_t n, hipStream_t s)
    {
        resize(n);
        if (n) YK_HIP(hipMemcpyAsync(ptr_, host, n * sizeof(T), hipMemcpyHostToDevice, s));
    }

private:
    T* ptr_ = nullptr;
    size_t n_ = 0;
};

// pinned host staging buffer (hipHostMalloc'd)
class PinnedBuffer {
public:
    PinnedBuffer() = default;
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    ~PinnedBuffer() { reset(); }
    void ensure(size_t bytes);
    void reset();
    void* get() const { return ptr_; }

private:
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
};

}  // namespace yk

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct yk_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    bool synchronous = true;
    yk::DeviceArray<double> reduce_scratch;
    yk::PinnedBuffer pinned;
    void finish();  // synchronizes when synchronous, surfaces async errors
    void activate() const;
};

namespace yk {

// ---------------------------------------------------------------------------
// geometry (FP64 host math; float results)
// ---------------------------------------------------------------------------
struct ViewD {
    double src[3], detS[3], detU[3], detV[3];
    double angle;
};

ViewD widen(const yk_view& v);

struct DerivedD {
    double theta, dtheta, fscale, sod, sdd_radial, sdd_plane, off_u, off_v, du, dv;
    double radial[3], normal[3];
    double uu, vv, uv;
};

struct CoeffD {
    double cu[4], cv[4], cd[4], cr[4];
    double wbase;
};

void build_views(const yk_system& system, std::vector<yk_view>& views);
void derive_geometry(const std::vector<yk_view>& views, int nu, int nv, double scan_range,
    std::vector<DerivedD>& out);
void fdk_coefficients(const std::vector<yk_view>& views, const std::vector<DerivedD>& derived,
    std::vector<CoeffD>& out);
void filter_weights(const yk_filter& f, int padded_n, double du, std::vector<float>& out);
// sum of |h[m]| over the padded_n-periodic impulse response h[m] =
// sum_k w[|k|] cos(2 pi k m / padded_n) of the filter weights (FP64): a
// filtered row never exceeds this times the largest |texel| of its input
double filter_response_l1(const std::vector<float>& weights, int padded_n);
void validate_filter(const yk_filter& f);
int padded_n(int nu);
void fft_twiddles(int n, std::vector<float2>& out);
int start_u(int nu, int n, double offset_u_pix);
int main_axis(const yk_view& v, int nu, int nv);
// world point -> detector pixel coordinates: u = (c[0..3].[X,1]) / (c[8..11].[X,1]),
// v = (c[4..7].[X,1]) / (c[8..11].[X,1]); same map as the FDK coefficients
void projection_map(const yk_view& v, double c[12]);

// ---------------------------------------------------------------------------
// device-side records (plain float POD passed to kernels)
// ---------------------------------------------------------------------------
struct alignas(16) BpCoeff {
    float cu[4];   // x y z w
    float cv[4];
    float cd[4];
    float cr[4];
    float wbase;
    float pad[3];
};

struct alignas(16) FilterView {
    float src[4];
    float detS[4];
    float detU[4];
    float detV[4];
    float sdd_radial;
    float wgeom;       // SDD_r / SOD / 2
    int start_u;
    int pad;
};

struct VolumeGeom {
    int nx, ny, nz;
    float vx, vy, vz;
    float ox, oy, oz;  // centre of voxel (0,0,0)
};

VolumeGeom volume_geom(const yk_volume_grid& g);
void validate_volume(const yk_volume_grid& g);

// ---------------------------------------------------------------------------
// kernel launchers (implemented in the .hip translation units)
// ---------------------------------------------------------------------------
struct FilterLaunch {
    const float* proj;          // [count][nv][nu]
    void* out;                  // [count][nv][nu] of `storage`
    float* row_scales;          // [count][nv] (half / bf16)
    const FilterView* views;    // device, `count` records
    const float* parker;        // device [count][nu] Parker weights x range/pi, or null
    const float* weights;       // device, N/2+1
    const float2* twiddles;     // device, N
    int nu, nv, count, n;
    int storage;
    float response_l1;          // filter_response_l1 of `weights` (half / bf16 row scales)
};
void launch_fdk_filter(const FilterLaunch& p, hipStream_t s);
// the cosine (and Parker) weighted projections of p's views into out ([count][nv][nu])
void launch_preweight(const FilterLaunch& p, float* out, hipStream_t s);

struct TileLaunch {
    const double* coeff64;      // device, per view: cu[4] cv[4] cd[4] (12 doubles)
    int n_views;
    VolumeGeom vol;
    int bx, by, bz;             // voxel block
    int nbx, nby, nbz;          // block counts
    int nu, nv;
    short4* tiles;              // [n_blocks][n_views]: u0, v0, width, height (quads)
    uint2* stage;               // [n_blocks][n_views]: staging shape (span reciprocal, chunks, rows)
    int* max_quads;             // device scalars: max width, max height, max staged quads (LDS), bad
};
void launch_bp_tiles(const TileLaunch& p, hipStream_t s);

struct BpLaunch {
    const void* proj;           // [count][nv][nu], storage type
    const unsigned* row_factors; // [count][nv] normalized row factors (half / bf16) or null
    const int* gscale_bits;     // device float bits: max row scale (half / bf16)
    const BpCoeff* coeffs;      // device, `count` records
    const short4* tiles;        // [n_blocks][n_tile_views] with column offset
    const uint2* stage;         // staging shapes, same indexing as tiles
    int tile_view_stride;       // = total views of the plan
    int tile_view_offset;       // first view in the tile table
    float* volume;
    VolumeGeom vol;
    int nu, nv, count;
    int storage;
    int accumulate;
    int kahan;
    int general;                // 0 regular; 1 / 2 u varies along z (tilted detector):
                                // a reciprocal per sample / Newton steps (fdk.cpp general_mode)
    int nbx, nby, nbz;
    int bz;                     // z slices per workgroup (32, 16 or 8)
    int lds_quads;              // LDS capacity in quads
};
void launch_fdk_bp(const BpLaunch& p, hipStream_t s);
// BpLaunch::storage of half texels already normalized by launch_normalize_rows
constexpr int kBpHalfNormalized = 16;
// rows x nu half / bfloat16 texels times their normalized row factors -> half
void launch_normalize_rows(const void* in, const unsigned* factors, int storage, int rows, int nu, void* out,
    hipStream_t s);
// global maximum row scale into *gbits, then per row the factor scale / max
// (packed half pair for half storage, float bits for bfloat16)
void launch_row_factors(const float* scales, int n, int storage, int* gbits, unsigned* factors,
    hipStream_t s);
int bp_block_z(int general, int storage, int kahan);
// bytes of one LDS quad of the backprojection: 16 (float) for float storage
// and for every storage of a general-geometry plan, 8 (half) otherwise
size_t bp_quad_bytes(int storage, int general);
// progress report of the iterative callers: waits for the stream's work, then
// calls cb(user, completed, total) on this thread; no-op without a callback
void report_progress(yk_progress_fn cb, void* user, int completed, int total, hipStream_t s);
// Fixed per-GPU scratch of the allocation-free backprojection: a header of
// kNoAllocHeader words (word 0: global row-scale bits) and room for
// kNoAllocRowFactors row factors (views x detector rows of one call).
constexpr size_t kNoAllocHeader = 64;
constexpr size_t kNoAllocRowFactors = size_t(YK_NOALLOC_MAX_ROWS);
// device address of that array on `device` (must be the current device),
// looked up with hipGetSymbolAddress on the first call for the device
unsigned* bp_noalloc_scratch(int device);

struct ProjViewDev {
    float src[4];               // voxel-normalized, volume-centred
    float detS[4];
    float detU[4];
    float detV[4];
    int axis;                   // 0 X, 1 Y, 2 Z
    int pad[3];
};

struct SiddonViewDev {
    float src[4];               // world
    float detS[4];
    float detU[4];
    float detV[4];
};

struct JosephLaunch {
    const float* volume;        // [z][y][x]
    const float* volume_t;      // [z][x][y] copy, for axis-X views (may be null)
    float* proj;                // [count][nv][nu]
    const ProjViewDev* views;   // device, `count`
    VolumeGeom vol;
    int nu, nv, count;
    int accumulate;
};
void launch_joseph_fp(const JosephLaunch& p, hipStream_t s);
void launch_transpose_xy(const float* in, float* out, int nx, int ny, int nz, hipStream_t s);

// Joseph projection for separable views from paired Z-fastest volume copies
struct JosephSepFpLaunch {
    const float* vol_zx;        // paired copy for X-main views (may be null if none)
    const float* vol_zy;        // paired copy for Y-main views (may be null if none)
    float* proj;                // [count][nv][nu]
    const ProjViewDev* views;   // device, `count`
    VolumeGeom vol;
    int nu, nv, count;
    int accumulate;
};
void launch_joseph_fp_sep(const JosephSepFpLaunch& p, hipStream_t s);
// Joseph projection of X- or Y-main views of any orientation from paired
// slab copies (vol_zx / vol_zy made by launch_joseph_pair_copy with slab)
void launch_joseph_fp_slab(const JosephSepFpLaunch& p, hipStream_t s);
// Paired guarded copy of a [z][y][x] volume for the Joseph projections:
// float2 entries (value at row r, value at row r + 1) of a copy with two zero
// planes before and after each minor axis.  Separable layout (slab false):
// slices along the main axis (x for x_main, else y), rows along the other
// of x and y, columns along z.  Slab layout: rows along z, columns along the
// other of x and y.  joseph_pair_size floats.
void launch_joseph_pair_copy(const float* in, float* out, int nx, int ny, int nz, bool x_main, bool slab,
    hipStream_t s);
size_t joseph_pair_size(int nx, int ny, int nz, bool x_main, bool slab);

struct JosephAdjointLaunch {
    const float* proj;          // [count][nv][nu]
    float* volume;              // [z][y][x]
    const ProjViewDev* views;   // device, `count`
    const float* coeff32;       // device, per view cu[4] cv[4] cd[4] (world projective map)
    float4* ray_ab;             // scratch [count][nv][nu]: a1, b1, a2, b2 of every ray
    float* ray_len;             // scratch [count][nv][nu]: L (0: ray skipped)
    VolumeGeom vol;
    int nu, nv, count;
    int accumulate;
};
void launch_joseph_adjoint(const JosephAdjointLaunch& p, hipStream_t s);

// zero entries closing each detector column of the separable ray table
constexpr int kSepPad = 8;

// Transpose of the Joseph projector for separable views (main axis X or Y,
// detV with no X/Y component); the ray table arrays are scratch of `count` views.
struct JosephSepLaunch {
    const float* proj;          // [count][nv][nu]
    float* volume;              // [z][y][x]
    const ProjViewDev* views;   // device, `count`
    const float* coeff32;       // device, per view cu[4] cv[4] cd[4] (world projective map)
    float2* ab1;                // scratch [count][nu]
    float2* ab2;                // scratch [count][nu][nv + kSepPad]
    float* pl;                  // scratch [count][nu][nv + kSepPad]
    VolumeGeom vol;
    int nu, nv, count;
    int accumulate;
};
void launch_joseph_adjoint_sep(const JosephSepLaunch& p, hipStream_t s);

struct SiddonLaunch {
    const float* volume;
    float* proj;
    const SiddonViewDev* views;
    VolumeGeom vol;
    int nu, nv, count;
    int accumulate;
};
void launch_siddon_fp(const SiddonLaunch& p, hipStream_t s);

struct SiddonAdjointLaunch {
    const float* proj;
    float* volume;
    const SiddonViewDev* views;
    const float* coeff32;  // the projective map in float, 12 per view
    VolumeGeom vol;
    int nu, nv, count;
    int accumulate;
};
void launch_siddon_adjoint(const SiddonAdjointLaunch& p, hipStream_t s);

// Transpose of the Siddon projector for separable views (detV along Z only);
// the ray table arrays are scratch of `count` views.
struct SiddonSepLaunch {
    const float* proj;          // [count][nv][nu]
    float* volume;              // [z][y][x]
    const SiddonViewDev* views; // device, `count`
    const float* coeff32;       // device, per view cu[4] cv[4] cd[4] (world projective map)
    float2* col_xy;             // scratch [count][nu]
    float2* ray_wz;             // scratch [count][nu][nv + kSepPad]
    VolumeGeom vol;
    int nu, nv, count;
    int accumulate;
};
void launch_siddon_adjoint_sep(const SiddonSepLaunch& p, hipStream_t s);

// per-view record of the operator-form FDK backprojections (world units)
struct OpBpViewDev {
    float cu[4], cv[4], cd[4];  // projective map u = Cu.[X,1] / Cd.[X,1], v likewise
    float cc[4];                // FDK weight denominator (-src / SOD, SOD)
    float src[4];
    float l2[6];                // |detS + u detU + v detV - src|^2: 1, u, v, uu, uv, vv
    float wbase;                // SOD^2 dtheta fscale
    float inv_sdd_plane;
};

struct OpBpLaunch {
    const float* proj;          // [count][nv][nu], unfiltered
    float* volume;              // [z][y][x]
    const OpBpViewDev* views;   // device, `count`
    VolumeGeom vol;
    int nu, nv, count;
    float scale;                // vx vy vz / (du dv) of the first view (matched weight)
    int mode;                   // 0 FDK weight, 1 pseudo-matched, 2 voxel-driven Joseph
    int accumulate;
};
void launch_fdk_operator_bp(const OpBpLaunch& p, hipStream_t s);

// vector helpers
void launch_dot(const float* a, const float* b, size_t n, double* partials, int n_partials,
    hipStream_t s);
void launch_axpy(float* y, const float* x, float a, size_t n, hipStream_t s);        // y += a x
void launch_xpby(float* y, const float* x, float b, size_t n, hipStream_t s);        // y = x + b y
void launch_scale_copy(float* y, const float* x, float a, size_t n, hipStream_t s);  // y = a x
void launch_fill(float* y, float v, size_t n, hipStream_t s);
// algebraic steps: mode 0 res = row > eps ? (b - res) / row : 0, mode 1 res = (b - res) row
void algebraic_residual(float* res, const float* b, const float* row, float eps, int mode, size_t n,
    hipStream_t s);
// x += lambda bp / (col + eps), then optional bounds
void algebraic_update(float* x, const float* bp, const float* col, float lambda, float eps, int use_min,
    float lo, int use_max, float hi, size_t n, hipStream_t s);
void launch_bounds_clamp(float* x, int use_min, float lo, int use_max, float hi, size_t n, hipStream_t s);
// x <= threshold -> 1e30, then optionally x = 1 / x
void weight_prepare(float* x, float threshold, int reciprocal, size_t n, hipStream_t s);
// smoothed TV value partials and gradient; dims 3 (volume) or 2 (each z slice alone)
void launch_tv_gradient(const float* x, float* grad, double* partials, int n_partials,
    const VolumeGeom& g, float eps, hipStream_t s, int dims = 3);
void launch_phantom(const float* ellipsoids, int n, const VolumeGeom& g, int clamp,
    float* out, hipStream_t s);

}  // namespace yk

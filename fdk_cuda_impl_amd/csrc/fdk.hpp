// FDK plan and projection-operator plan (library-owned handles of the C ABI).
#pragma once

#include <string>
#include <vector>

#include "yk_internal.hpp"

struct yk_fdk {
    yk_fdk(yk_ctx* ctx, const yk_fdk_desc& desc, const yk_view* views, int n_views, int nu,
        int nv, const yk_volume_grid& grid);

    void filter(const float* d_proj, int first, int count, void* d_out, int storage, float* d_scales,
        float* d_preweighted = nullptr);
    void backproject(const void* d_filtered, int storage, const float* d_scales, int first, int count,
        float* d_volume, int accumulate);
    // backproject() with no device allocation (scratch in the fixed per-GPU array)
    void backproject_noalloc(const void* d_filtered, int storage, const float* d_scales, int first, int count,
        float* d_volume, int accumulate);
    void execute(const float* projection, int ploc, size_t pcap, int offset, int count, float* volume,
        int vloc, size_t vcap, int clear_output);

    yk_ctx* ctx;
    yk_fdk_desc desc;
    int nu, nv, n_views;
    yk_volume_grid grid;
    yk::VolumeGeom vol{};
    double scan_range = 0.0;
    bool parker = false;
    bool general = false;
    int geo_mode = 0;              // 0 regular, 1 tilted (reciprocal per sample), 2 tilted (Newton steps)
    std::vector<yk_view> views;
    std::vector<yk::DerivedD> derived;
    std::vector<yk::CoeffD> coeffs;
    std::vector<float> weights;
    int n_pad = 0;
    float response_l1 = 0.f;       // filter_response_l1 of weights

    yk::DeviceArray<float> d_weights;
    yk::DeviceArray<float2> d_twiddles;
    yk::DeviceArray<yk::FilterView> d_fviews;
    yk::DeviceArray<yk::BpCoeff> d_coeffs;
    yk::DeviceArray<double> d_c64;
    yk::DeviceArray<float> d_parker;
    yk::DeviceArray<short4> d_tiles;
    yk::DeviceArray<uint2> d_stage;
    yk::DeviceArray<int> d_maxes;      // [0..3] footprint maxima, [4] global row-scale bits
    yk::DeviceArray<unsigned> d_rowfac; // normalized row factors of the views being backprojected
    yk::DeviceArray<unsigned short> d_norm_ws; // half texels times their row factors (half / bf16 storage)
    int nbx = 0, nby = 0, nbz = 0, bz = 0;
    int lds_quads = 0;

    int batch = 0;
    yk::DeviceArray<float> d_proj_ws;     // host projections: two staged chunks
    // Host buffers move through two library-owned pinned slots: the host
    // copies a chunk into a slot with several threads while the copy stream
    // moves the other slot to the device and the plan's stream filters and
    // backprojects the chunk before (stage_chunk_views()).
    yk::PinnedBuffer pinned[2];
    hipStream_t copy_stream = nullptr;   // default priority
    hipEvent_t ev_pinned[2] = {};        // slot's last transfer done (host may reuse it)
    hipEvent_t ev_consumed[2] = {};      // filter of the slot's device chunk done
    bool pinned_used[2] = {};
    bool consumed_used[2] = {};
    int stage_chunk_views() const;
    void ensure_staging(size_t slot_bytes);
    void upload_host(float* d_dst, const float* host, size_t bytes);
    void download_host(float* host, const float* d_src, size_t bytes);
    ~yk_fdk();
    yk::DeviceArray<unsigned char> d_filtered_ws;
    yk::DeviceArray<float> d_scales_ws;
    yk::DeviceArray<float> d_volume_ws;
    int received = 0;

    // calibrated-geometry report (diagnostics.cpp) and the last chunk's stages
    yk_geometry_diagnostics diagnostics{};
    bool capture_stages = false;
    yk::DeviceArray<float> d_preweighted_ws;
    yk_fdk_stage last_stage{};

private:
    void build_tiles();
};

struct yk_projector {
    yk_projector(yk_ctx* ctx, const yk_view* views, int n_views, int nu, int nv,
        const yk_volume_grid& grid, float scan_range);

    void project(int kind, const float* d_volume, float* d_proj, int first, int count, int accumulate);
    void backproject(int kind, const float* d_proj, float* d_volume, int first, int count, int accumulate);
    double dot(const float* a, const float* b, size_t n);

    yk_ctx* ctx;
    int nu, nv, n_views;
    yk_volume_grid grid;
    yk::VolumeGeom vol{};
    double scan_range;
    std::vector<yk_view> views;
    bool any_x_axis = false;
    yk::DeviceArray<yk::ProjViewDev> d_jviews;
    yk::DeviceArray<yk::SiddonViewDev> d_sviews;
    yk::DeviceArray<float> d_c32;      // the projective map in float (transpose candidates)
    yk::DeviceArray<float4> d_ray_ab;  // Joseph transpose ray table, one view chunk
    yk::DeviceArray<float> d_ray_len;
    std::vector<char> separable;       // per view: main axis X/Y and detV along Z only
    yk::DeviceArray<float2> d_sep_ab1; // separable Joseph transpose ray table, one view chunk
    yk::DeviceArray<float2> d_sep_ab2;
    yk::DeviceArray<float> d_sep_pl;
    std::vector<char> siddon_separable;  // per view: detV along Z only
    yk::DeviceArray<float2> d_sid_col;   // separable Siddon transpose ray table, one view chunk
    yk::DeviceArray<float2> d_sid_ray;
    // projector whose scratch arrays (volume copies, ray tables) this one uses;
    // null: its own.  Operators of one volume grid that run one after another
    // on one stream can share them.
    yk_projector* scratch_owner = nullptr;
    yk::DeviceArray<float> d_vol_t;
    yk::DeviceArray<float> d_vol_zx;   // Z-fastest volume copies of the separable Joseph projection
    yk::DeviceArray<float> d_vol_zy;
    yk::DeviceArray<double> d_partials;
    yk::DeviceArray<yk::OpBpViewDev> d_opviews;  // operator-form FDK backprojection records
    std::vector<double> op_scale;      // per first view: vx vy vz / (du dv)
};

// unified SIRT / SART / OS-SART (algebraic.cpp); throws through fail()
void run_algebraic(yk_projector* p, const yk_algebraic_desc& d, const float* d_b, float* d_x,
    yk_iterative_stats* stats);
// CGLS, robust-restart and ASTRA strategies (cgls.cpp); throws through fail()
void run_cgls(yk_projector* p, const yk_cgls_desc& d, const float* d_b, float* d_x, yk_iterative_stats* stats,
    double* residual_norms);

// calibrated-geometry diagnostics (diagnostics.cpp); throws INVALID_GEOMETRY
// for geometry FDK cannot use
void geometry_diagnostics(const std::vector<yk_view>& views, const std::vector<yk::DerivedD>& derived,
    yk_geometry_diagnostics& out);
// raw dumps (diagnostics.cpp)
std::string dump_name_a(const char* prefix, int a, const char* ext);
yk_dump* dump_create(const char* dir, int enabled);
void dump_destroy(yk_dump* d);
void dump_set_enabled(yk_dump* d, int on);
void dump_device_f32(yk_dump* d, const std::string& name, const float* src, size_t count, hipStream_t s);
void dump_text(yk_dump* d, const std::string& name, const char* text);

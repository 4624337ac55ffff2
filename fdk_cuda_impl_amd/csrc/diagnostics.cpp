// Calibrated-geometry diagnostics of the FDK path and the raw stage dump.
//   geometry_diagnostics: how far per-view (calibrated) geometry is from the
//   ideal circular scan (FdkPipeline::validateApproximateFdkGeometry_,
//   src/Reconstruction/Analytic/Circular/Flat/FDK/YkFdkPipeline.hpp:35-58,
//   664-866); FDK always uses the actual per-view geometry, this only reports.
//   yk_dump: raw float32 / text files of device buffers (DumpManager,
//   src/util/YkIoDump.hpp:9-72, YkIoDump.cpp).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "fdk.hpp"

using namespace yk;

namespace {

struct V3 {
    double x, y, z;
};
V3 v3(const float* p) { return { p[0], p[1], p[2] }; }
V3 v3(const double* p) { return { p[0], p[1], p[2] }; }
double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(const V3& a, const V3& b) { return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x }; }
double norm(const V3& a) { return std::sqrt(dot(a, a)); }
V3 unit(const V3& a)
{
    const double n = norm(a);
    return { a.x / n, a.y / n, a.z / n };
}

}  // namespace

void geometry_diagnostics(const std::vector<yk_view>& views, const std::vector<DerivedD>& derived,
    yk_geometry_diagnostics& d)
{
    constexpr double kRelativeTolerance = 1e-3, kAngleTolerance = 1e-5;
    if (views.empty() || derived.size() != views.size()) fail(YK_ERR_INVALID_GEOMETRY, "empty geometry");
    d = yk_geometry_diagnostics{};
    d.struct_size = sizeof(d);
    const DerivedD& f = derived.front();
    d.min_source_to_axis_mm = d.max_source_to_axis_mm = float(f.sod);
    d.min_source_to_detector_mm = d.max_source_to_detector_mm = float(f.sdd_radial);
    d.min_source_to_plane_mm = d.max_source_to_plane_mm = float(f.sdd_plane);
    d.min_principal_u_pix = d.max_principal_u_pix = float(f.off_u);
    d.min_principal_v_pix = d.max_principal_v_pix = float(f.off_v);
    const double du0 = f.du, dv0 = f.dv;
    if (!(du0 > 0.0) || !(dv0 > 0.0)) fail(YK_ERR_INVALID_GEOMETRY, "invalid detector pixel axes");
    bool approximate = false;
    for (size_t i = 0; i < views.size(); ++i) {
        const yk_view& v = views[i];
        const DerivedD& a = derived[i];
        const V3 U = v3(v.detU), V = v3(v.detV);
        const double du = norm(U), dv = norm(V);
        if (!(du > 0.0) || !(dv > 0.0) || !(a.sod > 0.0) || !(a.sdd_radial > 0.0) || !(a.sdd_plane > 0.0) ||
            !std::isfinite(v.angle[0]))
            fail(YK_ERR_INVALID_GEOMETRY, "per-view geometry that FDK cannot use");
        auto lo = [](float& m, double x) { m = std::min(m, float(x)); };
        auto hi = [](float& m, double x) { m = std::max(m, float(x)); };
        lo(d.min_source_to_axis_mm, a.sod);
        hi(d.max_source_to_axis_mm, a.sod);
        lo(d.min_source_to_detector_mm, a.sdd_radial);
        hi(d.max_source_to_detector_mm, a.sdd_radial);
        lo(d.min_source_to_plane_mm, a.sdd_plane);
        hi(d.max_source_to_plane_mm, a.sdd_plane);
        hi(d.max_abs_source_axis_mm, std::fabs(double(v.src[2])));
        hi(d.max_detector_skew, std::fabs(dot(U, V)) / (du * dv));
        hi(d.max_abs_principal_u_pix, std::fabs(a.off_u));
        hi(d.max_abs_principal_v_pix, std::fabs(a.off_v));
        lo(d.min_principal_u_pix, a.off_u);
        hi(d.max_principal_u_pix, a.off_u);
        lo(d.min_principal_v_pix, a.off_v);
        hi(d.max_principal_v_pix, a.off_v);
        // detector tilts relative to the ideal detector of this source (N toward
        // the source, V along the rotation axis, U x V = N), in the builder's
        // intrinsic U, V, N order
        const V3 n_ideal = { -a.radial[0], -a.radial[1], -a.radial[2] };
        const V3 u_ideal = unit(cross({ 0.0, 0.0, 1.0 }, n_ideal));
        const V3 v_ideal = unit(cross(n_ideal, u_ideal));
        const V3 mu = unit(U), mv = unit(V);
        V3 mn = unit(cross(mu, mv));
        if (dot(mn, n_ideal) < 0.0) mn = { -mn.x, -mn.y, -mn.z };
        const double tilt_v = std::asin(std::clamp(dot(u_ideal, mn), -1.0, 1.0));
        const double tilt_u = std::atan2(-dot(v_ideal, mn), dot(n_ideal, mn));
        const double tilt_n = std::atan2(-dot(u_ideal, mv), dot(u_ideal, mu));
        if (i == 0) {
            d.min_relative_tilt_u_rad = d.max_relative_tilt_u_rad = float(tilt_u);
            d.min_relative_tilt_v_rad = d.max_relative_tilt_v_rad = float(tilt_v);
            d.min_relative_tilt_n_rad = d.max_relative_tilt_n_rad = float(tilt_n);
        } else {
            lo(d.min_relative_tilt_u_rad, tilt_u);
            hi(d.max_relative_tilt_u_rad, tilt_u);
            lo(d.min_relative_tilt_v_rad, tilt_v);
            hi(d.max_relative_tilt_v_rad, tilt_v);
            lo(d.min_relative_tilt_n_rad, tilt_n);
            hi(d.max_relative_tilt_n_rad, tilt_n);
        }
        const double align = std::fabs(dot(v3(a.radial), v3(a.normal)));
        hi(d.max_normal_misalignment, 1.0 - align);
        approximate = approximate || std::fabs(du - du0) > du0 * kRelativeTolerance ||
                      std::fabs(dv - dv0) > dv0 * kRelativeTolerance;
    }
    for (size_t i = 1; i < views.size(); ++i)
        if (std::fabs(double(views[i].angle[0]) - double(views[i - 1].angle[0])) < kAngleTolerance)
            fail(YK_ERR_INVALID_GEOMETRY, "repeated adjacent acquisition angles: dtheta is undefined");
    auto spread = [](double low, double high) { return high > 0.0 ? (high - low) / high : 0.0; };
    approximate = approximate || spread(d.min_source_to_axis_mm, d.max_source_to_axis_mm) > kRelativeTolerance ||
                  spread(d.min_source_to_detector_mm, d.max_source_to_detector_mm) > kRelativeTolerance ||
                  d.max_abs_source_axis_mm > d.max_source_to_axis_mm * kRelativeTolerance ||
                  d.max_detector_skew > kRelativeTolerance || d.max_abs_principal_u_pix > kRelativeTolerance ||
                  d.max_abs_principal_v_pix > kRelativeTolerance || d.max_normal_misalignment > kRelativeTolerance;
    d.approximate = approximate ? 1 : 0;
}

// --- raw dump ----------------------------------------------------------------

struct yk_dump {
    std::string dir;
    bool enabled = true;
    void* pinned = nullptr;
    size_t pinned_bytes = 0;

    ~yk_dump()
    {
        if (pinned) (void)hipHostFree(pinned);
    }
    std::string path(const std::string& name) const { return dir.empty() ? name : dir + name; }
    void reserve(size_t bytes)
    {
        if (bytes <= pinned_bytes) return;
        if (pinned) YK_HIP(hipHostFree(pinned));
        pinned = nullptr;
        pinned_bytes = 0;
        YK_HIP(hipHostMalloc(&pinned, bytes, hipHostMallocDefault));
        pinned_bytes = bytes;
    }
    void write(const std::string& name, const void* data, size_t bytes, const char* mode)
    {
        const std::string p = path(name);
        FILE* f = std::fopen(p.c_str(), mode);
        if (!f) fail(YK_ERR_INVALID_CONFIG, "cannot open " + p + " for writing");
        const size_t w = bytes ? std::fwrite(data, 1, bytes, f) : 0;
        const bool ok = std::fclose(f) == 0 && w == bytes;
        if (!ok) fail(YK_ERR_INVALID_CONFIG, "short write to " + p);
    }
};

std::string dump_name_a(const char* prefix, int a, const char* ext)
{
    char num[32];
    std::snprintf(num, sizeof(num), "_a%04d", a);
    return std::string(prefix) + num + ext;
}

yk_dump* dump_create(const char* dir, int enabled)
{
    auto* d = new yk_dump();
    d->dir = dir ? dir : "";
    if (!d->dir.empty() && d->dir.back() != '/' && d->dir.back() != '\\') d->dir.push_back('/');
    d->enabled = enabled != 0;
    return d;
}

void dump_set_enabled(yk_dump* d, int on) { d->enabled = on != 0; }

void dump_device_f32(yk_dump* d, const std::string& name, const float* src, size_t count, hipStream_t s)
{
    if (!d->enabled) return;
    if (name.empty() || !src || count == 0) fail(YK_ERR_INVALID_BUFFER, "dump needs a name, a buffer and a count");
    const size_t bytes = count * sizeof(float);
    d->reserve(bytes);
    YK_HIP(hipMemcpyAsync(d->pinned, src, bytes, hipMemcpyDeviceToHost, s));
    YK_HIP(hipStreamSynchronize(s));
    d->write(name, d->pinned, bytes, "wb");
}

void dump_text(yk_dump* d, const std::string& name, const char* text)
{
    if (!d->enabled) return;
    if (name.empty() || !text) fail(YK_ERR_INVALID_BUFFER, "dump needs a name and a text");
    d->write(name, text, std::char_traits<char>::length(text), "w");
}

void dump_destroy(yk_dump* d) { delete d; }

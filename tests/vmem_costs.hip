// Vector-memory and LDS issue costs on gfx950: cycles per wave64 instruction
// per CU for global loads of 4, 8 and 16 bytes per lane (lanes on
// consecutive addresses, or one 128-byte line each) from data resident in
// the caches, and for LDS reads and ds_bpermute.  256 CUs x 8 workgroups of
// 256 lanes; each lane issues iterations x 8 independent instructions; time
// from HIP events over 5 launches; cycles = time x clock x CUs / wave
// instructions.  Build: fdk_cuda_impl_amd/build.py (build_vmem_costs), run
// fdk_cuda_impl_amd/yk_vmem_costs.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define YK_CHECK(x)                                                                        \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                  \
            std::exit(1);                                                                  \
        }                                                                                  \
    } while (0)

namespace {

typedef float f2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kIter = 2048;

// W dwords per lane, lanes STRIDE bytes apart; offsets wrap at mask + 1
// bytes (the buffer is larger than mask + 16)
template <int W, int STRIDE>
__global__ void __launch_bounds__(kThreads) k_gload(const float* __restrict__ buf, float* out, unsigned mask)
{
    const int lane = int(threadIdx.x) & 63;
    unsigned off = unsigned(lane * STRIDE) + (threadIdx.x >> 6) * 4096u;
    const char* b = reinterpret_cast<const char*>(buf);
    float acc = 0.f;
    for (int it = 0; it < kIter; ++it) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const unsigned o = (off + unsigned(k) * 512u) & mask;
            if constexpr (W == 1) {
                acc += *reinterpret_cast<const float*>(b + o);
            } else if constexpr (W == 2) {
                const f2 t = *reinterpret_cast<const f2*>(b + o);
                acc += t.x + t.y;
            } else {
                const f4 t = *reinterpret_cast<const f4*>(b + o);
                acc += (t.x + t.y) + (t.z + t.w);
            }
        }
        off += 64;
    }
    if (acc == 12345.f) out[0] = acc;
}

// LDS: 0 ds_bpermute_b32 (8 independent chains), 1 ds_read_b32, 2 ds_read_b64,
// 3 ds_read_b128 (lanes on consecutive addresses; 8 distinct offsets per
// iteration), 4 ds_read_b64 with every lane on one address
template <int KIND>
__global__ void __launch_bounds__(kThreads) k_lds(float* out)
{
    __shared__ float lds[8192];
    const int lane = int(threadIdx.x) & 63;
    for (int k = int(threadIdx.x); k < 8192; k += kThreads) lds[k] = float(k);
    __syncthreads();
    float ch[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) ch[k] = float(lane + k);
    int idx = ((lane * 7 + 3) & 63) * 4;
    float acc = 0.f;
    for (int it = 0; it < kIter; ++it) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if constexpr (KIND == 0) {
                ch[k] = __int_as_float(__builtin_amdgcn_ds_bpermute(idx + k * 4, __float_as_int(ch[k])));
            } else if constexpr (KIND == 1) {
                acc += lds[(lane + k * 64 + it) & 1023];
            } else if constexpr (KIND == 2) {
                const f2 t = *reinterpret_cast<const f2*>(&lds[(2 * lane + k * 128 + 2 * it) & 2046]);
                acc += t.x + t.y;
            } else if constexpr (KIND == 3) {
                const f4 t = *reinterpret_cast<const f4*>(&lds[(4 * lane + k * 256 + 4 * it) & 4092]);
                acc += (t.x + t.y) + (t.z + t.w);
            } else {
                const f2 t = *reinterpret_cast<const f2*>(&lds[(k * 2 + 2 * it) & 2046]);
                acc += t.x + t.y;
            }
        }
        idx = (idx + 4) & 252;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += ch[k];
    if (acc == 12345.f) out[0] = acc;
}

using Kernel = void (*)();

template <typename F>
void time_kind(const char* name, int cus, double ghz, F launch)
{
    hipEvent_t e0, e1;
    YK_CHECK(hipEventCreate(&e0));
    YK_CHECK(hipEventCreate(&e1));
    launch();
    YK_CHECK(hipGetLastError());
    YK_CHECK(hipEventRecord(e0));
    for (int r = 0; r < 5; ++r) launch();
    YK_CHECK(hipEventRecord(e1));
    YK_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    YK_CHECK(hipEventElapsedTime(&ms, e0, e1));
    ms /= 5.f;
    const double instructions = double(cus) * 8 * (kThreads / 64) * double(kIter) * 8;
    const double cycles = double(ms) * 1e-3 * ghz * 1e9;
    std::printf("%-44s %7.3f ms  %6.2f cycles per wave64 instruction per CU\n", name, ms,
                cycles * cus / instructions);
    YK_CHECK(hipEventDestroy(e0));
    YK_CHECK(hipEventDestroy(e1));
}

}  // namespace

int main()
{
    hipDeviceProp_t prop;
    YK_CHECK(hipGetDeviceProperties(&prop, 0));
    const int cus = prop.multiProcessorCount;
    const double ghz = prop.clockRate / 1e6;
    const dim3 grid(cus * 8), block(kThreads);
    float* buf = nullptr;
    float* out = nullptr;
    const size_t bytes = size_t(4) << 20;
    YK_CHECK(hipMalloc(&buf, bytes + 64));
    YK_CHECK(hipMalloc(&out, 64));
    YK_CHECK(hipMemset(buf, 0, bytes + 64));
    const unsigned m16k = (16u << 10) - 1, m1m = (1u << 20) - 1;
    std::printf("CUs %d, clock %.3f GHz\n", cus, ghz);
    time_kind("global_load_dword, consecutive lanes (L1)", cus, ghz,
        [&] { hipLaunchKernelGGL((k_gload<1, 4>), grid, block, 0, 0, buf, out, m16k); });
    time_kind("global_load_dwordx2, consecutive lanes (L1)", cus, ghz,
        [&] { hipLaunchKernelGGL((k_gload<2, 8>), grid, block, 0, 0, buf, out, m16k); });
    time_kind("global_load_dwordx4, consecutive lanes (L1)", cus, ghz,
        [&] { hipLaunchKernelGGL((k_gload<4, 16>), grid, block, 0, 0, buf, out, m16k); });
    time_kind("global_load_dwordx4, lanes 8 bytes apart (L1)", cus, ghz,
        [&] { hipLaunchKernelGGL((k_gload<4, 8>), grid, block, 0, 0, buf, out, m16k); });
    time_kind("global_load_dwordx2, consecutive lanes (L2)", cus, ghz,
        [&] { hipLaunchKernelGGL((k_gload<2, 8>), grid, block, 0, 0, buf, out, m1m); });
    time_kind("global_load_dwordx4, consecutive lanes (L2)", cus, ghz,
        [&] { hipLaunchKernelGGL((k_gload<4, 16>), grid, block, 0, 0, buf, out, m1m); });
    time_kind("global_load_dword, one 128-byte line per lane", cus, ghz,
        [&] { hipLaunchKernelGGL((k_gload<1, 128>), grid, block, 0, 0, buf, out, m1m); });
    time_kind("global_load_dwordx2, one 128-byte line per lane", cus, ghz,
        [&] { hipLaunchKernelGGL((k_gload<2, 128>), grid, block, 0, 0, buf, out, m1m); });
    time_kind("ds_bpermute_b32", cus, ghz, [&] { hipLaunchKernelGGL(k_lds<0>, grid, block, 0, 0, out); });
    time_kind("ds_read_b32, consecutive lanes", cus, ghz, [&] { hipLaunchKernelGGL(k_lds<1>, grid, block, 0, 0, out); });
    time_kind("ds_read_b64, consecutive lanes", cus, ghz, [&] { hipLaunchKernelGGL(k_lds<2>, grid, block, 0, 0, out); });
    time_kind("ds_read_b128, consecutive lanes", cus, ghz, [&] { hipLaunchKernelGGL(k_lds<3>, grid, block, 0, 0, out); });
    time_kind("ds_read_b64, one address for all lanes", cus, ghz, [&] { hipLaunchKernelGGL(k_lds<4>, grid, block, 0, 0, out); });
    YK_CHECK(hipFree(buf));
    YK_CHECK(hipFree(out));
    return 0;
}

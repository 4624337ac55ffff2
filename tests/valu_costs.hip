// Cycles per wave64 VALU instruction of the instruction kinds the
// backprojection, filter and projector kernels use, measured on the GPU it
// runs on: one kernel per kind, each lane issuing 2048 x 16 copies on 8
// independent registers (inline-asm ALU instructions: no memory access in
// the loop), 8 waves per SIMD; time from HIP events over 5 launches.
// Prints one line per kind; profiles/valu_instruction_costs_gfx950.txt holds
// the MI355X result.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

constexpr int kIter = 2048;

#define YK_CHECK(call)                                                                  \
    do {                                                                                \
        const hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                         \
            std::fprintf(stderr, "%s: %s\n", #call, hipGetErrorString(e_));             \
            std::exit(1);                                                               \
        }                                                                               \
    } while (0)

#define YK_ONE(INS, R) asm volatile(INS : "+v"(R) : "v"(x), "v"(y));
#define YK_EIGHT(INS)                                                                   \
    YK_ONE(INS, r0) YK_ONE(INS, r1) YK_ONE(INS, r2) YK_ONE(INS, r3)                     \
    YK_ONE(INS, r4) YK_ONE(INS, r5) YK_ONE(INS, r6) YK_ONE(INS, r7)
#define YK_KERNEL(NAME, INS)                                                            \
    __global__ void __launch_bounds__(256) NAME(unsigned* out, unsigned s)              \
    {                                                                                   \
        unsigned r0 = threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3;              \
        unsigned r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;                    \
        const unsigned x = 0x3f800000u + (threadIdx.x & 7u), y = 0x3c003c00u ^ s;       \
        for (int it = 0; it < kIter; ++it) { YK_EIGHT(INS) YK_EIGHT(INS) }              \
        out[blockIdx.x * 256 + threadIdx.x] = r0 ^ r1 ^ r2 ^ r3 ^ r4 ^ r5 ^ r6 ^ r7;   \
    }

YK_KERNEL(k_add_f32, "v_add_f32 %0, %0, %1")
YK_KERNEL(k_mul_f32, "v_mul_f32 %0, %0, %1")
YK_KERNEL(k_fmac_f32, "v_fmac_f32 %0, %1, %2")
YK_KERNEL(k_fma_f32, "v_fma_f32 %0, %1, %2, %0")
YK_KERNEL(k_fmamk_f32, "v_fmamk_f32 %0, %1, 0x3f800001, %0")
YK_KERNEL(k_add_u32, "v_add_u32 %0, %0, %1")
YK_KERNEL(k_fract_f32, "v_fract_f32 %0, %0")
YK_KERNEL(k_cvt_i32_f32, "v_cvt_i32_f32 %0, %0")
YK_KERNEL(k_cvt_f32_i32, "v_cvt_f32_i32 %0, %0")
YK_KERNEL(k_cvt_pkrtz, "v_cvt_pkrtz_f16_f32 %0, %0, %1")
YK_KERNEL(k_pk_fma_f16, "v_pk_fma_f16 %0, %1, %2, %0")
YK_KERNEL(k_pk_add_f16, "v_pk_add_f16 %0, %0, %2")
YK_KERNEL(k_dot2c, "v_dot2c_f32_f16 %0, %1, %2")
YK_KERNEL(k_mad_u24, "v_mad_u32_u24 %0, %0, %1, %2")
YK_KERNEL(k_mad_u16, "v_mad_u32_u16 %0, %0, %1, %2 op_sel:[1,0,0,0]")
YK_KERNEL(k_mul_lo_u32, "v_mul_lo_u32 %0, %0, %1")
YK_KERNEL(k_lshl_add, "v_lshl_add_u32 %0, %0, 3, %1")
YK_KERNEL(k_perm, "v_perm_b32 %0, %0, %1, %2")
YK_KERNEL(k_rcp_f32, "v_rcp_f32 %0, %0")
YK_KERNEL(k_fma_mix, "v_fma_mix_f32 %0, %0, %1, %2 op_sel_hi:[0,1,1]")
YK_KERNEL(k_cvt_f32_f16, "v_cvt_f32_f16 %0, %0")
YK_KERNEL(k_lshrrev, "v_lshrrev_b32 %0, 3, %0")
YK_KERNEL(k_and_b32, "v_and_b32 %0, %0, %1")
YK_KERNEL(k_or_b32, "v_or_b32 %0, %0, %1")
YK_KERNEL(k_mul_u24, "v_mul_u32_u24 %0, %0, %1")
YK_KERNEL(k_bfe_u32, "v_bfe_u32 %0, %0, 3, 9")
YK_KERNEL(k_max_f32, "v_max_f32 %0, %0, %1")
YK_KERNEL(k_sub_f32, "v_sub_f32 %0, %0, %1")
YK_KERNEL(k_ldexp_f32, "v_ldexp_f32 %0, %0, 1")
YK_KERNEL(k_floor_f32, "v_floor_f32 %0, %0")
YK_KERNEL(k_med3_f32, "v_med3_f32 %0, %0, %1, %2")
YK_KERNEL(k_sqrt_f32, "v_sqrt_f32 %0, %0")

using Kernel = void (*)(unsigned*, unsigned);

}  // namespace

int main()
{
    hipDeviceProp_t prop;
    YK_CHECK(hipGetDeviceProperties(&prop, 0));
    const int cus = prop.multiProcessorCount;
    const double ghz = prop.clockRate / 1e6;
    const int blocks = cus * 8;
    unsigned* out = nullptr;
    YK_CHECK(hipMalloc(&out, sizeof(unsigned) * size_t(blocks) * 256));
    const struct { const char* name; Kernel fn; } kinds[] = {
        { "v_add_f32", k_add_f32 }, { "v_mul_f32", k_mul_f32 }, { "v_fmac_f32", k_fmac_f32 },
        { "v_fma_f32", k_fma_f32 }, { "v_fmamk_f32", k_fmamk_f32 }, { "v_add_u32", k_add_u32 },
        { "v_fract_f32", k_fract_f32 }, { "v_cvt_i32_f32", k_cvt_i32_f32 }, { "v_cvt_f32_i32", k_cvt_f32_i32 },
        { "v_cvt_pkrtz_f16_f32", k_cvt_pkrtz }, { "v_pk_fma_f16", k_pk_fma_f16 }, { "v_pk_add_f16", k_pk_add_f16 },
        { "v_dot2c_f32_f16", k_dot2c }, { "v_mad_u32_u24", k_mad_u24 }, { "v_mad_u32_u16", k_mad_u16 },
        { "v_mul_lo_u32", k_mul_lo_u32 }, { "v_lshl_add_u32", k_lshl_add }, { "v_perm_b32", k_perm },
        { "v_rcp_f32", k_rcp_f32 }, { "v_fma_mix_f32", k_fma_mix }, { "v_cvt_f32_f16", k_cvt_f32_f16 },
        { "v_lshrrev_b32", k_lshrrev }, { "v_and_b32", k_and_b32 }, { "v_or_b32", k_or_b32 },
        { "v_mul_u32_u24", k_mul_u24 }, { "v_bfe_u32", k_bfe_u32 }, { "v_max_f32", k_max_f32 },
        { "v_sub_f32", k_sub_f32 }, { "v_ldexp_f32", k_ldexp_f32 }, { "v_floor_f32", k_floor_f32 },
        { "v_med3_f32", k_med3_f32 }, { "v_sqrt_f32", k_sqrt_f32 },
    };
    std::printf("CUs %d, clock %.3f GHz\n", cus, ghz);
    hipEvent_t e0, e1;
    YK_CHECK(hipEventCreate(&e0));
    YK_CHECK(hipEventCreate(&e1));
    for (const auto& k : kinds) {
        hipLaunchKernelGGL(k.fn, dim3(blocks), dim3(256), 0, 0, out, 1u);
        YK_CHECK(hipGetLastError());
        YK_CHECK(hipEventRecord(e0));
        for (int r = 0; r < 5; ++r) hipLaunchKernelGGL(k.fn, dim3(blocks), dim3(256), 0, 0, out, 1u);
        YK_CHECK(hipEventRecord(e1));
        YK_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        YK_CHECK(hipEventElapsedTime(&ms, e0, e1));
        ms /= 5.f;
        const double instructions = double(blocks) * 4 * kIter * 16;  // wave64 instructions
        const double cycles = double(ms) * 1e-3 * ghz * 1e9;
        std::printf("%-22s %7.3f ms  %.2f cycles per wave64 instruction per SIMD\n", k.name, ms,
                    cycles * cus * 4 / instructions);
    }
    YK_CHECK(hipFree(out));
    return 0;
}

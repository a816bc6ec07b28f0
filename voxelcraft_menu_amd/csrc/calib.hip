// calib.hip -- measures the VALU issue ceiling the march kernels are held against.
//
// The march kernels are VALU-issue bound (integer mask logic and f32 stepping, no dense
// contraction), so their roofline is vector-instruction issue, not HBM or MFMA. This probe
// runs a grid that fills every SIMD with waves whose loop bodies are independent 32-bit
// VALU operations, so nothing but issue limits it, and reports wave64 VALU instructions
// per second. The published MI355X FP32 vector peak (157.3 TFLOP/s = 78.6 T lane FMAs/s =
// 1.23 T wave64 instructions/s, two per four cycles per SIMD through dual issue) is the
// figure the fraction is taken against; this measurement shows how much of it a plain
// instruction stream reaches on the box.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kChains = 8;  // independent dependency chains per lane

// kind 0: f32 FMA chains (one v_fmamk_f32 per chain and trip); kind 1: 32-bit integer
// chains (v_lshrrev_b32 + v_xad_u32 per chain and trip, shift/xor/add as in mask logic).
template <int kKind>
__global__ __launch_bounds__(256) void k_valu_probe(uint32_t *out, uint32_t iters, uint32_t seed) {
    if constexpr (kKind == 0) {
        float a[kChains];
#pragma unroll
        for (int k = 0; k < kChains; k++) a[k] = (float)(threadIdx.x + k + seed) * 1e-3f;
        for (uint32_t i = 0; i < iters; i++) {
#pragma unroll
            for (int k = 0; k < kChains; k++) a[k] = __builtin_fmaf(a[k], 0.999f, 0.25f);
        }
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < kChains; k++) s += a[k];
        if (s == 1234.5f) out[blockIdx.x] = 1u;  // keeps the chains live
    } else {
        uint32_t a[kChains];
#pragma unroll
        for (int k = 0; k < kChains; k++) a[k] = threadIdx.x * 2654435761u + (uint32_t)k + seed;
        for (uint32_t i = 0; i < iters; i++) {
#pragma unroll
            for (int k = 0; k < kChains; k++) a[k] = (a[k] ^ (a[k] >> 7)) + 0x9E3779B9u;
        }
        uint32_t s = 0;
#pragma unroll
        for (int k = 0; k < kChains; k++) s ^= a[k];
        if (s == 0x12345678u) out[blockIdx.x] = 1u;
    }
}

}  // namespace

// Runs the probe (kind 0: f32 FMA, 1: integer xor/add) over `blocks` workgroups of 256
// lanes for `iters` loop trips, `reps` times; returns the mean milliseconds per launch and
// the VALU instructions per wave the loop body issues (chains x ops per chain x iters).
extern "C" int vxc_valu_probe(int kind, uint32_t blocks, uint32_t iters, uint32_t reps, double *ms_per_launch,
                              double *loop_insts_per_wave) {
    if (!ms_per_launch || !loop_insts_per_wave || blocks == 0 || reps == 0 || (kind != 0 && kind != 1)) return -3;
    uint32_t *out = nullptr;
    if (hipMalloc(&out, (size_t)blocks * 4) != hipSuccess) return -2;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        (void)hipFree(out);
        return -6;
    }
    auto launch = [&](uint32_t seed) {
        if (kind == 0) k_valu_probe<0><<<blocks, 256>>>(out, iters, seed);
        else k_valu_probe<1><<<blocks, 256>>>(out, iters, seed);
    };
    launch(1);  // warm-up
    (void)hipEventRecord(e0, nullptr);
    for (uint32_t r = 0; r < reps; r++) launch(r + 2);
    (void)hipEventRecord(e1, nullptr);
    hipError_t err = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (err != hipSuccess) return -6;
    *ms_per_launch = (double)ms / reps;
    *loop_insts_per_wave = (double)kChains * (kind == 0 ? 1.0 : 2.0) * iters;  // v_fmamk_f32; v_lshrrev + v_xad_u32
    return 0;
}

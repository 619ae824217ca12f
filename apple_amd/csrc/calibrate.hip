// Measurement support for the roofline: the issue rate of the VALU instruction classes the hot
// kernels use, and the Keccak-f[1600] rate of the one-lane form on register-resident states, on
// this chip, with the shader clock read during the run (s_memtime against the 100 MHz
// s_memrealtime).  No counterpart in the reference (its harness times kernels only).
//
// gfx950 issues a wave64 v_xor_b32 / v_bitop3_b32 / v_add_u32 in 2 SIMD cycles but
// v_alignbit_b32 / v_min_u32 / v_mad_u64_u32 in 4: an int32 "peak" of one lane-op per lane per
// cycle (the 4-cycle model) is exceeded by the Keccak kernels, so the roofline is stated in
// issue cycles per permutation instead.
#include "device_keccak.hpp"
#include "runtime_common.hpp"

using namespace azk;
using namespace azk::rt;

namespace {

constexpr uint32_t kChains = 8;

// iters x 2 x 8 instructions of one form in 8 independent chains per lane, registers only.
#define AZK_CAL_CHAINS(ASM, ...)                                                                    \
    asm volatile(ASM : "+v"(a[0]) : __VA_ARGS__); asm volatile(ASM : "+v"(a[1]) : __VA_ARGS__);   \
    asm volatile(ASM : "+v"(a[2]) : __VA_ARGS__); asm volatile(ASM : "+v"(a[3]) : __VA_ARGS__);   \
    asm volatile(ASM : "+v"(a[4]) : __VA_ARGS__); asm volatile(ASM : "+v"(a[5]) : __VA_ARGS__);   \
    asm volatile(ASM : "+v"(a[6]) : __VA_ARGS__); asm volatile(ASM : "+v"(a[7]) : __VA_ARGS__);

template <int OP>
__global__ __launch_bounds__(256) void k_cal_op(uint32_t *out, uint32_t iters, unsigned long long *clk)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    using W = typename std::conditional<OP == 5, uint64_t, uint32_t>::type;
    W a[kChains];
#pragma unroll
    for (uint32_t k = 0; k < kChains; ++k) a[k] = t + k;
    const uint32_t b = (t & 15) + 1, c = t * 5 + 7;
    unsigned long long c0 = 0, r0 = 0;
    if (t == 0) { c0 = __builtin_readcyclecounter(); r0 = wall_clock64(); }
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            if constexpr (OP == 0) { AZK_CAL_CHAINS("v_xor_b32 %0, %0, %1", "v"(b)) }
            if constexpr (OP == 1) { AZK_CAL_CHAINS("v_bitop3_b32 %0, %0, %1, %2 bitop3:0x96", "v"(b), "v"(c)) }
            if constexpr (OP == 2) { AZK_CAL_CHAINS("v_alignbit_b32 %0, %0, %1, %2", "v"(b), "v"(c)) }
            if constexpr (OP == 3) { AZK_CAL_CHAINS("v_add_u32 %0, %0, %1", "v"(b)) }
            if constexpr (OP == 4) { AZK_CAL_CHAINS("v_min_u32 %0, %0, %1", "v"(b)) }
            if constexpr (OP == 5) { AZK_CAL_CHAINS("v_mad_u64_u32 %0, s[40:41], %1, %2, %0", "v"(b), "v"(c) : "s40", "s41") }
        }
    }
    if (t == 0) { clk[0] = __builtin_readcyclecounter() - c0; clk[1] = wall_clock64() - r0; }
    uint32_t x = 0;
#pragma unroll
    for (uint32_t k = 0; k < kChains; ++k) x ^= (uint32_t)a[k];
    out[t] = x;
}

// `perms` one-lane permutations of a register-resident state per lane.
__global__ __launch_bounds__(256) void k_cal_keccak(uint32_t *out, uint32_t perms, unsigned long long *clk)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    u2 a[25];
#pragma unroll
    for (int k = 0; k < 25; ++k) a[k] = {t * 2654435761u + k, t ^ (k * 40503u)};
    unsigned long long c0 = 0, r0 = 0;
    if (t == 0) { c0 = __builtin_readcyclecounter(); r0 = wall_clock64(); }
    for (uint32_t i = 0; i < perms; ++i) kf_v(a);
    if (t == 0) { clk[0] = __builtin_readcyclecounter() - c0; clk[1] = wall_clock64() - r0; }
    uint32_t x = 0;
#pragma unroll
    for (int k = 0; k < 25; ++k) x ^= a[k].lo ^ a[k].hi;
    out[t] = x;
}

// `perms` dependent permutations of one state on one wave in the one-state-per-wave form; the
// 100 MHz clock is read by lane 0 around the chain.
__global__ __launch_bounds__(64) void k_lat_wave(uint32_t *out, uint32_t perms, unsigned long long *clk)
{
    const WaveLane c = wave_lane();
    uint32_t lo = c.word < 25 ? c.word * 2654435761u : 0u, hi = c.word < 25 ? c.word ^ 0x5bd1e995u : 0u;
    unsigned long long r0 = 0;
    if (threadIdx.x == 0) r0 = wall_clock64();
    for (uint32_t i = 0; i < perms; ++i) kf_w(lo, hi, c);
    if (threadIdx.x == 0) clk[0] = wall_clock64() - r0;
    out[threadIdx.x] = lo ^ hi;
}

// The same with 32 states on one wave in the two-lanes-per-state form.
__global__ __launch_bounds__(64) void k_lat_pair(uint32_t *out, uint32_t perms, unsigned long long *clk)
{
    const bool hi = threadIdx.x & 1u;
    uint32_t a[25];
#pragma unroll
    for (int k = 0; k < 25; ++k) a[k] = (threadIdx.x >> 1) * 2654435761u + k * (hi ? 40503u : 7919u);
    unsigned long long r0 = 0;
    if (threadIdx.x == 0) r0 = wall_clock64();
    for (uint32_t i = 0; i < perms; ++i) kf_h(a, hi);
    if (threadIdx.x == 0) clk[0] = wall_clock64() - r0;
    uint32_t x = 0;
#pragma unroll
    for (int k = 0; k < 25; ++k) x ^= a[k];
    out[threadIdx.x] = x;
}

__global__ void k_lat_empty() {}

typedef void (*cal_fn)(uint32_t *, uint32_t, unsigned long long *);

// Best of `reps` launches of `grid` x 256 lanes: elapsed ms and the clock (MHz) of that launch.
azk_status time_launch(hipStream_t st, cal_fn fn, uint32_t grid, uint32_t arg, uint32_t *out, unsigned long long *clk,
                       int reps, double &ms_best, double &mhz)
{
    Timer tm;
    azk_status s = tm.init();
    if (s) return s;
    ms_best = 1e30;
    for (int r = 0; r < reps; ++r) {
        HIP_TRY(hipEventRecord(tm.a, st));
        hipLaunchKernelGGL(fn, dim3(grid), dim3(256), 0, st, out, arg, clk);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(tm.b, st));
        HIP_TRY(hipEventSynchronize(tm.b));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, tm.a, tm.b));
        unsigned long long h[2] = {0, 0};
        HIP_TRY(hipMemcpyAsync(h, clk, 16, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (ms < ms_best) {
            ms_best = ms;
            mhz = h[1] ? (double)h[0] / ((double)h[1] / 100.0) : 0.0;
        }
    }
    return AZK_OK;
}

} // namespace

extern "C" azk_status azk_calibrate_valu(azk_context *ctx, azk_valu_calibration *out)
{
    CHECK_ARG(ctx && out);
    memset(out, 0, sizeof *out);
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> g(ctx->mu);
    constexpr uint32_t kGrid = 8192, kIters = 1024, kKeccakGrid = 4096, kPerms = 64;
    DeviceMem buf;
    azk_status s = buf.ensure((size_t)kGrid * 256 * 4 + 256);
    if (s) return s;
    uint32_t *words = (uint32_t *)buf.p;
    unsigned long long *clk = (unsigned long long *)((uint8_t *)buf.p + (size_t)kGrid * 256 * 4);
    hipStream_t st = ctx->stream;
    double ms = 0, mhz = 0;
    // warm the clock with the Keccak loop, then measure it
    s = time_launch(st, k_cal_keccak, kKeccakGrid, kPerms, words, clk, 6, ms, mhz);
    if (s) return s;
    const double lanes = (double)kKeccakGrid * 256;
    const double simds = (double)ctx->props.multiProcessorCount * 4;
    out->keccak_perms_per_s = lanes * kPerms / (ms * 1e-3);
    out->keccak_clock_mhz = mhz;
    out->keccak_cycles_per_wave_perm = ms * 1e-3 * simds * mhz * 1e6 / (lanes * kPerms / 64);
    const cal_fn ops[6] = {k_cal_op<0>, k_cal_op<1>, k_cal_op<2>, k_cal_op<3>, k_cal_op<4>, k_cal_op<5>};
    for (int k = 0; k < 6; ++k) {
        s = time_launch(st, ops[k], kGrid, kIters, words, clk, 3, ms, mhz);
        if (s) return s;
        const double instrs = (double)kGrid * 256 / 64 * kIters * 2 * kChains;  // wave instructions
        out->cycles_per_instr[k] = ms * 1e-3 * simds * mhz * 1e6 / instrs;
        out->clock_mhz[k] = mhz;
    }
    return AZK_OK;
}

extern "C" azk_status azk_calibrate_latency(azk_context *ctx, azk_latency_calibration *out)
{
    CHECK_ARG(ctx && out);
    memset(out, 0, sizeof *out);
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> g(ctx->mu);
    constexpr uint32_t kChain = 241, kNodes = 64, kReps = 5;
    DeviceMem buf;
    azk_status s = buf.ensure(64 * 4 + 64);
    if (s) return s;
    uint32_t *words = (uint32_t *)buf.p;
    unsigned long long *clk = (unsigned long long *)((uint8_t *)buf.p + 256);
    hipStream_t st = ctx->stream;
    const cal_fn chains[2] = {k_lat_wave, k_lat_pair};
    double *dst[2] = {&out->wave_perm_us, &out->pair_perm_us};
    for (int k = 0; k < 2; ++k) {
        double best = 1e30;
        for (uint32_t r = 0; r < kReps; ++r) {
            hipLaunchKernelGGL(chains[k], dim3(1), dim3(64), 0, st, words, kChain, clk);
            HIP_TRY(hipGetLastError());
            unsigned long long ticks = 0;
            HIP_TRY(hipMemcpyAsync(&ticks, clk, 8, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            best = std::min(best, (double)ticks / 100.0 / kChain);  // 100 MHz ticks -> us
        }
        *dst[k] = best;
    }
    // back-to-back empty kernels in a captured graph
    hipGraph_t gr = nullptr;
    hipGraphExec_t ge = nullptr;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (uint32_t i = 0; i < kNodes; ++i) hipLaunchKernelGGL(k_lat_empty, dim3(1), dim3(64), 0, st);
    hipError_t e = hipStreamEndCapture(st, &gr);
    if (e == hipSuccess) e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
    if (gr) (void)hipGraphDestroy(gr);
    if (e != hipSuccess) return fail(AZK_ERR_COMMAND_FAILED, "graph of empty kernels: %s", hipGetErrorString(e));
    Timer tm;
    s = tm.init();
    double best = 1e30;
    for (uint32_t r = 0; !s && r < kReps + 1; ++r) {
        hipError_t le = hipEventRecord(tm.a, st);
        if (le == hipSuccess) le = hipGraphLaunch(ge, st);
        if (le == hipSuccess) le = hipEventRecord(tm.b, st);
        if (le == hipSuccess) le = hipEventSynchronize(tm.b);
        if (le != hipSuccess) {
            s = fail(AZK_ERR_COMMAND_FAILED, "graph launch: %s", hipGetErrorString(le));
            break;
        }
        if (r) best = std::min(best, (double)tm.elapsed_ns() / 1e3 / kNodes);  // the first launch warms
    }
    (void)hipGraphExecDestroy(ge);
    if (s) return s;
    out->graph_kernel_us = best;
    return AZK_OK;
}

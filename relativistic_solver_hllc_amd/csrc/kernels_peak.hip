// Register-only MFMA throughput loops: the achievable matrix-core rate of the
// two instructions the contraction kernels are built on, measured on the
// device the library runs on (the denominator of their roofline fractions).
//
// Every wave keeps kPeakAcc independent accumulators and issues kPeakAcc MFMAs
// per iteration, so no MFMA waits for the result of the one before it; the
// operands never change and nothing touches memory until the accumulators are
// written once at the end (so that the compiler keeps the work). The loop body
// is kPeakUnroll rounds: the compiler carries the accumulators across the loop
// back-edge in VGPRs and copies them into and out of the AGPRs there, which
// with a one-round body halved the issue rate.
#include "mfma_split_core.hpp"

namespace rsh {

namespace {

constexpr int kPeakAcc = 8;
constexpr int kPeakUnroll = 32;  // rounds of kPeakAcc MFMAs per loop iteration; iters is a multiple

typedef double v4d __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void mfma_f64_peak_kernel(int iters, double seed, double* __restrict__ sink) {
    const int lane = threadIdx.x & 63;
    const double a = seed + lane, b = 1.0 / (seed + lane + 1);
    v4d acc[kPeakAcc];
#pragma unroll
    for (int i = 0; i < kPeakAcc; i++) acc[i] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int it = 0; it < iters; it += kPeakUnroll) {
#pragma unroll
        for (int u = 0; u < kPeakUnroll; u++)
#pragma unroll
            for (int i = 0; i < kPeakAcc; i++) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < kPeakAcc; i++) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
    sink[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void mfma_bf16_peak_kernel(int iters, float seed, double* __restrict__ sink) {
    const int lane = threadIdx.x & 63;
    v8bf a, b;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        a[e] = (__bf16)(seed + (float)(lane + e));
        b[e] = (__bf16)(1.0f / (seed + (float)(lane + e + 1)));
    }
    v4f acc[kPeakAcc];
#pragma unroll
    for (int i = 0; i < kPeakAcc; i++) acc[i] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int it = 0; it < iters; it += kPeakUnroll) {
#pragma unroll
        for (int u = 0; u < kPeakUnroll; u++)
#pragma unroll
            for (int i = 0; i < kPeakAcc; i++) acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i], 0, 0, 0);
    }
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < kPeakAcc; i++) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
    sink[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = (double)s;
}

}  // namespace

// kind 0: v_mfma_f64_16x16x4f64 (2 * 16 * 16 * 4 flops), 1: v_mfma_f32_16x16x32_bf16
// (2 * 16 * 16 * 32 flops). blocks x 256 threads; sink holds blocks * 256 doubles.
hipError_t launch_mfma_peak(int kind, int blocks, int iters, double* sink, hipStream_t st) {
    if (kind == 0)
        hipLaunchKernelGGL(mfma_f64_peak_kernel, dim3(blocks), dim3(256), 0, st, iters, 1.0, sink);
    else
        hipLaunchKernelGGL(mfma_bf16_peak_kernel, dim3(blocks), dim3(256), 0, st, iters, 1.0f, sink);
    return hipGetLastError();
}

double mfma_peak_flops_per_wave_iter(int kind) {
    return (double)kPeakAcc * 2.0 * 16 * 16 * (kind == 0 ? 4 : 32);
}

}  // namespace rsh

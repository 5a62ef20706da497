// Source iteration I <- (1 - a) B + a S(I) on cell-level state, for many
// uneven problems in one persistent ("grouped") launch per row-block class.
//
// Work list: items (problem, pass, cell tile) in device memory, consumed by
// persistent workgroups through an atomic counter, so the host can change the
// active set between launches by rewriting the list (a captured hipGraph stays
// valid). Each item runs the FP64 MFMA main loop on 16 W cells x RB row blocks
// and a fused epilogue:
//   I_out[k][c] = (1 - a) B[k][c] + a S[k][c]
//   resid       = max |I_in[k][c] - I_out[k][c]| / I_out[k][c]   (FP64; the
//                 residual of ReCalcIllum, Illum_async.cpp:472-510)
// reduced per wave and merged with atomicMax on the bit pattern (valid for
// non-negative doubles).
#include "mfma_f64_core.hpp"


namespace rsh {

__device__ inline double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Zeroes the residual slot `parity` of every problem and the item counters.
__global__ __launch_bounds__(256) void batch_zero_kernel(BatchProblem* __restrict__ probs, int nprobs,
                                                         unsigned* __restrict__ counters, int ncounters, int parity) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nprobs) probs[i].resid[parity] = 0ull;
    if (i < ncounters) counters[i] = 0u;
}

// Fused epilogue of one item for one lane: rows (rb0 + rb) * 16 + 4 r + jr.
// The state and source arrays hold whole row blocks: padding rows past M are
// zero and so are their panel rows, so they get nv = 0 and a residual term
// 0 / 0 = NaN, which fmax ignores; no row needs a branch. Accesses go through
// a buffer descriptor per row block: one per-lane 32-bit offset (cell and jr)
// and a uniform offset per row, so a load costs no VGPRs beyond its result.
// The loads (B and the previous state) of row block rb + kEpiAhead are issued,
// in program order, before the stores of row block rb: with kEpiAhead = 0 the
// 8 loads of a row block go out together (13 round trips per item at RB = 13,
// not 52); looking 1 or 2 row blocks further ahead needs more than the 168
// VGPRs of 3 waves per SIMD and measured slower at 2. Offsets stay below 2^31
// for n < 2^24 cells (checked on the host). Returns the lane's max relative
// change.
constexpr int kEpiAhead = 0;
constexpr int kBufWord3 = 0x00020000;  // raw (untyped) gfx9 buffer access, 32-bit format
typedef unsigned int v2u __attribute__((ext_vector_type(2)));

template <int RB>
__device__ __forceinline__ double iterate_epilogue(const double* bsrc, const double* isrc, double* odst, double oma,
                                                   double alb, int64_t n, int rb0, int jr, int64_t cell,
                                                   const v4d (&acc)[RB]) {
    constexpr int SLOTS = kEpiAhead + 1;
    const int rb_bytes = (int)(16 * n * 8);
    auto rsrc = [&](const double* a, int rb) {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(a) + (int64_t)(rb0 + rb) * 16 * n, (short)0,
                                                 rb_bytes, kBufWord3);
    };
    const int voff = (int)(((int64_t)jr * n + cell) * 8);
    const int nb = (int)(n * 8);
    double bv[SLOTS][4], ov[SLOTS][4];
    auto load = [&](int rb) {
        const __amdgpu_buffer_rsrc_t bd = rsrc(bsrc, rb), id = rsrc(isrc, rb);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            bv[rb % SLOTS][r] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(bd, voff, 4 * r * nb, 0));
            ov[rb % SLOTS][r] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(id, voff, 4 * r * nb, 0));
        }
    };
#pragma unroll
    for (int rb = 0; rb < kEpiAhead && rb < RB; rb++) load(rb);
    double res = 0.0;
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        if (rb + kEpiAhead < RB) load(rb + kEpiAhead);
        const __amdgpu_buffer_rsrc_t od = rsrc(odst, rb);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const double nv = oma * bv[rb % SLOTS][r] + alb * acc[rb][r];
            res = fmax(res, fabs(ov[rb % SLOTS][r] - nv) / nv);
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, nv), od, voff, 4 * r * nb, 0);
        }
    }
    return res;
}

template <int RB, int W, int SC>
__global__ __launch_bounds__(64 * W, W == 4 ? 3 : 2) void grouped_iter_f64_kernel(const BatchProblem* __restrict__ probs,
                                                                  const BatchItem* __restrict__ items,
                                                                  const int* __restrict__ n_items,
                                                                  unsigned* __restrict__ counter, int parity) {
    __shared__ v2d lds[2][F64TileShape<RB, W, SC>::UNITS];
    __shared__ int s_item;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int jr = lane >> 4;
    const int total = *n_items;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it >= total) break;
        const BatchItem item = items[it];
        const BatchProblem& P = probs[item.problem];
        const double* __restrict__ in = P.I[parity];
        double* __restrict__ out = P.I[parity ^ 1];
        const int64_t cell = item.tile * (16 * W) + wave * 16 + (lane & 15);
        v4d acc[RB];
        mfma_f64_accumulate<RB, W, SC, 1>(in, P.n, P.n, P.m, P.panel + (int64_t)item.pass * P.steps * RB * 64, P.steps,
                                          cell, lds, acc);
        double res = 0.0;
        if (cell < P.n)
            res = iterate_epilogue<RB>(P.B, in, out, P.one_minus_albedo, P.albedo, P.n, item.pass * RB, jr, cell, acc);
        res = wave_max(res);
        if (lane == 0 && res > 0.0)
            atomicMax(P.resid + parity, (unsigned long long)__double_as_longlong(res));
    }
}

hipError_t launch_batch_zero(BatchProblem* d_probs, int nprobs, unsigned* d_counters, int ncounters, int parity,
                             hipStream_t st) {
    const int n = nprobs > ncounters ? nprobs : ncounters;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_zero_kernel, dim3((n + 255) / 256), dim3(256), 0, st, d_probs, nprobs, d_counters,
                       ncounters, parity);
    return hipGetLastError();
}

// Row-block counts of the instances: every kRbSet value up to kGroupedF64MaxRb.
static_assert(kGroupedF64MaxRb == 13, "instance list below covers row blocks 1..13");

template <int W, int SC>
static hipError_t launch_grouped_rb(int rb, int grid, hipStream_t st, const BatchProblem* probs,
                                    const BatchItem* items, const int* n_items, unsigned* counter, int parity) {
#define RSH_RB_CASE(V)                                                                                  \
    case V:                                                                                             \
        hipLaunchKernelGGL((grouped_iter_f64_kernel<V, W, SC>), dim3(grid), dim3(64 * W), 0, st, probs, \
                           items, n_items, counter, parity);                                            \
        break;
    switch (rb) {
        RSH_RB_CASE(1) RSH_RB_CASE(2) RSH_RB_CASE(3) RSH_RB_CASE(4) RSH_RB_CASE(5) RSH_RB_CASE(6)
        RSH_RB_CASE(7) RSH_RB_CASE(8) RSH_RB_CASE(10) RSH_RB_CASE(13)
        default:
            return hipErrorInvalidValue;
    }
#undef RSH_RB_CASE
    return hipGetLastError();
}

hipError_t launch_grouped_iter_f64(int rb, int waves, int chunk_steps, int grid, const BatchProblem* d_probs,
                                   const BatchItem* d_items, const int* d_n_items, unsigned* d_counter, int parity,
                                   hipStream_t st) {
    if (waves == 8 && chunk_steps == 2)
        return launch_grouped_rb<8, 2>(rb, grid, st, d_probs, d_items, d_n_items, d_counter, parity);
    if (waves == 4 && chunk_steps == 2)
        return launch_grouped_rb<4, 2>(rb, grid, st, d_probs, d_items, d_n_items, d_counter, parity);
    if (waves == 4 && chunk_steps == 4)
        return launch_grouped_rb<4, 4>(rb, grid, st, d_probs, d_items, d_n_items, d_counter, parity);
    return hipErrorInvalidValue;
}

// Workgroups of a grouped launch that can be resident at once.
int grouped_iter_f64_resident(int rb, int waves, int chunk_steps, int device) {
    int per_cu = 0;
    hipError_t e = hipErrorInvalidValue;
#define RSH_OCC(V, W, SC)                                                                                         \
    if (rb == V && waves == W && chunk_steps == SC)                                                               \
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, grouped_iter_f64_kernel<V, W, SC>, 64 * W, 0);
#define RSH_OCC_ALL(V) RSH_OCC(V, 8, 2) RSH_OCC(V, 4, 2) RSH_OCC(V, 4, 4)
    RSH_OCC_ALL(1) RSH_OCC_ALL(2) RSH_OCC_ALL(3) RSH_OCC_ALL(4) RSH_OCC_ALL(5) RSH_OCC_ALL(6) RSH_OCC_ALL(7)
    RSH_OCC_ALL(8) RSH_OCC_ALL(10) RSH_OCC_ALL(13)
#undef RSH_OCC_ALL
#undef RSH_OCC
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    hipDeviceProp_t prop;
    int cus = 256;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        cus = prop.multiProcessorCount;
    return per_cu * cus;
}

}  // namespace rsh

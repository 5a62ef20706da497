// Source iteration on a split-precision state (cell-level, packed hi/lo[/mid]
// parts, per-cell power-of-two scale), grouped over many problems like the
// FP64 grouped kernel (kernels_iter.hip).
//
// Epilogue, per output (k, c): S in FP64 from the main and cross FP32
// sums and the exact scales; I_new = (1 - a) B + a S in FP64; the old
// value is the state the MFMA consumed (parts summed, rescaled exactly); the
// residual |I_old - I_new| / I_new is FP64 (its maximum kept as a ratio,
// RelMax); I_new is re-split into the packed
// layout of the other ping-pong buffer.
//
// Packed position of (row k, cell c): cell tile c / 16, K-step k / 32, vector
// lane ((k % 32) / 8) * 16 + c % 16, element k % 8. In the MFMA D layout a
// lane (g = l / 16, cl = l % 16) holds rows 16 rb + 4 g + r, r = 0..3, of row
// block rb: four consecutive elements (8 bytes at offset 8 (g % 2)) of vector
// lane (2 (rb % 2) + g / 2) * 16 + cl in K-step rb / 2. So every lane reads and
// writes 8 bytes per part, and lane pairs cover whole 16-byte vectors.
#include "mfma_split_core.hpp"

namespace rsh {

__device__ inline double wave_max_split(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

template <int RB, int W, int FMT, int NT>
__global__ __launch_bounds__(64 * W) void grouped_iter_split_kernel(const BatchProblem* __restrict__ probs,
                                                                    const BatchItem* __restrict__ items,
                                                                    const int* __restrict__ n_items,
                                                                    unsigned* __restrict__ counter, int parity) {
    typedef typename Fmt<FMT>::T T;
    constexpr int UNITS = SplitTileShape<RB, NT>::UNITS;
    __shared__ v4u lds[2][UNITS];
    __shared__ int s_item;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = lane >> 4;
    const int cl = lane & 15;
    const int total = *n_items;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it >= total) break;
        const BatchItem item = items[it];
        const BatchProblem& P = probs[item.problem];
        const v4u* __restrict__ st_in = static_cast<const v4u*>(P.state[parity]);
        unsigned char* __restrict__ st_out = static_cast<unsigned char*>(P.state[parity ^ 1]);
        const int64_t ct = item.tile * W + wave;
        const int64_t cell = ct * 16 + cl;
        v4f accm[RB], accx[RB];
        split_accumulate<RB, W, FMT, NT, 1, 1>(st_in, P.n, P.steps,
                                               static_cast<const v4u*>(P.spanel) + (int64_t)item.pass * P.steps * UNITS,
                                               ct, 0, lds, accm, accx);
        RelMax rm;
        if (cell < P.n) {
            const float cinvf = P.cell_inv[cell];
            const float csf = 1.0f / cinvf;  // power of two: exact
            const unsigned char* in_bytes = reinterpret_cast<const unsigned char*>(st_in);
            // descriptor fields in registers across the stores; global-space source
            const double oma = P.one_minus_albedo, alb = P.albedo;
            const int m = P.m;
            const int64_t n = P.n;
            const int64_t tiles = (n + 15) / 16;  // the source term is in output-tile order
            typedef const __attribute__((address_space(1))) v4f* gcv4f_p;
            const gcv4f_p bsrc = (gcv4f_p)P.Bs;
            const __attribute__((address_space(1))) float* rinv = (const __attribute__((address_space(1))) float*)P.row_inv;
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                const int rbg = item.pass * RB + rb;
                if (rbg * 16 >= m) continue;  // padding row block: no rows, no state
                const int s = rbg >> 1;
                const int vlane = (2 * (rbg & 1) + (g >> 1)) * 16 + cl;
                // The source array holds whole row blocks (padding rows zero), so the
                // row block's B loads go out together with no per-row branch; a
                // padding row (k >= M) gets row scale 0, hence S = 0, nv = 0, parts
                // 0 and a residual term 0 / 0 = NaN, which fmax drops.
                const v4f bv = bsrc[((int64_t)rbg * tiles + ct) * 64 + lane];  // rows 4g .. 4g + 3 of the cell
                float ri[4];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int k = rbg * 16 + 4 * g + r;
                    ri[r] = k < m ? rinv[k] : 0.0f;
                }
                T oldp[NT][4], newp[NT][4];
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    const int64_t off = ((((ct * P.steps + s) * NT + t) * 64 + vlane) * 16) + 8 * (g & 1);
                    const v2u raw = *reinterpret_cast<const v2u*>(in_bytes + off);
                    T tmp[4];
                    __builtin_memcpy(tmp, &raw, 8);
#pragma unroll
                    for (int r = 0; r < 4; r++) oldp[t][r] = tmp[r];
                }
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    // exact power-of-two scaling in FP32 and an explicit FMA, as
                    // in the resident-panel kernel (bit-identical results)
                    const float sc = ri[r] * cinvf;
                    const double S = split_value(accm[rb][r] * sc, accx[rb][r] * sc);
                    const double nv = __builtin_fma(alb, S, oma * (double)bv[r]);
                    float of = 0.0f;
#pragma unroll
                    for (int t = NT - 1; t >= 0; t--) of += (float)oldp[t][r];
                    const double old = (double)(of * cinvf);
                    rm.add(old - nv, nv, true);
                    T q[NT];
                    split_parts<FMT, NT>((float)nv * csf, q);
#pragma unroll
                    for (int t = 0; t < NT; t++) newp[t][r] = q[t];
                }
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    const int64_t off = ((((ct * P.steps + s) * NT + t) * 64 + vlane) * 16) + 8 * (g & 1);
                    v2u raw;
                    __builtin_memcpy(&raw, newp[t], 8);
                    *reinterpret_cast<v2u*>(st_out + off) = raw;
                }
            }
        }
        const double res = wave_max_split(rm.value());
        if (lane == 0 && res > 0.0)
            atomicMax(P.resid + parity, (unsigned long long)__double_as_longlong(res));
    }
}

// Dequantises a cell-level split state into FP64 [M][N].
template <int FMT, int NT>
__global__ __launch_bounds__(256) void unpack_state_kernel(const v4u* __restrict__ state, int64_t n, int m, int steps,
                                                           const float* __restrict__ cell_inv,
                                                           double* __restrict__ out) {
    typedef typename Fmt<FMT>::T T;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)m * n;
    if (idx >= total) return;
    const int k = (int)(idx / n);
    const int64_t c = idx - (int64_t)k * n;
    const int64_t ct = c / 16;
    const int s = k / 32, vlane = ((k % 32) / 8) * 16 + (int)(c % 16), e = k % 8;
    float v = 0.0f;
#pragma unroll
    for (int t = NT - 1; t >= 0; t--) {
        const v4u u = state[((ct * steps + s) * NT + t) * 64 + vlane];
        T tmp[8];
        __builtin_memcpy(tmp, &u, 16);
        v += (float)tmp[e];
    }
    out[idx] = (double)v * (double)cell_inv[c];
}

template <int W, int FMT, int NT>
static hipError_t launch_gs_rb(int rb, int grid, hipStream_t st, const BatchProblem* probs, const BatchItem* items,
                               const int* n_items, unsigned* counter, int parity) {
#define RSH_RB_CASE(V)                                                                                            \
    case V:                                                                                                       \
        hipLaunchKernelGGL((grouped_iter_split_kernel<V, W, FMT, NT>), dim3(grid), dim3(64 * W), 0, st, probs,    \
                           items, n_items, counter, parity);                                                      \
        break;
    switch (rb) {
        RSH_RB_CASE(1) RSH_RB_CASE(2) RSH_RB_CASE(3) RSH_RB_CASE(4) RSH_RB_CASE(5) RSH_RB_CASE(6)
        RSH_RB_CASE(7) RSH_RB_CASE(8) RSH_RB_CASE(10) RSH_RB_CASE(13) RSH_RB_CASE(16) RSH_RB_CASE(20)
        RSH_RB_CASE(25)
        default:
            return hipErrorInvalidValue;
    }
#undef RSH_RB_CASE
    return hipGetLastError();
}

constexpr int kSplitIterWaves = 4;

hipError_t launch_grouped_iter_split(int form, int rb, int grid, const BatchProblem* d_probs,
                                     const BatchItem* d_items, const int* d_n_items, unsigned* d_counter, int parity,
                                     hipStream_t st) {
    constexpr int W = kSplitIterWaves;
    const int fmt = split_fmt(form), nt = split_terms(form);
    if (fmt == 0 && nt == 2) return launch_gs_rb<W, 0, 2>(rb, grid, st, d_probs, d_items, d_n_items, d_counter, parity);
    if (fmt == 1 && nt == 2) return launch_gs_rb<W, 1, 2>(rb, grid, st, d_probs, d_items, d_n_items, d_counter, parity);
    return launch_gs_rb<W, 0, 3>(rb, grid, st, d_probs, d_items, d_n_items, d_counter, parity);
}

int grouped_iter_split_waves() { return kSplitIterWaves; }

int grouped_iter_split_resident(int form, int rb, int device) {
    constexpr int W = kSplitIterWaves;
    int per_cu = 0;
    hipError_t e = hipErrorInvalidValue;
    const int fmt = split_fmt(form), nt = split_terms(form);
#define RSH_OCC(V, FMT, NT)                                                                                       \
    if (rb == V && fmt == FMT && nt == NT)                                                                         \
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, grouped_iter_split_kernel<V, W, FMT, NT>, 64 * W, 0);
#define RSH_OCC_ALL(V) RSH_OCC(V, 0, 2) RSH_OCC(V, 1, 2) RSH_OCC(V, 0, 3)
    RSH_OCC_ALL(1) RSH_OCC_ALL(2) RSH_OCC_ALL(3) RSH_OCC_ALL(4) RSH_OCC_ALL(5) RSH_OCC_ALL(6) RSH_OCC_ALL(7)
    RSH_OCC_ALL(8) RSH_OCC_ALL(10) RSH_OCC_ALL(13) RSH_OCC_ALL(16) RSH_OCC_ALL(20) RSH_OCC_ALL(25)
#undef RSH_OCC_ALL
#undef RSH_OCC
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    hipDeviceProp_t prop;
    int cus = 256;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        cus = prop.multiProcessorCount;
    return per_cu * cus;
}

hipError_t launch_unpack_state(int form, const void* d_state, int64_t n, int m, const float* d_cell_inv,
                               double* d_out, hipStream_t st) {
    const int64_t total = (int64_t)m * n;
    if (total <= 0) return hipSuccess;
    const int steps = (m + 31) / 32;
    const dim3 grid((unsigned)((total + 255) / 256)), block(256);
    const v4u* s = static_cast<const v4u*>(d_state);
    const int fmt = split_fmt(form), nt = split_terms(form);
    if (fmt == 0 && nt == 2) hipLaunchKernelGGL((unpack_state_kernel<0, 2>), grid, block, 0, st, s, n, m, steps, d_cell_inv, d_out);
    else if (fmt == 1 && nt == 2) hipLaunchKernelGGL((unpack_state_kernel<1, 2>), grid, block, 0, st, s, n, m, steps, d_cell_inv, d_out);
    else hipLaunchKernelGGL((unpack_state_kernel<0, 3>), grid, block, 0, st, s, n, m, steps, d_cell_inv, d_out);
    return hipGetLastError();
}

}  // namespace rsh

// Split-precision MFMA main loop shared by the single-problem kernel and the
// grouped source-iteration kernel. See kernels_split.hip for layouts.
#pragma once

#include "rsh_internal.hpp"

namespace rsh {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
typedef unsigned int v2u __attribute__((ext_vector_type(2)));

constexpr int kTargetExp = 12;  // largest scaled magnitude ~ 2^12

template <int FMT>
struct Fmt;
template <>
struct Fmt<0> {  // BF16
    typedef __bf16 T;
    typedef v8bf V;
    static __device__ inline v4f mfma(V a, V b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};
template <>
struct Fmt<1> {  // FP16
    typedef _Float16 T;
    typedef v8h V;
    static __device__ inline v4f mfma(V a, V b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};

// Splits x into NT parts of format T with round-to-nearest at each stage;
// the residual subtraction is exact in FP32.
template <int FMT, int NT>
__device__ inline void split_parts(float x, typename Fmt<FMT>::T* parts) {
    typedef typename Fmt<FMT>::T T;
    float r = x;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const T p = (T)r;
        parts[t] = p;
        r = r - (float)p;
    }
}

template <int RB, int NT>
struct SplitTileShape {
    static constexpr int UNITS = RB * NT * 64;  // 16-byte units per LDS stage (one K-step)
};

// Accumulates, for the wave's cell tile ct (16 cells) and every row block of one
// pass, accm[rb] += A_hi B_hi and accx[rb] += cross terms over all K-steps.
// panel_pass points at the pass's panel (steps * RB * NT fragments). Ends with a
// workgroup barrier.
template <int RB, int W, int FMT, int NT, int F>
__device__ __forceinline__ void split_accumulate(const v4u* __restrict__ state, int64_t n, int steps,
                                                 const v4u* __restrict__ panel_pass, int64_t ct,
                                                 v4u (&lds)[2][SplitTileShape<RB, NT>::UNITS], v4f (&accm)[RB],
                                                 v4f (&accx)[RB]) {
    typedef typename Fmt<FMT>::T T;
    typedef typename Fmt<FMT>::V V;
    constexpr int UNITS = SplitTileShape<RB, NT>::UNITS;
    constexpr int THREADS = 64 * W;
    constexpr int PER = (UNITS + THREADS - 1) / THREADS;
    constexpr bool TAIL = (UNITS % THREADS) != 0;
    constexpr int VX = NT * F;  // state vectors per (cell tile, step)

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const bool tvalid = ct * 16 < n;

    const v4u* psrc = panel_pass + tid;
    const v4u* xsrc = state + (tvalid ? ct : 0) * (int64_t)steps * VX * 64 + lane;

#pragma unroll
    for (int i = 0; i < RB; i++) {
        accm[i] = v4f{0.f, 0.f, 0.f, 0.f};
        accx[i] = v4f{0.f, 0.f, 0.f, 0.f};
    }
    v4u preg[PER];
    v4u xr[VX];
    V bx[NT];

#define RSH_LOAD_PANEL(t)                                                    \
    _Pragma("unroll") for (int i = 0; i < PER; i++) {                        \
        if (!TAIL || i < PER - 1 || tid + i * THREADS < UNITS)              \
            preg[i] = psrc[(int64_t)(t) * UNITS + i * THREADS];              \
    }
#define RSH_STORE_PANEL(buf)                                                 \
    _Pragma("unroll") for (int i = 0; i < PER; i++) {                        \
        if (!TAIL || i < PER - 1 || tid + i * THREADS < UNITS)              \
            lds[buf][tid + i * THREADS] = preg[i];                           \
    }
#define RSH_LOAD_X(t)                                                        \
    if (tvalid) {                                                            \
        _Pragma("unroll") for (int v = 0; v < VX; v++) xr[v] = xsrc[((int64_t)(t) * VX + v) * 64]; \
    } else {                                                                 \
        _Pragma("unroll") for (int v = 0; v < VX; v++) xr[v] = v4u{0u, 0u, 0u, 0u}; \
    }

    RSH_LOAD_PANEL(0)
    RSH_LOAD_X(0)
    RSH_STORE_PANEL(0)
    __syncthreads();

    for (int t = 0; t < steps; t++) {
        const int buf = t & 1;
        const bool more = t + 1 < steps;
        // B operand of this step from the loaded state vectors
        if constexpr (F == 1) {
#pragma unroll
            for (int p = 0; p < NT; p++) bx[p] = __builtin_bit_cast(V, xr[p]);
        } else {
            T parts[NT][8];
#pragma unroll
            for (int e = 0; e < 8; e++) {
                float x = 0.0f;
#pragma unroll
                for (int f = 0; f < F; f++) {
                    float fv = 0.0f;
#pragma unroll
                    for (int p = NT - 1; p >= 0; p--) fv += (float)__builtin_bit_cast(V, xr[p * F + f])[e];
                    x += fv;
                }
                T q[NT];
                split_parts<FMT, NT>(x, q);
#pragma unroll
                for (int p = 0; p < NT; p++) parts[p][e] = q[p];
            }
#pragma unroll
            for (int p = 0; p < NT; p++) bx[p] = __builtin_bit_cast(V, parts[p]);
        }
        if (more) {
            RSH_LOAD_PANEL(t + 1)
            RSH_LOAD_X(t + 1)
        }
        const v4u* a_base = lds[buf] + lane;
        // A fragments of row block rb + PREF are read while rb's MFMAs run.
        constexpr int PREF = RB < 3 ? RB : 3;
        V apipe[PREF][NT];
#pragma unroll
        for (int q = 0; q < PREF; q++)
#pragma unroll
            for (int p = 0; p < NT; p++) apipe[q][p] = __builtin_bit_cast(V, a_base[(q * NT + p) * 64]);
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            V a[NT];
#pragma unroll
            for (int p = 0; p < NT; p++) a[p] = apipe[rb % PREF][p];
            if (rb + PREF < RB) {
#pragma unroll
                for (int p = 0; p < NT; p++)
                    apipe[rb % PREF][p] = __builtin_bit_cast(V, a_base[((rb + PREF) * NT + p) * 64]);
            }
            accm[rb] = Fmt<FMT>::mfma(a[0], bx[0], accm[rb]);
            if constexpr (NT == 2) {
                accx[rb] = Fmt<FMT>::mfma(a[0], bx[1], accx[rb]);
                accx[rb] = Fmt<FMT>::mfma(a[1], bx[0], accx[rb]);
            } else {
                accx[rb] = Fmt<FMT>::mfma(a[1], bx[1], accx[rb]);
                accx[rb] = Fmt<FMT>::mfma(a[0], bx[2], accx[rb]);
                accx[rb] = Fmt<FMT>::mfma(a[2], bx[0], accx[rb]);
                accx[rb] = Fmt<FMT>::mfma(a[0], bx[1], accx[rb]);
                accx[rb] = Fmt<FMT>::mfma(a[1], bx[0], accx[rb]);
            }
        }
        // Keep the reads PREF row blocks ahead: the scheduler would otherwise
        // sink each ds_read next to its MFMA. Pattern: PREF*NT reads, then per
        // row block its MFMAs followed by the reads of row block rb + PREF.
        __builtin_amdgcn_sched_group_barrier(0x100, PREF * NT, 0);
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            __builtin_amdgcn_sched_group_barrier(0x008, NT == 2 ? 3 : 6, 0);
            if (rb + PREF < RB) __builtin_amdgcn_sched_group_barrier(0x100, NT, 0);
        }
        if (more) RSH_STORE_PANEL(buf ^ 1)
        __syncthreads();
    }
#undef RSH_LOAD_PANEL
#undef RSH_STORE_PANEL
#undef RSH_LOAD_X

}

}  // namespace rsh

// FP64 MFMA main loop shared by the single-problem kernel and the grouped
// (batched, persistent) kernels. See kernels_f64.hip for the algorithm.
#pragma once

#include "rsh_internal.hpp"

namespace rsh {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

// Global-address-space views. A pointer the kernel reads from a descriptor in
// memory (the grouped kernels' problem table) is generic, and loads through it
// are "flat" loads, which also count on the LDS counter: every wait for an LDS
// read would then wait for the global loads in flight as well.
typedef const __attribute__((address_space(1))) double* gcdouble_p;
typedef __attribute__((address_space(1))) double* gdouble_p;
typedef const __attribute__((address_space(1))) v2d* gcv2d_p;
__device__ __forceinline__ gcdouble_p as_global(const double* p) { return (gcdouble_p)p; }
__device__ __forceinline__ gdouble_p as_global(double* p) { return (gdouble_p)p; }

template <int RB, int W, int SC>
struct F64TileShape {
    static constexpr int FRAG = 64;                   // doubles per fragment
    static constexpr int UNITS = SC * RB * FRAG / 2;  // 16-byte units per LDS stage
};

// Accumulates acc[rb] = sum_j panel(row block rb, j) * X[j][cell] over all
// steps of one pass for the wave's 16 cells (cells < n valid; direction rows
// of the input ldx cells apart). panel_pass points at the pass's panel (steps *
// RB fragments). Ends with a workgroup barrier, so the LDS buffers may be
// reused by the caller right after it returns.
template <int RB, int W, int SC, int FACES>
__device__ __forceinline__ void mfma_f64_accumulate(const double* __restrict__ illum, int64_t n, int64_t ldx, int m,
                                                    const double* __restrict__ panel_pass, int steps,
                                                    int64_t cell, v2d (&lds)[2][F64TileShape<RB, W, SC>::UNITS],
                                                    v4d (&acc)[RB]) {
    constexpr int FRAG = 64;
    constexpr int UNITS = F64TileShape<RB, W, SC>::UNITS;
    constexpr int THREADS = 64 * W;
    constexpr int PER = (UNITS + THREADS - 1) / THREADS;
    constexpr bool TAIL = (UNITS % THREADS) != 0;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int jr = lane >> 4;
    const bool cvalid = cell < n;

    // Panel of this pass: steps * RB fragments, stage t = units [t*UNITS, (t+1)*UNITS).
    const gcv2d_p psrc = (gcv2d_p)as_global(panel_pass) + tid;
    // Face pointer of (j = jr, cell); one j-step advances 4 directions.
    const gcdouble_p xp = as_global(illum) + ((int64_t)jr * ldx + (cvalid ? cell : 0)) * FACES;
    const int64_t xstep = 4 * ldx * FACES;

#pragma unroll
    for (int i = 0; i < RB; i++) acc[i] = v4d{0.0, 0.0, 0.0, 0.0};

    v2d preg[PER];
    double xr[SC][FACES];
    double xb[SC];

#define RSH_LOAD_PANEL(t)                                                     \
    _Pragma("unroll") for (int i = 0; i < PER; i++) {                         \
        if (!TAIL || i < PER - 1 || tid + i * THREADS < UNITS)               \
            preg[i] = psrc[(int64_t)(t) * UNITS + i * THREADS];               \
    }
#define RSH_STORE_PANEL(buf)                                                  \
    _Pragma("unroll") for (int i = 0; i < PER; i++) {                         \
        if (!TAIL || i < PER - 1 || tid + i * THREADS < UNITS)               \
            lds[buf][tid + i * THREADS] = preg[i];                            \
    }
#define RSH_LOAD_X(t)                                                         \
    _Pragma("unroll") for (int s = 0; s < SC; s++) {                          \
        const int step = (t) * SC + s;                                        \
        if (cvalid && 4 * step + jr < m) {                                    \
            const gcdouble_p p = xp + (int64_t)step * xstep;                  \
            if constexpr (FACES == 4) {                                       \
                const v2d a = ((gcv2d_p)p)[0];                                \
                const v2d b = ((gcv2d_p)p)[1];                                \
                xr[s][0] = a.x; xr[s][1 % FACES] = a.y;                       \
                xr[s][2 % FACES] = b.x; xr[s][3 % FACES] = b.y;               \
            } else {                                                          \
                xr[s][0] = p[0];                                              \
            }                                                                 \
        } else {                                                              \
            _Pragma("unroll") for (int f = 0; f < FACES; f++) xr[s][f] = 0.0; \
        }                                                                     \
    }
#define RSH_FACE_SUM()                                                        \
    _Pragma("unroll") for (int s = 0; s < SC; s++) {                          \
        double v = xr[s][0];                                                  \
        _Pragma("unroll") for (int f = 1; f < FACES; f++) v += xr[s][f];      \
        xb[s] = v;                                                            \
    }

    const int nchunks = steps / SC;
    RSH_LOAD_PANEL(0)
    RSH_LOAD_X(0)
    RSH_STORE_PANEL(0)
    RSH_FACE_SUM()
    __syncthreads();

    for (int t = 0; t < nchunks; t++) {
        const int buf = t & 1;
        const bool more = t + 1 < nchunks;
        if (more) {
            RSH_LOAD_PANEL(t + 1)
            RSH_LOAD_X(t + 1)
        }
        const double* a_base = reinterpret_cast<const double*>(lds[buf]) + lane;
        // A fragments are read PREF fragments ahead of their MFMA so that the
        // LDS latency hides behind the MFMAs already in the pipe.
        // Depth: 4 when the accumulators live in AGPRs (4-wave tiles), 2 when
        // they share the 256 VGPRs of an 8-wave tile.
        constexpr int NQ = SC * RB;
        constexpr int PREF_MAX = (W == 4) ? 4 : 2;
        constexpr int PREF = NQ < PREF_MAX ? NQ : PREF_MAX;
        double apipe[PREF];
#pragma unroll
        for (int q = 0; q < PREF; q++) apipe[q] = a_base[q * FRAG];
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const double a = apipe[q % PREF];
            if (q + PREF < NQ) apipe[q % PREF] = a_base[(q + PREF) * FRAG];
            acc[q % RB] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, xb[q / RB], acc[q % RB], 0, 0, 0);
        }
        if (more) {
            RSH_STORE_PANEL(buf ^ 1)
            RSH_FACE_SUM()
        }
        __syncthreads();
    }
#undef RSH_LOAD_PANEL
#undef RSH_STORE_PANEL
#undef RSH_LOAD_X
#undef RSH_FACE_SUM

}

}  // namespace rsh

// Split-precision MFMA main loop shared by the single-problem kernel and the
// grouped source-iteration kernel. See kernels_split.hip for layouts.
#pragma once

#include <type_traits>

#include "rsh_internal.hpp"

namespace rsh {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
typedef unsigned int v2u __attribute__((ext_vector_type(2)));

constexpr int kTargetExp = 12;  // largest scaled magnitude ~ 2^12

// Main-product accumulation. The MFMA adds its products to an FP32 accumulator
// much larger than they are by truncating the products' low bits (measured on
// gfx950: the error is one-signed, up to ~1e-6 relative after 12 K-steps when
// the first K-step holds most of a row's weight). So each K-step's hi*hi
// product is formed in a fresh (zero) accumulator and added to the running
// main sum with an ordinary round-to-nearest FP32 add (main_add); the cross
// terms, 2^-8 of the main sum, keep accumulating in the MFMA. Every split
// kernel does the same per output (main sum over K-steps in ascending order,
// cross products in the same order), so they stay bit-identical to each
// other. The epilogue combines main and cross sums in FP64 (split_value).
__device__ inline v4f main_add(v4f acc, v4f step) {
    v4f r;
#pragma unroll
    for (int e = 0; e < 4; e++) r[e] = acc[e] + step[e];
    return r;
}

__device__ inline double split_value(float m, float x) { return (double)m + (double)x; }

// Makes x (an MFMA operand, unchanged) depend on the values deps[0 .. N-1]: an
// empty asm per value that takes it and claims to rewrite x. The fresh main
// MFMAs depend on no accumulator, so the scheduler would otherwise issue them
// long before their adds and hold every product in registers meanwhile (in
// fully unrolled K loops it spills); making the A operand of row block i's
// main products depend on the sums row block i - 1 just added chains the row
// blocks, so that one row block's products are in flight at a time.
template <typename V, int N>
__device__ __forceinline__ void order_after(V& x, const v4f (&deps)[N]) {
    v4u u = __builtin_bit_cast(v4u, x);
#pragma unroll
    for (int q = 0; q < N; q++) asm("" : "+v"(u) : "v"(deps[q]));
    x = __builtin_bit_cast(V, u);
}
template <typename V>
__device__ __forceinline__ void order_after(V& x, const v4f& dep) {
    v4u u = __builtin_bit_cast(v4u, x);
    asm("" : "+v"(u) : "v"(dep));
    x = __builtin_bit_cast(V, u);
}

// Cache policy of the result stores: non-temporal (each result is written once
// and not read again by the kernel that writes it).
constexpr int kNontemporal = 2;

// Running maximum of |d| / v (the source iteration's relative change) over the
// results one lane owns, kept as the pair (num, den) of the largest ratio so
// far: a term wins when |d| den > num v (two products instead of an FP64
// division per result); value() divides once. As with fmax over the quotients,
// a NaN term, a padding row's 0 / 0 and a negative v never win, and v = 0 <
// |d| gives +inf.
struct RelMax {
    double num = 0.0, den = 1.0;
    __device__ __forceinline__ void add(double d, double v, bool ok) {
        const double a = fabs(d);
        const bool win = ok && v >= 0.0 && a * den > num * v;
        num = win ? a : num;
        den = win ? v : den;
    }
    __device__ __forceinline__ double value() const { return num / den; }
};

// Value of the other lane of the pair (2i, 2i + 1): DPP quad permutation
// [1, 0, 3, 2] (VALU, no LDS).
__device__ __forceinline__ unsigned swap_pair_lanes(unsigned v) {
    return (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
}
__device__ __forceinline__ float swap_pair_lanes(float v) {
    return __builtin_bit_cast(float, swap_pair_lanes(__builtin_bit_cast(unsigned, v)));
}
__device__ __forceinline__ double swap_pair_lanes(double v) {
    const v2u u = __builtin_bit_cast(v2u, v);
    return __builtin_bit_cast(double, v2u{swap_pair_lanes(u[0]), swap_pair_lanes(u[1])});
}

// Stores two adjacent output cells (first at byte offset voff + soff of the
// buffer, second after it) with one 16-byte (FP64) or 8-byte (FP32) store when
// both are inside the output (both), the first alone when only it is (one).
__device__ __forceinline__ void store_cell_pair(double first, double second, bool both, bool one,
                                                __amdgpu_buffer_rsrc_t rsrc, int voff, int soff) {
    if (both) {
        const v2u a = __builtin_bit_cast(v2u, first), b = __builtin_bit_cast(v2u, second);
        __builtin_amdgcn_raw_buffer_store_b128(v4u{a[0], a[1], b[0], b[1]}, rsrc, voff, soff, kNontemporal);
    } else if (one) {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, first), rsrc, voff, soff, kNontemporal);
    }
}
__device__ __forceinline__ void store_cell_pair(float first, float second, bool both, bool one,
                                                __amdgpu_buffer_rsrc_t rsrc, int voff, int soff) {
    if (both)
        __builtin_amdgcn_raw_buffer_store_b64(
            v2u{__builtin_bit_cast(unsigned, first), __builtin_bit_cast(unsigned, second)}, rsrc, voff, soff,
            kNontemporal);
    else if (one)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, first), rsrc, voff, soff, kNontemporal);
}

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

template <int RB, int NT, int RH = 1>
struct SplitTileShape {
    static constexpr int UNITS = RB * NT * 64;       // 16-byte units per LDS stage (one K-step)
    static constexpr int RBW = (RB + RH - 1) / RH;   // row blocks per wave
};

// State prefetch depth (K-steps ahead) of a tile: two where the register ring
// fits without spills, one otherwise.
template <int RB, int W, int NT, int F, int RH>
struct SplitPrefetch {
    static constexpr int RBW = (RB + RH - 1) / RH;
    static constexpr int XD = (RH == 2 && !(NT == 3 && F == 4)) ? 2 : 1;
};

// Accumulates, for the wave's cell tile ct (16 cells) and its share of the row
// blocks of one pass, accm[i] = sum over K-steps of A_hi B_hi (each step's
// product formed alone, then added, main_add) and accx[i] += cross terms. With RH > 1 the RB row blocks of a tile are split over RH waves:
// wave part h owns row blocks h*RBW .. h*RBW + RBW - 1 (a block past RB re-reads
// block RB - 1; its result is discarded by the caller). panel_pass points at the
// pass's panel (steps * RB * NT fragments). Ends with a workgroup barrier.
template <int RB, int W, int FMT, int NT, int F, int RH = 1>
__device__ __forceinline__ void split_accumulate(const v4u* __restrict__ state, int64_t n, int steps,
                                                 const v4u* __restrict__ panel_pass, int64_t ct, int h,
                                                 v4u (&lds)[2][SplitTileShape<RB, NT>::UNITS],
                                                 v4f (&accm)[SplitTileShape<RB, NT, RH>::RBW],
                                                 v4f (&accx)[SplitTileShape<RB, NT, RH>::RBW]) {
    typedef typename Fmt<FMT>::T T;
    typedef typename Fmt<FMT>::V V;
    constexpr int RBW = SplitTileShape<RB, NT, RH>::RBW;
    constexpr int UNITS = SplitTileShape<RB, NT>::UNITS;
    constexpr int THREADS = 64 * W;
    constexpr int PER = (UNITS + THREADS - 1) / THREADS;
    constexpr bool TAIL = (UNITS % THREADS) != 0;
    constexpr int VX = NT * F;  // state vectors per (cell tile, step)
    constexpr int XD = SplitPrefetch<RB, W, NT, F, RH>::XD;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const bool tvalid = ct * 16 < n;
    const int rb_first = h * RBW;

    const v4u* psrc = panel_pass + tid;
    const v4u* xsrc = state + (tvalid ? ct : 0) * (int64_t)steps * VX * 64 + lane;

#pragma unroll
    for (int i = 0; i < RBW; i++) {
        accm[i] = v4f{0.f, 0.f, 0.f, 0.f};
        accx[i] = v4f{0.f, 0.f, 0.f, 0.f};
    }
    v4u preg[PER];
    v4u xr[XD][VX];  // ring of state vectors: slot (t % XD) holds step t

    auto load_panel = [&](int t) {
#pragma unroll
        for (int i = 0; i < PER; i++)
            if (!TAIL || i < PER - 1 || tid + i * THREADS < UNITS) preg[i] = psrc[(int64_t)t * UNITS + i * THREADS];
    };
    auto store_panel = [&](int buf) {
#pragma unroll
        for (int i = 0; i < PER; i++)
            if (!TAIL || i < PER - 1 || tid + i * THREADS < UNITS) lds[buf][tid + i * THREADS] = preg[i];
    };
    auto load_x = [&](auto slot_tag, int t) {
        constexpr int sl = decltype(slot_tag)::value;
        if (tvalid) {
#pragma unroll
            for (int v = 0; v < VX; v++) xr[sl][v] = xsrc[((int64_t)t * VX + v) * 64];
        } else {
#pragma unroll
            for (int v = 0; v < VX; v++) xr[sl][v] = v4u{0u, 0u, 0u, 0u};
        }
    };

    // One K-step t. SLOT = t % XD is a compile-time constant (the loop below is
    // unrolled by XD) so the register ring is indexed statically. The panel of
    // step t + 1 is loaded before the state of step t + XD, so that waiting for
    // the panel never drains the state prefetch.
    auto kstep = [&](auto slot_tag, const int t) {
        constexpr int sl = decltype(slot_tag)::value;
        const int buf = t & 1;
        const bool more = t + 1 < steps;
        V bx[NT];
        if constexpr (F == 1) {
#pragma unroll
            for (int p = 0; p < NT; p++) bx[p] = __builtin_bit_cast(V, xr[sl][p]);
        } else {
            T parts[NT][8];
#pragma unroll
            for (int e = 0; e < 8; e++) {
                float x = 0.0f;
#pragma unroll
                for (int f = 0; f < F; f++) {
                    float fv = 0.0f;
#pragma unroll
                    for (int p = NT - 1; p >= 0; p--) fv += (float)__builtin_bit_cast(V, xr[sl][p * F + f])[e];
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
        if (more) load_panel(t + 1);
        if (t + XD < steps) load_x(slot_tag, t + XD);
        const v4u* a_base = lds[buf] + lane;
        // A fragments of row block i + PREF are read while i's MFMAs run (two
        // ahead where two waves share a SIMD and the state is averaged here:
        // three would not fit their 256 registers).
        constexpr int PMAX = (RH == 2 && F == 4) ? 2 : 3;
        constexpr int PREF = RBW < PMAX ? RBW : PMAX;
        V apipe[PREF][NT];
#define RSH_AFRAG(i, p) __builtin_bit_cast(V, a_base[((rb_first + (i) < RB ? rb_first + (i) : RB - 1) * NT + (p)) * 64])
#pragma unroll
        for (int q = 0; q < PREF; q++)
#pragma unroll
            for (int p = 0; p < NT; p++) apipe[q][p] = RSH_AFRAG(q, p);
#pragma unroll
        for (int i = 0; i < RBW; i++) {
            V a[NT];
#pragma unroll
            for (int p = 0; p < NT; p++) a[p] = apipe[i % PREF][p];
            if (i + PREF < RBW) {
#pragma unroll
                for (int p = 0; p < NT; p++) apipe[i % PREF][p] = RSH_AFRAG(i + PREF, p);
            }
            order_after(a[0], accm[i > 0 ? i - 1 : RBW - 1]);
            const v4f mstep = Fmt<FMT>::mfma(a[0], bx[0], v4f{0.f, 0.f, 0.f, 0.f});
            if constexpr (NT == 2) {
                accx[i] = Fmt<FMT>::mfma(a[0], bx[1], accx[i]);
                accx[i] = Fmt<FMT>::mfma(a[1], bx[0], accx[i]);
            } else {
                // part-major order (hi*lo, hi*mid, mid*mid, mid*hi, lo*hi), shared
                // with the resident-panel kernel
                accx[i] = Fmt<FMT>::mfma(a[0], bx[2], accx[i]);
                accx[i] = Fmt<FMT>::mfma(a[0], bx[1], accx[i]);
                accx[i] = Fmt<FMT>::mfma(a[1], bx[1], accx[i]);
                accx[i] = Fmt<FMT>::mfma(a[1], bx[0], accx[i]);
                accx[i] = Fmt<FMT>::mfma(a[2], bx[0], accx[i]);
            }
            accm[i] = main_add(accm[i], mstep);
        }
#undef RSH_AFRAG
        // Keep the reads PREF row blocks ahead: the scheduler would otherwise
        // sink each ds_read next to its MFMA. Pattern: PREF*NT reads, then per
        // row block its MFMAs followed by the reads of row block i + PREF.
        __builtin_amdgcn_sched_group_barrier(0x100, PREF * NT, 0);
#pragma unroll
        for (int i = 0; i < RBW; i++) {
            __builtin_amdgcn_sched_group_barrier(0x008, NT == 2 ? 3 : 6, 0);
            if (i + PREF < RBW) __builtin_amdgcn_sched_group_barrier(0x100, NT, 0);
        }
        if (more) store_panel(buf ^ 1);
        __syncthreads();
    };

    load_panel(0);
    load_x(std::integral_constant<int, 0>{}, 0);
    if constexpr (XD == 2) {
        if (steps > 1) load_x(std::integral_constant<int, 1>{}, 1);
    }
    store_panel(0);
    __syncthreads();
    if constexpr (XD == 2) {
        for (int t = 0; t < steps; t += 2) {
            kstep(std::integral_constant<int, 0>{}, t);
            if (t + 1 < steps) kstep(std::integral_constant<int, 1>{}, t + 1);
        }
    } else {
        for (int t = 0; t < steps; t++) kstep(std::integral_constant<int, 0>{}, t);
    }
}

}  // namespace rsh

// Split-precision contraction with the phase panel resident in LDS
// ("resident-panel" kernel), for cell-level state (one value per cell and
// direction, i.e. face-averaged intensities).
//
// The streaming kernel (kernels_split.hip) re-streams the whole panel through
// LDS for every 64 cells and synchronises its workgroup at every K-step; at
// 100 KB of LDS per workgroup only one workgroup fits a CU, so those barriers
// are exposed. Here each workgroup owns a fixed group of RBW row blocks (16 RBW
// output directions) and copies that slice of the panel for ALL K-steps into
// LDS once (RBW * steps * NT KB, <= 160 KB), then streams cells through it with
// no further barrier: persistent workgroups, one per CU.
//
// Work split. The G row groups of a problem are served by G workgroups that
// sit on the same XCD (workgroup b runs on XCD b % 8) and walk the same
// sequence of 256-cell chunks, so a chunk's state is read from HBM once and
// from that XCD's L2 by the other G - 1. Each XCD owns a contiguous 1/8 of the
// chunks; within it, the floor(CUs per XCD / G) sets of G workgroups take
// chunks round-robin.
//
// Per wave: two 16-cell tiles x RBW row blocks, so every A fragment read from
// LDS feeds two MFMAs per product; MFMAs are issued product-major so that no
// MFMA depends on the one before it; state vectors are prefetched several
// K-steps ahead, across chunk boundaries too, with ordinary (L2-allocating)
// loads, since the other row groups of the XCD read the same vectors. Products
// and their order per (row block, tile) are those of the streaming kernel (main hi*hi; cross hi*lo + lo*hi, and for three
// parts hi*lo, hi*mid, mid*mid, mid*hi, lo*hi), so the kernels give
// bit-identical results on the same state and panel.
#include "mfma_split_core.hpp"

namespace rsh {

namespace {

constexpr int kResWaves = 8;   // waves per workgroup (2 per SIMD)
constexpr int kResTiles = 2;   // 16-cell tiles per wave
constexpr int kResChunkTiles = kResWaves * kResTiles;  // 256 cells per chunk
constexpr int kXcds = 8;       // MI355X: 8 XCDs, workgroups dispatched round-robin
constexpr int kResLdsBytes = 160 * 1024;
// Buffer descriptor word 3 for raw (untyped) gfx9 buffer access: 32-bit data
// format, so that the hardware does not treat the descriptor as invalid.
constexpr int kBufWord3 = 0x00020000;
// Cache policy of the result stores: non-temporal (each result is written once
// and not read again by this kernel).
constexpr int kNontemporal = 2;

// State prefetch depth in K-steps (registers: XD x tiles x parts x 4 VGPRs).
// Measured at 1.57 M cells x 400 directions: two parts 2 / 4 / 6 steps ->
// 2.21 / 2.15 / 1.99 ms with FP64 output; three parts 2 / 3 / 4 -> 3.29 /
// 3.25 / 3.35 ms.
template <int NT>
constexpr int res_xd() { return NT == 2 ? 6 : 3; }

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// STEPS > 0: the K-step count is a compile-time constant and the K loop is
// fully unrolled, so the state prefetch crosses no loop back-edge (where the
// compiler would otherwise wait for every outstanding load). STEPS = 0: any M.
template <int RBW, int FMT, int NT, int STEPS>
__global__ __launch_bounds__(64 * kResWaves) void split_resident_kernel(
    const v4u* __restrict__ state, int64_t n, int steps_rt, const float* __restrict__ cell_inv,
    const v4u* __restrict__ panel, const float* __restrict__ row_inv, void* __restrict__ out, int64_t ld, int rows,
    int out_f32, int groups, int sets, int64_t chunks) {
    typedef typename Fmt<FMT>::V V;
    constexpr int XD = res_xd<NT>();
    const int steps = STEPS > 0 ? STEPS : steps_rt;
    extern __shared__ v4u lds[];  // [steps][RBW][NT][64], then the row group's 16 RBW row scales
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = lane >> 4;
    const int xcd = blockIdx.x % kXcds;
    const int local = blockIdx.x / kXcds;
    const int set = local / groups;
    const int rg = local % groups;
    if (set >= sets) return;  // whole workgroup leaves together: no barrier below is skipped by part of it

    // 1. the row group's panel slice, all K-steps, into LDS (contiguous in the panel)
    const int64_t units = (int64_t)steps * RBW * NT * 64;
    const v4u* __restrict__ src = panel + (int64_t)rg * units;
    for (int64_t u = threadIdx.x; u < units; u += 64 * kResWaves) lds[u] = src[u];
    const int row0 = rg * RBW * 16;
    // Row scales once per workgroup into LDS: scale loads from global memory in
    // every epilogue measurably broke the row groups' L2 sharing of the state
    // (2.4x the state bytes fetched from HBM).
    float* rscale = reinterpret_cast<float*>(lds + units);
    for (int r = threadIdx.x; r < RBW * 16; r += 64 * kResWaves)
        rscale[r] = row0 + r < rows ? row_inv[row0 + r] : 0.0f;
    __syncthreads();

    // 2. chunks of this XCD, round-robin over its sets
    const int64_t c_begin = chunks * xcd / kXcds, c_end = chunks * (xcd + 1) / kXcds;
    const int64_t tiles = (n + 15) / 16;
    const int64_t tile_units = (int64_t)steps * NT * 64;  // 16-byte vectors per 16-cell tile
    const int tile_bytes = (int)(tile_units * 16);

    // A chunk's state is read through one buffer descriptor: the uniform base and
    // the K-step offsets live in scalar registers, so a tile needs one VGPR
    // offset instead of a 64-bit address per K-step. Tiles past the end of the
    // state re-read the chunk's last tile (in range) and are not written.
    __amdgpu_buffer_rsrc_t xs;
    int xoff[kResTiles];
    auto bind_chunk = [&](int64_t chunk) {
        const int64_t ct_chunk = chunk * kResChunkTiles;
        const int ntile = (int)(tiles - ct_chunk < kResChunkTiles ? tiles - ct_chunk : kResChunkTiles);
        xs = __builtin_amdgcn_make_buffer_rsrc(const_cast<v4u*>(state + ct_chunk * tile_units), (short)0,
                                               ntile * tile_bytes, kBufWord3);
#pragma unroll
        for (int q = 0; q < kResTiles; q++) {
            const int tile = wave * kResTiles + q;
            xoff[q] = (tile < ntile ? tile : ntile - 1) * tile_bytes + lane * 16;
        }
    };
    v4u xr[XD][kResTiles][NT];
    auto load_x = [&](auto slot_tag, int t) {
        constexpr int sl = decltype(slot_tag)::value;
#pragma unroll
        for (int q = 0; q < kResTiles; q++)
#pragma unroll
            for (int p = 0; p < NT; p++)
                xr[sl][q][p] = __builtin_bit_cast(
                    v4u, __builtin_amdgcn_raw_buffer_load_b128(xs, xoff[q] + p * 1024, t * NT * 1024, 0));
    };
    // First XD K-steps of a chunk. For every chunk after the first they are
    // issued before the previous chunk's result stores: loads and stores retire
    // through one in-order counter, so a load issued after the stores could
    // only be waited for together with all of them.
    auto prefetch_head = [&]() {
        static_for<0, XD>([&](auto u) {
            if (u.value < steps) load_x(u, u.value);
        });
    };

    int64_t chunk = c_begin + set;
    if (chunk < c_end) {
        bind_chunk(chunk);
        prefetch_head();
    }
    for (; chunk < c_end; chunk += sets) {
        const int64_t ct0 = chunk * kResChunkTiles + wave * kResTiles;
        // Cell scales now, ahead of the next chunk's prefetch: a load issued
        // after those could only be waited for together with all of them.
        float cinv[kResTiles];
#pragma unroll
        for (int q = 0; q < kResTiles; q++) {
            const int64_t cell = (ct0 + q) * 16 + (lane & 15);
            cinv[q] = cell_inv[cell < n ? cell : n - 1];
        }
        v4f accm[RBW][kResTiles], accx[RBW][kResTiles];
#pragma unroll
        for (int i = 0; i < RBW; i++)
#pragma unroll
            for (int q = 0; q < kResTiles; q++) {
                accm[i][q] = v4f{0.f, 0.f, 0.f, 0.f};
                accx[i][q] = v4f{0.f, 0.f, 0.f, 0.f};
            }
        auto kstep = [&](auto slot_tag, const int t) {
            constexpr int sl = decltype(slot_tag)::value;
            // Scheduling fence per K-step: in the unrolled loop the scheduler
            // would otherwise hoist every K-step's loads and run out of registers.
            __builtin_amdgcn_sched_barrier(0);
            V bx[kResTiles][NT];
#pragma unroll
            for (int q = 0; q < kResTiles; q++)
#pragma unroll
                for (int p = 0; p < NT; p++) bx[q][p] = __builtin_bit_cast(V, xr[sl][q][p]);
            if (t + XD < steps) load_x(slot_tag, t + XD);
            const v4u* a_base = lds + (int64_t)t * RBW * NT * 64 + lane;
            // A fragments of one part at a time (part-major), so that only RBW
            // fragments are live. Product-major order: each round touches every
            // (row block, tile) accumulator once, so an MFMA never waits for the
            // previous result of its own accumulator.
            V a[RBW];
            auto fetch = [&](int p) {
#pragma unroll
                for (int i = 0; i < RBW; i++) a[i] = __builtin_bit_cast(V, a_base[(i * NT + p) * 64]);
            };
            auto round = [&](v4f (&acc)[RBW][kResTiles], int pb) {
#pragma unroll
                for (int i = 0; i < RBW; i++)
#pragma unroll
                    for (int q = 0; q < kResTiles; q++) acc[i][q] = Fmt<FMT>::mfma(a[i], bx[q][pb], acc[i][q]);
            };
            fetch(0);
            round(accm, 0);
            if constexpr (NT == 2) {
                round(accx, 1);
                fetch(1);
                round(accx, 0);
            } else {
                round(accx, 2);
                round(accx, 1);
                fetch(1);
                round(accx, 1);
                round(accx, 0);
                fetch(2);
                round(accx, 0);
            }
        };
        if constexpr (STEPS > 0) {
            static_for<0, STEPS>([&](auto t) { kstep(std::integral_constant<int, t.value % XD>{}, t.value); });
        } else {
            for (int t = 0; t < steps; t += XD)
                static_for<0, XD>([&](auto u) {
                    if (t + u.value < steps) kstep(u, t + u.value);
                });
        }
        __builtin_amdgcn_sched_barrier(0);
        if (chunk + sets < c_end) {
            bind_chunk(chunk + sets);
            prefetch_head();
        }
        __builtin_amdgcn_sched_barrier(0);

        // Epilogue: FP64 combination of main and cross terms, exact scale removal.
        // Row scales are re-read from LDS per chunk: an opaque zero offset keeps
        // the compiler from hoisting all 4 RBW of them out of the chunk loop,
        // where they would be spilled. Stores go through one buffer descriptor
        // per row block (uniform base, row offset r * ld in a scalar register,
        // one VGPR offset per tile), again to keep 64-bit addresses out of the
        // vector registers.
        int zoff;
        asm volatile("v_mov_b32 %0, 0" : "=v"(zoff));
        const float* __restrict__ rs = rscale + 4 * g + zoff;
        v4f rsv[RBW];  // this lane's 4 row scales of each row block, read together
#pragma unroll
        for (int i = 0; i < RBW; i++) rsv[i] = *reinterpret_cast<const v4f*>(rs + i * 16);
        double cd[kResTiles];
#pragma unroll
        for (int q = 0; q < kResTiles; q++) cd[q] = (double)cinv[q];
        auto store_all = [&](auto elem_tag, auto full_tag) {
            typedef decltype(elem_tag) OT;
            constexpr bool kFull = decltype(full_tag)::value;
            const int esz = (int)sizeof(OT);
            int voff[kResTiles];
            bool ok[kResTiles];
#pragma unroll
            for (int q = 0; q < kResTiles; q++) {
                const int64_t cell = (ct0 + q) * 16 + (lane & 15);
                ok[q] = cell < n;
                voff[q] = (int)((4 * g * ld + cell) * esz);
            }
#pragma unroll
            for (int i = 0; i < RBW; i++) {
                const int rb_row = row0 + i * 16;
                if (!kFull && rb_row >= rows) break;
                const __amdgpu_buffer_rsrc_t os = __builtin_amdgcn_make_buffer_rsrc(
                    static_cast<OT*>(out) + (int64_t)rb_row * ld, (short)0, (int)(16 * ld * esz), kBufWord3);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const double rsc = (double)rsv[i][r];
#pragma unroll
                    for (int q = 0; q < kResTiles; q++) {
                        if (!ok[q] || (!kFull && rb_row + 4 * g + r >= rows)) continue;
                        const double v = ((double)accm[i][q][r] + (double)accx[i][q][r]) * rsc * cd[q];
                        if constexpr (sizeof(OT) == 4)
                            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, (float)v), os,
                                                                  voff[q], (int)(r * ld * esz), kNontemporal);
                        else
                            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, v), os, voff[q],
                                                                  (int)(r * ld * esz), kNontemporal);
                    }
                }
            }
        };
        const bool full = row0 + RBW * 16 <= rows;
        if (out_f32) {
            if (full) store_all(0.0f, std::true_type{});
            else store_all(0.0f, std::false_type{});
        } else {
            if (full) store_all(0.0, std::true_type{});
            else store_all(0.0, std::false_type{});
        }
    }
}

template <int RBW, int FMT, int NT, int STEPS>
hipError_t launch_res(int blocks, size_t lds_bytes, hipStream_t st, const v4u* state, int64_t n, int steps,
                      const float* cell_inv, const v4u* panel, const float* row_inv, void* out, int64_t ld, int rows,
                      int out_f32, int groups, int sets, int64_t chunks) {
    auto kern = split_resident_kernel<RBW, FMT, NT, STEPS>;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(64 * kResWaves), lds_bytes, st, state, n, steps, cell_inv, panel,
                       row_inv, out, ld, rows, out_f32, groups, sets, chunks);
    return hipGetLastError();
}

// Largest row-block count per workgroup that the compiler keeps in 256
// registers without spills (two parts: 5; three parts: 4).
constexpr int res_max_rbw(int nt) { return nt == 2 ? 5 : 4; }
constexpr int kResMaxRbw = 5;

}  // namespace

// Row-group plan of the resident-panel kernel for output rows [0, rows) of an
// M-direction set: rb_per_pass = RBW row blocks per workgroup (largest slice
// that fits LDS, then balanced), passes = G row groups. False if the kernel
// does not apply (G row groups do not fit one XCD's CUs).
bool plan_split_resident(int form, int m, int rows, int cus, SplitPlan* pl, int* sets) {
    const int nt = split_terms(form);
    const int steps = (m + 31) / 32;
    const int rb_need = (rows + 15) / 16;
    const int per_rb = steps * nt * 1024 + 16 * 4;  // panel fragments + row scales
    int rbw_max = kResLdsBytes / per_rb;
    if (rbw_max > res_max_rbw(nt)) rbw_max = res_max_rbw(nt);
    if (rbw_max < 1 || rb_need < 1) return false;
    const int groups = (rb_need + rbw_max - 1) / rbw_max;
    const int rbw = (rb_need + groups - 1) / groups;
    const int per_xcd = cus / kXcds;
    if (groups > per_xcd) return false;
    pl->rb_per_pass = rbw;
    pl->passes = groups;
    pl->rb_total = rbw * groups;
    pl->steps = steps;
    *sets = per_xcd / groups;
    return true;
}

hipError_t launch_split_resident(int form, const void* d_state, int64_t n, const float* d_cell_inv,
                                 const void* d_panel, const float* d_row_inv, const SplitPlan& pl, int sets, int rows,
                                 void* d_out, int64_t ld, int out_f32, hipStream_t st) {
    if (n <= 0 || rows <= 0) return hipSuccess;
    const int nt = split_terms(form), fmt = split_fmt(form);
    const size_t lds_bytes = (size_t)pl.steps * pl.rb_per_pass * nt * 1024 + (size_t)pl.rb_per_pass * 16 * 4;
    if (lds_bytes > (size_t)kResLdsBytes || pl.rb_per_pass > res_max_rbw(nt) || sets < 1 || ld >= kSplitResidentMaxLd)
        return hipErrorInvalidValue;
    const int64_t tiles = (n + 15) / 16;
    const int64_t chunks = (tiles + kResChunkTiles - 1) / kResChunkTiles;
    const int blocks = kXcds * sets * pl.passes;
    const v4u* state = static_cast<const v4u*>(d_state);
    const v4u* panel = static_cast<const v4u*>(d_panel);
#define RSH_RES_S(RBW, FMT, NT, S)                                                                              \
    return launch_res<RBW, FMT, NT, S>(blocks, lds_bytes, st, state, n, pl.steps, d_cell_inv, panel, d_row_inv, \
                                       d_out, ld, rows, out_f32, pl.passes, sets, chunks)
    // Unrolled instances for the plans of M = 200, 400, 800 (7, 13, 25 K-steps).
#define RSH_RES(RBW, FMT, NT)                                                                          \
    {                                                                                                  \
        if constexpr (NT == 2 && RBW == 5) { if (pl.steps == 7) RSH_RES_S(RBW, FMT, NT, 7); }          \
        if constexpr (NT == 2 && RBW == 5) { if (pl.steps == 13) RSH_RES_S(RBW, FMT, NT, 13); }        \
        if constexpr (NT == 2 && RBW == 3) { if (pl.steps == 25) RSH_RES_S(RBW, FMT, NT, 25); }        \
        if constexpr (NT == 3 && RBW == 4) { if (pl.steps == 7) RSH_RES_S(RBW, FMT, NT, 7); }          \
        if constexpr (NT == 3 && RBW == 4) { if (pl.steps == 13) RSH_RES_S(RBW, FMT, NT, 13); }        \
        if constexpr (NT == 3 && RBW == 2) { if (pl.steps == 25) RSH_RES_S(RBW, FMT, NT, 25); }        \
        RSH_RES_S(RBW, FMT, NT, 0);                                                                    \
    }
#define RSH_RES_FORMS(RBW)                                  \
    case RBW:                                               \
        if (fmt == 0 && nt == 2) RSH_RES(RBW, 0, 2);        \
        if (fmt == 1 && nt == 2) RSH_RES(RBW, 1, 2);        \
        if constexpr (RBW <= res_max_rbw(3)) RSH_RES(RBW, 0, 3); \
        return hipErrorInvalidValue;
    switch (pl.rb_per_pass) {
        RSH_RES_FORMS(1) RSH_RES_FORMS(2) RSH_RES_FORMS(3) RSH_RES_FORMS(4) RSH_RES_FORMS(5)
        default:
            return hipErrorInvalidValue;
    }
#undef RSH_RES_FORMS
#undef RSH_RES
#undef RSH_RES_S
}

}  // namespace rsh

// Split-precision contraction with the phase panel resident in LDS
// ("resident-panel" kernel), for cell-level state (one value per cell and
// direction, i.e. face-averaged intensities).
//
// The streaming kernel (kernels_split.hip) re-streams the whole panel through
// LDS for every 64 cells and synchronises its workgroup at every K-step; at
// 100 KB of LDS per workgroup only one workgroup fits a CU, so those barriers
// are exposed. Here each workgroup owns a fixed group of RBW row blocks (16 RBW
// output directions) and copies that slice of the panel for ALL K-steps into
// LDS once, then streams cells through it with no further barrier: persistent
// workgroups, one per CU. Two parts per row block are resident (hi and lo;
// hi and mid for the three-part form, whose lo part is read per K-step from
// L2, where every row group's slice of it stays): 2 KB per (K-step, row block),
// so five row blocks fit at M = 400 (130 KB) and the 25 row blocks of M = 400
// split into five groups without padding.
//
// Work split. The G row groups of a problem are served by G workgroups that
// sit on the same XCD (workgroup b runs on XCD b % 8) and walk the same
// sequence of 256-cell chunks, so a chunk's state is read from HBM once and
// from that XCD's L2 by the other G - 1. Each XCD owns a contiguous 1/8 of the
// chunks; within it, the floor(CUs per XCD / G) sets of G workgroups take
// chunks round-robin.
//
// Per wave (four waves, one per SIMD): four 16-cell tiles x RBW row blocks, so
// every A fragment read from LDS feeds four MFMAs per product; MFMAs are issued
// product-major so that no MFMA depends on the one before it; state vectors
// are prefetched K-steps ahead, across chunk boundaries too, with ordinary
// (L2-allocating) loads, since the other row groups of the XCD read the same
// vectors. Products and their order per (row block, tile) are those of the
// streaming kernel (each K-step's main hi*hi formed alone and added in FP32;
// cross hi*lo + lo*hi, and for three parts hi*lo, hi*mid, mid*mid, mid*hi,
// lo*hi), so the kernels give bit-identical results on the same state and
// panel.
#include "mfma_split_core.hpp"

namespace rsh {

namespace {

constexpr int kResWaves = 8;   // waves per workgroup of the grouped iteration kernel (2 per SIMD)
constexpr int kResTiles = 2;   // 16-cell tiles per wave of the grouped iteration kernel
constexpr int kResChunkTiles = kResWaves * kResTiles;  // 256 cells per chunk
constexpr int kRes4Waves = 4;  // waves per workgroup of the single-problem kernel (1 per SIMD)
constexpr int kRes4Tiles = 4;  // 16-cell tiles per wave of the single-problem kernel
static_assert(kRes4Waves * kRes4Tiles == kResChunkTiles, "both resident kernels walk 256-cell chunks");
constexpr int kXcds = 8;       // MI355X: 8 XCDs, workgroups dispatched round-robin
constexpr int kResLdsBytes = 160 * 1024;
constexpr int kIterResSub = 4;  // 256-cell chunks per item of the grouped iteration kernel
// Buffer descriptor word 3 for raw (untyped) gfx9 buffer access: 32-bit data
// format, so that the hardware does not treat the descriptor as invalid.
constexpr int kBufWord3 = 0x00020000;

// State prefetch depth in K-steps of the single-problem kernel (registers:
// XD x 4 tiles x parts x 4 VGPRs); the three-part form also prefetches the lo
// part of its panel slice kResLoXd K-steps ahead (RBW x 4 VGPRs each).
template <int NT>
constexpr int res_xd() { return NT == 2 ? 3 : 2; }
constexpr int kResLoXd = 1;

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// Workgroup barrier that orders LDS only: it waits for this wave's LDS
// accesses, not for its global loads in flight (the state prefetch).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One K-step of MFMAs for TT tiles x RBW row blocks: A fragments of the parts
// resident in LDS from a_base (this lane's vector of the step's first
// fragment; NL parts per row block), the lo part of a three-part panel from
// alo when NL < NT, B from the state vectors in bx. A fragments of one part at
// a time (part-major), so that only RBW fragments are live. Product-major
// order: each round touches every (row block, tile) accumulator once, so an
// MFMA never waits for the previous result of its own accumulator. The main
// product hi*hi of the step is formed in fresh accumulators, one row block at a
// time (main_round), and added to accm with main_add. Products
// and their order per accumulator are those of the streaming kernel (cross:
// hi*lo + lo*hi; for three parts hi*lo, hi*mid, mid*mid, mid*hi, lo*hi).
//
// ahi (in/out): with AHEAD, ahi holds this step's hi fragments on entry, and
// the hi fragments of the K-step at a_next are read into it as soon as this
// step is done with its own (after the main products), so that the next step's
// first MFMAs do not wait for LDS. Without AHEAD, the step reads its own hi
// fragments first.
template <int RBW, int FMT, int NT, int TT, int NL, bool AHEAD>
__device__ __forceinline__ void res_mfma_step(const v4u* a_base, const v4u* a_next,
                                              typename Fmt<FMT>::V (&ahi)[RBW],
                                              const typename Fmt<FMT>::V (&bx)[TT][NT],
                                              const typename Fmt<FMT>::V* alo, v4f (&accm)[RBW][TT],
                                              v4f (&accx)[RBW][TT]) {
    typedef typename Fmt<FMT>::V V;
    V (&a)[RBW] = ahi;
    V am[RBW];  // second (and third) part
    auto fetch_into = [&](V(&dst)[RBW], const v4u* base, int p) {
#pragma unroll
        for (int i = 0; i < RBW; i++) dst[i] = __builtin_bit_cast(V, base[(i * NL + p) * 64]);
    };
    auto round = [&](const V(&af)[RBW], v4f (&acc)[RBW][TT], int pb) {
#pragma unroll
        for (int i = 0; i < RBW; i++)
#pragma unroll
            for (int q = 0; q < TT; q++) acc[i][q] = Fmt<FMT>::mfma(af[i], bx[q][pb], acc[i][q]);
    };
    // a holds the hi fragments. Software-pipelined over row blocks: row block
    // i's products are issued before row block i - 1's are added (so that the
    // adds do not wait for the MFMA just issued), and row block i's MFMAs are
    // ordered after row block i - 2's adds (order_after; two row blocks of
    // products in flight at most).
    auto main_round = [&]() {
        v4f ms[2][TT];
        static_for<0, RBW>([&](auto ii) {
            constexpr int i = decltype(ii)::value;
            order_after(a[i], accm[(i + 2 * RBW - 2) % RBW]);
#pragma unroll
            for (int q = 0; q < TT; q++) ms[i & 1][q] = Fmt<FMT>::mfma(a[i], bx[q][0], v4f{0.f, 0.f, 0.f, 0.f});
            if constexpr (i >= 1) {
#pragma unroll
                for (int q = 0; q < TT; q++) accm[i - 1][q] = main_add(accm[i - 1][q], ms[(i - 1) & 1][q]);
            }
        });
#pragma unroll
        for (int q = 0; q < TT; q++) accm[RBW - 1][q] = main_add(accm[RBW - 1][q], ms[(RBW - 1) & 1][q]);
    };
    if constexpr (!AHEAD) fetch_into(a, a_base, 0);
    fetch_into(am, a_base, 1);
    if constexpr (NT == 2) {
        round(a, accx, 1);
        main_round();
        if constexpr (AHEAD) fetch_into(a, a_next, 0);
        round(am, accx, 0);
    } else {
        round(a, accx, 2);
        round(a, accx, 1);
        main_round();
        if constexpr (AHEAD) fetch_into(a, a_next, 0);
        round(am, accx, 1);
        round(am, accx, 0);
        if constexpr (NL == 3) {
            fetch_into(am, a_base, 2);
        } else {
#pragma unroll
            for (int i = 0; i < RBW; i++) am[i] = alo[i];
        }
        round(am, accx, 0);
    }
}

// STEPS > 0: the K-step count is a compile-time constant and the K loop is
// fully unrolled, so the state prefetch crosses no loop back-edge (where the
// compiler would otherwise wait for every outstanding load). STEPS = 0: any M.
// NL = parts of the panel slice resident in LDS (2); a three-part panel's lo
// part comes from global memory (L2), kResLoXd K-steps ahead.
template <int RBW, int FMT, int NT, int STEPS>
__global__ __launch_bounds__(64 * kRes4Waves) void split_resident_kernel(
    const v4u* __restrict__ state, int64_t n, int steps_rt, const float* __restrict__ cell_inv,
    const v4u* __restrict__ panel, const float* __restrict__ row_inv, void* __restrict__ out, int64_t ld, int rows,
    int out_f32, int groups, int sets, int64_t chunks) {
    typedef typename Fmt<FMT>::V V;
    constexpr int TT = kRes4Tiles;
    constexpr int NL = 2;
    constexpr int XD = res_xd<NT>();
    constexpr int XL = NT == 3 ? kResLoXd : 1;
    const int steps = STEPS > 0 ? STEPS : steps_rt;
    extern __shared__ v4u lds[];  // [steps][RBW][NL][64], then the row group's 16 RBW row scales
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar tile bounds
    const int g = lane >> 4;
    const int xcd = blockIdx.x % kXcds;
    const int local = blockIdx.x / kXcds;
    const int set = local / groups;
    const int rg = local % groups;
    if (set >= sets) return;  // whole workgroup leaves together: no barrier below is skipped by part of it

    // 1. the row group's panel slice, all K-steps, parts 0 .. NL - 1, into LDS
    const int64_t slice = (int64_t)steps * RBW * NT * 64;  // the row group's panel, 16-byte vectors
    const v4u* __restrict__ src = panel + (int64_t)rg * slice;
    const int64_t units = (int64_t)steps * RBW * NL * 64;
    for (int64_t u = threadIdx.x; u < units; u += 64 * kRes4Waves) {
        const int64_t f = u >> 6;  // (s * RBW + rb) * NL + p
        lds[u] = src[((f / NL) * NT + f % NL) * 64 + (u & 63)];
    }
    const int row0 = rg * RBW * 16;
    // Row scales once per workgroup into LDS: scale loads from global memory in
    // every epilogue measurably broke the row groups' L2 sharing of the state.
    float* rscale = reinterpret_cast<float*>(lds + units);
    for (int r = threadIdx.x; r < RBW * 16; r += 64 * kRes4Waves)
        rscale[r] = row0 + r < rows ? row_inv[row0 + r] : 0.0f;
    __syncthreads();

    // The lo part of a three-part panel slice through one buffer descriptor; an
    // opaque zero in its offset, renewed per chunk, keeps the compiler from
    // hoisting all K-steps' loads out of the chunk loop (they would not fit).
    const __amdgpu_buffer_rsrc_t lo_rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<v4u*>(src), (short)0, (int)(slice * 16), kBufWord3);
    int lo_zero = 0;
    v4u plo[XL][RBW];
    auto load_lo = [&](auto slot_tag, int t) {
        if constexpr (NT == 3) {
            constexpr int sl = decltype(slot_tag)::value;
#pragma unroll
            for (int i = 0; i < RBW; i++)
                plo[sl][i] = __builtin_bit_cast(v4u, __builtin_amdgcn_raw_buffer_load_b128(
                                                         lo_rs, lane * 16 + lo_zero, ((t * RBW + i) * NT + 2) * 1024, 0));
        }
    };

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
    int xoff[TT];
    auto bind_chunk = [&](int64_t chunk) {
        const int64_t ct_chunk = chunk * kResChunkTiles;
        const int ntile = (int)(tiles - ct_chunk < kResChunkTiles ? tiles - ct_chunk : kResChunkTiles);
        xs = __builtin_amdgcn_make_buffer_rsrc(const_cast<v4u*>(state + ct_chunk * tile_units), (short)0,
                                               ntile * tile_bytes, kBufWord3);
#pragma unroll
        for (int q = 0; q < TT; q++) {
            const int tile = wave * TT + q;
            xoff[q] = (tile < ntile ? tile : ntile - 1) * tile_bytes + lane * 16;
        }
    };
    v4u xr[XD][TT][NT];
    auto load_x = [&](auto slot_tag, int t) {
        constexpr int sl = decltype(slot_tag)::value;
#pragma unroll
        for (int q = 0; q < TT; q++)
#pragma unroll
            for (int p = 0; p < NT; p++)
                xr[sl][q][p] = __builtin_bit_cast(
                    v4u, __builtin_amdgcn_raw_buffer_load_b128(xs, xoff[q] + p * 1024, t * NT * 1024, 0));
    };
    // First K-steps of a chunk. For every chunk after the first they are
    // issued before the previous chunk's result stores: loads and stores retire
    // through one in-order counter, so a load issued after the stores could
    // only be waited for together with all of them.
    auto prefetch_head = [&]() {
        asm volatile("v_mov_b32 %0, 0" : "=v"(lo_zero));
        static_for<0, XD>([&](auto u) {
            if (u.value < steps) load_x(u, u.value);
        });
        static_for<0, XL>([&](auto u) {
            if (u.value < steps) load_lo(u, u.value);
        });
    };

    int64_t chunk = c_begin + set;
    // hi fragments of the next K-step (res_mfma_step reads them a step ahead)
    V ahi[RBW];
#pragma unroll
    for (int i = 0; i < RBW; i++) ahi[i] = __builtin_bit_cast(V, lds[(i * NL) * 64 + lane]);
    if (chunk < c_end) {
        bind_chunk(chunk);
        prefetch_head();
    }
    for (; chunk < c_end; chunk += sets) {
        const int64_t ct0 = chunk * kResChunkTiles + wave * TT;
        // Cell scales now, ahead of the next chunk's prefetch: a load issued
        // after those could only be waited for together with all of them.
        float cinv[TT];
#pragma unroll
        for (int q = 0; q < TT; q++) {
            const int64_t cell = (ct0 + q) * 16 + (lane & 15);
            cinv[q] = cell_inv[cell < n ? cell : n - 1];
        }
        v4f accm[RBW][TT], accx[RBW][TT];
#pragma unroll
        for (int i = 0; i < RBW; i++)
#pragma unroll
            for (int q = 0; q < TT; q++) {
                accm[i][q] = v4f{0.f, 0.f, 0.f, 0.f};
                accx[i][q] = v4f{0.f, 0.f, 0.f, 0.f};
            }
        // K-step t with ring slot SL = t % XD and lo slot SLL = t % XL
        auto kstep = [&](auto slot_tag, auto lo_tag, const int t) {
            constexpr int sl = decltype(slot_tag)::value;
            constexpr int sll = decltype(lo_tag)::value;
            // Scheduling fence per K-step: in the unrolled loop the scheduler
            // would otherwise hoist every K-step's loads and run out of registers.
            __builtin_amdgcn_sched_barrier(0);
            V bx[TT][NT];
#pragma unroll
            for (int q = 0; q < TT; q++)
#pragma unroll
                for (int p = 0; p < NT; p++) bx[q][p] = __builtin_bit_cast(V, xr[sl][q][p]);
            V alo[RBW];
            if constexpr (NT == 3) {
#pragma unroll
                for (int i = 0; i < RBW; i++) alo[i] = __builtin_bit_cast(V, plo[sll][i]);
                if (t + XL < steps) load_lo(lo_tag, t + XL);
            }
            if (t + XD < steps) load_x(slot_tag, t + XD);
            const int tn = t + 1 < steps ? t + 1 : 0;  // the next chunk starts at K-step 0
            res_mfma_step<RBW, FMT, NT, TT, NL, true>(lds + (int64_t)t * RBW * NL * 64 + lane,
                                                lds + (int64_t)tn * RBW * NL * 64 + lane, ahi, bx, alo, accm, accx);
        };
        if constexpr (STEPS > 0) {
            static_for<0, STEPS>([&](auto t) {
                kstep(std::integral_constant<int, t.value % XD>{}, std::integral_constant<int, t.value % XL>{},
                      t.value);
            });
        } else {
            // XD and XL divide 6: blocks of six K-steps keep every ring slot a
            // compile-time constant
            static_assert(6 % XD == 0 && 6 % XL == 0, "ring depths must divide the unroll of six");
            for (int t = 0; t < steps; t += 6)
                static_for<0, 6>([&](auto u) {
                    if (t + u.value < steps)
                        kstep(std::integral_constant<int, u.value % XD>{},
                              std::integral_constant<int, u.value % XL>{}, t + u.value);
                });
        }
        __builtin_amdgcn_sched_barrier(0);
        if (chunk + sets < c_end) {
            bind_chunk(chunk + sets);
            prefetch_head();
        }
        __builtin_amdgcn_sched_barrier(0);

        // Epilogue: FP64 combination of the main and cross terms, exact scale
        // removal. Row scales are re-read from LDS per chunk: an opaque zero
        // offset keeps the compiler from hoisting them out of the chunk loop.
        // Lane pairs (cells 2c, 2c + 1 of a tile) swap half of their results
        // (swap_pair_lanes), so that a lane holds two adjacent cells of one row:
        // the even lane rows 4g, 4g + 1, the odd lane rows 4g + 2, 4g + 3. Each
        // store then writes two cells (16 bytes for FP64), half as many store
        // instructions as one cell per lane: with one cell per lane a chunk's
        // stores alone exceeded the wave's 63 outstanding memory operations.
        // Stores go through one buffer descriptor per row block (uniform base,
        // row offset h * ld in a scalar register, one VGPR offset per tile).
        int zoff;
        asm volatile("v_mov_b32 %0, 0" : "=v"(zoff));
        const float* __restrict__ rs = rscale + 4 * g + zoff;
        const int odd = lane & 1;
        auto store_all = [&](auto elem_tag, auto full_tag) {
            typedef decltype(elem_tag) OT;
            constexpr bool kFull = decltype(full_tag)::value;
            const int esz = (int)sizeof(OT);
            int voff[TT];
            bool ok1[TT], ok2[TT];
#pragma unroll
            for (int q = 0; q < TT; q++) {
                const int64_t cell0 = (ct0 + q) * 16 + (lane & 14);  // first cell of the lane pair
                ok1[q] = cell0 < n;
                ok2[q] = cell0 + 1 < n;
                voff[q] = (int)(((4 * g + 2 * odd) * ld + cell0) * esz);
            }
#pragma unroll
            for (int i = 0; i < RBW; i++) {
                const int rb_row = row0 + i * 16;
                if (!kFull && rb_row >= rows) break;
                const __amdgpu_buffer_rsrc_t os = __builtin_amdgcn_make_buffer_rsrc(
                    static_cast<OT*>(out) + (int64_t)rb_row * ld, (short)0, (int)(16 * ld * esz), kBufWord3);
                const v4f rsv = *reinterpret_cast<const v4f*>(rs + i * 16);  // this lane's 4 row scales
#pragma unroll
                for (int q = 0; q < TT; q++) {
                    OT v[4];
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        v[r] = (OT)(split_value(accm[i][q][r], accx[i][q][r]) * (double)rsv[r] * (double)cinv[q]);
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        const OT recv = swap_pair_lanes(odd ? v[h] : v[2 + h]);
                        const OT first = odd ? recv : v[h];      // cell 2c
                        const OT second = odd ? v[2 + h] : recv;  // cell 2c + 1
                        if (!kFull && rb_row + 4 * g + 2 * odd + h >= rows) continue;
                        if ((ct0 + q + 1) * 16 <= n)  // whole tile inside the output (uniform branch)
                            store_cell_pair(first, second, true, true, os, voff[q], (int)(h * ld * esz));
                        else
                            store_cell_pair(first, second, ok2[q], ok1[q], os, voff[q], (int)(h * ld * esz));
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
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(64 * kRes4Waves), lds_bytes, st, state, n, steps, cell_inv, panel,
                       row_inv, out, ld, rows, out_f32, groups, sets, chunks);
    return hipGetLastError();
}

// Source iteration I <- (1 - a) B + a S(I) on split state for a batch of
// problems, resident-panel form (the grouped counterpart of the kernel above).
// Work: items (problem, row group, kIterResSub chunks of 256 cells) in eight
// per-XCD lists; workgroup b takes items from list b % 8, so the row groups of
// one chunk run on one XCD at about the same time and share its L2 copy of the
// state. Per item the row group's panel slice (all K-steps) and row scales go
// into LDS, then the chunks are swept as in split_resident_kernel (state
// prefetch across chunk boundaries, issued before the stores).
// KS (K-streamed panel, for long K ranges): the slice is not held whole; a
// two-slot LDS ring holds one K-step each, the next K-step's part of the slice
// is loaded into registers while the current one is multiplied, written into
// the other slot after it and published by an LDS-only barrier per K-step. The
// LDS then no longer bounds the row group: at M = 800 it holds 4 (three parts)
// or 5 (two parts) row blocks instead of 2 or 3, so each state fragment read
// from L2 feeds twice the MFMAs. The epilogue is
// that of the streaming grouped kernel (kernels_iter_split.hip): S in FP64 from
// the two accumulators and the exact scales, I_new = (1 - a) B + a S, the
// residual against the state the MFMAs consumed, I_new re-split into the other
// ping-pong buffer. Same products, same order: bit-identical to it.
template <int RBW, int FMT, int NT, int STEPS, int W, bool KS>
__global__ __launch_bounds__(64 * W, W == 4 ? 2 : 1) void grouped_iter_split_res_kernel(
    const BatchProblem* __restrict__ probs, const BatchItem* __restrict__ items, const int* __restrict__ meta,
    unsigned* __restrict__ counters, int parity) {
    typedef typename Fmt<FMT>::V V;
    typedef typename Fmt<FMT>::T T;
    // Chunk of CT 16-cell tiles, TT per wave: 8 waves x 2 tiles (one workgroup
    // per CU), or 4 waves x 2 tiles (two workgroups per CU, whose epilogues and
    // barriers are independent, so one's epilogue overlaps the other's MFMAs).
    constexpr int TT = kResTiles;
    constexpr int CT = W * TT;
    constexpr int SUB = kIterResSub * kResChunkTiles / CT;  // chunks per item: an item is 1024 cells
    static_assert(SUB * CT == kIterResSub * kResChunkTiles, "items of 1024 cells");
    constexpr int XD = NT == 2 ? 3 : 2;  // shallower than the plain kernel: the epilogue needs the registers
    constexpr int steps = STEPS;
    constexpr int64_t units = (int64_t)steps * RBW * NT * 64;  // panel slice of a row group
    constexpr int SU = RBW * NT * 64;                          // one K-step of it
    constexpr int64_t lds_units = KS ? 2 * SU : units;
    // KS: 16-byte units of one K-step per thread (the last pass partial)
    constexpr int PER = (SU + 64 * W - 1) / (64 * W);
    extern __shared__ v4u lds[];  // [steps][RBW][NT][64] (KS: [2][RBW][NT][64]), then 16 RBW row scales
    __shared__ int s_item;
    float* rscale = reinterpret_cast<float*>(lds + lds_units);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar tile offsets
    const int g = lane >> 4;
    const int cl = lane & 15;
    const int xcd = blockIdx.x % kXcds;
    const int begin = meta[xcd], count = meta[kXcds + xcd];
    constexpr int64_t tile_units = (int64_t)steps * NT * 64;
    constexpr int tile_bytes = (int)(tile_units * 16);
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(counters + xcd, 1u);
        __syncthreads();  // also: every wave is done with the previous item's LDS panel
        const int it = s_item;
        __syncthreads();
        if (it >= count) break;
        const BatchItem item = items[begin + it];
        const BatchProblem& P = probs[item.problem];
        const int rg = item.pass;
        const int row0 = rg * RBW * 16;
        const v4u* __restrict__ src = static_cast<const v4u*>(P.spanel) + (int64_t)rg * units;
        // KS: the slice through a buffer descriptor (global loads, which an
        // LDS-only barrier does not wait for) and K-step 0 into ring slot 0.
        const __amdgpu_buffer_rsrc_t psrc =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<v4u*>(src), (short)0, (int)(units * 16), kBufWord3);
        v4u preg[PER];
        auto panel_load = [&](int step) {
#pragma unroll
            for (int i = 0; i < PER; i++) {
                const int u = (int)threadIdx.x + i * 64 * W;
                if (SU % (64 * W) == 0 || u < SU)
                    preg[i] = __builtin_bit_cast(
                        v4u, __builtin_amdgcn_raw_buffer_load_b128(psrc, u * 16, step * SU * 16, 0));
            }
        };
        auto panel_store = [&](int slot) {
#pragma unroll
            for (int i = 0; i < PER; i++) {
                const int u = (int)threadIdx.x + i * 64 * W;
                if (SU % (64 * W) == 0 || u < SU) lds[slot * SU + u] = preg[i];
            }
        };
        if constexpr (KS) {
            panel_load(0);
            panel_store(0);
        } else {
            for (int64_t u = threadIdx.x; u < units; u += 64 * W) lds[u] = src[u];
        }
        for (int r = threadIdx.x; r < RBW * 16; r += 64 * W)
            rscale[r] = row0 + r < P.m ? P.row_inv[row0 + r] : 0.0f;
        __syncthreads();

        const v4u* __restrict__ st_in = static_cast<const v4u*>(P.state[parity]);
        v4u* __restrict__ st_out = static_cast<v4u*>(P.state[parity ^ 1]);
        const int64_t n = P.n;
        const int64_t tiles = (n + 15) / 16;
        const int64_t chunks = (tiles + CT - 1) / CT;
        const int64_t c0 = item.tile * SUB;
        const int64_t c1 = c0 + SUB < chunks ? c0 + SUB : chunks;

        __amdgpu_buffer_rsrc_t xs;
        int xoff[TT];
        int64_t ntile_of_bound = 0;
        auto bind_chunk = [&](int64_t chunk) {
            const int64_t ct_chunk = chunk * CT;
            const int ntile = (int)(tiles - ct_chunk < CT ? tiles - ct_chunk : CT);
            ntile_of_bound = ntile;
            xs = __builtin_amdgcn_make_buffer_rsrc(const_cast<v4u*>(st_in + ct_chunk * tile_units), (short)0,
                                                   ntile * tile_bytes, kBufWord3);
#pragma unroll
            for (int q = 0; q < TT; q++) {
                const int tile = wave * TT + q;
                xoff[q] = (tile < ntile ? tile : ntile - 1) * tile_bytes + lane * 16;
            }
        };
        v4u xr[XD][TT][NT];
        auto load_x = [&](auto slot_tag, int t) {
            constexpr int sl = decltype(slot_tag)::value;
#pragma unroll
            for (int q = 0; q < TT; q++)
#pragma unroll
                for (int p = 0; p < NT; p++)
                    xr[sl][q][p] = __builtin_bit_cast(
                        v4u, __builtin_amdgcn_raw_buffer_load_b128(xs, xoff[q] + p * 1024, t * NT * 1024, 0));
        };
        auto prefetch_head = [&]() {
            static_for<0, XD>([&](auto u) {
                if (u.value < steps) load_x(u, u.value);
            });
        };
        RelMax rm;
        if (c0 < c1) {
            bind_chunk(c0);
            prefetch_head();
        }
        for (int64_t chunk = c0; chunk < c1; chunk++) {
            const int64_t ct0 = chunk * CT + wave * TT;
            float cinv[TT];
#pragma unroll
            for (int q = 0; q < TT; q++) {
                const int64_t cell = (ct0 + q) * 16 + cl;
                cinv[q] = P.cell_inv[cell < n ? cell : n - 1];
            }
            v4f accm[RBW][TT], accx[RBW][TT];
#pragma unroll
            for (int i = 0; i < RBW; i++)
#pragma unroll
                for (int q = 0; q < TT; q++) {
                    accm[i][q] = v4f{0.f, 0.f, 0.f, 0.f};
                    accx[i][q] = v4f{0.f, 0.f, 0.f, 0.f};
                }
            // KS: K-step t of this chunk sits in ring slot (t + pc) & 1 (STEPS
            // is odd, so the parity flips from chunk to chunk); the step after
            // the last one is K-step 0 of the next chunk.
            const int pc = (int)((chunk - c0) & 1);
            const bool more_chunks = chunk + 1 < c1;
            static_for<0, STEPS>([&](auto tt) {
                constexpr int t = decltype(tt)::value;
                constexpr int sl = t % XD;
                __builtin_amdgcn_sched_barrier(0);
                // KS: the next K-step's panel load goes out before this step's
                // state prefetch, so that waiting for it (before its LDS store)
                // leaves the state loads in flight (one in-order counter).
                const bool next = KS && (t + 1 < STEPS || more_chunks);
                if constexpr (KS) {
                    if (next) panel_load((t + 1) % STEPS);
                }
                V bx[TT][NT];
#pragma unroll
                for (int q = 0; q < TT; q++)
#pragma unroll
                    for (int p = 0; p < NT; p++) bx[q][p] = __builtin_bit_cast(V, xr[sl][q][p]);
                if constexpr (t + XD < STEPS) load_x(std::integral_constant<int, sl>{}, t + XD);
                V ahi[RBW];
                if constexpr (KS) {
                    static_assert(STEPS % 2 == 1, "the ring parity assumes an odd K-step count");
                    res_mfma_step<RBW, FMT, NT, TT, NT, false>(lds + ((t + pc) & 1) * SU + lane, nullptr, ahi, bx,
                                                               nullptr, accm, accx);
                    if (next) panel_store((t + 1 + pc) & 1);
                    lds_barrier();
                } else {
                    res_mfma_step<RBW, FMT, NT, TT, NT, false>(lds + (int64_t)t * RBW * NT * 64 + lane, nullptr, ahi,
                                                               bx, nullptr, accm, accx);
                }
            });
            // This chunk's state (old values) and the output chunk through buffer
            // descriptors (uniform base, 32-bit offsets), kept before the next
            // chunk's head is bound and requested ahead of this chunk's stores.
            const __amdgpu_buffer_rsrc_t xin = xs;
            const int ntile_cur = (int)ntile_of_bound;
            const __amdgpu_buffer_rsrc_t xout = __builtin_amdgcn_make_buffer_rsrc(
                st_out + chunk * CT * tile_units, (short)0, ntile_cur * tile_bytes, kBufWord3);
            __builtin_amdgcn_sched_barrier(0);
            if (chunk + 1 < c1) {
                bind_chunk(chunk + 1);
                prefetch_head();
            }
            __builtin_amdgcn_sched_barrier(0);
            int zoff;
            asm volatile("v_mov_b32 %0, 0" : "=v"(zoff));
            const float* __restrict__ rs = rscale + 4 * g + zoff;
            const double oma = P.one_minus_albedo, alb = P.albedo;  // in registers across the stores
            // Branch-free epilogue: every load is issued with an address that is
            // in range (cell clamped to the last cell, a padding row block to the
            // last real one), results of invalid positions are dropped by giving
            // their stores an offset past the buffer's range (the hardware
            // discards such stores) and by leaving them out of the residual. So
            // the whole epilogue is one basic block, and the scheduler overlaps the
            // loads of later (tile, row block) pairs with the arithmetic of earlier
            // ones instead of waiting for one memory round trip per pair.
            const int rb_last = (P.m + 15) / 16 - 1;  // last row block with rows
            constexpr int kDropped = 0x7FFF0000;      // store offset past any buffer's range
            // (tile q, row block i) pairs in order p = q * RBW + i; the loads of
            // pair p + LA are issued before the stores of pair p (the compiler
            // may not move a load above a store to another buffer: it cannot tell
            // that they do not alias). LA = 2 measured best at 4 and 5 row
            // blocks (with B in FP32; 4 was slower for three parts).
            constexpr int NPAIR = TT * RBW;
            constexpr int LA = RBW <= 3 ? 3 : 2;
            float bv[LA + 1][4];
            v2u oldraw[LA + 1][NT];
            auto pair_rb = [&](int i, bool& rb_ok) {
                const int rbg = rg * RBW + i;
                rb_ok = rbg <= rb_last;  // a padding row block has no rows and no state
                return rb_ok ? rbg : rb_last;
            };
            auto pair_voff = [&](int q, int rbe) {
                const int vlane = (2 * (rbe & 1) + (g >> 1)) * 16 + cl;
                return ((wave * TT + q) * (int)tile_units + vlane) * 16 + 8 * (g & 1);
            };
            auto load_pair = [&](auto pp) {
                constexpr int pi = decltype(pp)::value, q = pi / RBW, i = pi % RBW, buf = pi % (LA + 1);
                bool rb_ok;
                const int rbe = pair_rb(i, rb_ok);
                // the source term in output-tile order: this lane's 4 rows of its
                // cell are one 16-byte vector (the tile clamped into the array)
                const int64_t tq = ct0 + q < tiles ? ct0 + q : tiles - 1;
                const __amdgpu_buffer_rsrc_t bsrc = __builtin_amdgcn_make_buffer_rsrc(
                    const_cast<float*>(P.Bs) + (int64_t)rbe * tiles * 256, (short)0, (int)(tiles * 1024), kBufWord3);
                const v4f b4 = __builtin_bit_cast(
                    v4f, __builtin_amdgcn_raw_buffer_load_b128(bsrc, (int)((tq * 64 + lane) * 16), 0, 0));
#pragma unroll
                for (int r = 0; r < 4; r++) bv[buf][r] = b4[r];
                const int voff = pair_voff(q, rbe);
#pragma unroll
                for (int t = 0; t < NT; t++)
                    oldraw[buf][t] = __builtin_bit_cast(
                        v2u, __builtin_amdgcn_raw_buffer_load_b64(xin, voff, ((rbe >> 1) * NT + t) * 1024, 0));
            };
            static_for<0, (LA < NPAIR ? LA : NPAIR)>(load_pair);
            static_for<0, NPAIR>([&](auto pp) {
                constexpr int pi = decltype(pp)::value, q = pi / RBW, i = pi % RBW, buf = pi % (LA + 1);
                if constexpr (pi + LA < NPAIR) load_pair(std::integral_constant<int, pi + LA>{});
                const int tl = wave * TT + q;  // tile within the chunk
                const int64_t cell = (ct0 + q) * 16 + cl;
                bool rb_ok;
                const int rbe = pair_rb(i, rb_ok);
                // B rows exist up to the row block's end (the source array holds
                // whole row blocks, padding rows zero). A padding row (k >= M)
                // has zero panel rows, zero B and zero old state: nv = 0, parts
                // 0, residual term 0 / 0 = NaN, which fmax drops.
                const bool ok = tl < ntile_cur && cell < n && rb_ok;
                // The cell and row scales are powers of two, so scaling in FP32
                // is exact and commutes with every rounding below: the same
                // values as scaling in FP64, with four FP64 operations fewer per
                // result (FP64 runs at half the FP32 rate on the VALU).
                const float cif = cinv[q];
                const float csf = 1.0f / cif;  // power of two: exact
                T newp[NT][4];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const float sc = rs[i * 16 + r] * cif;
                    const double S = split_value(accm[i][q][r] * sc, accx[i][q][r] * sc);
                    // explicit FMA: the same rounding in every split iteration kernel
                    const double nv = __builtin_fma(alb, S, oma * (double)bv[buf][r]);
                    float of = 0.0f;
#pragma unroll
                    for (int t = NT - 1; t >= 0; t--) {
                        T tmp[4];
                        __builtin_memcpy(tmp, &oldraw[buf][t], 8);
                        of += (float)tmp[r];
                    }
                    const double old = (double)(of * cif);
                    rm.add(old - nv, nv, ok);
                    T qp[NT];
                    split_parts<FMT, NT>((float)nv * csf, qp);
#pragma unroll
                    for (int t = 0; t < NT; t++) newp[t][r] = qp[t];
                }
                const int soff = ok ? pair_voff(q, rbe) : kDropped;
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    v2u raw;
                    __builtin_memcpy(&raw, newp[t], 8);
                    __builtin_amdgcn_raw_buffer_store_b64(raw, xout, soff, ((rbe >> 1) * NT + t) * 1024, 0);
                }
            });
        }
        double res = rm.value();
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) res = fmax(res, __shfl_xor(res, off, 64));
        if (lane == 0 && res > 0.0)
            atomicMax(P.resid + parity, (unsigned long long)__double_as_longlong(res));
    }
}

// The grouped iteration kernel streams the panel slice through LDS by K-step
// (KS) at 25 K-steps (M = 800), where a whole slice would leave room for only
// 2 or 3 row blocks; below that it holds the slice.
constexpr bool res_k_streamed(int steps) { return steps >= 25; }

// Largest row-block count per workgroup: the grouped iteration kernel (eight
// waves, two per SIMD) keeps two parts in 256 registers up to five row blocks;
// three parts up to five at 7 K-steps (two dwords spill, in the prologue) and
// with the K-step ring at 25 (the eight-wave form spills three dwords, outside
// the K loop), four at 13, where a whole slice of five does not fit LDS. Five
// row blocks take the 13 of M = 200 in 3 row groups and the 50 of M = 800 in
// 10; four would pad them to 16 and 52. The single-problem kernel (one wave
// per SIMD, 512 registers) holds five.
constexpr int res_max_rbw(int nt, int steps) { return nt == 2 || res_k_streamed(steps) || steps <= 7 ? 5 : 4; }
constexpr int kRes4MaxRbw = 5;
constexpr int kRes4LdsParts = 2;  // panel parts resident in LDS in the single-problem kernel

}  // namespace

// Row-group plan of a resident-panel kernel for output rows [0, rows) of an
// M-direction set: rb_per_pass = RBW row blocks per workgroup (largest slice
// that fits LDS, then balanced), passes = G row groups. grouped: the plan of
// the grouped iteration kernel (all parts of the slice in LDS), else that of
// the single-problem kernel (two parts in LDS). False if the kernel does not
// apply (G row groups do not fit one XCD's CUs).
bool plan_split_resident(int form, int m, int rows, int cus, SplitPlan* pl, int* sets, bool grouped) {
    const int nt = split_terms(form);
    const int steps = (m + 31) / 32;
    const int rb_need = (rows + 15) / 16;
    const int lds_parts = grouped ? nt : kRes4LdsParts;
    const int lds_steps = grouped && res_k_streamed(steps) ? 2 : steps;    // a two-slot ring of K-steps
    const int per_rb = lds_steps * lds_parts * 1024 + 16 * 4;               // panel fragments + row scales
    int rbw_max = kResLdsBytes / per_rb;
    const int cap = grouped ? res_max_rbw(nt, steps) : kRes4MaxRbw;
    if (rbw_max > cap) rbw_max = cap;
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
    const size_t lds_bytes = (size_t)pl.steps * pl.rb_per_pass * kRes4LdsParts * 1024 + (size_t)pl.rb_per_pass * 16 * 4;
    if (lds_bytes > (size_t)kResLdsBytes || pl.rb_per_pass > kRes4MaxRbw || sets < 1 || ld >= kSplitResidentMaxLd)
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
#define RSH_RES(RBW, FMT, NT)                                                                   \
    {                                                                                           \
        if constexpr (RBW == 5) { if (pl.steps == 7) RSH_RES_S(RBW, FMT, NT, 7); }              \
        if constexpr (RBW == 5) { if (pl.steps == 13) RSH_RES_S(RBW, FMT, NT, 13); }            \
        if constexpr (RBW == 3) { if (pl.steps == 25) RSH_RES_S(RBW, FMT, NT, 25); }            \
        RSH_RES_S(RBW, FMT, NT, 0);                                                             \
    }
#define RSH_RES_FORMS(RBW)                           \
    case RBW:                                        \
        if (fmt == 0 && nt == 2) RSH_RES(RBW, 0, 2); \
        if (fmt == 1 && nt == 2) RSH_RES(RBW, 1, 2); \
        RSH_RES(RBW, 0, 3);
    switch (pl.rb_per_pass) {
        RSH_RES_FORMS(1) RSH_RES_FORMS(2) RSH_RES_FORMS(3) RSH_RES_FORMS(4) RSH_RES_FORMS(5)
        default:
            return hipErrorInvalidValue;
    }
#undef RSH_RES_FORMS
#undef RSH_RES
#undef RSH_RES_S
}

// Grouped iteration, resident-panel form: the specialised plans (M = 200, 400,
// 800). rb = RBW of the class, steps its K-steps; false if not specialised.
bool iter_split_res_supported(int form, int rb, int steps) {
    const int nt = split_terms(form);
    if (nt == 2) return rb == 5 && (steps == 7 || steps == 13 || steps == 25);
    return (rb == 5 && steps == 7) || (rb == 4 && steps == 13) || (rb == 5 && steps == 25);
}

size_t iter_split_res_lds(int form, int rb, int steps) {
    const int lds_steps = res_k_streamed(steps) ? 2 : steps;
    return (size_t)lds_steps * rb * split_terms(form) * 1024 + (size_t)rb * 16 * 4;
}

template <int RBW, int FMT, int NT, int STEPS>
static hipError_t iter_res_one(bool launch, int waves, int grid, size_t lds, hipStream_t st, const BatchProblem* probs,
                               const BatchItem* items, const int* meta, unsigned* counters, int parity) {
    constexpr bool KS = res_k_streamed(STEPS);
    auto kern = waves == 4 ? grouped_iter_split_res_kernel<RBW, FMT, NT, STEPS, 4, KS>
                           : grouped_iter_split_res_kernel<RBW, FMT, NT, STEPS, kResWaves, KS>;
    if (!launch)
        return hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * waves), lds, st, probs, items, meta, counters, parity);
    return hipGetLastError();
}

// launch = false: only sets the kernel's dynamic-LDS limit (call once, outside graph capture).
hipError_t launch_grouped_iter_split_res(bool launch, int form, int rb, int steps, int waves, int grid,
                                         const BatchProblem* d_probs, const BatchItem* d_items, const int* d_meta,
                                         unsigned* d_counters, int parity, hipStream_t st) {
    const int fmt = split_fmt(form), nt = split_terms(form);
    if (waves != 4 && waves != kResWaves) return hipErrorInvalidValue;
    const size_t lds = iter_split_res_lds(form, rb, steps);
#define RSH_IR(RBW, F, T, S)                                                                                       if (fmt == F && nt == T && rb == RBW && steps == S)                                                                return iter_res_one<RBW, F, T, S>(launch, waves, grid, lds, st, d_probs, d_items, d_meta, d_counters, parity);
    RSH_IR(5, 0, 2, 7) RSH_IR(5, 0, 2, 13) RSH_IR(5, 0, 2, 25)
    RSH_IR(5, 1, 2, 7) RSH_IR(5, 1, 2, 13) RSH_IR(5, 1, 2, 25)
    RSH_IR(5, 0, 3, 7) RSH_IR(4, 0, 3, 13) RSH_IR(5, 0, 3, 25)
#undef RSH_IR
    return hipErrorInvalidValue;
}

int iter_split_res_chunk_cells() { return kIterResSub * kResChunkTiles * 16; }

}  // namespace rsh

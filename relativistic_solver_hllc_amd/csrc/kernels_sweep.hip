// Transport sweep (short characteristics) on the GPU: every direction walks its
// cells in upwind order and updates the intensities on the cells' out faces.
// Replaces the direction loop of MPI_CalculateIllumAsync
// (solve_module/illum/Illum_async.cpp:725-784) with CalculateIllumeOnInnerFace
// (:172-199), BoundaryConditions (solve_module/illum/illum_utils.cpp:407-471),
// GetCurIllum (:259-430, classes 0, 1, 2, 3, 10) and ReCalcIllum (:472-510).
//
// Parallel form. The reference runs one direction per OpenMP thread, each a
// sequential walk of sorted_id_cell. Here one workgroup owns one direction and
// walks its cells level by level: a cell's level is one more than the levels of
// the cells it reads from (the neighbours across its in faces) and of the cells
// that write a face it also writes, so all cells of a level are independent and
// run in parallel. Every pass (at most one cell per thread) ends at a barrier
// that orders only LDS: a pass also writes its face values into an LDS ring of
// the last four passes, and a cell whose upstream neighbour was swept one to
// three passes earlier reads it there (the host marks each such in face with
// the pass distance and the neighbour's index in its pass); older upstream
// values come from global memory, where they are complete by then: a wave's
// stores of pass q are done once it has waited for a load issued after them,
// which it does in pass q + 1, before the barrier that ends pass q + 1, and
// such a value is read by pass q + 4 or later, whose loads are issued during
// pass q + 3, after the barrier that ends pass q + 2. The full fence (store
// completion) only precedes the residual pass.
//
// Face values. The reference keeps three node values per face and only ever
// uses their mean ((c0 + c1 + c2) / 3, both when a downstream cell reads its in
// face and when ReCalcIllum stores the face intensity), so this kernel stores
// that mean, computed in the same order, directly in the new illum array
// [M][N][4] (the ReCalcIllum output layout). The reference has the upstream
// cell write into the downstream cell's in-face entry; here each cell writes
// only its own four entries, as one 32-byte store (out faces computed, in faces
// copied from the upstream cell's own entry, which it reads anyway, or set by
// the boundary condition) - no partial-line writes into neighbours. The one
// exception is a face out on both sides (|n . omega| <= eps): the reference's
// later writer sets both entries, so the later cell also writes its
// neighbour's. The residual max |old - new| / new follows at the end of the
// direction, with the reference's comparison (a NaN ratio never wins).
//
// Data layout and schedule: per direction, the cells sorted by level (by cell
// number inside a level), each position's cell, face states, in-face ids of its
// records and pairs in one 32-byte record, and its up to nine s values in a
// fixed 80-byte slot, all in position order. A level is cut into passes of one
// cell per thread; a pass's position data is fetched a pass ahead, and the
// by-cell global reads (S, upstream entries older than the ring) of the next
// pass are issued before the current pass's stores and consumed after its
// barrier, together with that pass's ring reads. Threads past a pass's end
// repeat its first cell (identical values), so no branch surrounds a load and
// the wait counters stay exact. A face out on both sides whose neighbour was
// swept in the pass just before is the one write that the LDS-only barrier
// would not order: the host flags such passes, and the pass before completes
// its stores ahead of the barrier.
//
// Measured (196,608 cells x 400 directions): 5.9 ms per sweep (6.9 ms with
// global memory and a store-completing barrier between levels), of which
// ~0.8 ms the residual pass; the FP64 arithmetic class barely matters (pure
// transport, class 3, is 4 % faster than class 1), L2 misses 61 % of requests
// (the per-cell reads of S and the level's neighbour entries are scattered).
#include "rsh_internal.hpp"

namespace rsh {

namespace {

// BoundaryConditions, non-Sphere build: OutSource (-3) -> 1, Free (-1),
// Lock (-2), InnerSource (-4) -> 0 (types are validated on the host).
__device__ inline double boundary_value(int type_bound) { return type_bound == -3 ? 1.0 : 0.0; }

// GetCurIllum of one class, in the reference's expression order.
template <int CLS>
__device__ inline double cur_illum(bool in_source, double s, double I_0, double S) {
    double Q, alpha, betta;
    if constexpr (CLS == 2) {
        const double ss = s * 388189 * 1e5;
        alpha = 1;
        Q = 1;
        const double I = I_0 * exp(-alpha * ss) + Q * (1 - exp(-alpha * ss)) / alpha;
        return I < 0 ? 0 : I;
    } else {
        if constexpr (CLS == 0) {
            // |x - (0.5, 0, 0)| < 0.05 at the node, evaluated on the host
            Q = 0; alpha = 2; betta = 1;
            if (in_source) { Q = 0; alpha = 0.5; betta = 0.5; }
        } else if constexpr (CLS == 1) {
            Q = 1; alpha = 1; betta = alpha / 2;
        } else if constexpr (CLS == 3) {
            alpha = 0; betta = alpha / 2; Q = 0;
        } else {  // 10
            Q = 0; alpha = 0.5; betta = 100;
        }
        const double k = alpha + betta;
        double I;
        if (k > 1e-10) I = (exp(-k * s) * (I_0 * k + (exp(k * s) - 1) * (Q + S * betta))) / k;
        else I = (1 - s * k) * (I_0 + s * (Q + S * betta));
        return I < 0 ? 0 : I;
    }
}

// Inputs of one pass that do not depend on earlier results (fetched a pass
// ahead, while the current pass computes).
struct PassIn {
    SweepPos p;
    double s[9];
    int cnt;  // cells in the pass | (complete the previous pass's stores first) << 31
};

// The pass table is read through the scalar cache (constant address space,
// uniform index): its loads retire on their own counter, so reading the next
// pass's first position never waits for the vector loads and stores in flight.
typedef const int32_t __attribute__((address_space(4)))* const_i32_ptr;

// Global-address-space view of a pointer read from the direction table, so
// that its loads are vector-memory loads (a generic "flat" load also counts on
// the scalar-load counter, and waiting for a scalar load would then wait for it).
template <typename T>
__device__ inline const __attribute__((address_space(1))) T* gptr(const T* p) {
    return (const __attribute__((address_space(1))) T*)p;
}

__device__ inline void load_pass(const SweepDirDev& d, int pi, PassIn& x) {
    typedef double v2d __attribute__((ext_vector_type(2)));
    const const_i32_ptr pt = (const_i32_ptr)d.pass;
    const int b = pt[2 * pi];
    x.cnt = pt[2 * pi + 1];
    const int i = b + ((int)threadIdx.x < (x.cnt & 0x7fffffff) ? (int)threadIdx.x : 0);
    const auto* pp = gptr(d.pos) + i;
    x.p.cell = pp->cell;
    x.p.meta = pp->meta;
#pragma unroll
    for (int f = 0; f < 4; f++) x.p.pr[f] = pp->pr[f];
    x.p.ring[0] = pp->ring[0];
    x.p.ring[1] = pp->ring[1];
    // the nine s values only: a register loaded with the slot's unused tenth
    // value and then reused for a temporary would wait for the newest load
    static_assert(kSweepRecStride == 10, "nine s values in a 10-double slot");
    const auto* sp = gptr(reinterpret_cast<const v2d*>(d.s9 + (int64_t)i * kSweepRecStride));
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const v2d t = sp[q];
        x.s[2 * q] = t.x;
        x.s[2 * q + 1] = t.y;
    }
    x.s[8] = gptr(d.s9)[(int64_t)i * kSweepRecStride + 8];
}

// ReCalcIllum's per-entry term with the reference's comparison (a NaN ratio never wins).
__device__ inline void residual_term(double old, double cur, double& loc) {
    const double b = fabs(old - cur) / cur;
    if (b > loc) loc = b;
}

// Values a pass reads by cell from global memory: S and, for its in faces
// whose upstream cell was swept at least kSweepRingDelta + 1 passes earlier,
// that cell's entry (final by then, see the header). Issued for the next pass
// before the current pass's stores (a load issued after a store could only be
// waited for together with it) and consumed after the barrier, so their
// latency hides behind the stores and the barrier.
struct PassDyn {
    double S;
    double fin[4];
};

// LDS ring of the last kSweepRing passes' face values: [slot][face][index].
typedef double SweepRing[kSweepRing][4][kSweepBlock];

// Ring code of face f: (pass distance << 9) | upstream index in its pass, or 0
// when the face is not read from the ring.
__device__ inline uint32_t ring_code(const SweepPos& p, int f) {
    return (p.ring[f >> 1] >> (16 * (f & 1))) & 0xffffu;
}

__device__ inline void load_upstream(const PassIn& x, const double* __restrict__ Sk, const double* fv, PassDyn& y) {
    // Unconditional loads at always-valid addresses (no branch around a load,
    // so the wait counters stay exact): a face whose value comes from the ring,
    // or that needs none, reads the direction's first entry, the same address
    // in every lane, so that such lanes add no line requests (the value is
    // ignored).
    y.S = Sk[x.p.cell];
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const int p = x.p.pr[f];
        const bool up = ((x.p.meta >> f) & 1) && p >= 0;
        y.fin[f] = fv[up && ring_code(x.p, f) == 0 ? (int64_t)p : 0];  // upstream cell's own entry
    }
}

// Upstream values of the in faces of pass rp: from the ring for faces with a
// ring code (slot of an upstream pass = its index mod kSweepRing), else the
// global values fetched by load_upstream. Read after the barrier that ends
// pass rp - 1, before pass rp writes its own slot. All four ring reads are
// issued unconditionally, before any of them is used (a face without a ring
// code reads slot 0, face 0, index 0 and ignores it), so their latencies
// overlap.
__device__ inline void upstream_values(const SweepPos& p, const PassDyn& y, const SweepRing& ring, int rp,
                                       double (&v)[4]) {
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const uint32_t code = ring_code(p, f);
        const bool in_ring = code != 0 && ((p.meta >> f) & 1);  // out faces: class 0 flags
        const int slot = in_ring ? (rp - (int)(code >> 9)) & (kSweepRing - 1) : 0;
        v[f] = ring[slot][in_ring ? (p.pr[f] & 3) : 0][in_ring ? (int)(code & 511u) : 0];
    }
#pragma unroll
    for (int f = 0; f < 4; f++) v[f] = ring_code(p, f) != 0 ? v[f] : y.fin[f];
}

// Barrier that orders only LDS accesses: global stores may stay in flight
// across it (an upstream value read from global memory was stored at least
// kSweepRingDelta + 1 passes earlier, and every wave has waited for a load
// issued after its stores of pass q - 2 before it reaches the barrier of pass q).
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int CLS>
__global__ __launch_bounds__(kSweepBlock) void sweep_kernel(const SweepDirDev* __restrict__ dirs, int64_t n,
                                                            const double* __restrict__ S, int64_t ld,
                                                            const double* __restrict__ old_illum, double* new_illum,
                                                            double* __restrict__ norm) {
    typedef double v2d __attribute__((ext_vector_type(2)));
    const int k = blockIdx.x;
    const SweepDirDev d = dirs[k];
    double* fv = new_illum + (int64_t)k * n * 4;
    const double* __restrict__ Sk = S + (int64_t)k * ld;
    const double* __restrict__ ok = old_illum ? old_illum + (int64_t)k * n * 4 : nullptr;
    extern __shared__ double ring_mem[];
    SweepRing& ring = *reinterpret_cast<SweepRing*>(ring_mem);
    double loc = -1;
    // One pass: compute from (cur, dc) and fetch the next pass into (nxt, dn).
    // The loop below is unrolled by two with the buffers swapping roles, so no
    // register copy of a value still in flight (which would wait for it) occurs.
    auto step = [&](PassIn& cur, PassDyn& dc, PassIn& nxt, PassDyn& dn, int pi) {
        // Threads past the pass's cell count repeat its first cell: same inputs,
        // same values written, same residual terms (max is idempotent), so no
        // branch is needed around the loads and stores.
        const int64_t cell = cur.p.cell;
        const uint32_t meta = cur.p.meta;
        load_pass(d, pi + 1 < d.passes ? pi + 1 : pi, nxt);  // the last pass re-reads itself (unused)
        double own[4];
        {
            double ix[4];
            upstream_values(cur.p, dc, ring, pi, ix);
#pragma unroll
            for (int f = 0; f < 4; f++) {
                const int p = cur.p.pr[f];
                // in faces: BoundaryConditions value (its 3-node mean is the value itself for 0 and 1)
                // or the upstream cell's entry
                ix[f] = p < 0 ? boundary_value(p) : ix[f];
                own[f] = ix[f];
            }
            uint32_t left = ~meta & 15u;
#pragma unroll
            for (int r = 0; r < 3; r++) {
                if (!left) break;
                const int f = __builtin_ctz(left);
                left &= left - 1;
                double I[3];
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const int g = (meta >> (8 + 2 * (3 * r + j))) & 3;
                    const double ix0 = g == 0 ? ix[0] : g == 1 ? ix[1] : g == 2 ? ix[2] : ix[3];
                    const bool in_source = (ring_code(cur.p, f) >> j) & 1;  // out face: class 0 flags
                    I[j] = cur_illum<CLS>(in_source, cur.s[3 * r + j], ix0, dc.S);
                }
                const double v = (I[0] + I[1] + I[2]) / 3;
                own[0] = f == 0 ? v : own[0];
                own[1] = f == 1 ? v : own[1];
                own[2] = f == 2 ? v : own[2];
                own[3] = f == 3 ? v : own[3];
            }
        }
        load_upstream(nxt, Sk, fv, dn);
        {
            v2d* dst = reinterpret_cast<v2d*>(fv + cell * 4);
            dst[0] = v2d{own[0], own[1]};
            dst[1] = v2d{own[2], own[3]};
#pragma unroll
            for (int f = 0; f < 4; f++) {
                // out on both sides, this cell later: the neighbour's entry too
                // (rare; skipped by the whole wave otherwise - an unconditional
                // store into the cell's own entry measured 17 % slower)
                if ((meta >> (4 + f)) & 1) fv[cur.p.pr[f]] = own[f];
            }
            // this pass's values into its ring slot (threads past the pass's end
            // write their own, never-read indices)
#pragma unroll
            for (int f = 0; f < 4; f++) ring[pi & (kSweepRing - 1)][f][threadIdx.x] = own[f];
        }
        // the next pass rewrites an entry stored in this one (a face out on
        // both sides, flagged by the host): complete this pass's stores first,
        // so that the later write lands last (uniform branch, rare)
        if (nxt.cnt < 0) __builtin_amdgcn_s_waitcnt(0);
        // every pass ends at a barrier: no wave runs more than a pass ahead, so
        // a ring slot is rewritten only after every read of its previous pass
        lds_barrier();
    };
    PassIn pa, pb;
    PassDyn da, db;
    if (d.passes > 0) {
        load_pass(d, 0, pa);
        load_upstream(pa, Sk, fv, da);  // level 0: no upstream values
    }
    for (int pi = 0; pi < d.passes; pi += 2) {
        step(pa, da, pb, db, pi);
        if (pi + 1 < d.passes) step(pb, db, pa, da, pi + 1);
    }
    // ReCalcIllum's residual over the direction's final entries (a separate
    // streaming pass measured faster than terms accumulated during the sweep).
    if (!ok) return;
    __syncthreads();
    // one cell's four entries per thread as 16-byte loads (both arrays are
    // 16-byte aligned, checked on the host; 8-byte loads measured 0.42 ms
    // slower): the max is independent of the order, so this changes no result
    const v2d* __restrict__ ok2 = reinterpret_cast<const v2d*>(ok);
    const v2d* fv2 = reinterpret_cast<const v2d*>(fv);
    for (int64_t c = threadIdx.x; c < n; c += kSweepBlock) {
        const v2d a0 = ok2[2 * c], a1 = ok2[2 * c + 1], b0 = fv2[2 * c], b1 = fv2[2 * c + 1];
        residual_term(a0.x, b0.x, loc);
        residual_term(a0.y, b0.y, loc);
        residual_term(a1.x, b1.x, loc);
        residual_term(a1.y, b1.y, loc);
    }
    __shared__ double red[kSweepBlock];
    red[threadIdx.x] = loc;
    __syncthreads();
    for (int w = kSweepBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w && red[threadIdx.x + w] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) norm[k] = red[0];
}

}  // namespace

hipError_t launch_sweep(int cls, const SweepDirDev* d_dirs, int m, int64_t n,
                        const double* d_S, int64_t ld, const double* d_old, double* d_new, double* d_norm,
                        hipStream_t st) {
    if (m <= 0 || n <= 0) return hipSuccess;
    const dim3 grid((unsigned)m), block(kSweepBlock);
    constexpr int kRingBytes = (int)sizeof(SweepRing);
#define RSH_SWEEP(C)                                                                                             \
    do {                                                                                                         \
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(sweep_kernel<C>),                      \
                                           hipFuncAttributeMaxDynamicSharedMemorySize, kRingBytes);             \
        if (e != hipSuccess) return e;                                                                           \
        hipLaunchKernelGGL(sweep_kernel<C>, grid, block, kRingBytes, st, d_dirs, n, d_S, ld, d_old,              \
                           d_new, d_norm);                                                                       \
    } while (0)
    switch (cls) {
        case 0: RSH_SWEEP(0); break;
        case 1: RSH_SWEEP(1); break;
        case 2: RSH_SWEEP(2); break;
        case 3: RSH_SWEEP(3); break;
        case 10: RSH_SWEEP(10); break;
        default: return hipErrorInvalidValue;
    }
#undef RSH_SWEEP
    return hipGetLastError();
}

}  // namespace rsh

// Split-precision scattering integral on CDNA4 (gfx950).
//
// Operands are stored as sums of 16-bit parts, x ~ x_hi + x_lo (+ x_mid for the
// three-term BF16 form), and contracted on v_mfma_f32_16x16x32_{bf16,f16} with
// FP32 accumulation:
//     acc_main  += A_hi * B_hi
//     acc_cross += A_hi * B_lo + A_lo * B_hi      (+ hi*lo, lo*hi, mid*mid for x3)
// The cross terms are kept in their own accumulator so that their rounding is
// relative to their own (2^-8 smaller) magnitude; the epilogue adds the two in
// FP64 and removes the power-of-two operand scales exactly.
//
// Scaling: every cell c carries 2^e_c and every output row k carries 2^r_k,
// chosen so that the largest face value of the cell / phase entry of the row
// lands near 2^12. This keeps FP16 hi/lo parts inside the normal range for
// log-normal intensities; for BF16 it is harmless (exact).
//
// Lane layouts (gfx950, measured with one-hot operands):
//   A (16x32): lane l, element e -> A[l % 16][8 (l / 16) + e]
//   B (32x16): lane l, element e -> B[8 (l / 16) + e][l % 16]
//   D (16x16): lane l, item r    -> D[4 (l / 16) + r][l % 16]
//
// Packed intensity state ("split state"), MFMA-native:
//   state[((ct * steps + s) * V + v) * 64 + lane] = 16 bytes (8 parts of 16 bits)
//   ct = cell tile (16 cells), s = K-step (32 directions), v = part * F + face,
//   lane l, element e <-> cell 16 ct + l % 16, direction 32 s + 8 (l / 16) + e.
// One (ct, s, v) vector is 1 KB contiguous: a single coalesced wave load.
//
// Phase panel (A operand), per pass of RB row blocks:
//   panel[(((pass * steps + s) * RB + rb) * NT + t) * 64 + lane] = 16 bytes
//   lane l, element e <-> row k0 + 16 (pass RB + rb) + l % 16, direction 32 s + 8 (l / 16) + e.
#include "mfma_split_core.hpp"

namespace rsh {

// ------------------------------------------------------------------ cell scales
// inv_scale[c] = 2^-e_c with e_c = kTargetExp - 1 - ilogb(max_{j,f} |I[j][c][f]|);
// an optional second array of the same layout (the source term of an iteration,
// which bounds every later iterate together with I) joins the maximum.
__global__ __launch_bounds__(256) void cell_scale_kernel(const double* __restrict__ illum,
                                                         const double* __restrict__ extra, int64_t n, int m,
                                                         int faces, float* __restrict__ inv_scale) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double mx = 0.0;
    const int64_t stride = n * faces;
    const double* p = illum + c * faces;
    for (int j = 0; j < m; j++) {
        for (int f = 0; f < faces; f++) mx = fmax(mx, fabs(p[f]));
        p += stride;
    }
    if (extra) {
        const double* q = extra + c * faces;
        for (int j = 0; j < m; j++) {
            for (int f = 0; f < faces; f++) mx = fmax(mx, fabs(q[f]));
            q += stride;
        }
    }
    const int e = mx > 0.0 ? (kTargetExp - 1 - ilogb(mx)) : 0;
    inv_scale[c] = ldexpf(1.0f, -e);
}

// ------------------------------------------------------------------ state pack
// Thread = (ct, s, lane). Reads 8 directions x F faces of one cell, writes NT*F vectors.
template <int FMT, int NT, int F>
__global__ __launch_bounds__(256) void pack_state_kernel(const double* __restrict__ illum, int64_t n, int m,
                                                         int steps, const float* __restrict__ inv_scale,
                                                         v4u* __restrict__ state, int64_t total) {
    typedef typename Fmt<FMT>::T T;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int lane = (int)(idx & 63);
    const int64_t frag = idx >> 6;
    const int s = (int)(frag % steps);
    const int64_t ct = frag / steps;
    const int64_t cell = ct * 16 + (lane & 15);
    const int j0 = 32 * s + 8 * (lane >> 4);
    const bool cvalid = cell < n;
    const double scale = cvalid ? 1.0 / (double)inv_scale[cell] : 0.0;
    T buf[NT][F][8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int j = j0 + e;
        const bool ok = cvalid && j < m;
        const double* p = illum + ((int64_t)(ok ? j : 0) * n + (ok ? cell : 0)) * F;
#pragma unroll
        for (int f = 0; f < F; f++) {
            const float x = ok ? (float)(p[f] * scale) : 0.0f;
            T parts[NT];
            split_parts<FMT, NT>(x, parts);
#pragma unroll
            for (int t = 0; t < NT; t++) buf[t][f][e] = parts[t];
        }
    }
    v4u* out = state + (frag * (NT * F)) * 64 + lane;
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int f = 0; f < F; f++) out[(t * F + f) * 64] = __builtin_bit_cast(v4u, buf[t][f]);
}

// Face-averaged pack: reads the 4 faces of (direction, cell), averages them in
// FP64 (the sum of four doubles, times 1/4 exactly) and writes one cell-level
// vector per part.
template <int FMT, int NT>
__global__ __launch_bounds__(256) void pack_state_avg_kernel(const double* __restrict__ illum, int64_t n, int m,
                                                             int steps, const float* __restrict__ inv_scale,
                                                             v4u* __restrict__ state, int64_t total) {
    typedef typename Fmt<FMT>::T T;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int lane = (int)(idx & 63);
    const int64_t frag = idx >> 6;
    const int s = (int)(frag % steps);
    const int64_t ct = frag / steps;
    const int64_t cell = ct * 16 + (lane & 15);
    const int j0 = 32 * s + 8 * (lane >> 4);
    const bool cvalid = cell < n;
    const double scale = cvalid ? 1.0 / (double)inv_scale[cell] : 0.0;
    T buf[NT][8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int j = j0 + e;
        const bool ok = cvalid && j < m;
        const double* p = illum + ((int64_t)(ok ? j : 0) * n + (ok ? cell : 0)) * 4;
        const float x = ok ? (float)((((p[0] + p[1]) + p[2]) + p[3]) * 0.25 * scale) : 0.0f;
        T parts[NT];
        split_parts<FMT, NT>(x, parts);
#pragma unroll
        for (int t = 0; t < NT; t++) buf[t][e] = parts[t];
    }
    v4u* out = state + (frag * NT) * 64 + lane;
#pragma unroll
    for (int t = 0; t < NT; t++) out[t * 64] = __builtin_bit_cast(v4u, buf[t]);
}

// Fused cell scale + cell-level pack, one pass over the input: workgroup = one
// 16-cell tile. Every (direction, cell) of the tile is read once (F faces as
// 16-byte loads: 16 cells x F faces are contiguous per direction); its
// face average ((f0 + f1) + f2) + f3) * 1/4 in FP64 (F = 4; the value itself
// for F = 1) goes to LDS, its largest |face| (and |extra| when given, F = 1)
// into the cell's maximum. Then the cell scale is set exactly as
// cell_scale_kernel sets it and the tile's state vectors are written from LDS
// as pack_state_avg_kernel / pack_state_kernel<F = 1> write them: the same bits
// as those two passes, with the input read once instead of twice.
template <int FMT, int NT, int F>
__global__ __launch_bounds__(256) void pack_fused_kernel(const double* __restrict__ illum,
                                                         const double* __restrict__ extra, int64_t n, int m,
                                                         int steps, float* __restrict__ inv_scale,
                                                         v4u* __restrict__ state) {
    typedef typename Fmt<FMT>::T T;
    extern __shared__ double avg[];  // [steps * 32][16]: face averages, zero past m
    __shared__ double cmax[4][16];
    const int t = threadIdx.x;
    const int cl = t & 15;
    const int64_t ct = blockIdx.x;
    const int64_t cell = ct * 16 + cl;
    const bool cvalid = cell < n;
    const int64_t cell_c = cvalid ? cell : n - 1;  // tail threads re-read a valid cell, results unused
    const int jrows = steps * 32;
    double mx = 0.0;
    // Phase 1: thread t reads directions j = t / 16 + 16 i of cell cl.
    constexpr int UNR = 5;
    for (int j0 = t >> 4; j0 < jrows; j0 += 16 * UNR) {
        double a[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const int j = j0 + 16 * u;
            a[u] = 0.0;
            if (j < m) {
                const double* p = illum + ((int64_t)j * n + cell_c) * F;
                if constexpr (F == 4) {
                    const double2 lo = *reinterpret_cast<const double2*>(p);
                    const double2 hi = *reinterpret_cast<const double2*>(p + 2);
                    a[u] = (((lo.x + lo.y) + hi.x) + hi.y) * 0.25;
                    mx = fmax(mx, fmax(fmax(fabs(lo.x), fabs(lo.y)), fmax(fabs(hi.x), fabs(hi.y))));
                } else {
                    a[u] = p[0];
                    mx = fmax(mx, fabs(p[0]));
                    if (extra) mx = fmax(mx, fabs(extra[(int64_t)j * n + cell_c]));
                }
            }
        }
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const int j = j0 + 16 * u;
            if (j < jrows) avg[j * 16 + cl] = cvalid ? a[u] : 0.0;
        }
    }
    // Phase 2: the cell's maximum over the 16 threads that read it.
    mx = fmax(mx, __shfl_xor(mx, 16, 64));
    mx = fmax(mx, __shfl_xor(mx, 32, 64));
    if ((t & 63) < 16) cmax[t >> 6][cl] = mx;
    __syncthreads();
    const double cm = fmax(fmax(cmax[0][cl], cmax[1][cl]), fmax(cmax[2][cl], cmax[3][cl]));
    const int e = cm > 0.0 ? (kTargetExp - 1 - ilogb(cm)) : 0;
    const float inv = ldexpf(1.0f, -e);
    if (t < 16 && cvalid) inv_scale[cell] = inv;
    const double scale = cvalid ? 1.0 / (double)inv : 0.0;
    // Phase 3: state vectors of the tile: (K-step s, lane) slots, lane l holding
    // directions 32 s + 8 (l / 16) + e of cell l % 16.
    v4u* out = state + (int64_t)ct * steps * NT * 64;
    for (int slot = t; slot < steps * 64; slot += 256) {
        const int s = slot >> 6, lane = slot & 63;
        const int j0 = 32 * s + 8 * (lane >> 4);
        // lane % 16 == t % 16: this thread's own cell and scale
        T buf[NT][8];
#pragma unroll
        for (int e8 = 0; e8 < 8; e8++) {
            const float x = (float)(avg[(j0 + e8) * 16 + (lane & 15)] * scale);
            T parts[NT];
            split_parts<FMT, NT>(x, parts);
#pragma unroll
            for (int q = 0; q < NT; q++) buf[q][e8] = parts[q];
        }
#pragma unroll
        for (int q = 0; q < NT; q++) out[(s * NT + q) * 64 + lane] = __builtin_bit_cast(v4u, buf[q]);
    }
}

// Wide form of the fused pack for M <= 32 kPackWideMaxSteps: a workgroup of
// 16 CPB threads packs CPB / 16 tiles (CPB cells) and keeps the face averages
// in registers instead of LDS. Thread slot u = t + 16 CPB i stands for
// (K-step s = u / (4 CPB), row group g = (u / CPB) % 4, cell u % CPB), i.e.
// the 8 directions 32 s + 8 g + e of one cell that one state vector lane
// holds: the thread reads exactly the values it packs, so no transposition
// through LDS is needed, and a load instruction reads CPB consecutive cells of
// a direction row (32 CPB contiguous bytes; 512 in the 16-cell LDS form).
// Same arithmetic, same bits as pack_fused_kernel.
constexpr int kPackWideMaxSteps = 16;  // 4 slots per thread: 8 x 4 doubles in registers
constexpr int kPackCells = 64;  // cells per workgroup (32 measured the same, 16 2 % slower)
template <int FMT, int NT, int F, int MAXS, int CPB>
__global__ __launch_bounds__(16 * CPB) void pack_wide_kernel(const double* __restrict__ illum,
                                                             const double* __restrict__ extra, int64_t n, int m,
                                                             int steps, float* __restrict__ inv_scale,
                                                             v4u* __restrict__ state) {
    typedef typename Fmt<FMT>::T T;
    constexpr int THREADS = 16 * CPB;
    __shared__ double cmax[16][CPB];
    const int t = threadIdx.x;
    const int cc = t % CPB;  // the thread's cell in the group (every slot of a thread has the same cell)
    const int64_t tiles = (n + 15) / 16;
    const int64_t cell = (int64_t)blockIdx.x * CPB + cc;
    const bool cvalid = cell < n;
    const int64_t cell_c = cvalid ? cell : n - 1;  // invalid cells re-read a valid one, results unused
    const int nslots = steps * 4 * CPB;
    double avg[MAXS][8];
    double mx = 0.0;
#pragma unroll
    for (int i = 0; i < MAXS; i++) {
        const int u = t + THREADS * i;
        const int j0 = 32 * (u / (4 * CPB)) + 8 * ((u / CPB) % 4);
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int j = j0 + e;
            double a = 0.0;
            if (u < nslots && j < m) {
                const double* p = illum + ((int64_t)j * n + cell_c) * F;
                if constexpr (F == 4) {
                    const double2 lo = *reinterpret_cast<const double2*>(p);
                    const double2 hi = *reinterpret_cast<const double2*>(p + 2);
                    a = (((lo.x + lo.y) + hi.x) + hi.y) * 0.25;
                    mx = fmax(mx, fmax(fmax(fabs(lo.x), fabs(lo.y)), fmax(fabs(hi.x), fabs(hi.y))));
                } else {
                    a = p[0];
                    mx = fmax(mx, fabs(p[0]));
                    if (extra) mx = fmax(mx, fabs(extra[(int64_t)j * n + cell_c]));
                }
            }
            avg[i][e] = cvalid ? a : 0.0;
        }
    }
    // the cell's maximum over the 16 threads that read it
    cmax[t / CPB][cc] = mx;
    __syncthreads();
    double cm = 0.0;
#pragma unroll
    for (int w = 0; w < 16; w++) cm = fmax(cm, cmax[w][cc]);
    const int e = cm > 0.0 ? (kTargetExp - 1 - ilogb(cm)) : 0;
    const float inv = ldexpf(1.0f, -e);
    if (t < CPB && cvalid) inv_scale[cell] = inv;
    const double scale = cvalid ? 1.0 / (double)inv : 0.0;
    const int64_t tile = (int64_t)blockIdx.x * (CPB / 16) + cc / 16;
    if (tile >= tiles) return;  // no barrier below
    v4u* out = state + tile * steps * NT * 64;
#pragma unroll
    for (int i = 0; i < MAXS; i++) {
        const int u = t + THREADS * i;
        if (u >= nslots) break;
        const int sk = u / (4 * CPB), lane = ((u / CPB) % 4) * 16 + cc % 16;
        T buf[NT][8];
#pragma unroll
        for (int e8 = 0; e8 < 8; e8++) {
            const float x = (float)(avg[i][e8] * scale);
            T parts[NT];
            split_parts<FMT, NT>(x, parts);
#pragma unroll
            for (int q = 0; q < NT; q++) buf[q][e8] = parts[q];
        }
#pragma unroll
        for (int q = 0; q < NT; q++) out[(sk * NT + q) * 64 + lane] = __builtin_bit_cast(v4u, buf[q]);
    }
}

// ------------------------------------------------------------------ phase panel
// One block per output row: row scale from the row maximum, then the fragments.
template <int FMT, int NT>
__global__ __launch_bounds__(256) void split_panel_kernel(const double* __restrict__ P, int m, int k0, int k1,
                                                          double factor, int rb_per_pass, int steps,
                                                          v4u* __restrict__ panel, float* __restrict__ row_inv) {
    typedef typename Fmt<FMT>::T T;
    const int r = blockIdx.x;  // local row, padded range [0, passes * RB * 16)
    const int k = k0 + r;
    const bool valid = k < k1;
    __shared__ double red[256];
    double mx = 0.0;
    if (valid)
        for (int j = threadIdx.x; j < m; j += blockDim.x) mx = fmax(mx, fabs(P[(int64_t)k * m + j] * factor));
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + st]);
        __syncthreads();
    }
    const int e = red[0] > 0.0 ? (kTargetExp - 1 - ilogb(red[0])) : 0;
    const double scale = ldexp(1.0, e);
    if (threadIdx.x == 0) row_inv[r] = ldexpf(1.0f, -e);
    const int rb_g = r / 16, rl = r % 16;
    const int pass = rb_g / rb_per_pass, rb = rb_g % rb_per_pass;
    // each (s, lane-group g) slot covers directions 32 s + 8 g + e
    for (int slot = threadIdx.x; slot < steps * 4; slot += blockDim.x) {
        const int s = slot >> 2, g = slot & 3;
        T buf[NT][8];
#pragma unroll
        for (int e8 = 0; e8 < 8; e8++) {
            const int j = 32 * s + 8 * g + e8;
            const float x = (valid && j < m) ? (float)(P[(int64_t)k * m + j] * factor * scale) : 0.0f;
            T parts[NT];
            split_parts<FMT, NT>(x, parts);
#pragma unroll
            for (int t = 0; t < NT; t++) buf[t][e8] = parts[t];
        }
        const int lane = g * 16 + rl;
        v4u* out = panel + ((((int64_t)pass * steps + s) * rb_per_pass + rb) * NT) * 64 + lane;
#pragma unroll
        for (int t = 0; t < NT; t++) out[t * 64] = __builtin_bit_cast(v4u, buf[t]);
    }
}

// ------------------------------------------------------------------ contraction
// W waves per workgroup; RH waves share one 16-cell tile (each owns RB / RH of
// the row blocks), so a workgroup covers 16 W / RH cells.
template <int RB, int W, int FMT, int NT, int F, int RH>
__global__ __launch_bounds__(64 * W) void split_mfma_kernel(const v4u* __restrict__ state, int64_t n, int steps,
                                                            const float* __restrict__ cell_inv,
                                                            const v4u* __restrict__ panel,
                                                            const float* __restrict__ row_inv, void* __restrict__ out,
                                                            int64_t ld, int rows, int out_f32) {
    constexpr int RBW = SplitTileShape<RB, NT, RH>::RBW;
    __shared__ v4u lds[2][SplitTileShape<RB, NT>::UNITS];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int h = wave % RH;
    const int64_t ct = (int64_t)blockIdx.x * (W / RH) + wave / RH;
    const int64_t cell = ct * 16 + (lane & 15);
    const int rb0 = blockIdx.y * RB;
    v4f accm[RBW], accx[RBW];
    split_accumulate<RB, W, FMT, NT, F, RH>(state, n, steps,
                                            panel + (int64_t)blockIdx.y * steps * SplitTileShape<RB, NT>::UNITS, ct,
                                            h, lds, accm, accx);
    if (cell >= n) return;
    const double cinv = (double)cell_inv[cell];
    const int g = lane >> 4;
#pragma unroll
    for (int i = 0; i < RBW; i++) {
        const int rb = h * RBW + i;
        if (rb >= RB) break;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = (rb0 + rb) * 16 + 4 * g + r;
            if (row < rows) {
                const double v = split_value(accm[i][r], accx[i][r]) * (double)row_inv[row] * cinv;
                if (out_f32)
                    __builtin_nontemporal_store((float)v, static_cast<float*>(out) + (int64_t)row * ld + cell);
                else
                    __builtin_nontemporal_store(v, static_cast<double*>(out) + (int64_t)row * ld + cell);
            }
        }
    }
}

// ------------------------------------------------------------------ launchers

SplitPlan plan_split(int m, int rows) {
    SplitPlan pl{};
    const F64Plan fp = plan_f64(m, rows);  // same row-block / pass selection
    pl.rb_per_pass = fp.rb_per_pass;
    pl.passes = fp.passes;
    pl.rb_total = fp.rb_total;
    pl.steps = (m + 31) / 32;
    return pl;
}

int split_terms(int form) { return form == RSH_FORM_BF16X3 ? 3 : 2; }
int split_fmt(int form) { return form == RSH_FORM_FP16X2 ? 1 : 0; }

int64_t split_state_units(int form, int faces, int64_t n, int m) {
    const int64_t tiles = (n + 15) / 16;
    const int64_t steps = (m + 31) / 32;
    return tiles * steps * split_terms(form) * faces * 64;
}

int64_t split_panel_units(int form, const SplitPlan& pl) {
    return (int64_t)pl.passes * pl.steps * pl.rb_per_pass * split_terms(form) * 64;
}

hipError_t launch_cell_scale(const double* d_illum, const double* d_extra, int64_t n, int m, int faces, float* d_inv,
                             hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(cell_scale_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d_illum, d_extra, n, m,
                       faces, d_inv);
    return hipGetLastError();
}

hipError_t launch_pack_state(int form, const double* d_illum, int faces, int64_t n, int m, const float* d_inv,
                             void* d_state, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int steps = (m + 31) / 32;
    const int64_t total = (n + 15) / 16 * steps * 64;
    const dim3 grid((unsigned)((total + 255) / 256)), block(256);
    v4u* out = static_cast<v4u*>(d_state);
#define RSH_PACK(FMT, NT, F) hipLaunchKernelGGL((pack_state_kernel<FMT, NT, F>), grid, block, 0, st, d_illum, n, m, steps, d_inv, out, total)
    const int fmt = split_fmt(form), nt = split_terms(form);
    if (fmt == 0 && nt == 2) { if (faces == 4) RSH_PACK(0, 2, 4); else RSH_PACK(0, 2, 1); }
    else if (fmt == 1 && nt == 2) { if (faces == 4) RSH_PACK(1, 2, 4); else RSH_PACK(1, 2, 1); }
    else { if (faces == 4) RSH_PACK(0, 3, 4); else RSH_PACK(0, 3, 1); }
#undef RSH_PACK
    return hipGetLastError();
}

hipError_t launch_pack_state_avg(int form, const double* d_illum, int64_t n, int m, const float* d_inv,
                                 void* d_state, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int steps = (m + 31) / 32;
    const int64_t total = (n + 15) / 16 * steps * 64;
    const dim3 grid((unsigned)((total + 255) / 256)), block(256);
    v4u* out = static_cast<v4u*>(d_state);
    const int fmt = split_fmt(form), nt = split_terms(form);
    if (fmt == 0 && nt == 2)
        hipLaunchKernelGGL((pack_state_avg_kernel<0, 2>), grid, block, 0, st, d_illum, n, m, steps, d_inv, out, total);
    else if (fmt == 1 && nt == 2)
        hipLaunchKernelGGL((pack_state_avg_kernel<1, 2>), grid, block, 0, st, d_illum, n, m, steps, d_inv, out, total);
    else
        hipLaunchKernelGGL((pack_state_avg_kernel<0, 3>), grid, block, 0, st, d_illum, n, m, steps, d_inv, out, total);
    return hipGetLastError();
}

bool pack_fused_supported(int m) { return (size_t)((m + 31) / 32) * 32 * 16 * sizeof(double) <= 160 * 1024; }

hipError_t launch_pack_fused(int form, const double* d_illum, const double* d_extra, int faces, int64_t n, int m,
                             float* d_inv, void* d_state, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (!pack_fused_supported(m) || (faces != 1 && faces != 4) || (faces == 4 && d_extra))
        return hipErrorInvalidValue;
    const int steps = (m + 31) / 32;
    v4u* out = static_cast<v4u*>(d_state);
    const int fmt = split_fmt(form), nt = split_terms(form);
    if (steps <= kPackWideMaxSteps) {
        const dim3 wgrid((unsigned)((n + kPackCells - 1) / kPackCells)), wblock(16 * kPackCells);
#define RSH_PW(FMT, NT, F)                                                                                        \
    {                                                                                                             \
        hipLaunchKernelGGL((pack_wide_kernel<FMT, NT, F, kPackWideMaxSteps / 4, kPackCells>), wgrid, wblock, 0, st, \
                           d_illum, d_extra, n, m, steps, d_inv, out);                                            \
        return hipGetLastError();                                                                                 \
    }
        if (fmt == 0 && nt == 2) { if (faces == 4) RSH_PW(0, 2, 4) else RSH_PW(0, 2, 1) }
        if (fmt == 1 && nt == 2) { if (faces == 4) RSH_PW(1, 2, 4) else RSH_PW(1, 2, 1) }
        if (faces == 4) RSH_PW(0, 3, 4) else RSH_PW(0, 3, 1)
#undef RSH_PW
    }
    const size_t lds = (size_t)steps * 32 * 16 * sizeof(double);
    const dim3 grid((unsigned)((n + 15) / 16)), block(256);
#define RSH_PF(FMT, NT, F)                                                                                       \
    {                                                                                                            \
        auto k = pack_fused_kernel<FMT, NT, F>;                                                                  \
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),                                    \
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                \
        if (e != hipSuccess) return e;                                                                           \
        hipLaunchKernelGGL(k, grid, block, lds, st, d_illum, d_extra, n, m, steps, d_inv, out);                   \
        return hipGetLastError();                                                                                \
    }
    if (fmt == 0 && nt == 2) { if (faces == 4) RSH_PF(0, 2, 4) else RSH_PF(0, 2, 1) }
    if (fmt == 1 && nt == 2) { if (faces == 4) RSH_PF(1, 2, 4) else RSH_PF(1, 2, 1) }
    if (faces == 4) RSH_PF(0, 3, 4) else RSH_PF(0, 3, 1)
#undef RSH_PF
}

hipError_t launch_split_panel(int form, const double* d_P, int m, int k0, int k1, double factor, const SplitPlan& pl,
                              void* d_panel, float* d_row_inv, hipStream_t st) {
    const int rows_padded = pl.rb_total * 16;
    if (rows_padded == 0) return hipSuccess;
    v4u* out = static_cast<v4u*>(d_panel);
#define RSH_PANEL(FMT, NT) hipLaunchKernelGGL((split_panel_kernel<FMT, NT>), dim3(rows_padded), dim3(256), 0, st, d_P, m, k0, k1, factor, pl.rb_per_pass, pl.steps, out, d_row_inv)
    const int fmt = split_fmt(form), nt = split_terms(form);
    if (fmt == 0 && nt == 2) RSH_PANEL(0, 2);
    else if (fmt == 1) RSH_PANEL(1, 2);
    else RSH_PANEL(0, 3);
#undef RSH_PANEL
    return hipGetLastError();
}

template <int W, int FMT, int NT, int F, int RH>
static hipError_t launch_split_rb(int rb, dim3 grid, hipStream_t st, const v4u* state, int64_t n, int steps,
                                  const float* cell_inv, const v4u* panel, const float* row_inv, void* out,
                                  int64_t ld, int rows, int out_f32) {
#define RSH_RB_CASE(V)                                                                                             \
    case V:                                                                                                        \
        hipLaunchKernelGGL((split_mfma_kernel<V, W, FMT, NT, F, RH>), grid, dim3(64 * W), 0, st, state, n, steps, \
                           cell_inv, panel, row_inv, out, ld, rows, out_f32);                                      \
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

bool split_tile_supported(int waves, int row_split) {
    return (waves == 4 && row_split == 1) || (waves == 8 && row_split == 2);
}

hipError_t launch_split_mfma(int form, const void* d_state, int faces, int64_t n, const float* d_cell_inv,
                             const void* d_panel, const float* d_row_inv, const SplitPlan& pl, int rows, void* d_out,
                             int64_t ld, int out_f32, int waves, int row_split, hipStream_t st) {
    if (n <= 0 || rows <= 0) return hipSuccess;
    if (!split_tile_supported(waves, row_split)) return hipErrorInvalidValue;
    const int64_t tiles = (n + 15) / 16;
    const int tiles_per_wg = waves / row_split;
    const dim3 grid((unsigned)((tiles + tiles_per_wg - 1) / tiles_per_wg), (unsigned)pl.passes);
    const v4u* state = static_cast<const v4u*>(d_state);
    const v4u* panel = static_cast<const v4u*>(d_panel);
    const int fmt = split_fmt(form), nt = split_terms(form), rb = pl.rb_per_pass;
#define RSH_GO(W, FMT, NT, F, RH) return launch_split_rb<W, FMT, NT, F, RH>(rb, grid, st, state, n, pl.steps, d_cell_inv, panel, d_row_inv, d_out, ld, rows, out_f32)
#define RSH_FORMS(W, RH)                                                           \
    if (fmt == 0 && nt == 2) { if (faces == 4) RSH_GO(W, 0, 2, 4, RH); RSH_GO(W, 0, 2, 1, RH); } \
    if (fmt == 1 && nt == 2) { if (faces == 4) RSH_GO(W, 1, 2, 4, RH); RSH_GO(W, 1, 2, 1, RH); } \
    if (faces == 4) RSH_GO(W, 0, 3, 4, RH);                                         \
    RSH_GO(W, 0, 3, 1, RH);
    if (waves == 8) { RSH_FORMS(8, 2) }
    RSH_FORMS(4, 1)
#undef RSH_FORMS
#undef RSH_GO
}

}  // namespace rsh

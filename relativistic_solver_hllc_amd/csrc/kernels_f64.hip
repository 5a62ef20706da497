// FP64 kernels of the scattering integral on CDNA4 (gfx950):
//   * phase-matrix build  P[k][j] = p(w_k . w_j) w_j / den_k   (once per direction set)
//   * panel re-layout     P rows [k0,k1) -> MFMA A-fragment order (cached per range)
//   * direct kernel       one thread per (cell, output direction), reference order
//   * MFMA GEMM           S[k][c] = sum_j P[k][j] * sum_f I[j][c][f]
//
// The contraction is S (R x N) = P (R x M) * X (M x N) with X[j][c] the face sum
// of illum[j][c][0..F). On MI355X it is compute-bound in FP64 (2M flops per
// (cell, direction) pair against 32 B of face input), so it runs on
// v_mfma_f64_16x16x4f64 with:
//   - one wave per 16 consecutive cells, holding ALL output rows of a pass in
//     accumulators (25 row blocks = 400 directions -> 100 doubles per lane), so
//     every face value is read from HBM exactly once per pass;
//   - the face sum fused into the B-operand load: lane (jr, cl) reads the F
//     contiguous faces of cell cl for direction 4s + jr (a 16-lane group reads
//     16*F*8 contiguous bytes), sums them in FP64 and feeds the sum to MFMA;
//   - the phase panel streamed through LDS in double-buffered chunks shared by
//     all waves of the workgroup (A operand, one conflict-free ds_read_b64 per
//     fragment), prefetched from L2 one chunk ahead;
//   - the epilogue storing each 16x16 tile as four 128-byte row segments of the
//     direction-major output with non-temporal stores (S is not re-read).
#include "mfma_f64_core.hpp"

namespace rsh {

// No FMA contraction: keeps the reference's rounding (3 (1 + mu^2)) / 4 exactly.
__host__ __device__ inline double phase_kernel(int kind, double g, double mu) {
#pragma clang fp contract(off)
    if (kind == RSH_PHASE_THOMSON) return (3. * (1 + mu * mu)) / 4;  // cuda/cuda_utils.cu:15
    const double d = 1.0 + g * g - 2.0 * g * mu;
    return (1.0 - g * g) / (d * sqrt(d));
}

// ---------------------------------------------------------------------------
// Phase matrix. One 256-thread block per output row k; the FP64 row sum for
// normalisation is a block tree reduction.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void build_phase_kernel(const double* __restrict__ dirs,
                                                          const double* __restrict__ w, int m,
                                                          double full_area, int kind, double g,
                                                          int normalize, double* __restrict__ P) {
#pragma clang fp contract(off)
    const int k = blockIdx.x;
    const double xk = dirs[3 * k], yk = dirs[3 * k + 1], zk = dirs[3 * k + 2];
    double* row = P + (int64_t)k * m;
    double part = 0;
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
        // mu in the order of the reference's Gamma loop (cuda_utils.cu:9-13)
        const double mu = (dirs[3 * j] * xk + dirs[3 * j + 1] * yk) + dirs[3 * j + 2] * zk;
        const double v = phase_kernel(kind, g, mu) * w[j];
        row[j] = v;
        part += v;
    }
    __shared__ double red[256];
    red[threadIdx.x] = part;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    const double den = normalize ? red[0] : full_area;
    for (int j = threadIdx.x; j < m; j += blockDim.x) row[j] = row[j] / den;
}

hipError_t launch_build_phase(const double* d_dirs, const double* d_w, int m, double full_area, int kind,
                              double g, int normalize, double* d_P, hipStream_t st) {
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(build_phase_kernel, dim3(m), dim3(256), 0, st, d_dirs, d_w, m, full_area, kind, g,
                       normalize, d_P);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Panel re-layout (see rsh_internal.hpp for the fragment-major layout).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void panel_f64_kernel(const double* __restrict__ P, int m, int k0, int k1,
                                                        double scale, int rb_per_pass, int steps, int64_t total,
                                                        double* __restrict__ panel) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int lane = (int)(idx & 63);
    int64_t frag = idx >> 6;
    const int rb_local = (int)(frag % rb_per_pass);
    frag /= rb_per_pass;
    const int s = (int)(frag % steps);
    const int pass = (int)(frag / steps);
    const int k = k0 + (pass * rb_per_pass + rb_local) * 16 + (lane & 15);
    const int j = 4 * s + (lane >> 4);
    double v = 0.0;
    if (k < k1 && j < m) v = P[(int64_t)k * m + j] * scale;
    panel[idx] = v;
}

static const int kRbSet[] = {1, 2, 3, 4, 5, 6, 7, 8, 10, 13, 16, 20, 25};

F64Plan plan_f64(int m, int rows, int max_rb) {
    F64Plan pl{};
    const int nrb = (rows + 15) / 16;
    if (max_rb < 1 || max_rb > kMaxRowBlocksPerPass) max_rb = kMaxRowBlocksPerPass;
    pl.passes = nrb > 0 ? (nrb + max_rb - 1) / max_rb : 0;
    const int per = pl.passes > 0 ? (nrb + pl.passes - 1) / pl.passes : 0;
    pl.rb_per_pass = 0;
    for (int v : kRbSet)
        if (v >= per) { pl.rb_per_pass = v; break; }
    pl.rb_total = pl.rb_per_pass * pl.passes;
    const int steps = (m + 3) / 4;
    pl.steps = (steps + kPanelStepAlign - 1) / kPanelStepAlign * kPanelStepAlign;
    return pl;
}

int64_t panel_f64_doubles(const F64Plan& pl) { return (int64_t)pl.steps * pl.rb_total * 64; }

hipError_t launch_panel_f64(const double* d_P, int m, int k0, int k1, double scale, const F64Plan& pl,
                            double* d_panel, hipStream_t st) {
    const int64_t total = panel_f64_doubles(pl);
    if (total == 0) return hipSuccess;
    const int64_t blocks = (total + 255) / 256;
    hipLaunchKernelGGL(panel_f64_kernel, dim3((unsigned)blocks), dim3(256), 0, st, d_P, m, k0, k1, scale,
                       pl.rb_per_pass, pl.steps, total, d_panel);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Direct kernel: thread = (cell, output direction). With ref_order set it
// evaluates Gamma inline and accumulates ((Gamma * Ibar) * w_j) for j
// ascending, then divides by F, exactly as GetS (Illum_async.cpp:212-231) and
// d_GetS_MPI_Stream (cuda_kernel.cu:98-110) with no FMA contraction, so its
// result is bit-identical to a non-contracted sequential CPU evaluation.
// Otherwise it accumulates P[k][j] * Ibar in the same order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void direct_f64_kernel(const double* __restrict__ illum, int faces, int64_t n,
                                                         int m, const double* __restrict__ dirs,
                                                         const double* __restrict__ w, double full_area,
                                                         const double* __restrict__ P, int ref_order, int k0,
                                                         double* __restrict__ out, int64_t ld) {
#pragma clang fp contract(off)
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = k0 + (int)blockIdx.y;
    if (c >= n) return;
    const double xk = dirs[3 * k], yk = dirs[3 * k + 1], zk = dirs[3 * k + 2];
    const int64_t stride = n * faces;
    const double* src = illum + c * faces;
    double s = 0;
    for (int j = 0; j < m; j++) {
        double I = 0;
        for (int f = 0; f < faces; f++) I += src[f];
        I /= faces;
        if (ref_order) {
            const double mu = (dirs[3 * j] * xk + dirs[3 * j + 1] * yk) + dirs[3 * j + 2] * zk;
            s += phase_kernel(RSH_PHASE_THOMSON, 0.0, mu) * I * w[j];
        } else {
            s += P[(int64_t)k * m + j] * I;
        }
        src += stride;
    }
    out[(int64_t)blockIdx.y * ld + c] = ref_order ? s / full_area : s;
}

hipError_t launch_direct_f64(const double* d_illum, int faces, int64_t n, int m, const double* d_dirs,
                             const double* d_w, double full_area, const double* d_P, int ref_order, int k0,
                             int k1, double* d_out, int64_t ld, hipStream_t st) {
    if (n <= 0 || k1 <= k0) return hipSuccess;
    const int64_t bx = (n + 255) / 256;
    hipLaunchKernelGGL(direct_f64_kernel, dim3((unsigned)bx, (unsigned)(k1 - k0)), dim3(256), 0, st, d_illum,
                       faces, n, m, d_dirs, d_w, full_area, d_P, ref_order, k0, d_out, ld);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP64 MFMA GEMM.
//
// Operand layouts of v_mfma_f64_16x16x4f64 (D = A * B + C, 16x16x4):
//   A: lane l holds A[l % 16][l / 16]      (row = output direction, col = j)
//   B: lane l holds B[l / 16][l % 16]      (row = j, col = cell)
//   D: lane l, item r holds D[l / 16 + 4 r][l % 16]   (rows interleaved by 4)
// (measured with one-hot operands on gfx950; rsh_selftest_mfma_f64 re-checks it.)
//
// Workgroup = W waves = 16 W consecutive cells; grid.y = pass over row blocks.
// Loop over chunks of SC j-steps: the panel chunk (SC * RB fragments) is staged
// through registers into one of two LDS buffers while the other is consumed.
// ---------------------------------------------------------------------------
template <int RB, int W, int SC, int FACES>
__global__ __launch_bounds__(64 * W) void mfma_f64_kernel(const double* __restrict__ illum, int64_t n, int64_t ldx,
                                                          int m, const double* __restrict__ panel, int steps,
                                                          double* __restrict__ out, int64_t ld, int rows) {
    __shared__ v2d lds[2][F64TileShape<RB, W, SC>::UNITS];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int jr = lane >> 4;
    const int64_t cell = (int64_t)blockIdx.x * (16 * W) + wave * 16 + (lane & 15);
    const int rb0 = blockIdx.y * RB;
    v4d acc[RB];
    mfma_f64_accumulate<RB, W, SC, FACES>(illum, n, ldx, m, panel + (int64_t)blockIdx.y * steps * RB * 64, steps,
                                          cell, lds, acc);
    if (cell >= n) return;
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = (rb0 + rb) * 16 + jr + 4 * r;
            if (row < rows) __builtin_nontemporal_store(acc[rb][r], out + (int64_t)row * ld + cell);
        }
    }
}

template <int W, int SC, int FACES>
static hipError_t launch_mfma_rb(int rb, dim3 grid, hipStream_t st, const double* illum, int64_t n, int64_t ldx, int m,
                                 const double* panel, int steps, double* out, int64_t ld, int rows) {
#define RSH_RB_CASE(V)                                                                                           \
    case V:                                                                                                      \
        hipLaunchKernelGGL((mfma_f64_kernel<V, W, SC, FACES>), grid, dim3(64 * W), 0, st, illum, n, ldx, m, panel, \
                           steps, out, ld, rows);                                                                \
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

bool mfma_f64_tile_supported(int waves, int chunk_steps) {
    // (8 waves, 4 steps) is left out: at 25 row blocks it exceeds 256 VGPRs and spills.
    return (waves == 4 && (chunk_steps == 2 || chunk_steps == 4)) || (waves == 8 && chunk_steps == 2);
}

hipError_t launch_mfma_f64(const double* d_illum, int faces, int64_t n, int64_t ldx, int m, const double* d_panel,
                           const F64Plan& pl, int rows, double* d_out, int64_t ld, int waves, int chunk_steps,
                           hipStream_t st) {
    if (n <= 0 || rows <= 0) return hipSuccess;
    if (ldx < n) return hipErrorInvalidValue;
    if (!mfma_f64_tile_supported(waves, chunk_steps) || (faces != 1 && faces != 4)) return hipErrorInvalidValue;
    if (pl.steps % chunk_steps != 0) return hipErrorInvalidValue;
    const int64_t bx = (n + 16 * waves - 1) / (16 * waves);
    const dim3 grid((unsigned)bx, (unsigned)pl.passes);
    const int rb = pl.rb_per_pass;
#define RSH_TILE(W, SC)                                                                                \
    if (waves == W && chunk_steps == SC) {                                                             \
        if (faces == 4)                                                                                \
            return launch_mfma_rb<W, SC, 4>(rb, grid, st, d_illum, n, ldx, m, d_panel, pl.steps, d_out, ld, rows); \
        return launch_mfma_rb<W, SC, 1>(rb, grid, st, d_illum, n, ldx, m, d_panel, pl.steps, d_out, ld, rows);    \
    }
    RSH_TILE(8, 2)
    RSH_TILE(4, 2)
    RSH_TILE(4, 4)
#undef RSH_TILE
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Layout probe: one MFMA on a single wave with A, B given in natural order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void mfma_layout_probe_kernel(const double* __restrict__ A,  // [16][4]
                                                               const double* __restrict__ B,  // [4][16]
                                                               double* __restrict__ D) {      // [64][4]
    const int l = threadIdx.x;
    const double a = A[(l % 16) * 4 + l / 16];
    const double b = B[(l / 16) * 16 + l % 16];
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    for (int r = 0; r < 4; r++) D[l * 4 + r] = acc[r];
}

hipError_t launch_mfma_layout_probe(const double* d_a, const double* d_b, double* d_d, hipStream_t st) {
    hipLaunchKernelGGL(mfma_layout_probe_kernel, dim3(1), dim3(64), 0, st, d_a, d_b, d_d);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Seeded log-normal fill (SplitMix64 hash of (seed, index), Box-Muller).
// ---------------------------------------------------------------------------
__device__ inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ __launch_bounds__(256) void fill_lognormal_kernel(double* __restrict__ out, int64_t count, double mu,
                                                             double sigma, uint64_t seed) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const uint64_t h1 = splitmix64(seed ^ splitmix64(2 * (uint64_t)i));
        const uint64_t h2 = splitmix64(seed ^ splitmix64(2 * (uint64_t)i + 1));
        const double u1 = ((double)(h1 >> 11) + 0.5) * 0x1.0p-53;
        const double u2 = (double)(h2 >> 11) * 0x1.0p-53;
        const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        out[i] = exp(mu + sigma * z);
    }
}

hipError_t launch_fill_lognormal(double* d, int64_t count, double mu, double sigma, uint64_t seed,
                                 hipStream_t st) {
    if (count <= 0) return hipSuccess;
    int64_t blocks = (count + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_lognormal_kernel, dim3((unsigned)blocks), dim3(256), 0, st, d, count, mu, sigma,
                       seed);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void fill_uniform_kernel(double* __restrict__ out, int64_t count, double lo,
                                                           double hi, uint64_t seed) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const uint64_t h = splitmix64(seed ^ splitmix64(2 * (uint64_t)i + 1));
        out[i] = lo + (hi - lo) * ((double)(h >> 11) * 0x1.0p-53);
    }
}

hipError_t launch_fill_uniform(double* d, int64_t count, double lo, double hi, uint64_t seed, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    int64_t blocks = (count + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_uniform_kernel, dim3((unsigned)blocks), dim3(256), 0, st, d, count, lo, hi, seed);
    return hipGetLastError();
}

typedef float v4f_src __attribute__((ext_vector_type(4)));

// The split forms' source term: FP64 [m][n] rounded to FP32 (nearest) in the
// layout of the MFMA output tiles, so that an epilogue lane reads its four rows
// of one cell as one 16-byte vector: element (k, c) at
// ((k / 16 * tiles + c / 16) * 64 + 16 ((k % 16) / 4) + c % 16) * 4 + k % 4,
// tiles = ceil(n / 16); rows up to rows_alloc (whole row blocks), cells up to
// 16 tiles; positions past m or n hold 0. Thread = one 16-byte vector.
__global__ __launch_bounds__(256) void source_tiles_kernel(const double* __restrict__ src, int m, int64_t n,
                                                           int64_t tiles, int64_t vectors, v4f_src* __restrict__ dst) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < vectors; v += stride) {
        const int lane = (int)(v & 63);
        const int64_t rest = v >> 6;
        const int64_t tile = rest % tiles, rb = rest / tiles;
        const int64_t c = tile * 16 + (lane & 15);
        const int k0 = (int)(rb * 16 + 4 * (lane >> 4));
        v4f_src o;
#pragma unroll
        for (int r = 0; r < 4; r++) o[r] = (k0 + r < m && c < n) ? (float)src[(int64_t)(k0 + r) * n + c] : 0.0f;
        dst[v] = o;
    }
}

hipError_t launch_source_tiles(const double* src, int m, int64_t n, int rows_alloc, float* dst, hipStream_t st) {
    const int64_t tiles = (n + 15) / 16;
    const int64_t vectors = (int64_t)(rows_alloc / 16) * tiles * 64;
    if (vectors <= 0) return hipSuccess;
    int64_t blocks = (vectors + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(source_tiles_kernel, dim3((unsigned)blocks), dim3(256), 0, st, src, m, n, tiles, vectors,
                       reinterpret_cast<v4f_src*>(dst));
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fill / scale utilities (kernels, so that they are safe inside captured graphs).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void fill_f64_kernel(double* __restrict__ d, int64_t count, double value) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) d[i] = value;
}

__global__ __launch_bounds__(256) void scale_f64_kernel(double* __restrict__ d, int64_t count, double factor) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) d[i] *= factor;
}

__global__ __launch_bounds__(256) void copy_f64_kernel(v2d* __restrict__ dst, const v2d* __restrict__ src,
                                                       int64_t pairs) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < pairs; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

__global__ void copy_tail_kernel(double* __restrict__ dst, const double* __restrict__ src) { dst[0] = src[0]; }

// Relative error of every result against an FP64 reference, reduced per
// workgroup: partial[b] = (max, sum, count above threshold, flat index r * n + c
// of the max) over the workgroup's share of rows x n. A NaN result (or a NaN
// relative error) counts as an infinite error. Row r of got starts at
// r * ld_got, of ref at r * ld_ref.
template <typename T>
__global__ __launch_bounds__(256) void rel_error_kernel(const T* __restrict__ got, int64_t ld_got,
                                                        const double* __restrict__ ref, int64_t ld_ref, int64_t rows,
                                                        int64_t n, double threshold, double* __restrict__ partial) {
    double mx = -1.0, sum = 0.0, above = 0.0, at = -1.0;
    for (int64_t r = blockIdx.y; r < rows; r += gridDim.y) {
        const T* g = got + r * ld_got;
        const double* f = ref + r * ld_ref;
        for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
            const double want = f[c];
            const double d = fabs((double)g[c] - want);
            double e = d == 0.0 ? 0.0 : d / fabs(want);
            if (e != e) e = HUGE_VAL;
            if (e > mx) {
                mx = e;
                at = (double)(r * n + c);
            }
            sum += e;
            above += e > threshold ? 1.0 : 0.0;
        }
    }
    __shared__ double red[4][256];
    red[0][threadIdx.x] = mx;
    red[1][threadIdx.x] = sum;
    red[2][threadIdx.x] = above;
    red[3][threadIdx.x] = at;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            const int o = threadIdx.x + st;
            if (red[0][o] > red[0][threadIdx.x] ||
                (red[0][o] == red[0][threadIdx.x] && red[3][o] >= 0 && (red[3][threadIdx.x] < 0 || red[3][o] < red[3][threadIdx.x]))) {
                red[0][threadIdx.x] = red[0][o];
                red[3][threadIdx.x] = red[3][o];
            }
            red[1][threadIdx.x] += red[1][o];
            red[2][threadIdx.x] += red[2][o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* o = partial + 4 * ((int64_t)blockIdx.y * gridDim.x + blockIdx.x);
        o[0] = red[0][0];
        o[1] = red[1][0];
        o[2] = red[2][0];
        o[3] = red[3][0];
    }
}

hipError_t launch_rel_error(const void* got, int got_f32, int64_t ld_got, const double* ref, int64_t ld_ref,
                            int64_t rows, int64_t n, double threshold, double* partial, int gx, int gy,
                            hipStream_t st) {
    if (got_f32)
        hipLaunchKernelGGL(rel_error_kernel<float>, dim3(gx, gy), dim3(256), 0, st, static_cast<const float*>(got),
                           ld_got, ref, ld_ref, rows, n, threshold, partial);
    else
        hipLaunchKernelGGL(rel_error_kernel<double>, dim3(gx, gy), dim3(256), 0, st, static_cast<const double*>(got),
                           ld_got, ref, ld_ref, rows, n, threshold, partial);
    return hipGetLastError();
}

static unsigned grid_for(int64_t count) {
    int64_t b = (count + 255) / 256;
    return (unsigned)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

hipError_t launch_fill_f64(double* d, int64_t count, double value, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_f64_kernel, dim3(grid_for(count)), dim3(256), 0, st, d, count, value);
    return hipGetLastError();
}

hipError_t launch_scale_f64(double* d, int64_t count, double factor, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(scale_f64_kernel, dim3(grid_for(count)), dim3(256), 0, st, d, count, factor);
    return hipGetLastError();
}

hipError_t launch_copy_f64(double* dst, const double* src, int64_t count, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    const int64_t pairs = count / 2;
    if (pairs > 0) {
        int64_t b = (pairs + 255) / 256;
        if (b > 8192) b = 8192;
        hipLaunchKernelGGL(copy_f64_kernel, dim3((unsigned)b), dim3(256), 0, st, reinterpret_cast<v2d*>(dst),
                           reinterpret_cast<const v2d*>(src), pairs);
    }
    if (count & 1)
        hipLaunchKernelGGL(copy_tail_kernel, dim3(1), dim3(1), 0, st, dst + count - 1, src + count - 1);
    return hipGetLastError();
}

}  // namespace rsh

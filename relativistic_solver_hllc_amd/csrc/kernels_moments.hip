// Radiation moments of the face intensities: energy, flux ("stream"), pressure
// tensor ("impuls") and the divergences of flux and pressure tensor, per cell,
// over all M directions. Replaces d_MakeIllumParam (cuda/cuda_kernel.cu:113-226)
// and its helpers IntegarteDirection / IntegarteDirection3Faces /
// IntegarteDirection9Faces (cuda/cuda_utils.cu:18-167).
//
// Every moment is a weighted sum over directions of one face intensity:
//   T_f[q] = sum_j psi[j][q] I[j][c][f],  psi[j][q] = phi_q(omega_j) w_j / F,
//   phi = (1, x, y, z, xx, xy, xz, yy, yz, zz)
// (ten functions: the pressure tensor is symmetric). That is a [10 x M] x
// [M x 4N] product with M = 400 against 4N = 6.3 M columns: 2.5 flop per byte
// of the intensity, so the kernel is HBM-bound (the intensity is read exactly
// once, 32 M bytes per cell) and runs on the FP64 VALU: psi is uniform across
// the wave and comes in through scalar loads as the SGPR operand of each
// v_fma_f64, the intensity is the vector operand, loaded with 16-byte loads.
//
// Thread mapping: FPT consecutive faces of one cell per thread (TPC = FACES /
// FPT threads per cell, adjacent lanes). The cell's face terms are combined
// with lane shuffles, then the first lane of the cell writes
//   energy    = (sum_f T_f[0]) / FACES
//   stream[k] = (sum_f T_f[1+k]) / FACES
//   impuls    = (sum_f T_f[4..9]) / FACES, expanded to 3x3 row-major
//   divstream = (sum_f (T_f[1..3] . n_f) area_f) / volume
//   divimpuls[h] = (sum_f (T_f[h,:] . n_f) area_f) / volume
// with the reference's layouts: normals[c][f][3], areas[c][f], volume[c].
#include "rsh_internal.hpp"

namespace rsh {

typedef double v2d_m __attribute__((ext_vector_type(2)));

// 128-thread workgroups and four rows in flight per thread: 72-74 % of the
// 8 TB/s HBM peak at 1.57 M cells x 400 directions (a pure streaming read of
// the same bytes: 88 %). Blocks of 64-1024 threads with 2-16 rows in flight
// measured 69-79 %, with the same configuration differing by up to 10 % from
// one process to the next, so none is reliably faster.
constexpr int kMomentBlock = 128;
constexpr int kMomentPipe = 4;

// Non-temporal loads: every face value is read exactly once, so it should not
// displace anything in the caches (plain loads measured 4-6 % slower).
template <int FPT>
__device__ inline void load_faces(const double* __restrict__ p, double (&v)[FPT]) {
    if constexpr (FPT == 4) {
        const v2d_m a = __builtin_nontemporal_load(reinterpret_cast<const v2d_m*>(p));
        const v2d_m b = __builtin_nontemporal_load(reinterpret_cast<const v2d_m*>(p + 2));
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    } else if constexpr (FPT == 2) {
        const v2d_m a = __builtin_nontemporal_load(reinterpret_cast<const v2d_m*>(p));
        v[0] = a.x; v[1] = a.y;
    } else {
        v[0] = __builtin_nontemporal_load(p);
    }
}

template <int FPT>
__device__ inline void accumulate(const double* __restrict__ ps, const double (&v)[FPT],
                                  double (&acc)[FPT][kMomentFuncs]) {
#pragma unroll
    for (int q = 0; q < kMomentFuncs; q++) {
        const double w = ps[q];
#pragma unroll
        for (int f = 0; f < FPT; f++) acc[f][q] = __builtin_fma(w, v[f], acc[f][q]);
    }
}

template <int TPC>
__device__ inline double cell_sum(double x) {
#pragma unroll
    for (int off = 1; off < TPC; off <<= 1) x += __shfl_xor(x, off, 64);
    return x;
}

template <int FACES, int FPT, int PIPE>
__global__ __launch_bounds__(kMomentBlock) void moments_f64_kernel(
    const double* __restrict__ illum, int64_t n, int m, const double* __restrict__ psi,
    const double* __restrict__ normals, const double* __restrict__ areas, const double* __restrict__ volume,
    double* __restrict__ energy, double* __restrict__ stream, double* __restrict__ impuls,
    double* __restrict__ divstream, double* __restrict__ divimpuls) {
    constexpr int TPC = FACES / FPT;
    const int64_t gid = (int64_t)blockIdx.x * kMomentBlock + threadIdx.x;
    const int64_t cell = gid / TPC;
    const int fg = (int)(gid % TPC);
    const bool active = cell < n;
    // Lanes past the last cell read the last cell (in bounds) and write nothing,
    // so every lane of a cell group takes part in the shuffles.
    const int64_t c = active ? cell : n - 1;
    const double* __restrict__ src = illum + c * FACES + fg * FPT;
    const int64_t stride = n * FACES;

    double acc[FPT][kMomentFuncs];
#pragma unroll
    for (int f = 0; f < FPT; f++)
#pragma unroll
        for (int q = 0; q < kMomentFuncs; q++) acc[f][q] = 0.0;

    // Ring of PIPE direction rows in flight: row j + PIPE is requested while
    // row j is consumed (indices past the end re-read the last row, in bounds).
    double ring[PIPE][FPT];
#pragma unroll
    for (int u = 0; u < PIPE; u++) load_faces<FPT>(src + (int64_t)(u < m ? u : m - 1) * stride, ring[u]);
    int j = 0;
    for (; j + PIPE <= m; j += PIPE) {
#pragma unroll
        for (int u = 0; u < PIPE; u++) {
            double v[FPT];
#pragma unroll
            for (int f = 0; f < FPT; f++) v[f] = ring[u][f];
            const int jn = j + u + PIPE;
            load_faces<FPT>(src + (int64_t)(jn < m ? jn : m - 1) * stride, ring[u]);
            accumulate<FPT>(psi + (j + u) * kMomentFuncs, v, acc);
        }
    }
#pragma unroll
    for (int u = 0; u < PIPE; u++)
        if (j + u < m) accumulate<FPT>(psi + (j + u) * kMomentFuncs, ring[u], acc);

    double e = 0.0, s[3] = {0.0, 0.0, 0.0}, t[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double ds = 0.0, di[3] = {0.0, 0.0, 0.0};
    const bool want_div = divstream != nullptr || divimpuls != nullptr;
#pragma unroll
    for (int f = 0; f < FPT; f++) {
        const double* a = acc[f];
        e += a[0];
        s[0] += a[1]; s[1] += a[2]; s[2] += a[3];
#pragma unroll
        for (int q = 0; q < 6; q++) t[q] += a[4 + q];
        if (want_div) {
            const int64_t cf = c * FACES + fg * FPT + f;
            const double nx = normals[3 * cf], ny = normals[3 * cf + 1], nz = normals[3 * cf + 2];
            const double ar = areas[cf];
            ds += (a[1] * nx + a[2] * ny + a[3] * nz) * ar;
            di[0] += (a[4] * nx + a[5] * ny + a[6] * nz) * ar;
            di[1] += (a[5] * nx + a[7] * ny + a[8] * nz) * ar;
            di[2] += (a[6] * nx + a[8] * ny + a[9] * nz) * ar;
        }
    }
    if constexpr (TPC > 1) {
        e = cell_sum<TPC>(e);
#pragma unroll
        for (int k = 0; k < 3; k++) s[k] = cell_sum<TPC>(s[k]);
#pragma unroll
        for (int q = 0; q < 6; q++) t[q] = cell_sum<TPC>(t[q]);
        if (want_div) {
            ds = cell_sum<TPC>(ds);
#pragma unroll
            for (int k = 0; k < 3; k++) di[k] = cell_sum<TPC>(di[k]);
        }
    }
    if (!active || fg != 0) return;
    constexpr double inv_faces = 1.0 / FACES;  // FACES is 1 or 4: exact
    if (energy) energy[cell] = e * inv_faces;
    if (stream) {
#pragma unroll
        for (int k = 0; k < 3; k++) stream[3 * cell + k] = s[k] * inv_faces;
    }
    if (impuls) {
        // row-major 3x3 of the symmetric tensor (xx xy xz / xy yy yz / xz yz zz)
        constexpr int map[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
#pragma unroll
        for (int q = 0; q < 9; q++) impuls[9 * cell + q] = t[map[q]] * inv_faces;
    }
    if (want_div) {
        const double vol = volume[cell];
        if (divstream) divstream[cell] = ds / vol;
        if (divimpuls) {
#pragma unroll
            for (int k = 0; k < 3; k++) divimpuls[3 * cell + k] = di[k] / vol;
        }
    }
}

hipError_t launch_moments_f64(const double* d_illum, int faces, int64_t n, int m, const double* d_psi,
                              const double* d_normals, const double* d_areas, const double* d_volume,
                              double* d_energy, double* d_stream, double* d_impuls, double* d_divstream,
                              double* d_divimpuls, int faces_per_thread, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int fpt = faces == 1 ? 1 : faces_per_thread;
    const int64_t threads = n * (faces / fpt);
    const dim3 grid((unsigned)((threads + kMomentBlock - 1) / kMomentBlock)), block(kMomentBlock);
#define RSH_MOM(FA, FP)                                                                                        \
    hipLaunchKernelGGL((moments_f64_kernel<FA, FP, kMomentPipe>), grid, block, 0, st, d_illum, n, m, d_psi,   \
                       d_normals, d_areas, d_volume, d_energy, d_stream, d_impuls, d_divstream, d_divimpuls)
    if (faces == 1) RSH_MOM(1, 1);
    else if (fpt == 4) RSH_MOM(4, 4);
    else if (fpt == 2) RSH_MOM(4, 2);
    else if (fpt == 1) RSH_MOM(4, 1);
    else return hipErrorInvalidValue;
#undef RSH_MOM
    return hipGetLastError();
}

}  // namespace rsh

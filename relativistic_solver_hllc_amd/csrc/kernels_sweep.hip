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
// run in parallel, with a workgroup barrier between levels. The face values a
// level writes reach the next level through global memory inside one CU.
//
// Face values. The reference keeps three node values per face and only ever
// uses their mean ((c0 + c1 + c2) / 3, both when a downstream cell reads its in
// face and when ReCalcIllum stores the face intensity), so this kernel stores
// that mean, computed in the same order, directly in the new illum array
// [M][N][4] (the ReCalcIllum output layout); every face entry is written exactly
// once per direction (boundary in faces by the cell, every other face by the
// cell for which it is an out face, "out on both sides" faces by both in the
// order of their levels). The residual max |old - new| / new follows at the end
// of the direction, with the reference's comparison (a NaN ratio never wins).
#include "rsh_internal.hpp"

namespace rsh {

namespace {

constexpr int kSweepBlock = 256;

// BoundaryConditions, non-Sphere build: OutSource (-3) -> 1, Free (-1),
// Lock (-2), InnerSource (-4) -> 0 (types are validated on the host).
__device__ inline double boundary_value(int type_bound) { return type_bound == -3 ? 1.0 : 0.0; }

// GetCurIllum of one class, in the reference's expression order.
template <int CLS>
__device__ inline double cur_illum(const double* __restrict__ x, double s, double I_0, double S) {
    double Q, alpha, betta;
    if constexpr (CLS == 2) {
        const double ss = s * 388189 * 1e5;
        alpha = 1;
        Q = 1;
        const double I = I_0 * exp(-alpha * ss) + Q * (1 - exp(-alpha * ss)) / alpha;
        return I < 0 ? 0 : I;
    } else {
        if constexpr (CLS == 0) {
            Q = 0; alpha = 2; betta = 1;
            const double dx = x[0] - 0.5, dy = x[1] - 0.0, dz = x[2] - 0.0;
            if (sqrt(dx * dx + dy * dy + dz * dz) < 0.05) { Q = 0; alpha = 0.5; betta = 0.5; }
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

template <int CLS>
__global__ __launch_bounds__(kSweepBlock) void sweep_kernel(const SweepDirDev* __restrict__ dirs, int64_t n,
                                                            const int32_t* __restrict__ pairs,
                                                            const double* __restrict__ nodes,
                                                            const double* __restrict__ S, int64_t ld,
                                                            const double* __restrict__ old_illum, double* new_illum,
                                                            double* __restrict__ norm) {
    const int k = blockIdx.x;
    const SweepDirDev d = dirs[k];
    double* fv = new_illum + (int64_t)k * n * 4;
    const double* __restrict__ Sk = S + (int64_t)k * ld;
    for (int l = 0; l < d.levels; l++) {
        const int b = d.level_ptr[l], e = d.level_ptr[l + 1];
        for (int i = b + (int)threadIdx.x; i < e; i += kSweepBlock) {
            const int64_t cell = d.cells[i];
            const int st = d.state[cell];
            int64_t r = d.x0_off[cell];
            const double Sv = Sk[cell];
            int pr[4];
#pragma unroll
            for (int f = 0; f < 4; f++) pr[f] = pairs[cell * 4 + f];
#pragma unroll
            for (int f = 0; f < 4; f++) {
                const int p = pr[f];
                if ((st >> f) & 1) {
                    if (p < 0) {
                        const double i0 = boundary_value(p);
                        fv[cell * 4 + f] = (i0 + i0 + i0) / 3;
                    }
                    continue;
                }
                double I[3];
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const double s = d.s[r];
                    const int g = d.in_face[r];
                    r++;
                    const int pin = g == 0 ? pr[0] : g == 1 ? pr[1] : g == 2 ? pr[2] : pr[3];
                    const double ix0 = pin < 0 ? boundary_value(pin) : fv[cell * 4 + g];
                    const double* x = CLS == 0 ? nodes + ((cell * 4 + f) * 3 + j) * 3 : nullptr;
                    I[j] = cur_illum<CLS>(x, s, ix0, Sv);
                }
                const double v = (I[0] + I[1] + I[2]) / 3;
                fv[cell * 4 + f] = v;
                if (p >= 0) fv[p] = v;
            }
        }
        __syncthreads();
    }
    if (!old_illum) return;
    const double* __restrict__ ok = old_illum + (int64_t)k * n * 4;
    double loc = -1;
    for (int64_t id = threadIdx.x; id < n * 4; id += kSweepBlock) {
        const double cur = fv[id];
        const double b = fabs(ok[id] - cur) / cur;
        if (b > loc) loc = b;
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

hipError_t launch_sweep(int cls, const SweepDirDev* d_dirs, int m, int64_t n, const int32_t* d_pairs,
                        const double* d_nodes, const double* d_S, int64_t ld, const double* d_old, double* d_new,
                        double* d_norm, hipStream_t st) {
    if (m <= 0 || n <= 0) return hipSuccess;
    const dim3 grid((unsigned)m), block(kSweepBlock);
#define RSH_SWEEP(C)                                                                                             \
    hipLaunchKernelGGL(sweep_kernel<C>, grid, block, 0, st, d_dirs, n, d_pairs, d_nodes, d_S, ld, d_old, d_new, \
                       d_norm)
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

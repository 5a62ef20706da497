// CPU oracle for the transport sweep (short characteristics on a tetrahedral
// mesh), the step that produces the face intensities the scattering integral
// consumes.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP sweep is
// compared against; nothing in the product path links or calls it. It is used
// by tests/ and by bench.py's cpu_baseline leg.
//
// It restates, sequentially and with the reference's data structures (paths
// relative to solver_hllc_illum_cuda/ of Lartsev-Artem/relativistic_solver_hllc):
//
//   orc_sweep               the direction loop body of MPI_CalculateIllumAsync
//                           solve_module/illum/Illum_async.cpp:725-784: cells in
//                           sorted_id_cell order, out faces (face_state bit clear),
//                           three nodes per out face reading vec_x0 records in
//                           traversal order, the face's 3 node values stored for
//                           the cell and for its neighbour (pairs)
//                           + CalculateIllumeOnInnerFace  Illum_async.cpp:172-199
//                           + BoundaryConditions          solve_module/illum/illum_utils.cpp:407-471
//                           + GetCurIllum classes 0,1,2,3,10  Illum_async.cpp:259-430
//                           + ReCalcIllum (face value = mean of its 3 node values,
//                             norm = max |old - new| / new)  Illum_async.cpp:472-510
//
// The per-face coefficient array (inter_coef, 3 node values per face) is kept
// in full, as the reference does; it starts as NaN for every direction so that
// a value read before it is written shows up in the result (the reference
// reuses one per-thread array across directions).
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp).

#include <omp.h>

#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

namespace {

constexpr int kBase = 4;

// BoundaryConditions (illum_utils.cpp:407-471) with the non-Sphere build:
// eBound_OutSource (-3) -> 1, FreeBound (-1), LockBound (-2), InnerSource (-4) -> 0.
// Returns false for an unknown type (the reference aborts there).
bool boundary_value(int type_bound, double* i0) {
    switch (type_bound) {
        case -3: *i0 = 1.0; return true;
        case -1: *i0 = 0.0; return true;
        case -2: *i0 = 0.0; return true;
        case -4: *i0 = 0.0; return true;
        default: return false;
    }
}

// GetCurIllum (Illum_async.cpp:259-430), classes 0, 1, 2, 3 and 10 in the
// reference's expression order.
double cur_illum(int cls, const double* x, double s, double I_0, double int_scattering) {
    switch (cls) {
        case 0: {
            double Q = 0, alpha = 2, betta = 1;
            const double S = int_scattering;
            const double dx = x[0] - 0.5, dy = x[1] - 0.0, dz = x[2] - 0.0;
            if (std::sqrt(dx * dx + dy * dy + dz * dz) < 0.05) { Q = 0; alpha = 0.5; betta = 0.5; }
            const double k = alpha + betta;
            double I;
            if (k > 1e-10) I = (std::exp(-k * s) * (I_0 * k + (std::exp(k * s) - 1) * (Q + S * betta))) / k;
            else I = (1 - s * k) * (I_0 + s * (Q + S * betta));
            return I < 0 ? 0 : I;
        }
        case 1: {
            const double S = int_scattering, Q = 1, alpha = 1;
            const double betta = alpha / 2;
            const double k = alpha + betta;
            double I;
            if (k > 1e-10) I = (std::exp(-k * s) * (I_0 * k + (std::exp(k * s) - 1) * (Q + S * betta))) / k;
            else I = (1 - s * k) * (I_0 + s * (Q + S * betta));
            return I < 0 ? 0 : I;
        }
        case 2: {
            const double ss = s * 388189 * 1e5;
            const double alpha = 1, Q = 1;
            double I;
            if (alpha > 1e-15) I = I_0 * std::exp(-alpha * ss) + Q * (1 - std::exp(-alpha * ss)) / alpha;
            else I = I_0 * (1 - alpha * ss) + Q * ss;
            return I < 0 ? 0 : I;
        }
        case 3: {
            const double S = int_scattering, alpha = 0;
            const double betta = alpha / 2, Q = 0;
            const double k = alpha + betta;
            double I;
            if (k > 1e-10) I = (std::exp(-k * s) * (I_0 * k + (std::exp(k * s) - 1) * (Q + S * betta))) / k;
            else I = (1 - s * k) * (I_0 + s * (Q + S * betta));
            return I < 0 ? 0 : I;
        }
        case 10: {
            const double Q = 0, alpha = 0.5, betta = 100;
            const double S = int_scattering;
            const double k = alpha + betta;
            double I;
            if (k > 1e-10) I = (std::exp(-k * s) * (I_0 * k + (std::exp(k * s) - 1) * (Q + S * betta))) / k;
            else I = (1 - s * k) * (I_0 + s * (Q + S * betta));
            return I < 0 ? 0 : I;
        }
        default:
            return std::numeric_limits<double>::quiet_NaN();
    }
}

struct Coef {
    double v[3];
};

}  // namespace

extern "C" {

// One sweep over directions [0, m):
//   order      [m][n]  sorted_id_cell (a topological order per direction)
//   face_state [m][n]  bit f set = face f is an in face (n_f . omega < -eps)
//   x0_s, x0_face      vec_x0 records of all directions, concatenated; direction
//                      k's records start at x0_begin[k] (in traversal order)
//   pairs      [n][4]  neighbour face id (cell * 4 + face) or boundary type < 0
//   nodes      [n][4][3][3] node coordinates (only class 0 reads them; may be NULL otherwise)
//   scattering [m][n]  grid.scattering (direction-major)
//   illum      [m][n][4] in: previous face values, out: new ones (ReCalcIllum)
// Returns the max over directions of ReCalcIllum's norm, or -2 on malformed
// input (record count mismatch, unknown boundary type or class).
double orc_sweep(int64_t n, int32_t m, const int32_t* order, const int32_t* face_state, const double* x0_s,
                 const uint8_t* x0_face, const int64_t* x0_begin, const int32_t* pairs, const double* nodes,
                 const double* scattering, int32_t cls, double* illum, int32_t threads) {
    if (threads > 0) omp_set_num_threads(threads);
    double norm = -1;
    bool bad = false;
#pragma omp parallel
    {
        std::vector<Coef> inter_coef((size_t)(n * kBase));
        double norm_loc = -1;
#pragma omp for schedule(dynamic, 1)
        for (int32_t d = 0; d < m; d++) {
            const double nan = std::numeric_limits<double>::quiet_NaN();
            for (auto& c : inter_coef) c.v[0] = c.v[1] = c.v[2] = nan;
            int64_t posX0 = x0_begin[d];
            const int32_t* sorted = order + (int64_t)d * n;
            const int32_t* states = face_state + (int64_t)d * n;
            for (int64_t h = 0; h < n; ++h) {
                const int32_t num_cell = sorted[h];
                const int64_t face_block_id = (int64_t)num_cell * kBase;
                for (int num_out_face = 0; num_out_face < kBase; ++num_out_face) {
                    const int32_t out_id_face = pairs[face_block_id + num_out_face];
                    if ((states[num_cell] >> num_out_face) & 1) {
                        if (out_id_face < 0) {
                            double i0;
                            if (!boundary_value(out_id_face, &i0)) {
#pragma omp atomic write
                                bad = true;
                            }
                            inter_coef[face_block_id + num_out_face] = Coef{{i0, i0, i0}};
                        }
                        continue;
                    }
                    Coef I;
                    for (int num_node = 0; num_node < 3; ++num_node) {
                        const double* x = nodes ? nodes + ((face_block_id + num_out_face) * 3 + num_node) * 3 : nullptr;
                        const double s = x0_s[posX0];
                        const int num_in_face = x0_face[posX0];
                        posX0++;
                        // CalculateIllumeOnInnerFace
                        const int32_t neigh_id = pairs[face_block_id + num_in_face];
                        double I_x0;
                        if (neigh_id < 0) {
                            if (!boundary_value(neigh_id, &I_x0)) {
#pragma omp atomic write
                                bad = true;
                            }
                            inter_coef[face_block_id + num_in_face] = Coef{{I_x0, I_x0, I_x0}};
                        } else {
                            const Coef& coef = inter_coef[face_block_id + num_in_face];
                            I_x0 = (coef.v[0] + coef.v[1] + coef.v[2]) / 3;
                        }
                        const double zero[3] = {0, 0, 0};
                        I.v[num_node] = cur_illum(cls, x ? x : zero, s, I_x0, scattering[(int64_t)d * n + num_cell]);
                    }
                    inter_coef[face_block_id + num_out_face] = I;
                    if (out_id_face >= 0) inter_coef[out_id_face] = I;
                }
            }
            if (posX0 != x0_begin[d + 1]) {
#pragma omp atomic write
                bad = true;
            }
            // ReCalcIllum
            double* Illum = illum + (int64_t)d * n * kBase;
            for (int64_t id = 0; id < n * kBase; id++) {
                const Coef& Il = inter_coef[id];
                const double curI = (Il.v[0] + Il.v[1] + Il.v[2]) / 3;
                const double buf_norm = std::fabs(Illum[id] - curI) / curI;
                if (buf_norm > norm_loc) norm_loc = buf_norm;
                Illum[id] = curI;
            }
        }
#pragma omp critical
        {
            if (norm_loc > norm) norm = norm_loc;
        }
    }
    return bad ? -2.0 : norm;
}

}  // extern "C"

// CPU oracle for the radiation-transport scattering integral.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP library is
// compared against; nothing in the product path links or calls it. It is used
// by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.
//
// It restates, in plain C++17/OpenMP with 64-bit indices, the reference's CPU
// scattering path and the helpers around it (paths relative to
// solver_hllc_illum_cuda/ of Lartsev-Artem/relativistic_solver_hllc):
//
//   orc_thomson_gamma      Gamma lambda      solve_module/illum/Illum_async.cpp:203-206,
//                                            cuda/cuda_utils.cu:7-16
//   orc_scatter_reference  GetS + CalculateIntCPU
//                                            solve_module/illum/Illum_async.cpp:201-256
//                                            (same loop nest: OpenMP over output
//                                            directions, then cells, then input
//                                            directions; same arithmetic order)
//   orc_face_average       GetDirectionIllumFromFace
//                                            solve_module/illum/illum_utils.cpp:11-28
//   orc_get_send/disp      GetSend / GetDisp solve_module/solve_utils.cpp:177-202
//   orc_recalc_norm        ReCalcIllum residual
//                                            solve_module/illum/Illum_async.cpp:472-510
//   orc_read_directions    ReadSphereDirectionDecartToSpherical
//                                            file_module/reader_txt.cpp:25-47
//
// Extensions with no reference counterpart (defined by the task statement):
// Henyey-Greenstein phase kernel, FP64 row normalisation, a general phase
// matrix contraction, cell-level (one "face") input, the Fibonacci direction
// generator and the source iteration I <- (1-a) B + a S.
//
// Every function is pinned by tests/test_oracle.py against exact-arithmetic
// golden vectors (tests/golden/) and closed forms; see the test file header.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp). The parity
// build keeps -ffp-contract=off so the reference's sequential rounding order
// is reproduced exactly; the baseline builds (liborc_o3.so, liborc_ofast.so)
// use -O3 -march=x86-64-v3 and -Ofast -march=x86-64-v3.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#include <omp.h>

extern "C" {

// ---------------------------------------------------------------------------
// Phase kernels. Both are normalised so that their mean over the sphere is 1
// ((1/4pi) * integral p dOmega = 1); the reference divides the quadrature sum by
// full_area = sum of weights (= 4pi), Illum_async.cpp:231, cuda_kernel.cu:49.
// ---------------------------------------------------------------------------

// Thomson: Gamma(mu) = 3 (1 + mu^2) / 4, Illum_async.cpp:205, cuda_utils.cu:15.
double orc_thomson_gamma(double mu) { return (3. * (1 + mu * mu)) / 4; }

// Henyey-Greenstein with the same normalisation: (1 - g^2) / (1 + g^2 - 2 g mu)^{3/2}.
double orc_hg_gamma(double g, double mu) {
    const double d = 1.0 + g * g - 2.0 * g * mu;
    return (1.0 - g * g) / (d * std::sqrt(d));
}

// kind 0 = Thomson, 1 = Henyey-Greenstein.
double orc_phase(int kind, double g, double mu) {
    return kind == 0 ? orc_thomson_gamma(mu) : orc_hg_gamma(g, mu);
}

// Dot product in the order of the CUDA Gamma loop (cuda_utils.cu:9-13):
// ((0 + a0 b0) + a1 b1) + a2 b2.
static inline double dot3(const double* a, const double* b) {
    double s = 0;
    for (int i = 0; i < 3; i++) s += a[i] * b[i];
    return s;
}

// ---------------------------------------------------------------------------
// Reference CPU path, GetS + CalculateIntCPU (Illum_async.cpp:201-256).
//
// illum : [n_dirs][n_cells][faces] doubles (faces = base = 4 in the reference,
//         global_def.h:23 with prj_config.h:18; Illum_async.cpp:217 indexes
//         illum[num_direction * N_cell + base*cell + i]).
// dirs  : [n_dirs][3], w : [n_dirs], full_area as read from the direction file.
// Output directions k in [k_disp, k_disp + k_count) are written to
// S[(k - k_disp) * n_cells + cell] (Illum_async.cpp:251).
// cell_begin/cell_end restrict the cells computed (the sampled-cell checks and
// the timed CPU slices use this); the rest of S is left untouched.
// ---------------------------------------------------------------------------
void orc_scatter_reference(const double* illum, int64_t n_cells, int32_t n_dirs, int32_t faces,
                           const double* dirs, const double* w, double full_area,
                           int32_t k_disp, int32_t k_count, int64_t cell_begin, int64_t cell_end,
                           double* S, int32_t n_threads) {
    if (n_threads > 0) omp_set_num_threads(n_threads);
    const int64_t n_face_stride = n_cells * (int64_t)faces;  // N_cell in Illum_async.cpp:210
#pragma omp parallel for schedule(static)
    for (int32_t kk = 0; kk < k_count; ++kk) {  // Illum_async.cpp:247
        const double* direction = dirs + 3 * (int64_t)(k_disp + kk);
        for (int64_t cell = cell_begin; cell < cell_end; cell++) {  // :249
            double s = 0;
            for (int32_t j = 0; j < n_dirs; j++) {  // GetS loop, :212
                double I = 0;
                const double* src = illum + (int64_t)j * n_face_stride + (int64_t)faces * cell;
                for (int32_t f = 0; f < faces; f++) I += src[f];  // :215-218
                I /= faces;                                       // :219
                s += orc_thomson_gamma(dot3(dirs + 3 * (int64_t)j, direction)) * I * w[j];  // :221
            }
            S[(int64_t)kk * n_cells + cell] = s / full_area;  // :231, :251
        }
    }
}

// Cell-averaged intensity of one direction (illum_utils.cpp:11-28).
void orc_face_average(const double* illum, int64_t n_cells, int32_t faces, int32_t dir, double* out) {
    const double* src = illum + (int64_t)dir * n_cells * faces;
    for (int64_t c = 0; c < n_cells; c++) {
        double I = 0;
        for (int32_t f = 0; f < faces; f++) I += src[c * faces + f];
        out[c] = I / faces;
    }
}

// ---------------------------------------------------------------------------
// General phase matrix P[k][j] for output rows k in [k0, k1), all M inputs.
//   normalize == 0 : P = p(mu_kj) * w_j / full_area     (reference convention)
//   normalize == 1 : P = p(mu_kj) * w_j / sum_j' p(mu_kj') w_j'  (FP64 row norm,
//                    so that each row sums to 1 -> energy conservation)
// ---------------------------------------------------------------------------
void orc_phase_matrix(int32_t n_dirs, const double* dirs, const double* w, double full_area,
                      int32_t kind, double g, int32_t normalize, int32_t k0, int32_t k1, double* P) {
    for (int32_t k = k0; k < k1; k++) {
        double* row = P + (int64_t)(k - k0) * n_dirs;
        double sum = 0;
        for (int32_t j = 0; j < n_dirs; j++) {
            const double v = orc_phase(kind, g, dot3(dirs + 3 * (int64_t)j, dirs + 3 * (int64_t)k)) * w[j];
            row[j] = v;
            sum += v;
        }
        const double den = normalize ? sum : full_area;
        for (int32_t j = 0; j < n_dirs; j++) row[j] /= den;
    }
}

// S[k][c] = sum_j P[k][j] * Ibar[c][j], Ibar = face average of illum
// ([M][N][faces]); j ascending, one FP64 accumulator per output.
void orc_scatter_matrix(const double* illum, int64_t n_cells, int32_t n_dirs, int32_t faces,
                        const double* P, int32_t n_rows, int64_t cell_begin, int64_t cell_end,
                        double* S, int32_t n_threads) {
    if (n_threads > 0) omp_set_num_threads(n_threads);
    const int64_t stride = n_cells * (int64_t)faces;
#pragma omp parallel for schedule(static)
    for (int32_t r = 0; r < n_rows; ++r) {
        const double* prow = P + (int64_t)r * n_dirs;
        for (int64_t c = cell_begin; c < cell_end; c++) {
            double s = 0;
            for (int32_t j = 0; j < n_dirs; j++) {
                const double* src = illum + (int64_t)j * stride + (int64_t)faces * c;
                double I = 0;
                for (int32_t f = 0; f < faces; f++) I += src[f];
                I /= faces;
                s += prow[j] * I;
            }
            S[(int64_t)r * n_cells + c] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// Direction partition over MPI ranks, restated from GetSend / GetDisp
// (solve_utils.cpp:177-202): first n % np ranks take one extra direction;
// displacements follow the reference's own construction.
// ---------------------------------------------------------------------------
void orc_get_send(int32_t np, int32_t n, int32_t* send) {
    for (int32_t i = 0; i < np; i++) send[i] = n / np;
    if (n % np)
        for (int32_t i = 0; i < n % np; i++) ++send[i];
}

void orc_get_disp(int32_t np, int32_t n, int32_t* disp) {
    for (int32_t i = 0; i < np; i++) disp[i] = 0;
    const int32_t a = n % np;
    for (int32_t i = 1; i < np; i++) disp[i] = i * (n / np) + a;
    if (n % np)
        for (int32_t i = 1; i < a; i++) disp[a - i] -= i;
}

// ---------------------------------------------------------------------------
// Residual of one direction, restated from ReCalcIllum (Illum_async.cpp:472-510):
// norm = max over cells and faces of |I_old - I_new| / I_new, starting at -1;
// the new values replace the old ones in place (:491-495).
// ---------------------------------------------------------------------------
double orc_recalc_norm(double* illum_old, const double* illum_new, int64_t count) {
    double norm = -1;
    for (int64_t i = 0; i < count; i++) {
        const double cur = illum_new[i];
        const double b = std::fabs(illum_old[i] - cur) / cur;
        if (b > norm) norm = b;
        illum_old[i] = cur;
    }
    return norm;
}

// ---------------------------------------------------------------------------
// Source iteration on cell-level state (task extension; mirrors the residual
// definition above): I <- (1 - a) B + a S(I), S from the phase matrix P.
// I and B are [M][N]. resid[t] receives the max relative change of iteration t.
// ---------------------------------------------------------------------------
void orc_source_iteration(double* I, const double* B, int64_t n_cells, int32_t n_dirs, const double* P,
                          double albedo, int32_t n_iter, double* resid, int32_t n_threads) {
    std::vector<double> S((size_t)n_dirs * n_cells);
    for (int32_t t = 0; t < n_iter; t++) {
        orc_scatter_matrix(I, n_cells, n_dirs, 1, P, n_dirs, 0, n_cells, S.data(), n_threads);
        for (size_t i = 0; i < S.size(); i++) S[i] = (1.0 - albedo) * B[i] + albedo * S[i];
        resid[t] = orc_recalc_norm(I, S.data(), (int64_t)S.size());
    }
}

// ---------------------------------------------------------------------------
// Fibonacci-sphere directions with equal weights 4pi/M and F = sum of weights.
// z_i = 1 - (2i + 1)/M, phi_i = i * pi (3 - sqrt 5).
// ---------------------------------------------------------------------------
void orc_fibonacci(int32_t m, double* xyz, double* w, double* full_area) {
    const double pi = 3.14159265358979323846;
    const double ga = pi * (3.0 - std::sqrt(5.0));
    double F = 0;
    for (int32_t i = 0; i < m; i++) {
        const double z = 1.0 - (2.0 * i + 1.0) / m;
        const double r = std::sqrt(std::fmax(0.0, 1.0 - z * z));
        const double ph = ga * i;
        xyz[3 * i + 0] = r * std::cos(ph);
        xyz[3 * i + 1] = r * std::sin(ph);
        xyz[3 * i + 2] = z;
        w[i] = 4.0 * pi / m;
        F += w[i];
    }
    *full_area = F;
}

// ---------------------------------------------------------------------------
// Direction text file (reader_txt.cpp:25-47): M, then M lines "w x y z", then F.
// Returns 0 on success, 1 on open/parse failure, 2 if M > cap.
// ---------------------------------------------------------------------------
int orc_read_directions(const char* path, int32_t cap, int32_t* m, double* xyz, double* w, double* full_area) {
    FILE* f = std::fopen(path, "r");
    if (!f) return 1;
    int n = 0;
    if (std::fscanf(f, "%d", &n) != 1 || n < 0) { std::fclose(f); return 1; }
    if (n > cap) { std::fclose(f); return 2; }
    for (int i = 0; i < n; i++) {
        if (std::fscanf(f, "%lf %lf %lf %lf", &w[i], &xyz[3 * i], &xyz[3 * i + 1], &xyz[3 * i + 2]) != 4) {
            std::fclose(f);
            return 1;
        }
    }
    if (std::fscanf(f, "%lf", full_area) != 1) { std::fclose(f); return 1; }
    std::fclose(f);
    *m = n;
    return 0;
}

int orc_max_threads(void) { return omp_get_max_threads(); }

// ---------------------------------------------------------------------------
// Radiation moments of the face intensities, one cell at a time, in the
// arithmetic order of the reference's GPU kernels d_MakeIllumParam =
// d_MakeEnergy + d_MakeDivStream + d_MakeDivImpuls (cuda/cuda_kernel.cu:113-226)
// with their device helpers (cuda/cuda_utils.cu:18-167), built with
// CUDA_FULL_ARRAYS so that the cell stream and impuls are also produced:
//
//   energy[c]        IntegarteDirection       cuda_utils.cu:18-37
//                    (face average per direction, * area, summed, / full_area)
//   Stream_f[c][k]   IntegarteDirection3Faces cuda_utils.cu:132-165
//                    (I * area, then dir[k] * that, summed, / full_area)
//   stream[c]        CUDA_CONVERT_FATE_TO_CELL cuda_kernel.cu:9-17 (face average)
//   divstream[c]     d_MakeDivStream          cuda_kernel.cu:149-178
//                    (sum_f (sum_k Stream_f[k] n_f[k]) * area_f, / volume)
//   Impuls_f[c][hk]  IntegarteDirection9Faces cuda_utils.cu:99-130
//                    (dir[h] * dir[k] * (I * area), summed, / full_area)
//   impuls[c]        face average, 9 values, index h*3+k (row-major; the
//                    tensor is symmetric)
//   divimpuls[c][h]  d_MakeDivImpuls          cuda_kernel.cu:180-217
//                    (sum_f (sum_k Impuls_f[h*3+k] n_f[k]) * area_f, / volume)
//
// Layouts (the reference's device arrays, cuda_struct.cu:63-78,85-107):
// illum[j][c][f], normals[c][f][3] (outward, per cell face), areas[c][f],
// volume[c]; energy[c], stream[c][3], impuls[c][9], divstream[c],
// divimpuls[c][3]. Any output pointer may be null (not computed). The
// divergences need normals, areas and volume; pass null geometry only with
// null divergence outputs.
// ---------------------------------------------------------------------------
void orc_moments(const double* illum, int64_t n_cells, int32_t n_dirs, int32_t faces, const double* dirs,
                 const double* w, double full_area, const double* normals, const double* areas,
                 const double* volume, double* energy, double* stream, double* impuls, double* divstream,
                 double* divimpuls, int32_t threads) {
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n_cells; c++) {
        if (energy) {
            double res = 0;
            for (int32_t i = 0; i < n_dirs; i++) {
                double I = 0;
                for (int32_t f = 0; f < faces; f++) I += illum[faces * (n_cells * i + c) + f];
                I /= faces;
                res += I * w[i];
            }
            energy[c] = res / full_area;
        }
        if (stream || divstream) {
            std::vector<double> S(3 * (size_t)faces, 0.0);
            for (int32_t i = 0; i < n_dirs; i++) {
                const double* d = dirs + 3 * i;
                for (int32_t f = 0; f < faces; f++) {
                    const double I = illum[faces * (n_cells * i + c) + f];
                    const double i_a = I * w[i];
                    S[3 * f + 0] += d[0] * i_a;
                    S[3 * f + 1] += d[1] * i_a;
                    S[3 * f + 2] += d[2] * i_a;
                }
            }
            for (int32_t f = 0; f < faces; f++)
                for (int k = 0; k < 3; k++) S[3 * f + k] /= full_area;
            if (stream) {
                for (int k = 0; k < 3; k++) {
                    double v = 0;
                    for (int32_t f = 0; f < faces; f++) v += S[3 * f + k];
                    stream[3 * c + k] = v / faces;
                }
            }
            if (divstream) {
                double div = 0;
                for (int32_t f = 0; f < faces; f++) {
                    double sum = 0;
                    for (int k = 0; k < 3; k++) sum += S[3 * f + k] * normals[3 * (c * faces + f) + k];
                    div += sum * areas[c * faces + f];
                }
                divstream[c] = div / volume[c];
            }
        }
        if (impuls || divimpuls) {
            std::vector<double> T(9 * (size_t)faces, 0.0);
            for (int32_t i = 0; i < n_dirs; i++) {
                const double* d = dirs + 3 * i;
                for (int32_t f = 0; f < faces; f++)
                    for (int h = 0; h < 3; h++)
                        for (int k = 0; k < 3; k++) {
                            const double I = illum[faces * (n_cells * i + c) + f];
                            T[9 * f + 3 * h + k] += d[h] * d[k] * (I * w[i]);
                        }
            }
            for (int32_t f = 0; f < faces; f++)
                for (int q = 0; q < 9; q++) T[9 * f + q] /= full_area;
            if (impuls) {
                for (int q = 0; q < 9; q++) {
                    double v = 0;
                    for (int32_t f = 0; f < faces; f++) v += T[9 * f + q];
                    impuls[9 * c + q] = v / faces;
                }
            }
            if (divimpuls) {
                double div[3] = {0, 0, 0};
                for (int32_t f = 0; f < faces; f++)
                    for (int h = 0; h < 3; h++) {
                        double sum = 0;
                        for (int k = 0; k < 3; k++) sum += T[9 * f + 3 * h + k] * normals[3 * (c * faces + f) + k];
                        div[h] += sum * areas[c * faces + f];
                    }
                for (int h = 0; h < 3; h++) divimpuls[3 * c + h] = div[h] / volume[c];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Geometry binary files read by InitDevice (cuda/cuda_struct.cu:137-147):
//   simple-bin: int32 count, then count raw elements
//               (ReadSimpleFileBin, file_module/reader_bin.h:13-31);
//   normals:    int32 cells, then per cell `faces` Vector3 (3 doubles)
//               (ReadNormalFile, file_module/reader_bin.cpp:49-69).
// orc_read_simple_bin returns the element count (or -1 on failure) and copies
// at most cap elements of elem_size bytes.
// ---------------------------------------------------------------------------
int64_t orc_read_simple_bin(const char* path, int32_t elem_size, void* out, int64_t cap) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return -1;
    int32_t n = 0;
    if (std::fread(&n, sizeof(int32_t), 1, f) != 1 || n < 0) { std::fclose(f); return -1; }
    const int64_t take = n < cap ? n : cap;
    if ((int64_t)std::fread(out, (size_t)elem_size, (size_t)take, f) != take) { std::fclose(f); return -1; }
    std::fclose(f);
    return n;
}

}  // extern "C"

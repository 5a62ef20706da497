// Internal declarations shared by the HIP kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/rsh_scatter.h"

namespace rsh {

// Thread-local error message behind rsh_last_error().
void set_error(const std::string& msg);
int fail(int status, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define RSH_HIP_TRY(expr)                                     \
    do {                                                      \
        hipError_t _e = (expr);                               \
        if (_e != hipSuccess) return ::rsh::hip_fail(_e, #expr); \
    } while (0)

// ------------------------------------------------------------------ layout
// FP64 phase panel (A operand of v_mfma_f64_16x16x4f64), fragment-major:
//   panel[((pass * steps + s) * rb_per_pass + rb) * 64 + lane]
//       = P[k0 + 16 (pass * rb_per_pass + rb) + lane % 16][4 s + lane / 16] * scale
// for j-steps s (4 input directions each) and row blocks rb (16 output
// directions each); zero outside [k0,k1) x [0,M). One (s, rb) fragment is 512
// contiguous bytes, so a wave reads it with one coalesced load / LDS read.
constexpr int kPanelStepAlign = 4;      // j-steps padded to a multiple of this
constexpr int kMaxRowBlocksPerPass = 25; // 400 output directions per pass
// The grouped iteration kernel's fused epilogue (B and the previous state read
// per output) is latency-bound at one workgroup per CU; passes of at most 13 row
// blocks (208 rows, at 4 waves x 2 j-steps per LDS stage) let two workgroups
// share a CU, so one's epilogue overlaps the other's MFMA work, for 4 %
// padding at M = 200, 400 and 800 and a re-read of the state per pass.
constexpr int kGroupedF64MaxRb = 13;
constexpr int kGroupedF64Waves = 4;
constexpr int kGroupedF64ChunkSteps = 2;
constexpr int64_t kGroupedF64MaxCells = int64_t(1) << 24;  // epilogue offsets: 16 rows x n x 8 B < 2^31

struct F64Plan {
    int rb_per_pass;   // row blocks per workgroup pass (template value)
    int passes;        // grid.y
    int rb_total;      // rb_per_pass * passes (panel rows / 16)
    int steps;         // padded j-steps (multiple of kPanelStepAlign)
};

F64Plan plan_f64(int m, int rows, int max_rb = kMaxRowBlocksPerPass);
int64_t panel_f64_doubles(const F64Plan& pl);

// ------------------------------------------------------------------ kernels
hipError_t launch_build_phase(const double* d_dirs, const double* d_w, int m, double full_area,
                              int kind, double g, int normalize, double* d_P, hipStream_t st);

hipError_t launch_panel_f64(const double* d_P, int m, int k0, int k1, double scale, const F64Plan& pl,
                            double* d_panel, hipStream_t st);

hipError_t launch_direct_f64(const double* d_illum, int faces, int64_t n, int m, const double* d_dirs,
                             const double* d_w, double full_area, const double* d_P, int ref_order,
                             int k0, int k1, double* d_out, int64_t ld, hipStream_t st);

bool mfma_f64_tile_supported(int waves, int chunk_steps);
hipError_t launch_mfma_f64(const double* d_illum, int faces, int64_t n, int64_t ldx, int m, const double* d_panel,
                           const F64Plan& pl, int rows, double* d_out, int64_t ld, int waves,
                           int chunk_steps, hipStream_t st);

// Relative-error reduction of results against an FP64 reference (gx x gy
// workgroups, 3 doubles of partial result each: max, sum, count above).
hipError_t launch_rel_error(const void* got, int got_f32, int64_t ld_got, const double* ref, int64_t ld_ref,
                            int64_t rows, int64_t n, double threshold, double* partial, int gx, int gy,
                            hipStream_t st);

// Register-only MFMA throughput loops (kernels_peak.hip): kind 0 FP64
// 16x16x4, kind 1 BF16 16x16x32; blocks of 4 waves, `iters` iterations each.
hipError_t launch_mfma_peak(int kind, int blocks, int iters, double* sink, hipStream_t st);
double mfma_peak_flops_per_wave_iter(int kind);

// ------------------------------------------------------------------ split forms
struct SplitPlan {
    int rb_per_pass;
    int passes;
    int rb_total;
    int steps;  // K-steps of 32 directions
};
SplitPlan plan_split(int m, int rows);
int split_terms(int form);  // 2 or 3 parts per operand
int split_fmt(int form);    // 0 = BF16, 1 = FP16
int64_t split_state_units(int form, int faces, int64_t n, int m);  // 16-byte units
int64_t split_panel_units(int form, const SplitPlan& pl);
hipError_t launch_cell_scale(const double* d_illum, const double* d_extra, int64_t n, int m, int faces,
                             float* d_inv, hipStream_t st);
hipError_t launch_pack_state(int form, const double* d_illum, int faces, int64_t n, int m, const float* d_inv,
                             void* d_state, hipStream_t st);
hipError_t launch_split_panel(int form, const double* d_P, int m, int k0, int k1, double factor,
                              const SplitPlan& pl, void* d_panel, float* d_row_inv, hipStream_t st);
// Transport sweep (kernels_sweep.hip). Per direction, cells are grouped in
// levels (a cell reads only from earlier levels) and each level is cut into
// passes of kSweepBlock cells, one cell per thread of the direction's
// workgroup. Everything a pass needs that does not depend on earlier results is
// stored per position in level order, so that it can be fetched a pass ahead.
constexpr int kSweepBlock = 512;
constexpr int kSweepRecStride = 10;  // record slots per position (9 used; 80 bytes = five 16-byte loads)
// The face values of the last kSweepRing passes stay in an LDS ring (one slot
// of kSweepBlock x 4 doubles per pass); an upstream value written at most
// kSweepRingDelta passes before the reading pass is read from there, older
// ones from global memory (complete by then: see kernels_sweep.hip).
constexpr int kSweepRing = 4;
constexpr int kSweepRingDelta = 3;
struct SweepPos {
    int32_t cell;
    // bits 0-3: in faces; 4-7: faces out on both sides whose neighbour came
    // earlier (this cell also writes the neighbour's entry); 8 + 2q: in-face id
    // of record slot q (q = 3 r + j: node j of the r-th out face)
    uint32_t meta;
    int32_t pr[4];  // pairs of the cell
    // ring code of in face f in bits 16 (f % 2) of ring[f / 2]: 0 = read the
    // upstream entry from global memory, else (passes back << 9) | index of the
    // upstream cell in its pass; for an out face f the same 16 bits hold, in
    // bit j, class 0's source test for node j of that face (|x - (0.5, 0, 0)|
    // < 0.05, evaluated once when the direction is set)
    uint32_t ring[2];
};
struct SweepDirDev {
    const SweepPos* pos;   // [n] by position in level order
    const double* s9;      // [n][kSweepRecStride] s of record slot q (0 where unused)
    const int32_t* pass;   // [passes][2]: first position, count | (complete earlier stores first) << 31
    int32_t passes;
    int32_t levels;
};
hipError_t launch_sweep(int cls, const SweepDirDev* d_dirs, int m, int64_t n,
                        const double* d_S, int64_t ld, const double* d_old, double* d_new, double* d_norm,
                        hipStream_t st);
// Resident-panel kernel (cell-level state only): plan and launch. Its result
// stores address 16 output rows through one 32-bit buffer range, so the output
// leading dimension must stay below kSplitResidentMaxLd cells.
constexpr int64_t kSplitResidentMaxLd = int64_t(1) << 24;
// grouped: the plan of the grouped iteration kernel (every panel part in LDS),
// else that of the single-problem kernel (two parts in LDS).
bool plan_split_resident(int form, int m, int rows, int cus, SplitPlan* pl, int* sets, bool grouped);
hipError_t launch_split_resident(int form, const void* d_state, int64_t n, const float* d_cell_inv,
                                 const void* d_panel, const float* d_row_inv, const SplitPlan& pl, int sets, int rows,
                                 void* d_out, int64_t ld, int out_f32, hipStream_t st);
// Cell scale and cell-level state in one pass over the input (faces = 4: the
// face average; faces = 1: the values, with d_extra joining the scale's
// maximum); bit-identical to launch_cell_scale followed by
// launch_pack_state_avg (faces = 4) or launch_pack_state (faces = 1). Needs the
// tile's averages in LDS: M <= 1280.
bool pack_fused_supported(int m);
hipError_t launch_pack_fused(int form, const double* d_illum, const double* d_extra, int faces, int64_t n, int m,
                             float* d_inv, void* d_state, hipStream_t st);
// Packs the face average of 4-face FP64 input as a cell-level state.
hipError_t launch_pack_state_avg(int form, const double* d_illum, int64_t n, int m, const float* d_inv,
                                 void* d_state, hipStream_t st);
bool split_tile_supported(int waves, int row_split);
hipError_t launch_split_mfma(int form, const void* d_state, int faces, int64_t n, const float* d_cell_inv,
                             const void* d_panel, const float* d_row_inv, const SplitPlan& pl, int rows,
                             void* d_out, int64_t ld, int out_f32, int waves, int row_split, hipStream_t st);

// ------------------------------------------------------------------ batches
// Device descriptor of one problem of a batch (source iteration on cell-level state).
struct BatchProblem {
    const double* panel;        // FP64 phase panel (faces = 1), pass-major
    double* I[2];               // ping-pong FP64 state [M][N]
    const double* B;            // FP64 form: source term [M][N]
    unsigned long long* resid;  // 2 parity slots, bit pattern of a non-negative double
    int64_t n;
    int m;
    int steps;                  // j-steps (FP64) or K-steps (split) of the panel
    double albedo;
    double one_minus_albedo;
    // split forms
    void* state[2];             // ping-pong packed split state (faces = 1)
    const float* cell_inv;      // per-cell inverse scale of the state
    const void* spanel;         // split phase panel
    const float* row_inv;       // per-row inverse scale of the split panel
    const float* Bs;            // source term rounded to FP32, MFMA output-tile order (launch_source_tiles)
};
struct BatchItem {
    int problem;
    int pass;
    int64_t tile;  // cell tile of 16 * waves cells
};
hipError_t launch_batch_zero(BatchProblem* d_probs, int nprobs, unsigned* d_counters, int ncounters, int parity,
                             hipStream_t st);
hipError_t launch_grouped_iter_f64(int rb, int waves, int chunk_steps, int grid, const BatchProblem* d_probs,
                                   const BatchItem* d_items, const int* d_n_items, unsigned* d_counter, int parity,
                                   hipStream_t st);
int grouped_iter_f64_resident(int rb, int waves, int chunk_steps, int device);
hipError_t launch_grouped_iter_split(int form, int rb, int grid, const BatchProblem* d_probs,
                                     const BatchItem* d_items, const int* d_n_items, unsigned* d_counter, int parity,
                                     hipStream_t st);
int grouped_iter_split_resident(int form, int rb, int device);
// Grouped iteration on split state, resident-panel form (kernels_split_resident.hip):
// items (problem, row group, chunk of iter_split_res_chunk_cells() cells) in 8
// per-XCD lists; meta = [8 list begins, 8 list counts]; counters = 8 per class.
// waves: 8 (two tiles of 16 cells per wave, two waves per SIMD) or 4 (four
// tiles per wave, one wave per SIMD); same items, panel and results.
bool iter_split_res_supported(int form, int rb, int steps);
size_t iter_split_res_lds(int form, int rb, int steps);
hipError_t launch_grouped_iter_split_res(bool launch, int form, int rb, int steps, int waves, int grid,
                                         const BatchProblem* d_probs, const BatchItem* d_items, const int* d_meta,
                                         unsigned* d_counters, int parity, hipStream_t st);
int iter_split_res_chunk_cells();
int grouped_iter_split_waves();
hipError_t launch_unpack_state(int form, const void* d_state, int64_t n, int m, const float* d_cell_inv,
                               double* d_out, hipStream_t st);

// Radiation moments: psi[j][q] = phi_q(omega_j) w_j / F for
// phi = (1, x, y, z, xx, xy, xz, yy, yz, zz).
constexpr int kMomentFuncs = 10;
hipError_t launch_moments_f64(const double* d_illum, int faces, int64_t n, int m, const double* d_psi,
                              const double* d_normals, const double* d_areas, const double* d_volume,
                              double* d_energy, double* d_stream, double* d_impuls, double* d_divstream,
                              double* d_divimpuls, int faces_per_thread, hipStream_t st);

hipError_t launch_fill_uniform(double* d, int64_t count, double lo, double hi, uint64_t seed, hipStream_t st);
// Source term of a split slot: FP32 in MFMA output-tile order (whole row blocks
// of rows_alloc rows and 16-cell tiles, padding 0).
hipError_t launch_source_tiles(const double* src, int m, int64_t n, int rows_alloc, float* dst, hipStream_t st);
hipError_t launch_fill_lognormal(double* d, int64_t count, double mu, double sigma, uint64_t seed,
                                 hipStream_t st);

hipError_t launch_fill_f64(double* d, int64_t count, double value, hipStream_t st);
hipError_t launch_scale_f64(double* d, int64_t count, double factor, hipStream_t st);
hipError_t launch_copy_f64(double* dst, const double* src, int64_t count, hipStream_t st);

hipError_t launch_mfma_layout_probe(const double* d_a, const double* d_b, double* d_d, hipStream_t st);

}  // namespace rsh

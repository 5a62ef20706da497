/*
 * rsh_scatter.h -- C ABI of the MI355X scattering-integral library
 * (librsh_scatter.so, built from relativistic_solver_hllc_amd/csrc/).
 *
 * The library evaluates the radiation-transport scattering integral of
 * Lartsev-Artem/relativistic_solver_hllc,
 *
 *     S[k][c] = sum_j P[k][j] * Ibar[c][j],   Ibar[c][j] = (1/F) sum_f I[j][c][f]
 *
 * with P[k][j] = p(w_k . w_j) * w_j / F (reference convention, p = Thomson
 * 3(1+mu^2)/4, cuda/cuda_utils.cu:7-16) or the FP64 row-normalised form
 * P[k][j] = p w_j / sum_j p w_j (Henyey-Greenstein or Thomson).
 *
 * Reference paths below are relative to solver_hllc_illum_cuda/.
 *
 * Conventions (all entry points):
 *   - every function returns an int status (rsh_status); 0 = success.
 *     rsh_last_error() returns a thread-local message for the last failure.
 *   - sizes and offsets are int64_t (the reference's int offsets overflow at
 *     1.57M cells x 400 directions x 4 faces, e.g. cuda/cuda_kernel.cu:40).
 *   - data layouts are the reference's: face intensities illum[j][c][f]
 *     (direction-major, then cell, then face; cuda/cuda_kernel.cu:40-43,
 *     solve_module/illum/Illum_async.cpp:217) and the scattering integral
 *     S[k - k_begin][c] (direction-major; cuda/cuda_kernel.cu:34).
 *   - "stream" arguments are hipStream_t values passed as void*; the
 *     rsh_problem API instead takes the reference's stream ids (rsh_stream_id).
 *   - no function in this header allocates host memory the caller must free.
 */
#ifndef RSH_SCATTER_H
#define RSH_SCATTER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RSH_ABI_VERSION 1

typedef enum {
    RSH_OK = 0,
    RSH_E_ARG = 1,         /* invalid argument (sizes, ranges, null pointers)   */
    RSH_E_HIP = 2,         /* HIP runtime error                                  */
    RSH_E_NOMEM = 3,       /* device or pinned host allocation failed            */
    RSH_E_STATE = 4,       /* call out of order (e.g. split state not packed)    */
    RSH_E_IO = 5,          /* file could not be opened or parsed                 */
    RSH_E_UNSUPPORTED = 6  /* configuration outside what the kernels implement   */
} rsh_status;

typedef enum { RSH_PHASE_THOMSON = 0, RSH_PHASE_HG = 1 } rsh_phase_kind;

/* Arithmetic form of the contraction. */
typedef enum {
    RSH_FORM_FP64 = 0,     /* FP64 MFMA (v_mfma_f64_16x16x4f64), FP64 in/out     */
    RSH_FORM_BF16X2 = 1,   /* hi+lo BF16 operands, FP32 MFMA accumulation        */
    RSH_FORM_FP16X2 = 2,   /* hi+lo FP16 operands (power-of-two scaled), FP32 acc */
    RSH_FORM_BF16X3 = 3    /* hi+mid+lo BF16 operands, FP32 accumulation         */
} rsh_form;

/* FP64 kernel family. */
typedef enum {
    RSH_F64_AUTO = 0,      /* MFMA GEMM with the tuned tile                      */
    RSH_F64_DIRECT = 1,    /* one thread per (cell, direction), reference order  */
    RSH_F64_MFMA = 2       /* MFMA GEMM, tile given by rsh_tile                  */
} rsh_f64_kernel;

/* Stream ids of the reference (solve_module/solve_global_struct.h:234-240). */
typedef enum {
    RSH_STREAM_SCATTERING_1 = 0,
    RSH_STREAM_SCATTERING_2 = 1,
    RSH_STREAM_PARAMS = 2,
    RSH_STREAM_COUNT = 3
} rsh_stream_id;

/* Tile of the FP64 MFMA kernel: waves per workgroup (16 cells per wave) and
 * j-steps (4 input directions each) per LDS stage of the phase panel. */
typedef struct {
    int32_t waves;        /* 4 or 8          */
    int32_t chunk_steps;  /* 2 or 4          */
} rsh_tile;

int rsh_abi_version(void);
const char* rsh_last_error(void);

/* ---------------------------------------------------------------- host only */

/* Fibonacci-sphere direction set: xyz[m][3], w[m] = 4 pi / m, *full_area = sum w.
 * No reference counterpart (the reference reads directions from a file). */
int rsh_fibonacci_directions(int32_t m, double* xyz, double* w, double* full_area);

/* Direction text file: "M", M lines "area x y z", "full_area".
 * Replaces ReadSphereDirectionDecartToSpherical (file_module/reader_txt.cpp:25-47).
 * cap = capacity of xyz (3*cap doubles) and w; RSH_E_ARG if the file holds more. */
int rsh_read_directions(const char* path, int32_t cap, int32_t* m, double* xyz, double* w,
                        double* full_area);
/* Writes the same format with round-trip (17 significant digit) precision. */
int rsh_write_directions(const char* path, int32_t m, const double* xyz, const double* w,
                         double full_area);

/* Simple-bin file: int32 element count, then the raw elements.
 * Replaces ReadSimpleFileBin / WriteSimpleFileBin (file_module/reader_bin.h:13-31,
 * file_module/writer_bin.h:9-27) and, with elem_size = faces * 24 (faces
 * Vector3 per cell), ReadNormalFile (file_module/reader_bin.cpp:49-69); these
 * are normals.bin, squares.bin and volume.bin read by InitDevice
 * (cuda/cuda_struct.cu:137-147). rsh_read_simple_bin stores the file's count
 * in *count and copies min(count, cap) elements (out may be NULL when cap = 0,
 * to query the count). */
int rsh_read_simple_bin(const char* path, int32_t elem_size, void* out, int64_t cap, int64_t* count);
int rsh_write_simple_bin(const char* path, int32_t elem_size, const void* data, int64_t count);

/* Block partition of n directions over np ranks.
 * Replaces GetSend / GetDisp (solve_module/solve_utils.cpp:177-202). */
int rsh_partition(int32_t np, int32_t n, int32_t* send, int32_t* disp);

/* ---------------------------------------------------------------- device */

int rsh_device_count(int32_t* n);

/* Select the device for the calling thread.
 * Replaces SetDevice (cuda/cuda_solve.cu:46-67) minus stream creation, which
 * rsh_create does per problem. */
int rsh_set_device(int32_t device);

/* Direction set resident on one device: directions, weights and the FP64
 * phase matrix P[M][M], built on the device once (phase-matrix build kernel).
 * Replaces grid_directions_init_copy_host_to_device (cuda/cuda_struct.cu:17-38)
 * together with the per-thread Gamma evaluation of cuda/cuda_kernel.cu:45,106.
 *   normalize = 0 : P = p w_j / full_area   (reference convention)
 *   normalize = 1 : P = p w_j / sum_j p w_j (FP64 row normalisation)       */
typedef struct rsh_dirset_s* rsh_dirset;
int rsh_dirset_create(int32_t device, int32_t m, const double* xyz, const double* w, double full_area,
                      int32_t phase, double g, int32_t normalize, rsh_dirset* out);
int rsh_dirset_destroy(rsh_dirset ds);
/* Copies the FP64 phase matrix rows [k0,k1) (row-major, M columns) to host. */
int rsh_dirset_phase_matrix(rsh_dirset ds, int32_t k0, int32_t k1, double* host_out);

/* FP64 scattering integral on caller-owned device memory.
 *   d_illum : [M][n_cells][faces] doubles (faces = 4 as in the reference, or 1
 *             for cell-averaged input)
 *   d_out   : rows k0..k1-1 written to d_out[(k - k0) * ld_out + c]
 * kernel = rsh_f64_kernel; tile may be NULL (tuned / default tile).
 * Replaces d_GetS / d_GetS_MPI / d_GetS_MPI_Stream (cuda/cuda_kernel.cu:23-111). */
int rsh_scatter_f64_dev(rsh_dirset ds, const double* d_illum, int32_t faces, int64_t n_cells,
                        int32_t k0, int32_t k1, double* d_out, int64_t ld_out, void* stream,
                        int32_t kernel, const rsh_tile* tile);

/* Radiation moments over all M directions of ds, per cell, on caller-owned
 * device memory:
 *   energy[c]       = sum_j w_j Ibar_j / F
 *   stream[c][3]    = face average of sum_j w_j omega_j I_jf / F
 *   impuls[c][9]    = face average of sum_j w_j omega_j omega_j^T I_jf / F
 *                     (row-major; symmetric)
 *   divstream[c]    = sum_f (Stream_f . n_f) area_f / volume
 *   divimpuls[c][3] = sum_f (Impuls_f n_f) area_f / volume
 * with d_illum [M][n_cells][faces], d_normals [n_cells][faces][3] (outward
 * normal of each cell face), d_areas [n_cells][faces], d_volume [n_cells].
 * Any output may be NULL; the geometry may be NULL when both divergences are.
 * faces_per_thread: 0 = default, else 1, 2 or 4 (faces of one cell handled by
 * one thread). Requires full_area > 0 in ds.
 * Replaces d_MakeIllumParam = d_MakeEnergy + d_MakeDivStream + d_MakeDivImpuls
 * (cuda/cuda_kernel.cu:113-226, helpers cuda/cuda_utils.cu:18-167) with
 * CUDA_FULL_ARRAYS (cell stream and impuls also written), and d_MakeStream /
 * d_MakeImpuls (cuda/cuda_kernel.cu:123-147). */
int rsh_moments_f64_dev(rsh_dirset ds, const double* d_illum, int32_t faces, int64_t n_cells,
                        const double* d_normals, const double* d_areas, const double* d_volume,
                        double* d_energy, double* d_stream, double* d_impuls, double* d_divstream,
                        double* d_divimpuls, int32_t faces_per_thread, void* stream);

/* Split-precision forms (RSH_FORM_BF16X2, RSH_FORM_FP16X2, RSH_FORM_BF16X3).
 * No reference counterpart: the reference is FP64 only.
 * The intensities are first packed once into a resident, MFMA-native split
 * state (hi/lo [/mid] 16-bit parts per face, per-cell power-of-two scale),
 * then the contraction reads only that state:
 *   rsh_split_sizes        bytes of the state and of the per-cell scale array
 *   rsh_pack_split_dev     FP64 illum [M][n_cells][faces] -> state + cell scales
 *   rsh_pack_split_avg_dev FP64 illum [M][n_cells][4] -> cell-level state (faces = 1);
 *                          cell-level packs (this one, and rsh_pack_split_dev with
 *                          faces = 1) read the input once for M <= 1280
 *                          of the face average (summed and halved twice in FP64)
 *   rsh_scatter_split_dev  S rows [k0,k1) from the state, FP32 MFMA accumulation,
 *                          written as FP64 (out_f32 = 0) or FP32 (out_f32 = 1)
 *                          to d_out[(k - k0) * ld_out + c]; waves selects the
 *                          kernel: 4 or 8 = streaming-panel kernel with 4 waves
 *                          per workgroup (one per 16-cell tile) or 8 (two per
 *                          tile, each with half of the rows); 1 = resident-panel
 *                          kernel (faces = 1 only: each persistent workgroup
 *                          keeps its rows' panel for all directions in LDS);
 *                          0 = resident-panel kernel where it applies, else 8.
 *                          All kernels give bit-identical results. */
int rsh_split_sizes(int32_t form, int32_t faces, int64_t n_cells, int32_t m, int64_t* state_bytes,
                    int64_t* scale_bytes);
int rsh_pack_split_dev(int32_t form, const double* d_illum, int32_t faces, int64_t n_cells, int32_t m,
                       void* d_state, float* d_cell_scale, void* stream);
int rsh_pack_split_avg_dev(int32_t form, const double* d_illum, int64_t n_cells, int32_t m, void* d_state,
                           float* d_cell_scale, void* stream);
int rsh_scatter_split_dev(rsh_dirset ds, int32_t form, const void* d_state, const float* d_cell_scale,
                          int32_t faces, int64_t n_cells, int32_t k0, int32_t k1, void* d_out, int64_t ld_out,
                          int32_t out_f32, void* stream, int32_t waves);

/* Allocation-free variant of rsh_scatter_f64_dev (MFMA kernel, tuned or
 * default tile). No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under stream capture, tight serving
 * loops, harnesses that forbid hipMalloc). Instead of the direction set's
 * panel cache it builds the phase panel of [k0, k1) into a fixed per-device
 * scratch array of rsh_noalloc_scratch_bytes() bytes (8 MiB: M up to about
 * 1000 directions; larger panels return RSH_E_UNSUPPORTED) and calls no
 * hipMalloc / hipFree. Results are bit-identical to rsh_scatter_f64_dev's.
 * Every call on a device rewrites that device's scratch: calls that overlap on
 * one device (other streams or host threads) race and must be ordered by the
 * caller. */
int rsh_scatter_f64_dev_noalloc(rsh_dirset ds, const double* d_illum, int32_t faces, int64_t n_cells,
                                int32_t k0, int32_t k1, double* d_out, int64_t ld_out, void* stream);
int64_t rsh_noalloc_scratch_bytes(void);

/* Seeded log-normal fill exp(mu + sigma * N(0,1)) on the device (counter-based
 * generator: the value at index i depends only on (seed, i)). */
int rsh_fill_lognormal_dev(double* d_out, int64_t count, double mu, double sigma, uint64_t seed,
                           void* stream);

/* Seeded uniform fill lo + (hi - lo) * U[0,1) (same counter-based generator). */
int rsh_fill_uniform_dev(double* d_out, int64_t count, double lo, double hi, uint64_t seed, void* stream);

/* Checks the lane layout of v_mfma_f64_16x16x4f64 on the device against a
 * host product; *mismatches = number of differing outputs (0 expected). */
int rsh_selftest_mfma_f64(int32_t device, int32_t* mismatches);

/* ------------------------------------------------ device memory utilities */
/* Allocation, copies, streams and events through the library's own HIP
 * runtime, so host programs (and the Python mirror) need no other GPU stack.
 * No reference counterpart beyond the CUDA runtime calls the reference makes
 * inline (cuda/cuda_memory.cuh:5-18).
 * memcpy kind: 0 = host->device, 1 = device->host, 2 = device->device;
 * stream NULL makes the copy synchronous. */
int rsh_dev_alloc(int32_t device, int64_t bytes, void** out);
int rsh_dev_free(void* ptr);
int rsh_host_alloc_pinned(int64_t bytes, void** out);
int rsh_host_free_pinned(void* ptr);
int rsh_dev_memcpy(void* dst, const void* src, int64_t bytes, int32_t kind, void* stream);
int rsh_dev_memcpy2d(void* dst, int64_t dpitch, const void* src, int64_t spitch, int64_t width,
                     int64_t height, int32_t kind);
int rsh_dev_fill_f64(double* d, int64_t count, double value, void* stream);
int rsh_dev_scale_f64(double* d, int64_t count, double factor, void* stream);
/* Plain streaming copy kernel (16-byte loads/stores); reads and writes
 * count*8 bytes, used to calibrate HBM byte counters. */
int rsh_dev_copy_f64(double* dst, const double* src, int64_t count, void* stream);
/* Relative error of every result against an FP64 reference on the device:
 * max and mean over rows x n_cells of |got - ref| / |ref| (got FP64, or FP32
 * when got_f32; row r at d_got + r * ld_got and d_ref + r * ld_ref elements),
 * how many exceed `threshold`, and (argmax, may be NULL) the flat index
 * r * n_cells + c of the largest; a NaN counts as an infinite error.
 * Synchronous on `stream`. No reference counterpart: it states the split
 * forms' error per result over a whole output instead of on sampled cells. */
int rsh_dev_rel_error(const void* d_got, int32_t got_f32, int64_t ld_got, const double* d_ref, int64_t ld_ref,
                      int64_t rows, int64_t n_cells, double threshold, double* max_rel, double* mean_rel,
                      int64_t* count_above, int64_t* argmax, void* stream);
/* Achievable matrix-core rate of the device, TFLOP/s, from register-only loops
 * of independent MFMAs (8 waves per CU, no memory traffic): kind 0 =
 * v_mfma_f64_16x16x4f64 (the FP64 kernels' instruction), 1 =
 * v_mfma_f32_16x16x32_bf16 (the BF16 split forms'). Three warm-up launches,
 * then the fastest of five timed ones (~1 ms each). Measures the peak that
 * the roofline fractions divide by; no reference counterpart. */
int rsh_mfma_peak(int32_t device, int32_t kind, double* tflops);
int rsh_dev_synchronize(int32_t device);
int rsh_dev_mem_info(int32_t device, int64_t* free_bytes, int64_t* total_bytes);
int rsh_stream_create(int32_t device, void** out);
int rsh_stream_destroy(void* stream);
int rsh_stream_synchronize(void* stream);
int rsh_event_create(void** out);
int rsh_event_destroy(void* ev);
int rsh_event_record(void* ev, void* stream);
int rsh_event_elapsed_ms(void* ev_start, void* ev_stop, float* ms);

/* ------------------------------------------------ tile autotuner */
/* Times every supported FP64 MFMA tile on a synthetic problem (n_cells cells,
 * m directions, all m output rows) and records the fastest for the row-block
 * class of m; later single-problem FP64 launches (RSH_F64_AUTO) of that class
 * use it. Batches use the grouped kernel's own tile (4 waves, 2 j-steps per
 * stage, at most 13 row blocks per pass). No reference counterpart. */
int rsh_autotune_f64(int32_t device, int32_t m, int32_t faces, int64_t n_cells, rsh_tile* best, float* best_ms);
int rsh_tuned_tile_f64(int32_t device, int32_t m, rsh_tile* tile);

/* ------------------------------------------------ batched source iteration */
/* Source iteration I <- (1 - a) B + a S(I) on cell-level state [M][n_cells]
 * for many uneven problems at once, with the residual of ReCalcIllum
 * (solve_module/illum/Illum_async.cpp:472-510): max |I_old - I_new| / I_new, in
 * FP64. The reference runs one problem per process and loops on the host
 * (Illum_async.cpp:663-1060); here problems share persistent grouped launches
 * (one per row-block class) captured as a hipGraph, the state stays resident,
 * and problems join and leave between runs (continuous batching).
 *   rsh_batch_footprint  device bytes a problem occupies (admission by HBM;
 *                        slots hold whole row blocks of 16 directions; each
 *                        buffer counted as rsh_batch_reserve's arena carves it,
 *                        rounded up to 256 bytes); RSH_E_UNSUPPORTED for what a
 *                        batch cannot hold (FP64 form with n_cells >= 2^24)
 *   rsh_batch_add        copies I0 and B (device, [M][n_cells] FP64) into a slot
 *                        (FP64 form: n_cells < 2^24, else RSH_E_UNSUPPORTED),
 *                        ordered after all work submitted to the null stream
 *                        (arrays written on other streams must be complete);
 *                        ds holds the directions and phase matrix; form selects
 *                        the FP64 state or a packed split state (RSH_FORM_BF16X2,
 *                        RSH_FORM_FP16X2, RSH_FORM_BF16X3; residual still FP64;
 *                        a split slot keeps B rounded to FP32, the update
 *                        (1 - a) B + a S is formed in FP64)
 *   rsh_batch_run        enqueues `iterations` iterations of every active slot
 *                        on the batch's stream (asynchronous)
 *   rsh_batch_residuals  per slot: residual of its last iteration (-1 if none)
 *                        and iterations done (-1 = inactive); synchronises
 *   rsh_batch_get_state  current state of a slot into d_out ([M][n_cells] FP64) */
typedef struct rsh_batch_s* rsh_batch;
int rsh_batch_footprint(int32_t form, int64_t n_cells, int32_t m, int64_t* bytes);
int rsh_batch_create(int32_t device, int32_t capacity, int32_t use_graph, rsh_batch* out);
int rsh_batch_destroy(rsh_batch b);
/* Kernel of the split-form slots added from now on: 0 = auto (the device's
 * tuned choice for the slot's M, else the resident-panel grouped kernel with
 * 8 waves where its plan fits one XCD and is specialised: M = 200, 400, 800
 * with n_cells < 2^24; else the streaming-panel one), 1 = resident panel,
 * 8 waves of 2 cell tiles (rsh_batch_add fails where it does not apply),
 * 2 = streaming panel, 3 = resident panel, 4 waves of 2 cell tiles, two
 * workgroups per CU where the LDS allows (fails as 1). All give bit-identical
 * states and residuals. No reference counterpart
 * (batching is an extension). */
int rsh_batch_set_split_kernel(rsh_batch b, int32_t kind);

/* Autotuner of the grouped iteration launches. No reference counterpart (the
 * reference's launch shape is fixed: 32 x 16 threads, cuda/cuda_solve.cu:98-99).
 *   rsh_batch_autotune    times every candidate configuration of the batch
 *                         class of (form, m) on a seeded problem of n_cells
 *                         cells, `iterations` graph-replayed iterations each
 *                         in three rounds (the fastest round of each counts),
 *                         and records the fastest for the device: FP64 classes
 *                         (row blocks per pass) choose the MFMA tile (waves,
 *                         j-steps per LDS stage) among (4,2), (4,4), (8,2);
 *                         split forms choose, per M, the resident-panel
 *                         grouped kernel with 8 ({1,0}) or 4 waves ({3,0}) or
 *                         the streaming-panel one ({2,0}).
 *                         Classes (FP64) and slots (split) created afterwards
 *                         in any batch on the device use the choice; every
 *                         candidate gives bit-identical states.
 *   rsh_batch_tuned       the recorded choice (the default if none: FP64 (4,2);
 *                         split {0,0} = resident panel where it applies)
 *   rsh_batch_set_f64_tile forces the FP64 tile of the classes this batch
 *                         creates from now on (NULL or {0,0}: tuned or default)
 *   rsh_batch_slot_config what a slot runs with: its class's FP64 tile, or
 *                         {1,0} / {3,0} / {2,0} for the split kernels */
int rsh_batch_autotune(int32_t device, int32_t form, int32_t m, int64_t n_cells, int32_t iterations, rsh_tile* best,
                       float* best_ms);
int rsh_batch_tuned(int32_t device, int32_t form, int32_t m, rsh_tile* choice);
int rsh_batch_set_f64_tile(rsh_batch b, const rsh_tile* tile);
int rsh_batch_slot_config(rsh_batch b, int32_t slot, rsh_tile* config);

/* Reserves one device arena of `bytes` for the batch's slot buffers (the batch
 * must be empty): later adds carve their buffers from it first-fit instead of
 * calling hipMalloc / hipFree (which synchronise the device). An add that finds
 * no free block large enough returns RSH_E_NOMEM. bytes = 0 releases it. */
int rsh_batch_reserve(rsh_batch b, int64_t bytes);
/* Like rsh_batch_add, with I0 ~ log-normal(0, 1) (seed 2 seed + 1) and
 * B ~ U(0.5, 2) (seed 2 seed + 2) generated on the device in the slot itself
 * (rsh_fill_lognormal_dev / rsh_fill_uniform_dev); asynchronous, ordered on the
 * batch stream. */
int rsh_batch_add_seeded(rsh_batch b, rsh_dirset ds, int64_t n_cells, uint64_t seed, double albedo,
                         int32_t form, int32_t* slot);
int rsh_batch_add(rsh_batch b, rsh_dirset ds, int64_t n_cells, const double* d_I0, const double* d_B,
                  double albedo, int32_t form, int32_t* slot);
int rsh_batch_remove(rsh_batch b, int32_t slot);
int rsh_batch_run(rsh_batch b, int32_t iterations);
int rsh_batch_sync(rsh_batch b);
int rsh_batch_residuals(rsh_batch b, double* resid, int32_t* iterations);
int rsh_batch_get_state(rsh_batch b, int32_t slot, double* d_out);
int rsh_batch_stream(rsh_batch b, void** stream);

/* ------------------------------------------------ serving loop */
/* Continuous-batching server of scattering-solve requests. No reference
 * counterpart (the reference solves one problem per process). A request is a
 * seeded source-iteration problem: n_cells cells, n_dirs Fibonacci directions,
 * phase function, albedo; I0 ~ log-normal(0,1), B ~ U(0.5, 2) generated on the
 * device from the seed; it iterates until its FP64 residual < tolerance (or
 * max_iterations). Each rsh_server_step admits pending requests in arrival
 * order while slots and the HBM budget allow, runs iters_per_step iterations of
 * every resident request in grouped graph launches, and retires finished ones.
 * rsh_server_run_trace replays arrival times in real time and reports each
 * request's latency (completion - arrival). */
typedef enum {
    RSH_REQ_PENDING = 0,
    RSH_REQ_RUNNING = 1,
    RSH_REQ_CONVERGED = 2,
    RSH_REQ_MAXED = 3,
    RSH_REQ_CANCELLED = 4,
    RSH_REQ_FAILED = 5  /* could not be admitted even with no other request resident */
} rsh_request_status;

typedef struct {
    int64_t n_cells;
    int32_t n_dirs;
    int32_t phase;
    double g;
    double albedo;
    double tolerance;
    int32_t max_iterations;
    int32_t form;
    uint64_t seed;
} rsh_request_desc;

typedef struct {
    int32_t status;
    int32_t iterations;
    double residual;
    double t_submit, t_admit, t_done;  /* seconds on the server's monotonic clock */
    int64_t footprint_bytes;
} rsh_request_info;

typedef struct rsh_server_s* rsh_server;
int rsh_server_create(int32_t device, int64_t hbm_budget_bytes, int32_t max_slots, int32_t iters_per_step,
                      int32_t keep_results, rsh_server* out);
int rsh_server_destroy(rsh_server srv);
int rsh_server_submit(rsh_server srv, const rsh_request_desc* desc, int32_t* request_id);
int rsh_server_cancel(rsh_server srv, int32_t request_id);
int rsh_server_step(rsh_server srv, int32_t* running, int32_t* finished_now);
int rsh_server_poll(rsh_server srv, int32_t request_id, rsh_request_info* info);
/* Final state [n_dirs][n_cells] FP64 of a finished request (keep_results = 1). */
int rsh_server_result(rsh_server srv, int32_t request_id, double* d_out);
int rsh_server_stats(rsh_server srv, int64_t* used_bytes, int32_t* pending, int32_t* running, int64_t* completed);
double rsh_server_clock(void);
int rsh_server_run_trace(rsh_server srv, int32_t n, const rsh_request_desc* requests, const double* arrival,
                         double* latency, double* wall_seconds);

/* -------------------------------------------- reference-shaped problem API */

/* One problem = one mesh (n_cells, faces per cell) + one direction set + the
 * local output range [k_begin, k_end) of this rank.
 * rsh_create replaces InitDevice (cuda/cuda_struct.cu:134-178,187-190): it
 * uploads directions, builds the phase matrix, allocates device illum
 * [faces*M*N] and scattering [(k_end-k_begin)*N] and creates the three streams
 * (default priority). Host buffers stay caller-owned (the C++ shim in
 * rsh_cuda_solve.hpp allocates the pinned grid.Illum / grid.scattering). */
typedef struct {
    int32_t device;
    int64_t n_cells;
    int32_t faces;
    int32_t n_dirs;
    const double* dirs_xyz;
    const double* weights;
    double full_area;
    int32_t k_begin, k_end;
    int32_t phase;
    double g;
    int32_t normalize;
} rsh_problem_desc;

typedef struct rsh_problem_s* rsh_problem;

int rsh_create(const rsh_problem_desc* desc, rsh_problem* out);

/* Replaces ClearDevice (cuda/cuda_struct.cu:191-195). */
int rsh_destroy(rsh_problem p);

/* Asynchronous host->device copy of count doubles from host + src_off to
 * device illum + dst_off on the given stream id.  Replaces both overloads of
 * CudaSendIllumAsync (cuda/cuda_solve.cu:69-83; the reference always uses
 * stream 0). The copy is ordered before every later rsh_scatter on any
 * stream of the problem (the reference leaves this unordered across streams). */
int rsh_upload_illum(rsh_problem p, const double* host, int64_t count, int64_t dst_off,
                     int64_t src_off, int32_t stream);

/* Synchronous copy of count doubles to device illum at offset 0.
 * Replaces CopyIllumOnDevice (cuda/cuda_solve.cu:85-88). */
int rsh_copy_illum_sync(rsh_problem p, const double* host, int64_t count);

/* Scattering integral for output directions [k0, k1) (within the problem's
 * [k_begin, k_end)) on the given stream id; when host_out is not NULL the rows
 * are copied asynchronously to host_out + (k0 - k_begin) * n_cells.
 * Replaces CalculateIntScatteringAsyncMPIStream / ...AsyncMPI / ...Async /
 * CalculateIntScattering (cuda/cuda_solve.cu:89-170). The reference's copy
 * offset "+start" (cuda_solve.cu:107) is corrected to "+start*N". */
int rsh_scatter(rsh_problem p, int32_t k0, int32_t k1, int32_t stream, double* host_out);

/* Upload of the whole host illum [M][n_cells][faces] and the scattering
 * integral for [k0, k1), pipelined by cell range: the faces go up in `chunks`
 * cell ranges (2-D copies on stream RSH_STREAM_SCATTERING_1), each range's S is
 * computed as soon as it has landed (RSH_STREAM_SCATTERING_2) and, when
 * host_out is not NULL, copied back (RSH_STREAM_PARAMS) while the next range
 * uploads. Returns when everything is done; the device illum then holds the
 * whole upload, as after rsh_copy_illum_sync. The synchronous sequence of
 * CalculateIntScattering (cuda/cuda_solve.cu:153-170: upload, compute, copy
 * back) with the three steps overlapped; chunks = 1 is that sequence. */
int rsh_scatter_from_host(rsh_problem p, const double* host_illum, int32_t k0, int32_t k1, double* host_out,
                          int32_t chunks);

/* Waits for a stream id; stream < 0 waits for the whole device.
 * Replaces CudaSyncStream and CudaWait (cuda/cuda_solve.cu:37-45). */
int rsh_sync(rsh_problem p, int32_t stream);

/* Device buffers of the problem (for zero-copy use by callers). */
int rsh_problem_buffers(rsh_problem p, double** d_illum, double** d_scattering);

/* hipStream_t of a stream id. */
int rsh_problem_stream(rsh_problem p, int32_t stream, void** hip_stream);

/* Cell geometry for the moment divergences, host arrays copied synchronously:
 * normals [n_cells][faces][3], areas [n_cells][faces], volume [n_cells].
 * Replaces the start == 0 part of InitDevice (cuda/cuda_struct.cu:63-71,
 * 85-107, 169-171), which reads them from normals.bin / squares.bin /
 * volume.bin. */
int rsh_problem_set_geometry(rsh_problem p, const double* normals, const double* areas,
                             const double* volume);

/* Radiation moments of the device illum (all M directions) on a stream id,
 * ordered after every earlier upload. Results stay on the device
 * (rsh_problem_moment_buffers); each non-NULL host pointer receives its array
 * by an asynchronous copy on the same stream (energy [N], stream [N][3],
 * impuls [N][9], divstream [N], divimpuls [N][3]); wait with rsh_sync.
 * The divergences need rsh_problem_set_geometry first.
 * Replaces CalculateAllParamStream (cuda/cuda_solve.cu:297-320) and, with a
 * following rsh_sync, CalculateAllParam / CalculateEnergy / CalculateStream /
 * CalculateImpuls / CalculateDivStream / CalculateDivImpuls
 * (cuda/cuda_solve.cu:172-295). */
int rsh_moments(rsh_problem p, int32_t stream, double* energy, double* stream_out, double* impuls,
                double* divstream, double* divimpuls);

/* Device arrays of the last rsh_moments (NULL before the first call). */
int rsh_problem_moment_buffers(rsh_problem p, double** d_energy, double** d_stream, double** d_impuls,
                               double** d_divstream, double** d_divimpuls);

/* Solution dump in simple-bin files named main_dir + name (plain string
 * concatenation, as the reference does): Illum.bin = face average of the first
 * direction [N]; after rsh_moments also energy.bin [N], stream.bin [N] Vector3,
 * impuls.bin [N] 3x3, and with geometry divstream.bin [N], divimpuls.bin [N]
 * Vector3 (count = N, elements of 8, 24 or 72 bytes). Waits for all work on the
 * device first. Replaces the ILLUM part of WriteFileSolution
 * (file_module/writer_bin.cpp:15-44); its CUDA branch writes only N doubles of
 * stream / impuls / divimpuls (WRITE_FILE with n = grid.size), here the whole
 * arrays are written with the element layout of its CPU branch
 * (WRITE_FILE_VECTOR), which CompareFiles (utils/rebuild_solve.cpp:601-669)
 * reads as Vector3 / Matrix3. */
int rsh_write_solution(rsh_problem p, const char* main_dir);

/* Milliseconds of the last rsh_scatter kernel on each stream id (hipEvents).
 * Replaces CudaEventPrint (cuda/cuda_solve.cu:16-36), whose events the
 * reference never creates. ms must hold RSH_STREAM_COUNT values; streams
 * without a timed launch report -1. */
int rsh_event_times(rsh_problem p, float* ms);

/* ------------------------------------------------------------------ sweep
 * Transport sweep by short characteristics on a tetrahedral mesh: the step
 * that produces the face intensities [M][N][4] the scattering integral reads.
 * Face f of a cell is the face opposite its vertex f; pairs[c*4+f] is the
 * neighbour's face id (cell*4 + face) or a boundary type. */

/* Boundary types of eBoundaryTypes (global_def.h:45-51); BoundaryConditions
 * (solve_module/illum/illum_utils.cpp:407-471) gives 1 on OUT_SOURCE and 0 on
 * the others. */
enum { RSH_BOUND_FREE = -1, RSH_BOUND_LOCK = -2, RSH_BOUND_OUT_SOURCE = -3, RSH_BOUND_INNER_SOURCE = -4 };

/* Unit cube [0,1]^3 in n_edge^3 cubes of 6 tetrahedra (N = 6 n_edge^3 cells,
 * conforming Kuhn split): vertices [N][4][3], pairs [N][4] (faces on x = 0 are
 * OUT_SOURCE, other boundary faces FREE). Synthetic mesh with no reference
 * counterpart (the reference reads VTK meshes); n_edge <= 100. */
int rsh_mesh_cube_tetra(int32_t n_edge, double* vertices, int32_t* pairs);

/* Cell geometry from vertices [N][4][3]; each output may be NULL: outward unit
 * normals [N][4][3] and areas [N][4] (normals.bin / squares.bin layouts),
 * volume [N], and the three nodes of each face [N][4][3][3] (X.bin layout,
 * BasePointTetra, global_def.h:92-99; node j at weight 2/3 on face vertex j). */
int rsh_tetra_geometry(int64_t n, const double* vertices, double* normals, double* areas, double* volume,
                       double* nodes);

/* Per-direction sweep structures for unit direction omega, as the reference's
 * make_illum_struct writes them (make_illum_struct/
 * struct_short_characteristics_main.cpp:128-165): face_state [N] (bit f set =
 * in face, n_f . omega < -1e-10, GetInAndOutFaces), order [N] (an upwind
 * cell order, graph<k>.bin) and, for every node of every out face in that
 * order, the distance s back along -omega to the cell's in-face plane and that
 * face's id (vec_x0 / LocX0<k>.bin records). cap = capacity of x0_s / x0_face;
 * *n_records receives the count needed (RSH_E_NOMEM if above cap). */
int rsh_sweep_prepare_direction(int64_t n, const double* vertices, const int32_t* pairs, const double* omega,
                                int32_t* order, int32_t* face_state, double* x0_s, uint8_t* x0_face, int64_t cap,
                                int64_t* n_records);

/* The reference's MAKE step in its own numbering. Its meshes come from VTK:
 * face f of a cell holds points {0,1,3}, {1,2,3}, {2,0,3}, {0,2,1} (f = 0..3),
 * i.e. the face opposite point 2, 0, 1, 3. */

/* The reference's input arrays from vertices [N][4][3] and pairs [N][4] in this
 * library's numbering (face f opposite vertex f); each output may be NULL:
 * vertex_matrix [N][4][4] (vertex.bin: Eigen Matrix4d, column j = (x, y, z, 1)
 * of point j), faces [N][4][3][3] (grid.bin: the face's points A, B, C in VTK
 * order), outward unit normals [N][4][3] (normals.bin) and pairs_vtk [N][4]
 * (pairs.bin, face ids renumbered, boundary types kept). Synthetic-mesh helper:
 * the reference writes these from its VTK reader (WriteCellFaces, WriteVertex,
 * make_illum_struct/struct_short_characteristics_calculations.cpp:9-56). */
int rsh_mesh_vtk_arrays(int64_t n, const double* vertices, const int32_t* pairs, double* vertex_matrix, double* faces,
                        double* normals, int32_t* pairs_vtk);

/* Face nodes [N][4][3][3] in the reference's placement and order (X.bin,
 * BasePointTetra): node j of face f at weight 2/3 on one point of the face and
 * 1/6 on the other two. Replaces MakeArrayX
 * (make_illum_struct/struct_short_characteristics_logic_function.cpp:10-72). */
int rsh_make_nodes(int64_t n, const double* vertex_matrix, double* nodes);

/* One direction of the reference's MAKE loop (make_illum_struct/
 * struct_short_characteristics_main.cpp:128-165 with GetNodes and
 * CalculateNodeValue, struct_short_characteristics_logic_function.cpp:75-296):
 * cells in the given order (graph<k>.bin), face_state [N] from each cell's own
 * normals (bit f = in face, n_f . omega < -1e-10), and for every node of every
 * out face the first in face whose plane the line along omega meets strictly
 * inside its triangle: x0_s = distance to that point, x0_face = the face,
 * x0_local [2] = the point in the face's local coordinates (0, 0) behind a
 * boundary face (LocX0<k>.bin's cell_local records). RSH_E_ARG if order is not
 * a permutation or a node finds no such face (where the reference stops with
 * "Not found inner face"); RSH_E_NOMEM if cap is below *n_records. */
int rsh_make_direction(int64_t n, const double* vertex_matrix, const double* faces, const double* normals,
                       const int32_t* pairs, const int32_t* order, const double* omega, int32_t* face_state,
                       double* x0_s, double* x0_local, uint8_t* x0_face, int64_t cap, int64_t* n_records);

typedef struct rsh_sweep_s* rsh_sweep;

/* Sweep object for n_cells cells and n_dirs directions on a device: pairs
 * [N][4] (checked: symmetric neighbours or boundary types), nodes [N][4][3][3]
 * (needed only by class 0; may be NULL). Replaces the pairs / X.bin part of
 * ReadIllumGeometry (file_module/reader_bin.cpp:164-203). */
int rsh_sweep_create(int32_t device, int64_t n_cells, int32_t n_dirs, const int32_t* pairs, const double* nodes,
                     rsh_sweep* out);
int rsh_sweep_destroy(rsh_sweep s);

/* Direction k's structures in the reference's formats (sorted_id_cell,
 * face_states, vec_x0's s and in_face_id in traversal order), checked (order a
 * permutation that is upwind, every record naming an in face, record count =
 * 3 x out faces) and uploaded; the cell levels for the parallel sweep are
 * derived here. Replaces the per-direction part of ReadIllumGeometry. */
int rsh_sweep_set_direction(rsh_sweep s, int32_t k, const int32_t* order, const int32_t* face_state,
                            const double* x0_s, const uint8_t* x0_face, int64_t n_records);
/* Number of dependency levels of direction k (cells of one level run in parallel). */
int rsh_sweep_levels(rsh_sweep s, int32_t k, int32_t* levels);

/* One sweep over all directions on a stream (NULL = null stream), device
 * arrays: scattering [M][ld] (grid.scattering), new face intensities
 * d_illum_new [M][N][4]; with d_illum_old (distinct, [M][N][4]) also
 * d_norm [M] = ReCalcIllum's max |old - new| / new per direction. Both
 * illum arrays are 16-byte aligned (RSH_E_ARG otherwise; device allocations are).
 * cls = solve_mode.class_vtk: 0, 1, 2, 3 or 10 (GetCurIllum,
 * Illum_async.cpp:259-430; 11 couples to the gas-dynamics state and is not
 * offered). Replaces the direction loop of MPI_CalculateIllumAsync
 * (solve_module/illum/Illum_async.cpp:725-784) and ReCalcIllum (:472-510). */
int rsh_sweep_run(rsh_sweep s, int32_t cls, const double* d_scattering, int64_t ld, const double* d_illum_old,
                  double* d_illum_new, double* d_norm, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RSH_SCATTER_H */

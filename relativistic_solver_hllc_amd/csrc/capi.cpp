// C-ABI layer of librsh_scatter.so (declarations in include/rsh_scatter.h).
//
// Owns device state: direction sets (directions, weights, FP64 phase matrix,
// cached MFMA panels) and problems (device illum / scattering buffers, the
// three streams of the reference, upload-ordering events, timing events).
// Every entry point returns an rsh_status; HIP failures become RSH_E_HIP with
// the HIP error string in rsh_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>
#include <vector>

#include "rsh_objects.hpp"

namespace rsh {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int status, const std::string& msg) {
    set_error(msg);
    return status;
}

int hip_fail(hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
    set_error(m);
    return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? RSH_E_NOMEM : RSH_E_HIP;
}

}  // namespace rsh

using namespace rsh;

// ------------------------------------------------------------------ objects

namespace rsh {

int check_device(int dev) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    if (dev < 0 || dev >= n) return fail(RSH_E_ARG, "device index " + std::to_string(dev) + " out of range [0," +
                                                       std::to_string(n) + ")");
    return RSH_OK;
}

// Build an entry of the panel cache: on stream st, for rows [k0, k1).
int build_panel(rsh_dirset ds, int k0, int k1, int faces, int form, hipStream_t st, const SplitPlan* resident,
                int f64_max_rb, PanelEntry& pe) {
    pe.device = ds->device;
    hipError_t e = hipSuccess;
    if (form == RSH_FORM_FP64) {
        pe.plan = plan_f64(ds->m, k1 - k0, f64_max_rb);
        const int64_t nd = panel_f64_doubles(pe.plan);
        e = hipMalloc(&pe.d_panel, std::max<int64_t>(nd, 1) * sizeof(double));
        if (e == hipSuccess)
            e = launch_panel_f64(ds->d_P, ds->m, k0, k1, 1.0 / faces, pe.plan, static_cast<double*>(pe.d_panel), st);
    } else {
        pe.splan = resident ? *resident : plan_split(ds->m, k1 - k0);
        const int64_t units = split_panel_units(form, pe.splan);
        e = hipMalloc(&pe.d_panel, std::max<int64_t>(units, 1) * 16);
        if (e == hipSuccess) e = hipMalloc(&pe.d_row_inv, std::max(pe.splan.rb_total * 16, 1) * sizeof(float));
        if (e == hipSuccess)
            e = launch_split_panel(form, ds->d_P, ds->m, k0, k1, 1.0 / faces, pe.splan, pe.d_panel, pe.d_row_inv, st);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&pe.ready, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(pe.ready, st);
    if (e != hipSuccess) return hip_fail(e, "phase panel build");
    return RSH_OK;
}

// Panel for rows [k0,k1), face count and form, built on stream st if absent.
int get_panel(rsh_dirset ds, int k0, int k1, int faces, int form, hipStream_t st, PanelRef* out,
              const SplitPlan* resident, int f64_max_rb) {
    std::lock_guard<std::mutex> lock(ds->mu);
    // layout: FP64 -> pass cap; split -> 0 streaming plan, 100 + row blocks per group for a resident plan
    auto key = std::make_tuple(k0, k1, faces, form,
                               form == RSH_FORM_FP64 ? f64_max_rb : (resident ? 100 + resident->rb_per_pass : 0));
    auto it = ds->panels.find(key);
    if (it != ds->panels.end()) {
        // Make st wait for the build only while it may still be running.
        if (hipEventQuery(it->second->ready) == hipErrorNotReady)
            RSH_HIP_TRY(hipStreamWaitEvent(st, it->second->ready, 0));
        *out = it->second;
        return RSH_OK;
    }
    if (ds->panels.size() >= kPanelCacheCap) {
        // Bounded cache: evict the entries nobody else holds (batch slots and
        // calls still enqueuing keep theirs); each is freed after the device
        // is idle, by its destructor.
        for (auto e = ds->panels.begin(); e != ds->panels.end();) {
            if (e->second.use_count() == 1)
                e = ds->panels.erase(e);
            else
                ++e;
        }
    }
    auto pe = std::make_shared<PanelEntry>();
    int rc = build_panel(ds, k0, k1, faces, form, st, resident, f64_max_rb, *pe);
    if (rc) return rc;
    ds->panels.emplace(key, pe);
    *out = pe;
    return RSH_OK;
}

}  // namespace rsh

PanelEntry::~PanelEntry() {
    if (d_panel || d_row_inv) {
        // Work that read the panel may still be queued on any stream of the
        // device: free only once the device is idle.
        int cur = -1;
        const bool have = hipGetDevice(&cur) == hipSuccess;
        if (hipSetDevice(device) == hipSuccess) {
            hipDeviceSynchronize();
            hipFree(d_panel);
            hipFree(d_row_inv);
        }
        if (have && cur != device) hipSetDevice(cur);
    }
    if (ready) hipEventDestroy(ready);
}

namespace rsh {

// Tuning table: (device, row blocks per pass) -> FP64 MFMA tile.
static std::mutex g_tune_mu;
static std::map<std::pair<int, int>, rsh_tile> g_tune;

rsh_tile f64_tile_for(int device, int rb) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = g_tune.find({device, rb});
    if (it != g_tune.end()) return it->second;
    return rsh_tile{8, 2};
}

void set_tuned_tile(int device, int rb, rsh_tile t) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune[{device, rb}] = t;
}

}  // namespace rsh

namespace {

// n cells of the input, whose direction rows are ldx cells apart (ldx = n for
// a whole array, larger for a cell range of one).
int scatter_f64(rsh_dirset ds, const double* d_illum, int faces, int64_t n, int64_t ldx, int k0, int k1,
                double* d_out, int64_t ld, hipStream_t st, int kernel, const rsh_tile* tile) {
    if (!ds) return fail(RSH_E_ARG, "null direction set");
    if (faces != 1 && faces != 4) return fail(RSH_E_UNSUPPORTED, "faces must be 1 or 4");
    if (n < 0 || k0 < 0 || k1 > ds->m || k0 > k1) return fail(RSH_E_ARG, "invalid cell count or direction range");
    if (n == 0 || k0 == k1) return RSH_OK;
    if (!d_illum || !d_out) return fail(RSH_E_ARG, "null device pointer");
    if (ld < n) return fail(RSH_E_ARG, "ld_out smaller than n_cells");
    if (ldx < n) return fail(RSH_E_ARG, "input row stride smaller than n_cells");
    RSH_HIP_TRY(hipSetDevice(ds->device));
    if (kernel == RSH_F64_DIRECT) {
        if (ldx != n) return fail(RSH_E_UNSUPPORTED, "the reference-order kernel reads whole arrays only");
        RSH_HIP_TRY(launch_direct_f64(d_illum, faces, n, ds->m, ds->d_dirs, ds->d_w, ds->full_area, ds->d_P,
                                      ds->ref_order() ? 1 : 0, k0, k1, d_out, ld, st));
        return RSH_OK;
    }
    if (kernel != RSH_F64_AUTO && kernel != RSH_F64_MFMA) return fail(RSH_E_ARG, "unknown FP64 kernel id");
    rsh_tile t = (tile && kernel == RSH_F64_MFMA) ? *tile
                                                  : f64_tile_for(ds->device, plan_f64(ds->m, k1 - k0).rb_per_pass);
    if (!mfma_f64_tile_supported(t.waves, t.chunk_steps))
        return fail(RSH_E_UNSUPPORTED, "unsupported MFMA tile (waves, chunk_steps)");
    PanelRef pe;  // held until the launch is enqueued
    int rc = get_panel(ds, k0, k1, faces, RSH_FORM_FP64, st, &pe);
    if (rc) return rc;
    RSH_HIP_TRY(launch_mfma_f64(d_illum, faces, n, ldx, ds->m, static_cast<const double*>(pe->d_panel), pe->plan,
                                k1 - k0, d_out, ld, t.waves, t.chunk_steps, st));
    return RSH_OK;
}

int split_default_waves() { return 8; }

}  // namespace

namespace rsh {
int device_cus(int device) {
    static std::mutex mu;
    static std::map<int, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(device);
    if (it != cache.end()) return it->second;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) cus = 256;
    cache[device] = cus;
    return cus;
}
}  // namespace rsh

namespace {

bool valid_split_form(int form) {
    return form == RSH_FORM_BF16X2 || form == RSH_FORM_FP16X2 || form == RSH_FORM_BF16X3;
}

int scatter_split(rsh_dirset ds, int form, const void* d_state, const float* d_cell_inv, int faces, int64_t n,
                  int k0, int k1, void* d_out, int64_t ld, int out_f32, int waves, hipStream_t st) {
    if (waves != 0 && waves != 1 && waves != 4 && waves != 8)
        return fail(RSH_E_UNSUPPORTED, "split kernel: 0 (auto), 1 (resident panel), 4 or 8 (streaming)");
    if (!ds) return fail(RSH_E_ARG, "null direction set");
    if (!valid_split_form(form)) return fail(RSH_E_ARG, "not a split form");
    if (faces != 1 && faces != 4) return fail(RSH_E_UNSUPPORTED, "faces must be 1 or 4");
    if (n < 0 || k0 < 0 || k1 > ds->m || k0 > k1) return fail(RSH_E_ARG, "invalid cell count or direction range");
    if (n == 0 || k0 == k1) return RSH_OK;
    if (!d_state || !d_cell_inv || !d_out) return fail(RSH_E_ARG, "null device pointer");
    if (ld < n) return fail(RSH_E_ARG, "ld_out smaller than n_cells");
    RSH_HIP_TRY(hipSetDevice(ds->device));
    SplitPlan rplan{};
    int sets = 0;
    const bool resident_ok = faces == 1 && ld < kSplitResidentMaxLd &&
                             plan_split_resident(form, ds->m, k1 - k0, device_cus(ds->device), &rplan, &sets, false);
    if (waves == 1 && !resident_ok)
        return fail(RSH_E_UNSUPPORTED, "resident-panel kernel needs cell-level state (faces = 1), ld_out below 2^24 "
                                       "and a panel slice that fits LDS with at most one row group per CU of an XCD");
    // Auto: the resident-panel kernel when its row groups keep at least 3/4 of
    // the CUs busy (M = 800 needs 17 groups of 3 row blocks: 17 of 32 CUs per
    // XCD), else the streaming kernel with 8 waves.
    const int per_xcd = device_cus(ds->device) / 8;
    if (waves == 0)
        waves = (resident_ok && 4 * sets * rplan.passes >= 3 * per_xcd) ? 1 : split_default_waves();
    PanelRef pe;  // held until the launch is enqueued
    if (waves == 1) {
        int rc = get_panel(ds, k0, k1, faces, form, st, &pe, &rplan);
        if (rc) return rc;
        RSH_HIP_TRY(launch_split_resident(form, d_state, n, d_cell_inv, pe->d_panel, pe->d_row_inv, pe->splan, sets,
                                          k1 - k0, d_out, ld, out_f32 ? 1 : 0, st));
        return RSH_OK;
    }
    int rc = get_panel(ds, k0, k1, faces, form, st, &pe);
    if (rc) return rc;
    RSH_HIP_TRY(launch_split_mfma(form, d_state, faces, n, d_cell_inv, pe->d_panel, pe->d_row_inv, pe->splan,
                                  k1 - k0, d_out, ld, out_f32 ? 1 : 0, waves, waves == 8 ? 2 : 1, st));
    return RSH_OK;
}

}  // namespace

extern "C" {

int rsh_abi_version(void) { return RSH_ABI_VERSION; }

const char* rsh_last_error(void) { return g_last_error.c_str(); }

// ------------------------------------------------------------------ host only

int rsh_fibonacci_directions(int32_t m, double* xyz, double* w, double* full_area) {
    if (m <= 0 || !xyz || !w || !full_area) return fail(RSH_E_ARG, "invalid Fibonacci arguments");
    const double pi = 3.14159265358979323846;
    const double ga = pi * (3.0 - std::sqrt(5.0));
    double F = 0;
    for (int32_t i = 0; i < m; i++) {
        const double z = 1.0 - (2.0 * i + 1.0) / m;
        const double r = std::sqrt(std::fmax(0.0, 1.0 - z * z));
        const double ph = ga * i;
        xyz[3 * i] = r * std::cos(ph);
        xyz[3 * i + 1] = r * std::sin(ph);
        xyz[3 * i + 2] = z;
        w[i] = 4.0 * pi / m;
        F += w[i];
    }
    *full_area = F;
    return RSH_OK;
}

int rsh_read_directions(const char* path, int32_t cap, int32_t* m, double* xyz, double* w, double* full_area) {
    if (!path || !m || !full_area) return fail(RSH_E_ARG, "null argument");
    FILE* f = std::fopen(path, "r");
    if (!f) return fail(RSH_E_IO, std::string("cannot open direction file ") + path);
    long n = -1;
    int rc = RSH_OK;
    if (std::fscanf(f, "%ld", &n) != 1 || n < 0) rc = fail(RSH_E_IO, "bad direction count");
    else if (n > cap || (n > 0 && (!xyz || !w))) rc = fail(RSH_E_ARG, "direction file holds more than cap entries");
    for (long i = 0; rc == RSH_OK && i < n; i++)
        if (std::fscanf(f, "%lf %lf %lf %lf", &w[i], &xyz[3 * i], &xyz[3 * i + 1], &xyz[3 * i + 2]) != 4)
            rc = fail(RSH_E_IO, "truncated direction line " + std::to_string(i));
    if (rc == RSH_OK && std::fscanf(f, "%lf", full_area) != 1) rc = fail(RSH_E_IO, "missing full_area");
    std::fclose(f);
    if (rc == RSH_OK) *m = (int32_t)n;
    return rc;
}

int rsh_write_directions(const char* path, int32_t m, const double* xyz, const double* w, double full_area) {
    if (!path || m < 0 || (m > 0 && (!xyz || !w))) return fail(RSH_E_ARG, "invalid arguments");
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(RSH_E_IO, std::string("cannot create ") + path);
    std::fprintf(f, "%d\n", m);
    for (int32_t i = 0; i < m; i++)
        std::fprintf(f, "%.17g %.17g %.17g %.17g\n", w[i], xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    std::fprintf(f, "%.17g\n", full_area);
    const bool ok = std::fclose(f) == 0;
    return ok ? RSH_OK : fail(RSH_E_IO, "write failed");
}

int rsh_read_simple_bin(const char* path, int32_t elem_size, void* out, int64_t cap, int64_t* count) {
    if (!path || !count || elem_size <= 0 || cap < 0 || (cap > 0 && !out)) return fail(RSH_E_ARG, "invalid arguments");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(RSH_E_IO, std::string("cannot open ") + path);
    int32_t n = 0;
    int rc = RSH_OK;
    if (std::fread(&n, sizeof(int32_t), 1, f) != 1 || n < 0) {
        rc = fail(RSH_E_IO, std::string("bad element count in ") + path);
    } else {
        const int64_t take = std::min<int64_t>(n, cap);
        if (take > 0 && (int64_t)std::fread(out, (size_t)elem_size, (size_t)take, f) != take)
            rc = fail(RSH_E_IO, std::string("truncated file ") + path);
    }
    std::fclose(f);
    if (rc == RSH_OK) *count = n;
    return rc;
}

int rsh_write_simple_bin(const char* path, int32_t elem_size, const void* data, int64_t count) {
    if (!path || elem_size <= 0 || count < 0 || count > INT32_MAX || (count > 0 && !data))
        return fail(RSH_E_ARG, "invalid arguments (the format's count is a 32-bit int)");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(RSH_E_IO, std::string("cannot create ") + path);
    const int32_t n = (int32_t)count;
    bool ok = std::fwrite(&n, sizeof(int32_t), 1, f) == 1;
    if (ok && count > 0) ok = (int64_t)std::fwrite(data, (size_t)elem_size, (size_t)count, f) == count;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? RSH_OK : fail(RSH_E_IO, std::string("write failed: ") + path);
}

int rsh_partition(int32_t np, int32_t n, int32_t* send, int32_t* disp) {
    if (np <= 0 || n < 0 || !send || !disp) return fail(RSH_E_ARG, "invalid partition arguments");
    const int32_t q = n / np, r = n % np;
    int32_t off = 0;
    for (int32_t i = 0; i < np; i++) {
        send[i] = q + (i < r ? 1 : 0);
        disp[i] = off;
        off += send[i];
    }
    return RSH_OK;
}

// ------------------------------------------------------------------ device

int rsh_device_count(int32_t* n) {
    if (!n) return fail(RSH_E_ARG, "null argument");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e == hipErrorNoDevice) c = 0;
    else if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    *n = c;
    return RSH_OK;
}

int rsh_set_device(int32_t device) {
    int rc = check_device(device);
    if (rc) return rc;
    RSH_HIP_TRY(hipSetDevice(device));
    return RSH_OK;
}

int rsh_dirset_create(int32_t device, int32_t m, const double* xyz, const double* w, double full_area,
                      int32_t phase, double g, int32_t normalize, rsh_dirset* out) {
    if (!out || m <= 0 || !xyz || !w) return fail(RSH_E_ARG, "invalid direction set arguments");
    if (phase != RSH_PHASE_THOMSON && phase != RSH_PHASE_HG) return fail(RSH_E_ARG, "unknown phase function");
    if (phase == RSH_PHASE_HG && !(g > -1.0 && g < 1.0)) return fail(RSH_E_ARG, "HG asymmetry g must be in (-1,1)");
    if (!normalize && !(full_area > 0)) return fail(RSH_E_ARG, "full_area must be positive");
    *out = nullptr;
    int rc = check_device(device);
    if (rc) return rc;
    RSH_HIP_TRY(hipSetDevice(device));
    auto ds = std::make_unique<rsh_dirset_s>();
    ds->device = device;
    ds->m = m;
    ds->full_area = full_area;
    ds->phase = phase;
    ds->g = g;
    ds->normalize = normalize ? 1 : 0;
    ds->xyz.assign(xyz, xyz + 3 * (size_t)m);
    ds->w.assign(w, w + m);
    auto cleanup = [&](int status) {
        hipFree(ds->d_dirs);
        hipFree(ds->d_w);
        hipFree(ds->d_P);
        hipFree(ds->d_psi);
        return status;
    };
    hipError_t e;
    if ((e = hipMalloc(&ds->d_dirs, 3 * (size_t)m * sizeof(double))) != hipSuccess ||
        (e = hipMalloc(&ds->d_w, (size_t)m * sizeof(double))) != hipSuccess ||
        (e = hipMalloc(&ds->d_P, (size_t)m * m * sizeof(double))) != hipSuccess)
        return cleanup(hip_fail(e, "hipMalloc(direction set)"));
    if ((e = hipMemcpy(ds->d_dirs, xyz, 3 * (size_t)m * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(ds->d_w, w, (size_t)m * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
        return cleanup(hip_fail(e, "hipMemcpy(direction set)"));
    if ((e = launch_build_phase(ds->d_dirs, ds->d_w, m, full_area, phase, g, ds->normalize, ds->d_P, nullptr)) !=
            hipSuccess ||
        (e = hipDeviceSynchronize()) != hipSuccess)
        return cleanup(hip_fail(e, "build_phase_kernel"));
    if (full_area > 0) {
        // Moment weights psi[j][q] = phi_q(omega_j) * (w_j / F), phi = (1, x, y, z, xx, xy, xz, yy, yz, zz).
        std::vector<double> psi((size_t)m * kMomentFuncs);
        for (int j = 0; j < m; j++) {
            const double a = w[j] / full_area, x = xyz[3 * j], y = xyz[3 * j + 1], z = xyz[3 * j + 2];
            const double phi[kMomentFuncs] = {1.0, x, y, z, x * x, x * y, x * z, y * y, y * z, z * z};
            for (int q = 0; q < kMomentFuncs; q++) psi[(size_t)j * kMomentFuncs + q] = phi[q] * a;
        }
        if ((e = hipMalloc(&ds->d_psi, psi.size() * sizeof(double))) != hipSuccess ||
            (e = hipMemcpy(ds->d_psi, psi.data(), psi.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
            return cleanup(hip_fail(e, "moment weights"));
    }
    *out = ds.release();
    return RSH_OK;
}

int rsh_dirset_destroy(rsh_dirset ds) {
    if (!ds) return RSH_OK;
    hipSetDevice(ds->device);
    hipDeviceSynchronize();
    ds->panels.clear();  // entries still held by batch slots live on until those slots go
    hipFree(ds->d_dirs);
    hipFree(ds->d_w);
    hipFree(ds->d_P);
    hipFree(ds->d_psi);
    delete ds;
    return RSH_OK;
}

int rsh_dirset_phase_matrix(rsh_dirset ds, int32_t k0, int32_t k1, double* host_out) {
    if (!ds || !host_out || k0 < 0 || k1 > ds->m || k0 > k1) return fail(RSH_E_ARG, "invalid arguments");
    RSH_HIP_TRY(hipSetDevice(ds->device));
    RSH_HIP_TRY(hipMemcpy(host_out, ds->d_P + (int64_t)k0 * ds->m, (size_t)(k1 - k0) * ds->m * sizeof(double),
                          hipMemcpyDeviceToHost));
    return RSH_OK;
}

int rsh_scatter_f64_dev(rsh_dirset ds, const double* d_illum, int32_t faces, int64_t n_cells, int32_t k0,
                        int32_t k1, double* d_out, int64_t ld_out, void* stream, int32_t kernel,
                        const rsh_tile* tile) {
    return scatter_f64(ds, d_illum, faces, n_cells, n_cells, k0, k1, d_out, ld_out, (hipStream_t)stream, kernel, tile);
}

static int moments_f64(rsh_dirset ds, const double* d_illum, int faces, int64_t n, const double* d_normals,
                       const double* d_areas, const double* d_volume, double* d_energy, double* d_stream,
                       double* d_impuls, double* d_divstream, double* d_divimpuls, int fpt, hipStream_t st) {
    if (!ds || (faces != 1 && faces != 4) || n < 0) return fail(RSH_E_ARG, "invalid moment arguments");
    if (fpt == 0) fpt = 2;
    if (fpt != 1 && fpt != 2 && fpt != 4) return fail(RSH_E_ARG, "faces_per_thread must be 0, 1, 2 or 4");
    if (!ds->d_psi) return fail(RSH_E_ARG, "moments need a direction set with full_area > 0");
    if ((d_divstream || d_divimpuls) && (!d_normals || !d_areas || !d_volume))
        return fail(RSH_E_ARG, "divergences need normals, areas and volume");
    if (n == 0) return RSH_OK;
    if (!d_illum) return fail(RSH_E_ARG, "null device illum");
    if (reinterpret_cast<uintptr_t>(d_illum) % 16 != 0) fpt = 1;  // 16-byte loads need 16-byte alignment
    RSH_HIP_TRY(hipSetDevice(ds->device));
    RSH_HIP_TRY(launch_moments_f64(d_illum, faces, n, ds->m, ds->d_psi, d_normals, d_areas, d_volume, d_energy,
                                   d_stream, d_impuls, d_divstream, d_divimpuls, fpt, st));
    return RSH_OK;
}

int rsh_moments_f64_dev(rsh_dirset ds, const double* d_illum, int32_t faces, int64_t n_cells,
                        const double* d_normals, const double* d_areas, const double* d_volume, double* d_energy,
                        double* d_stream, double* d_impuls, double* d_divstream, double* d_divimpuls,
                        int32_t faces_per_thread, void* stream) {
    return moments_f64(ds, d_illum, faces, n_cells, d_normals, d_areas, d_volume, d_energy, d_stream, d_impuls,
                       d_divstream, d_divimpuls, faces_per_thread, (hipStream_t)stream);
}

int rsh_split_sizes(int32_t form, int32_t faces, int64_t n_cells, int32_t m, int64_t* state_bytes,
                    int64_t* scale_bytes) {
    if (!valid_split_form(form) || (faces != 1 && faces != 4) || n_cells < 0 || m <= 0 || !state_bytes ||
        !scale_bytes)
        return fail(RSH_E_ARG, "invalid split size arguments");
    *state_bytes = split_state_units(form, faces, n_cells, m) * 16;
    *scale_bytes = n_cells * (int64_t)sizeof(float);
    return RSH_OK;
}

int rsh_pack_split_dev(int32_t form, const double* d_illum, int32_t faces, int64_t n_cells, int32_t m,
                       void* d_state, float* d_cell_scale, void* stream) {
    if (!valid_split_form(form) || (faces != 1 && faces != 4) || n_cells < 0 || m <= 0)
        return fail(RSH_E_ARG, "invalid split pack arguments");
    if (n_cells == 0) return RSH_OK;
    if (!d_illum || !d_state || !d_cell_scale) return fail(RSH_E_ARG, "null device pointer");
    hipStream_t st = (hipStream_t)stream;
    if (faces == 1 && pack_fused_supported(m)) {  // one pass over the input
        RSH_HIP_TRY(launch_pack_fused(form, d_illum, nullptr, 1, n_cells, m, d_cell_scale, d_state, st));
        return RSH_OK;
    }
    RSH_HIP_TRY(launch_cell_scale(d_illum, nullptr, n_cells, m, faces, d_cell_scale, st));
    RSH_HIP_TRY(launch_pack_state(form, d_illum, faces, n_cells, m, d_cell_scale, d_state, st));
    return RSH_OK;
}

int rsh_pack_split_avg_dev(int32_t form, const double* d_illum, int64_t n_cells, int32_t m, void* d_state,
                           float* d_cell_scale, void* stream) {
    if (!valid_split_form(form) || n_cells < 0 || m <= 0) return fail(RSH_E_ARG, "invalid split pack arguments");
    if (n_cells == 0) return RSH_OK;
    if (!d_illum || !d_state || !d_cell_scale) return fail(RSH_E_ARG, "null device pointer");
    hipStream_t st = (hipStream_t)stream;
    if (pack_fused_supported(m)) {  // one pass over the input
        RSH_HIP_TRY(launch_pack_fused(form, d_illum, nullptr, 4, n_cells, m, d_cell_scale, d_state, st));
        return RSH_OK;
    }
    RSH_HIP_TRY(launch_cell_scale(d_illum, nullptr, n_cells, m, 4, d_cell_scale, st));
    RSH_HIP_TRY(launch_pack_state_avg(form, d_illum, n_cells, m, d_cell_scale, d_state, st));
    return RSH_OK;
}

int rsh_scatter_split_dev(rsh_dirset ds, int32_t form, const void* d_state, const float* d_cell_scale,
                          int32_t faces, int64_t n_cells, int32_t k0, int32_t k1, void* d_out, int64_t ld_out,
                          int32_t out_f32, void* stream, int32_t waves) {
    return scatter_split(ds, form, d_state, d_cell_scale, faces, n_cells, k0, k1, d_out, ld_out, out_f32, waves,
                         (hipStream_t)stream);
}

int rsh_fill_uniform_dev(double* d_out, int64_t count, double lo, double hi, uint64_t seed, void* stream) {
    if (count < 0 || (count > 0 && !d_out)) return fail(RSH_E_ARG, "invalid fill arguments");
    RSH_HIP_TRY(launch_fill_uniform(d_out, count, lo, hi, seed, (hipStream_t)stream));
    return RSH_OK;
}

int rsh_fill_lognormal_dev(double* d_out, int64_t count, double mu, double sigma, uint64_t seed, void* stream) {
    if (count < 0 || (count > 0 && !d_out)) return fail(RSH_E_ARG, "invalid fill arguments");
    RSH_HIP_TRY(launch_fill_lognormal(d_out, count, mu, sigma, seed, (hipStream_t)stream));
    return RSH_OK;
}

int rsh_selftest_mfma_f64(int32_t device, int32_t* mismatches) {
    if (!mismatches) return fail(RSH_E_ARG, "null argument");
    int rc = check_device(device);
    if (rc) return rc;
    RSH_HIP_TRY(hipSetDevice(device));
    double A[16 * 4], B[4 * 16], D[64 * 4];
    for (int i = 0; i < 64; i++) {
        A[i] = 1.0 + i;         // A[r][k] = 1 + 4r + k
        B[i] = 1.0 / (1 + i);   // B[k][c]
    }
    double *dA = nullptr, *dB = nullptr, *dD = nullptr;
    RSH_HIP_TRY(hipMalloc(&dA, sizeof A));
    RSH_HIP_TRY(hipMalloc(&dB, sizeof B));
    RSH_HIP_TRY(hipMalloc(&dD, sizeof D));
    hipMemcpy(dA, A, sizeof A, hipMemcpyHostToDevice);
    hipMemcpy(dB, B, sizeof B, hipMemcpyHostToDevice);
    hipError_t e = launch_mfma_layout_probe(dA, dB, dD, nullptr);
    if (e == hipSuccess) e = hipMemcpy(D, dD, sizeof D, hipMemcpyDeviceToHost);
    hipFree(dA);
    hipFree(dB);
    hipFree(dD);
    if (e != hipSuccess) return hip_fail(e, "mfma layout probe");
    int bad = 0;
    for (int l = 0; l < 64; l++)
        for (int r = 0; r < 4; r++) {
            const int row = l / 16 + 4 * r, col = l % 16;  // D layout of v_mfma_f64_16x16x4f64
            double ref = 0;
            for (int k = 0; k < 4; k++) ref += A[row * 4 + k] * B[k * 16 + col];
            if (std::fabs(D[l * 4 + r] - ref) > 1e-12 * std::fabs(ref)) bad++;
        }
    *mismatches = bad;
    return RSH_OK;
}

// ------------------------------------------------------------------ device memory utilities

int rsh_dev_alloc(int32_t device, int64_t bytes, void** out) {
    if (!out || bytes < 0) return fail(RSH_E_ARG, "invalid allocation arguments");
    *out = nullptr;
    int rc = check_device(device);
    if (rc) return rc;
    RSH_HIP_TRY(hipSetDevice(device));
    RSH_HIP_TRY(hipMalloc(out, (size_t)std::max<int64_t>(bytes, 1)));
    return RSH_OK;
}

int rsh_dev_free(void* ptr) {
    if (ptr) RSH_HIP_TRY(hipFree(ptr));
    return RSH_OK;
}

int rsh_host_alloc_pinned(int64_t bytes, void** out) {
    if (!out || bytes < 0) return fail(RSH_E_ARG, "invalid allocation arguments");
    *out = nullptr;
    RSH_HIP_TRY(hipHostMalloc(out, (size_t)std::max<int64_t>(bytes, 1), hipHostMallocDefault));
    return RSH_OK;
}

int rsh_host_free_pinned(void* ptr) {
    if (ptr) RSH_HIP_TRY(hipHostFree(ptr));
    return RSH_OK;
}

static int memcpy_kind(int32_t kind, hipMemcpyKind* k) {
    if (kind == 0) *k = hipMemcpyHostToDevice;
    else if (kind == 1) *k = hipMemcpyDeviceToHost;
    else if (kind == 2) *k = hipMemcpyDeviceToDevice;
    else return fail(RSH_E_ARG, "memcpy kind must be 0 (H2D), 1 (D2H) or 2 (D2D)");
    return RSH_OK;
}

int rsh_dev_memcpy(void* dst, const void* src, int64_t bytes, int32_t kind, void* stream) {
    hipMemcpyKind k;
    if (int rc = memcpy_kind(kind, &k)) return rc;
    if (bytes < 0) return fail(RSH_E_ARG, "negative byte count");
    if (bytes == 0) return RSH_OK;
    if (!dst || !src) return fail(RSH_E_ARG, "null pointer");
    if (stream) RSH_HIP_TRY(hipMemcpyAsync(dst, src, (size_t)bytes, k, (hipStream_t)stream));
    else RSH_HIP_TRY(hipMemcpy(dst, src, (size_t)bytes, k));
    return RSH_OK;
}

int rsh_dev_memcpy2d(void* dst, int64_t dpitch, const void* src, int64_t spitch, int64_t width, int64_t height,
                     int32_t kind) {
    hipMemcpyKind k;
    if (int rc = memcpy_kind(kind, &k)) return rc;
    if (width < 0 || height < 0 || dpitch < width || spitch < width) return fail(RSH_E_ARG, "invalid 2D copy");
    if (width == 0 || height == 0) return RSH_OK;
    if (!dst || !src) return fail(RSH_E_ARG, "null pointer");
    RSH_HIP_TRY(hipMemcpy2D(dst, (size_t)dpitch, src, (size_t)spitch, (size_t)width, (size_t)height, k));
    return RSH_OK;
}

int rsh_dev_fill_f64(double* d, int64_t count, double value, void* stream) {
    if (count < 0 || (count > 0 && !d)) return fail(RSH_E_ARG, "invalid fill arguments");
    RSH_HIP_TRY(launch_fill_f64(d, count, value, (hipStream_t)stream));
    return RSH_OK;
}

int rsh_dev_scale_f64(double* d, int64_t count, double factor, void* stream) {
    if (count < 0 || (count > 0 && !d)) return fail(RSH_E_ARG, "invalid scale arguments");
    RSH_HIP_TRY(launch_scale_f64(d, count, factor, (hipStream_t)stream));
    return RSH_OK;
}

int rsh_dev_copy_f64(double* dst, const double* src, int64_t count, void* stream) {
    if (count < 0 || (count > 0 && (!dst || !src))) return fail(RSH_E_ARG, "invalid copy arguments");
    if (((uintptr_t)dst | (uintptr_t)src) & 15) return fail(RSH_E_ARG, "copy kernel needs 16-byte aligned pointers");
    RSH_HIP_TRY(launch_copy_f64(dst, src, count, (hipStream_t)stream));
    return RSH_OK;
}

int rsh_dev_rel_error(const void* d_got, int32_t got_f32, int64_t ld_got, const double* d_ref, int64_t ld_ref,
                      int64_t rows, int64_t n_cells, double threshold, double* max_rel, double* mean_rel,
                      int64_t* count_above, int64_t* argmax, void* stream) {
    if (!max_rel || !mean_rel || !count_above || rows < 0 || n_cells < 0 || ld_got < n_cells || ld_ref < n_cells)
        return fail(RSH_E_ARG, "invalid relative-error arguments");
    *max_rel = 0.0;
    *mean_rel = 0.0;
    *count_above = 0;
    if (argmax) *argmax = -1;
    if (rows == 0 || n_cells == 0) return RSH_OK;
    if (!d_got || !d_ref) return fail(RSH_E_ARG, "null device pointer");
    const int gx = (int)std::min<int64_t>((n_cells + 255) / 256, 64);
    const int gy = (int)std::min<int64_t>(rows, 128);
    double* d_part = nullptr;
    RSH_HIP_TRY(hipMalloc(&d_part, (size_t)gx * gy * 4 * sizeof(double)));
    hipError_t e = launch_rel_error(d_got, got_f32 ? 1 : 0, ld_got, d_ref, ld_ref, rows, n_cells, threshold, d_part,
                                    gx, gy, (hipStream_t)stream);
    std::vector<double> part((size_t)gx * gy * 4);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e == hipSuccess) e = hipMemcpy(part.data(), d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost);
    hipFree(d_part);
    if (e != hipSuccess) return hip_fail(e, "relative-error reduction");
    double mx = -1.0, sum = 0.0, above = 0.0, at = -1.0;
    for (size_t b = 0; b < part.size(); b += 4) {
        if (part[b] > mx) {
            mx = part[b];
            at = part[b + 3];
        }
        sum += part[b + 1];
        above += part[b + 2];
    }
    *max_rel = mx < 0.0 ? 0.0 : mx;
    *mean_rel = sum / ((double)rows * (double)n_cells);
    *count_above = (int64_t)above;
    if (argmax) *argmax = (int64_t)at;
    return RSH_OK;
}

int rsh_mfma_peak(int32_t device, int32_t kind, double* tflops) {
    if (!tflops || (kind != 0 && kind != 1)) return fail(RSH_E_ARG, "kind: 0 = FP64 16x16x4, 1 = BF16 16x16x32");
    *tflops = 0.0;
    int rc = check_device(device);
    if (rc) return rc;
    RSH_HIP_TRY(hipSetDevice(device));
    const int blocks = 2 * device_cus(device);   // 8 waves per CU: 2 per SIMD
    const int iters = kind == 0 ? 2048 : 8192;   // ~1 ms per launch at the vendor peak (multiple of the loop's unroll)
    const int reps = 5;
    double* sink = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&sink, (size_t)blocks * 256 * sizeof(double));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    // warm-up launches bring the engine clock up; then the fastest of reps timed launches
    for (int r = 0; r < 3 && e == hipSuccess; r++) e = launch_mfma_peak(kind, blocks, iters, sink, st);
    float ms = 0.f;
    for (int r = 0; r < reps && e == hipSuccess; r++) {
        e = hipEventRecord(e0, st);
        if (e == hipSuccess) e = launch_mfma_peak(kind, blocks, iters, sink, st);
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float t = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e == hipSuccess && (r == 0 || t < ms)) ms = t;
    }
    hipFree(sink);
    if (st) hipStreamDestroy(st);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (e != hipSuccess) return hip_fail(e, "MFMA peak loop");
    const double flops = (double)blocks * 4 * iters * mfma_peak_flops_per_wave_iter(kind);
    *tflops = flops / (ms * 1e-3) / 1e12;
    return RSH_OK;
}

int rsh_dev_synchronize(int32_t device) {
    int rc = check_device(device);
    if (rc) return rc;
    RSH_HIP_TRY(hipSetDevice(device));
    RSH_HIP_TRY(hipDeviceSynchronize());
    return RSH_OK;
}

int rsh_dev_mem_info(int32_t device, int64_t* free_bytes, int64_t* total_bytes) {
    if (!free_bytes || !total_bytes) return fail(RSH_E_ARG, "null argument");
    int rc = check_device(device);
    if (rc) return rc;
    RSH_HIP_TRY(hipSetDevice(device));
    size_t f = 0, t = 0;
    RSH_HIP_TRY(hipMemGetInfo(&f, &t));
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
    return RSH_OK;
}

int rsh_stream_create(int32_t device, void** out) {
    if (!out) return fail(RSH_E_ARG, "null argument");
    int rc = check_device(device);
    if (rc) return rc;
    RSH_HIP_TRY(hipSetDevice(device));
    hipStream_t s = nullptr;
    RSH_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));  // default priority
    *out = (void*)s;
    return RSH_OK;
}

int rsh_stream_destroy(void* stream) {
    if (stream) RSH_HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return RSH_OK;
}

int rsh_stream_synchronize(void* stream) {
    RSH_HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return RSH_OK;
}

int rsh_event_create(void** out) {
    if (!out) return fail(RSH_E_ARG, "null argument");
    hipEvent_t e = nullptr;
    RSH_HIP_TRY(hipEventCreate(&e));
    *out = (void*)e;
    return RSH_OK;
}

int rsh_event_destroy(void* ev) {
    if (ev) RSH_HIP_TRY(hipEventDestroy((hipEvent_t)ev));
    return RSH_OK;
}

int rsh_event_record(void* ev, void* stream) {
    if (!ev) return fail(RSH_E_ARG, "null event");
    RSH_HIP_TRY(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
    return RSH_OK;
}

int rsh_event_elapsed_ms(void* ev_start, void* ev_stop, float* ms) {
    if (!ev_start || !ev_stop || !ms) return fail(RSH_E_ARG, "null argument");
    RSH_HIP_TRY(hipEventSynchronize((hipEvent_t)ev_stop));
    RSH_HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)ev_start, (hipEvent_t)ev_stop));
    return RSH_OK;
}

// ------------------------------------------------------------------ problem API

int rsh_create(const rsh_problem_desc* d, rsh_problem* out) {
    if (!d || !out) return fail(RSH_E_ARG, "null argument");
    *out = nullptr;
    if (d->n_cells < 0 || d->n_dirs <= 0 || (d->faces != 1 && d->faces != 4))
        return fail(RSH_E_ARG, "invalid problem sizes (faces must be 1 or 4)");
    if (d->k_begin < 0 || d->k_end > d->n_dirs || d->k_begin > d->k_end)
        return fail(RSH_E_ARG, "invalid local direction range");
    rsh_dirset ds = nullptr;
    int rc = rsh_dirset_create(d->device, d->n_dirs, d->dirs_xyz, d->weights, d->full_area, d->phase, d->g,
                               d->normalize, &ds);
    if (rc) return rc;
    auto p = std::make_unique<rsh_problem_s>();
    p->device = d->device;
    p->n = d->n_cells;
    p->faces = d->faces;
    p->k_begin = d->k_begin;
    p->k_end = d->k_end;
    p->ds = ds;
    auto cleanup = [&](int status) {
        hipFree(p->d_illum);
        hipFree(p->d_S);
        for (int i = 0; i < RSH_STREAM_COUNT; i++) {
            if (p->streams[i]) hipStreamDestroy(p->streams[i]);
            if (p->upload_ev[i]) hipEventDestroy(p->upload_ev[i]);
            if (p->t0[i]) hipEventDestroy(p->t0[i]);
            if (p->t1[i]) hipEventDestroy(p->t1[i]);
        }
        rsh_dirset_destroy(ds);
        return status;
    };
    const size_t illum_bytes = (size_t)d->faces * d->n_dirs * d->n_cells * sizeof(double);
    const size_t s_bytes = (size_t)(d->k_end - d->k_begin) * d->n_cells * sizeof(double);
    hipError_t e;
    if ((e = hipMalloc(&p->d_illum, std::max<size_t>(illum_bytes, 8))) != hipSuccess ||
        (e = hipMalloc(&p->d_S, std::max<size_t>(s_bytes, 8))) != hipSuccess)
        return cleanup(hip_fail(e, "hipMalloc(problem buffers)"));
    for (int i = 0; i < RSH_STREAM_COUNT; i++) {
        // Default priority on purpose: the reference creates these three streams with
        // high / mid / low priorities (cuda_solve.cu:60-66); work on streams of
        // different priorities has been seen to hang on this platform, and the
        // ordering the callers rely on comes from events, not priorities.
        if ((e = hipStreamCreateWithFlags(&p->streams[i], hipStreamNonBlocking)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&p->upload_ev[i], hipEventDisableTiming)) != hipSuccess ||
            (e = hipEventCreate(&p->t0[i])) != hipSuccess || (e = hipEventCreate(&p->t1[i])) != hipSuccess)
            return cleanup(hip_fail(e, "stream/event creation"));
    }
    *out = p.release();
    return RSH_OK;
}

int rsh_destroy(rsh_problem p) {
    if (!p) return RSH_OK;
    hipSetDevice(p->device);
    for (int i = 0; i < RSH_STREAM_COUNT; i++) hipStreamSynchronize(p->streams[i]);
    hipFree(p->d_illum);
    hipFree(p->d_S);
    hipFree(p->d_geo);
    hipFree(p->d_mom);
    for (int i = 0; i < RSH_STREAM_COUNT; i++) {
        hipStreamDestroy(p->streams[i]);
        hipEventDestroy(p->upload_ev[i]);
        hipEventDestroy(p->t0[i]);
        hipEventDestroy(p->t1[i]);
    }
    rsh_dirset_destroy(p->ds);
    delete p;
    return RSH_OK;
}

static bool valid_stream(int s) { return s >= 0 && s < RSH_STREAM_COUNT; }

int rsh_upload_illum(rsh_problem p, const double* host, int64_t count, int64_t dst_off, int64_t src_off,
                     int32_t stream) {
    if (!p || !valid_stream(stream)) return fail(RSH_E_ARG, "invalid problem or stream id");
    const int64_t total = (int64_t)p->faces * p->ds->m * p->n;
    if (count < 0 || dst_off < 0 || src_off < 0 || dst_off + count > total)
        return fail(RSH_E_ARG, "upload range outside device illum");
    if (count == 0) return RSH_OK;
    if (!host) return fail(RSH_E_ARG, "null host pointer");
    std::lock_guard<std::mutex> lock(p->mu);
    RSH_HIP_TRY(hipSetDevice(p->device));
    RSH_HIP_TRY(hipMemcpyAsync(p->d_illum + dst_off, host + src_off, (size_t)count * sizeof(double),
                               hipMemcpyHostToDevice, p->streams[stream]));
    RSH_HIP_TRY(hipEventRecord(p->upload_ev[stream], p->streams[stream]));
    p->upload_pending[stream] = true;
    return RSH_OK;
}

int rsh_copy_illum_sync(rsh_problem p, const double* host, int64_t count) {
    if (!p) return fail(RSH_E_ARG, "null problem");
    const int64_t total = (int64_t)p->faces * p->ds->m * p->n;
    if (count < 0 || count > total) return fail(RSH_E_ARG, "copy larger than device illum");
    if (count == 0) return RSH_OK;
    if (!host) return fail(RSH_E_ARG, "null host pointer");
    RSH_HIP_TRY(hipSetDevice(p->device));
    for (int i = 0; i < RSH_STREAM_COUNT; i++) RSH_HIP_TRY(hipStreamSynchronize(p->streams[i]));
    RSH_HIP_TRY(hipMemcpy(p->d_illum, host, (size_t)count * sizeof(double), hipMemcpyHostToDevice));
    return RSH_OK;
}

int rsh_scatter(rsh_problem p, int32_t k0, int32_t k1, int32_t stream, double* host_out) {
    if (!p || !valid_stream(stream)) return fail(RSH_E_ARG, "invalid problem or stream id");
    if (k0 < p->k_begin || k1 > p->k_end || k0 > k1)
        return fail(RSH_E_ARG, "direction range outside the problem's local range");
    std::lock_guard<std::mutex> lock(p->mu);
    RSH_HIP_TRY(hipSetDevice(p->device));
    hipStream_t st = p->streams[stream];
    // Order every earlier upload (on any stream) before this launch.
    for (int i = 0; i < RSH_STREAM_COUNT; i++)
        if (p->upload_pending[i] && i != stream) RSH_HIP_TRY(hipStreamWaitEvent(st, p->upload_ev[i], 0));
    double* dst = p->d_S + (int64_t)(k0 - p->k_begin) * p->n;
    RSH_HIP_TRY(hipEventRecord(p->t0[stream], st));
    int rc = scatter_f64(p->ds, p->d_illum, p->faces, p->n, p->n, k0, k1, dst, p->n, st, RSH_F64_AUTO, nullptr);
    if (rc) return rc;
    RSH_HIP_TRY(hipEventRecord(p->t1[stream], st));
    p->timed[stream] = true;
    if (host_out && k1 > k0 && p->n > 0)
        RSH_HIP_TRY(hipMemcpyAsync(host_out + (int64_t)(k0 - p->k_begin) * p->n, dst,
                                   (size_t)(k1 - k0) * p->n * sizeof(double), hipMemcpyDeviceToHost, st));
    return RSH_OK;
}

int rsh_scatter_from_host(rsh_problem p, const double* host_illum, int32_t k0, int32_t k1, double* host_out,
                          int32_t chunks) {
    if (!p || !host_illum) return fail(RSH_E_ARG, "null problem or host illum");
    if (k0 < p->k_begin || k1 > p->k_end || k0 > k1)
        return fail(RSH_E_ARG, "direction range outside the problem's local range");
    if (chunks < 1 || chunks > 64) return fail(RSH_E_ARG, "chunks must be 1..64");
    std::lock_guard<std::mutex> lock(p->mu);
    RSH_HIP_TRY(hipSetDevice(p->device));
    // The device illum is overwritten: earlier work on the problem's streams first.
    for (int i = 0; i < RSH_STREAM_COUNT; i++) RSH_HIP_TRY(hipStreamSynchronize(p->streams[i]));
    const int64_t n = p->n;
    const int faces = p->faces, m = p->ds->m;
    if (n == 0) return RSH_OK;
    hipStream_t up = p->streams[RSH_STREAM_SCATTERING_1], comp = p->streams[RSH_STREAM_SCATTERING_2],
                down = p->streams[RSH_STREAM_PARAMS];
    std::vector<hipEvent_t> ev((size_t)2 * chunks, nullptr);
    auto cleanup = [&](int status) {
        // On an error, let the copies already queued finish before the caller reuses its buffers.
        if (status != RSH_OK)
            for (int i = 0; i < RSH_STREAM_COUNT; i++) (void)hipStreamSynchronize(p->streams[i]);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        return status;
    };
    for (auto& e : ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return cleanup(fail(RSH_E_HIP, "event creation"));
    const size_t row = (size_t)n * faces * sizeof(double);      // one direction of the illum
    const size_t srow = (size_t)n * sizeof(double);             // one direction of S
    double* s_base = p->d_S + (int64_t)(k0 - p->k_begin) * n;
    double* h_base = host_out ? host_out + (int64_t)(k0 - p->k_begin) * n : nullptr;
    hipError_t e = hipSuccess;
    for (int c = 0; c < chunks; c++) {
        const int64_t c0 = n * c / chunks, c1 = n * (c + 1) / chunks, nc = c1 - c0;
        if (nc == 0) continue;
        if ((e = hipMemcpy2DAsync(p->d_illum + c0 * faces, row, host_illum + c0 * faces, row,
                                  (size_t)nc * faces * sizeof(double), (size_t)m, hipMemcpyHostToDevice, up)) !=
                hipSuccess ||
            (e = hipEventRecord(ev[(size_t)(2 * c)], up)) != hipSuccess ||
            (e = hipStreamWaitEvent(comp, ev[(size_t)(2 * c)], 0)) != hipSuccess)
            return cleanup(hip_fail(e, "chunk upload"));
        if (k1 == k0) continue;
        const int rc = scatter_f64(p->ds, p->d_illum + c0 * faces, faces, nc, n, k0, k1, s_base + c0, n, comp,
                                   RSH_F64_AUTO, nullptr);
        if (rc) return cleanup(rc);
        if (!h_base) continue;
        if ((e = hipEventRecord(ev[(size_t)(2 * c + 1)], comp)) != hipSuccess ||
            (e = hipStreamWaitEvent(down, ev[(size_t)(2 * c + 1)], 0)) != hipSuccess ||
            (e = hipMemcpy2DAsync(h_base + c0, srow, s_base + c0, srow, (size_t)nc * sizeof(double),
                                  (size_t)(k1 - k0), hipMemcpyDeviceToHost, down)) != hipSuccess)
            return cleanup(hip_fail(e, "chunk copy back"));
    }
    for (int i = 0; i < RSH_STREAM_COUNT; i++)
        if ((e = hipStreamSynchronize(p->streams[i])) != hipSuccess) return cleanup(hip_fail(e, "synchronize"));
    for (int i = 0; i < RSH_STREAM_COUNT; i++) p->upload_pending[i] = false;
    return cleanup(RSH_OK);
}

int rsh_problem_set_geometry(rsh_problem p, const double* normals, const double* areas, const double* volume) {
    if (!p) return fail(RSH_E_ARG, "null problem");
    const int64_t n = p->n, nf = n * p->faces;
    if (n > 0 && (!normals || !areas || !volume)) return fail(RSH_E_ARG, "null geometry array");
    std::lock_guard<std::mutex> lock(p->mu);
    RSH_HIP_TRY(hipSetDevice(p->device));
    if (!p->d_geo) RSH_HIP_TRY(hipMalloc(&p->d_geo, (size_t)std::max<int64_t>(4 * nf + n, 1) * sizeof(double)));
    // A moment launch in flight may still read the old geometry.
    for (int i = 0; i < RSH_STREAM_COUNT; i++) RSH_HIP_TRY(hipStreamSynchronize(p->streams[i]));
    if (n > 0) {
        RSH_HIP_TRY(hipMemcpy(p->d_geo, normals, (size_t)3 * nf * sizeof(double), hipMemcpyHostToDevice));
        RSH_HIP_TRY(hipMemcpy(p->d_geo + 3 * nf, areas, (size_t)nf * sizeof(double), hipMemcpyHostToDevice));
        RSH_HIP_TRY(hipMemcpy(p->d_geo + 4 * nf, volume, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    }
    return RSH_OK;
}

// Offsets of the moment arrays inside d_mom, in doubles per cell.
static constexpr int kMomOff[5] = {0, 1, 4, 13, 14};
static constexpr int kMomWidth[5] = {1, 3, 9, 1, 3};
static constexpr int kMomPerCell = 17;

int rsh_moments(rsh_problem p, int32_t stream, double* energy, double* stream_out, double* impuls,
                double* divstream, double* divimpuls) {
    if (!p || !valid_stream(stream)) return fail(RSH_E_ARG, "invalid problem or stream id");
    const int64_t n = p->n, nf = n * p->faces;
    std::lock_guard<std::mutex> lock(p->mu);
    RSH_HIP_TRY(hipSetDevice(p->device));
    if (!p->d_geo && (divstream || divimpuls))
        return fail(RSH_E_ARG, "divergences need rsh_problem_set_geometry first");
    if (!p->d_mom) RSH_HIP_TRY(hipMalloc(&p->d_mom, (size_t)std::max<int64_t>(kMomPerCell * n, 1) * sizeof(double)));
    hipStream_t st = p->streams[stream];
    for (int i = 0; i < RSH_STREAM_COUNT; i++)
        if (p->upload_pending[i] && i != stream) RSH_HIP_TRY(hipStreamWaitEvent(st, p->upload_ev[i], 0));
    double* dev[5];
    for (int i = 0; i < 5; i++) dev[i] = p->d_mom + kMomOff[i] * n;
    const bool geo = p->d_geo != nullptr;
    int rc = moments_f64(p->ds, p->d_illum, p->faces, n, geo ? p->d_geo : nullptr, geo ? p->d_geo + 3 * nf : nullptr,
                         geo ? p->d_geo + 4 * nf : nullptr, dev[0], dev[1], dev[2], geo ? dev[3] : nullptr,
                         geo ? dev[4] : nullptr, 0, st);
    if (rc) return rc;
    p->mom_valid = true;
    p->div_valid = geo;
    double* host[5] = {energy, stream_out, impuls, divstream, divimpuls};
    for (int i = 0; i < 5; i++)
        if (host[i] && n > 0)
            RSH_HIP_TRY(hipMemcpyAsync(host[i], dev[i], (size_t)kMomWidth[i] * n * sizeof(double),
                                       hipMemcpyDeviceToHost, st));
    return RSH_OK;
}

int rsh_write_solution(rsh_problem p, const char* main_dir) {
    if (!p || !main_dir) return fail(RSH_E_ARG, "null problem or directory");
    const int64_t n = p->n, nf = n * p->faces;
    if (n > INT32_MAX) return fail(RSH_E_ARG, "the solution format's count is a 32-bit int");
    std::lock_guard<std::mutex> lock(p->mu);
    RSH_HIP_TRY(hipSetDevice(p->device));
    RSH_HIP_TRY(hipDeviceSynchronize());
    const std::string dir(main_dir);
    // Illum.bin: face average of the first direction, summed in face order
    // then divided (GetDirectionIllumFromFace(size, 0, ...)).
    std::vector<double> faces((size_t)std::max<int64_t>(nf, 1)), cell((size_t)std::max<int64_t>(n, 1));
    if (nf > 0) RSH_HIP_TRY(hipMemcpy(faces.data(), p->d_illum, (size_t)nf * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t c = 0; c < n; c++) {
        double acc = 0.0;
        for (int f = 0; f < p->faces; f++) acc += faces[(size_t)(c * p->faces + f)];
        cell[(size_t)c] = acc / p->faces;
    }
    int rc = rsh_write_simple_bin((dir + "Illum.bin").c_str(), 8, cell.data(), n);
    if (rc || !p->mom_valid) return rc;
    static const char* names[5] = {"energy.bin", "stream.bin", "impuls.bin", "divstream.bin", "divimpuls.bin"};
    std::vector<double> host((size_t)std::max<int64_t>(kMomPerCell * n, 1));
    if (n > 0)
        RSH_HIP_TRY(hipMemcpy(host.data(), p->d_mom, (size_t)(kMomPerCell * n) * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < 5; i++) {
        if (i >= 3 && !p->div_valid) break;
        rc = rsh_write_simple_bin((dir + names[i]).c_str(), 8 * kMomWidth[i], host.data() + kMomOff[i] * n, n);
        if (rc) return rc;
    }
    return RSH_OK;
}

int rsh_problem_moment_buffers(rsh_problem p, double** d_energy, double** d_stream, double** d_impuls,
                               double** d_divstream, double** d_divimpuls) {
    if (!p) return fail(RSH_E_ARG, "null problem");
    double** outs[5] = {d_energy, d_stream, d_impuls, d_divstream, d_divimpuls};
    for (int i = 0; i < 5; i++)
        if (outs[i]) *outs[i] = p->d_mom ? p->d_mom + kMomOff[i] * p->n : nullptr;
    return RSH_OK;
}

int rsh_sync(rsh_problem p, int32_t stream) {
    if (!p) return fail(RSH_E_ARG, "null problem");
    RSH_HIP_TRY(hipSetDevice(p->device));
    if (stream < 0) {
        RSH_HIP_TRY(hipDeviceSynchronize());
        return RSH_OK;
    }
    if (!valid_stream(stream)) return fail(RSH_E_ARG, "invalid stream id");
    RSH_HIP_TRY(hipStreamSynchronize(p->streams[stream]));
    return RSH_OK;
}

int rsh_problem_buffers(rsh_problem p, double** d_illum, double** d_scattering) {
    if (!p) return fail(RSH_E_ARG, "null problem");
    if (d_illum) *d_illum = p->d_illum;
    if (d_scattering) *d_scattering = p->d_S;
    return RSH_OK;
}

int rsh_problem_stream(rsh_problem p, int32_t stream, void** hip_stream) {
    if (!p || !hip_stream || !valid_stream(stream)) return fail(RSH_E_ARG, "invalid arguments");
    *hip_stream = (void*)p->streams[stream];
    return RSH_OK;
}

int rsh_event_times(rsh_problem p, float* ms) {
    if (!p || !ms) return fail(RSH_E_ARG, "null argument");
    RSH_HIP_TRY(hipSetDevice(p->device));
    for (int i = 0; i < RSH_STREAM_COUNT; i++) {
        ms[i] = -1.0f;
        if (!p->timed[i]) continue;
        RSH_HIP_TRY(hipEventSynchronize(p->t1[i]));
        RSH_HIP_TRY(hipEventElapsedTime(&ms[i], p->t0[i], p->t1[i]));
    }
    return RSH_OK;
}

}  // extern "C"

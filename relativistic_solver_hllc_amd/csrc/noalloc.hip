// Allocation-free FP64 scattering integral (rsh_scatter_f64_dev_noalloc).
//
// rsh_scatter_f64_dev keeps the MFMA phase panels of a direction set in a
// cache whose entries it allocates on first use. Callers that cannot allocate
// device memory (code under stream capture, tight serving loops, harnesses
// that forbid hipMalloc) use this entry point instead: it builds the panel of
// the requested output range into a fixed file-scope device array, `scratch`,
// and runs the same MFMA kernel on it, so it calls no hipMalloc / hipFree.
//
// Each device holds its own copy of `scratch` (kNoAllocScratchBytes) until the
// process ends; its device address is looked up once per device with
// hipGetSymbolAddress and used for every kernel argument. Every call on a
// device rebuilds the panel in that device's array: calls that overlap on one
// device (different streams, or other host threads) race on it and must be
// ordered by the caller.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "rsh_objects.hpp"

namespace rsh {

namespace {

constexpr size_t kNoAllocScratchBytes = size_t(8) << 20;  // FP64 panel of up to ~1000 directions
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};

// Device address of `scratch` on `device` (the current device), looked up on
// the first call on that device.
int scratch_for(int device, void** out) {
    if (device < 0 || device >= kNoAllocMaxDevices) return fail(RSH_E_UNSUPPORTED, "device index beyond the table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(RSH_E_HIP, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return RSH_OK;
}

}  // namespace

}  // namespace rsh

using namespace rsh;

extern "C" {

int64_t rsh_noalloc_scratch_bytes(void) { return (int64_t)kNoAllocScratchBytes; }

int rsh_scatter_f64_dev_noalloc(rsh_dirset ds, const double* d_illum, int32_t faces, int64_t n_cells, int32_t k0,
                                int32_t k1, double* d_out, int64_t ld_out, void* stream) {
    if (!ds) return fail(RSH_E_ARG, "null direction set");
    if (faces != 1 && faces != 4) return fail(RSH_E_UNSUPPORTED, "faces must be 1 or 4");
    if (n_cells < 0 || k0 < 0 || k1 > ds->m || k0 > k1) return fail(RSH_E_ARG, "invalid cell count or direction range");
    if (n_cells == 0 || k0 == k1) return RSH_OK;
    if (!d_illum || !d_out) return fail(RSH_E_ARG, "null device pointer");
    if (ld_out < n_cells) return fail(RSH_E_ARG, "ld_out smaller than n_cells");
    const F64Plan pl = plan_f64(ds->m, k1 - k0);
    const int64_t bytes = panel_f64_doubles(pl) * (int64_t)sizeof(double);
    if (bytes > (int64_t)kNoAllocScratchBytes)
        return fail(RSH_E_UNSUPPORTED, "phase panel of " + std::to_string(bytes) + " bytes exceeds the " +
                                           std::to_string(kNoAllocScratchBytes) +
                                           "-byte scratch of the allocation-free entry point");
    RSH_HIP_TRY(hipSetDevice(ds->device));
    void* panel = nullptr;
    int rc = scratch_for(ds->device, &panel);
    if (rc) return rc;
    const rsh_tile t = f64_tile_for(ds->device, pl.rb_per_pass);
    hipStream_t st = (hipStream_t)stream;
    RSH_HIP_TRY(launch_panel_f64(ds->d_P, ds->m, k0, k1, 1.0 / faces, pl, static_cast<double*>(panel), st));
    RSH_HIP_TRY(launch_mfma_f64(d_illum, faces, n_cells, n_cells, ds->m, static_cast<const double*>(panel), pl,
                                k1 - k0, d_out, ld_out, t.waves, t.chunk_steps, st));
    return RSH_OK;
}

}  // extern "C"

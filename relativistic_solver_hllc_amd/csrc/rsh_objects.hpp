// Host-side objects behind the C ABI handles (direction sets, problems).
#pragma once

#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

#include "rsh_internal.hpp"

// A cached phase panel. Shared by the direction set's cache, batch slots that
// point their device problem table at it, and calls that are enqueuing work on
// it; the device memory is released (after the device is idle) when the last
// holder lets go.
struct PanelEntry {
    int device = 0;
    void* d_panel = nullptr;
    float* d_row_inv = nullptr;  // split forms: per-row inverse scale
    rsh::F64Plan plan{};
    rsh::SplitPlan splan{};
    hipEvent_t ready = nullptr;  // recorded after the panel build
    PanelEntry() = default;
    PanelEntry(const PanelEntry&) = delete;
    PanelEntry& operator=(const PanelEntry&) = delete;
    ~PanelEntry();
};
typedef std::shared_ptr<const PanelEntry> PanelRef;

struct rsh_dirset_s {
    int device = 0;
    int m = 0;
    double full_area = 0;
    int phase = 0;
    double g = 0;
    int normalize = 0;
    std::vector<double> xyz, w;
    double* d_dirs = nullptr;
    double* d_w = nullptr;
    double* d_P = nullptr;  // [m][m] FP64 phase matrix
    double* d_psi = nullptr;  // [m][kMomentFuncs] moment weights (full_area > 0 only)
    std::mutex mu;
    // key: (k0, k1, faces, form, layout); layout: FP64 pass cap, split 0 = streaming plan,
    // 100 + row blocks per group = resident-panel plan
    std::map<std::tuple<int, int, int, int, int>, std::shared_ptr<PanelEntry>> panels;
    bool ref_order() const { return phase == RSH_PHASE_THOMSON && !normalize; }
};

namespace rsh {
// Phase panel of (k0, k1, faces, form), built on stream st if absent. A split
// panel with resident != nullptr uses that row-group plan (layout 1); an FP64
// panel uses passes of at most f64_max_rb row blocks. The caller holds *out for
// as long as it enqueues work that reads the panel (a batch slot: until the
// slot is removed); the cache never frees an entry someone else holds.
int get_panel(rsh_dirset ds, int k0, int k1, int faces, int form, hipStream_t st, PanelRef* out,
              const SplitPlan* resident = nullptr, int f64_max_rb = kMaxRowBlocksPerPass);
// Cache capacity of a direction set's panel cache (entries not held elsewhere
// are evicted beyond it).
constexpr size_t kPanelCacheCap = 64;
// Compute units of a device (cached).
int device_cus(int device);
// Tile of the FP64 MFMA kernel chosen for a row-block class (tuned or default).
rsh_tile f64_tile_for(int device, int rb);
void set_tuned_tile(int device, int rb, rsh_tile t);
int check_device(int dev);
}  // namespace rsh

struct rsh_problem_s {
    int device = 0;
    int64_t n = 0;
    int faces = 4;
    int k_begin = 0, k_end = 0;
    rsh_dirset ds = nullptr;
    double* d_illum = nullptr;
    double* d_S = nullptr;
    double* d_geo = nullptr;  // normals [n][faces][3], areas [n][faces], volume [n]
    double* d_mom = nullptr;  // energy [n], stream [n][3], impuls [n][9], divstream [n], divimpuls [n][3]
    bool mom_valid = false, div_valid = false;  // what the last rsh_moments computed
    hipStream_t streams[RSH_STREAM_COUNT] = {};
    hipEvent_t upload_ev[RSH_STREAM_COUNT] = {};
    bool upload_pending[RSH_STREAM_COUNT] = {};
    hipEvent_t t0[RSH_STREAM_COUNT] = {}, t1[RSH_STREAM_COUNT] = {};
    bool timed[RSH_STREAM_COUNT] = {};
    std::mutex mu;
};


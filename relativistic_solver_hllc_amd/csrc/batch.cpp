// Batches of uneven source-iteration problems (C ABI: rsh_batch_*), and the
// FP64 tile autotuner (rsh_autotune_f64).
//
// A batch owns, per problem slot, the ping-pong cell-level state and the
// source term; the phase panel comes from the problem's direction set. Slots
// are grouped by row-block class (row blocks per pass of the MFMA kernel; the
// FP64 iteration caps passes at 13 row blocks, so M = 200, 400 and 800 share
// one class with 1, 2 and 4 passes); every class has a device work list of
// (problem, pass, cell tile) items that one persistent launch consumes.
// One batch iteration = a zeroing kernel + one grouped launch per class. Two
// iterations (parity 0 then 1) are captured once into a hipGraph; changing the
// active set rewrites the work lists and counts in device memory, which leaves
// the captured graph valid (only a new class forces a re-capture).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

#include "rsh_objects.hpp"

using namespace rsh;

namespace {

constexpr int kMaxClasses = 16;
constexpr int kXcdLists = 8;  // per-XCD work lists of the resident-panel split classes

// Grouped-launch tuning table, filled by rsh_batch_autotune: FP64 classes by
// (device, form, row blocks per pass) -> MFMA tile; split forms by (device,
// form, M) -> kernel (tile.waves: 1 resident panel, 2 streaming panel).
std::mutex g_gtune_mu;
std::map<std::tuple<int, int, int>, rsh_tile> g_gtune;

int tune_key(int form, int m) { return form == RSH_FORM_FP64 ? plan_f64(m, m, kGroupedF64MaxRb).rb_per_pass : m; }

bool grouped_tuned(int device, int form, int key, rsh_tile* t) {
    std::lock_guard<std::mutex> lock(g_gtune_mu);
    auto it = g_gtune.find({device, form, key});
    if (it == g_gtune.end()) return false;
    *t = it->second;
    return true;
}

void set_grouped_tuned(int device, int form, int key, rsh_tile t) {
    std::lock_guard<std::mutex> lock(g_gtune_mu);
    g_gtune[{device, form, key}] = t;
}

struct Slot {
    bool active = false;
    rsh_dirset ds = nullptr;
    int64_t n = 0;
    int m = 0;
    int form = RSH_FORM_FP64;
    double albedo = 0;
    double* I[2] = {nullptr, nullptr};  // FP64 form
    void* state[2] = {nullptr, nullptr}; // split forms
    float* cell_inv = nullptr;
    double* B = nullptr;                        // FP64 form: source term
    float* Bs = nullptr;                        // split forms: source term rounded to FP32
    size_t fbytes = 0, sbytes = 0, cbytes = 0;  // sizes of B / I (M rows, FP64), state, cell_inv
    size_t abytes = 0;                          // allocated size of B (Bs) and (FP64) I: padded rows
    F64Plan plan{};
    SplitPlan splan{};
    bool res = false;  // split form on the resident-panel grouped kernel
    int res_waves = 8; // its waves per workgroup (8 or 4)
    PanelRef panel;    // the phase panel the device problem table points at
    int iterations = 0;
    int64_t bytes = 0;
};

struct ClassState {
    int form = RSH_FORM_FP64;
    int rb = 0;
    rsh_tile tile{8, 2};
    int grid = 0;
    BatchItem* d_items = nullptr;
    int64_t capacity = 0;
    int* d_n_items = nullptr;
    int n_items = 0;
    int counter_index = 0;
    // resident-panel split classes: K-steps, per-XCD list begins and counts
    bool res = false;
    int steps = 0;
    int* d_meta = nullptr;
};

}  // namespace

struct rsh_batch_s {
    int device = 0;
    int capacity = 0;
    bool use_graph = true;
    std::vector<Slot> slots;
    std::vector<BatchProblem> h_probs;
    BatchProblem* d_probs = nullptr;
    unsigned long long* d_resid = nullptr;  // 2 * capacity
    unsigned* d_counters = nullptr;         // kMaxClasses * kXcdLists
    int split_kernel = 0;                   // 0 auto, 1 resident panel, 2 streaming panel, 3 resident, 4 waves
    rsh_tile f64_tile{0, 0};                // FP64 grouped tile of new classes; {0, 0} = tuned or default
    std::map<std::pair<int, int>, ClassState> classes;  // by (form, row-block class)
    bool lists_dirty = true;
    bool graph_valid = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t stream = nullptr;
    int parity = 0;
    // Optional device arena (rsh_batch_reserve): slot buffers are carved from
    // it first-fit (256-byte aligned) instead of hipMalloc / hipFree, which
    // synchronise the device and cost milliseconds at GB sizes.
    char* arena = nullptr;
    int64_t arena_size = 0;
    std::map<int64_t, int64_t> free_list;  // offset -> size, coalesced
    // Grow-only FP64 staging for seeded split-form slots: the FP64 intensities
    // the state is packed from and the FP64 source term (the cell scale bounds
    // both) before it is rounded into the slot's FP32 copy.
    double* staging = nullptr;
    int64_t staging_bytes = 0;
};

namespace {

constexpr int64_t kArenaAlign = 256;

hipError_t dev_alloc(rsh_batch b, size_t bytes, void** p) {
    *p = nullptr;
    if (bytes == 0) return hipSuccess;
    if (!b->arena) return hipMalloc(p, bytes);
    const int64_t need = ((int64_t)bytes + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
    for (auto it = b->free_list.begin(); it != b->free_list.end(); ++it) {
        if (it->second < need) continue;
        const int64_t off = it->first, size = it->second;
        b->free_list.erase(it);
        if (size > need) b->free_list[off + need] = size - need;
        *p = b->arena + off;
        return hipSuccess;
    }
    return hipErrorOutOfMemory;
}

void dev_free(rsh_batch b, void* p, size_t bytes) {
    if (!p) return;
    char* c = static_cast<char*>(p);
    if (!b->arena || c < b->arena || c >= b->arena + b->arena_size) {
        hipFree(p);
        return;
    }
    int64_t off = c - b->arena;
    int64_t size = ((int64_t)bytes + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
    auto next = b->free_list.lower_bound(off);
    if (next != b->free_list.end() && off + size == next->first) {  // merge with the following block
        size += next->second;
        next = b->free_list.erase(next);
    }
    if (next != b->free_list.begin()) {  // merge with the preceding block
        auto prev = std::prev(next);
        if (prev->first + prev->second == off) {
            prev->second += size;
            return;
        }
    }
    b->free_list[off] = size;
}

void free_slot(rsh_batch b, Slot& s) {
    dev_free(b, s.I[0], s.abytes);
    dev_free(b, s.I[1], s.abytes);
    dev_free(b, s.state[0], s.sbytes);
    dev_free(b, s.state[1], s.sbytes);
    dev_free(b, s.cell_inv, s.cbytes);
    dev_free(b, s.B, s.abytes);
    dev_free(b, s.Bs, s.abytes);
    s.I[0] = s.I[1] = nullptr;
    s.state[0] = s.state[1] = nullptr;
    s.cell_inv = nullptr;
    s.B = nullptr;
    s.Bs = nullptr;
}

int slot_rb(const Slot& s) { return s.form == RSH_FORM_FP64 ? s.plan.rb_per_pass : s.splan.rb_per_pass; }
// Class key: (form, row blocks per pass) for the FP64 and streaming split
// kernels; resident-panel split classes also separate by K-steps (one kernel
// instance per plan).
int slot_class(const Slot& s) {
    return s.res ? 1000 + (s.res_waves == 4 ? 100000 : 0) + 64 * s.splan.rb_per_pass + s.splan.steps : slot_rb(s);
}
int slot_passes(const Slot& s) { return s.form == RSH_FORM_FP64 ? s.plan.passes : s.splan.passes; }

void destroy_graph(rsh_batch b) {
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_valid = false;
}

// Rebuilds the device descriptor table and every class's work list.
int rebuild(rsh_batch b) {
    RSH_HIP_TRY(hipStreamSynchronize(b->stream));
    std::map<std::pair<int, int>, std::vector<BatchItem>> lists;
    for (auto& kv : b->classes) lists[kv.first];  // keep existing classes (possibly empty)
    // Items of more expensive problems first (longest-processing-time order).
    std::vector<int> order;
    for (int i = 0; i < b->capacity; i++)
        if (b->slots[i].active) order.push_back(i);
    std::sort(order.begin(), order.end(), [&](int x, int y) { return b->slots[x].m > b->slots[y].m; });
    bool new_class = false;
    std::map<std::pair<int, int>, std::vector<std::vector<BatchItem>>> xlists;  // resident classes: per XCD
    std::map<std::pair<int, int>, int> xnext;                                   // round-robin XCD of the next chunk
    for (int i : order) {
        const Slot& s = b->slots[i];
        const int rb = slot_rb(s);
        const auto key = std::make_pair(s.form, slot_class(s));
        if (!b->classes.count(key)) {
            if ((int)b->classes.size() >= kMaxClasses) return fail(RSH_E_UNSUPPORTED, "too many row-block classes");
            ClassState cs;
            cs.form = s.form;
            cs.rb = rb;
            if (s.form == RSH_FORM_FP64) {
                // forced for this batch, else tuned for the class, else the default
                cs.tile = rsh_tile{kGroupedF64Waves, kGroupedF64ChunkSteps};
                rsh_tile tuned{};
                if (b->f64_tile.waves > 0)
                    cs.tile = b->f64_tile;
                else if (grouped_tuned(b->device, s.form, rb, &tuned))
                    cs.tile = tuned;
                cs.grid = grouped_iter_f64_resident(rb, cs.tile.waves, cs.tile.chunk_steps, b->device);
            } else if (s.res) {
                cs.res = true;
                cs.steps = s.splan.steps;
                cs.tile = rsh_tile{s.res_waves, 1};
                // one workgroup per CU (8 waves), or two (4 waves; they co-reside
                // where the panel is streamed by K-step and the LDS is small)
                cs.grid = device_cus(b->device) / kXcdLists * kXcdLists * (s.res_waves == 4 ? 2 : 1);
                RSH_HIP_TRY(launch_grouped_iter_split_res(false, cs.form, rb, cs.steps, cs.tile.waves, 0, nullptr,
                                                          nullptr, nullptr, nullptr, 0, nullptr));
                RSH_HIP_TRY(hipMalloc(&cs.d_meta, 2 * kXcdLists * sizeof(int)));
            } else {
                cs.tile = rsh_tile{grouped_iter_split_waves(), 1};
                cs.grid = grouped_iter_split_resident(s.form, rb, b->device);
            }
            cs.counter_index = (int)b->classes.size() * kXcdLists;
            RSH_HIP_TRY(hipMalloc(&cs.d_n_items, sizeof(int)));
            b->classes[key] = cs;
            new_class = true;
        }
        if (s.res) {
            // items (problem, row group, chunk); a chunk's row groups go to one XCD's list
            auto& X = xlists[key];
            X.resize(kXcdLists);
            const int64_t cc = iter_split_res_chunk_cells();
            const int64_t chunks = (s.n + cc - 1) / cc;
            for (int64_t c = 0; c < chunks; c++) {
                auto& L = X[(size_t)(xnext[key]++ % kXcdLists)];
                for (int rg = 0; rg < s.splan.passes; rg++) L.push_back(BatchItem{i, rg, c});
            }
            continue;
        }
        const int tiles_w = 16 * b->classes[key].tile.waves;
        const int64_t tiles = (s.n + tiles_w - 1) / tiles_w;
        auto& L = lists[key];
        for (int pass = 0; pass < slot_passes(s); pass++)
            for (int64_t t = 0; t < tiles; t++) L.push_back(BatchItem{i, pass, t});
    }
    for (auto& kv : b->classes) {
        ClassState& cs = kv.second;
        auto& L = lists[kv.first];
        int meta[2 * kXcdLists] = {};
        if (cs.res) {
            auto& X = xlists[kv.first];
            X.resize(kXcdLists);
            for (int x = 0; x < kXcdLists; x++) {
                meta[x] = (int)L.size();
                meta[kXcdLists + x] = (int)X[(size_t)x].size();
                L.insert(L.end(), X[(size_t)x].begin(), X[(size_t)x].end());
            }
        }
        if ((int64_t)L.size() > cs.capacity) {
            hipFree(cs.d_items);
            cs.d_items = nullptr;
            cs.capacity = std::max<int64_t>((int64_t)L.size(), 1024);
            RSH_HIP_TRY(hipMalloc(&cs.d_items, cs.capacity * sizeof(BatchItem)));
            new_class = true;  // kernel argument (items pointer) changed
        }
        if (!L.empty())
            RSH_HIP_TRY(hipMemcpy(cs.d_items, L.data(), L.size() * sizeof(BatchItem), hipMemcpyHostToDevice));
        cs.n_items = (int)L.size();
        RSH_HIP_TRY(hipMemcpy(cs.d_n_items, &cs.n_items, sizeof(int), hipMemcpyHostToDevice));
        if (cs.res) RSH_HIP_TRY(hipMemcpy(cs.d_meta, meta, sizeof(meta), hipMemcpyHostToDevice));
    }
    RSH_HIP_TRY(hipMemcpy(b->d_probs, b->h_probs.data(), b->capacity * sizeof(BatchProblem), hipMemcpyHostToDevice));
    if (new_class) destroy_graph(b);
    b->lists_dirty = false;
    return RSH_OK;
}

int enqueue_iteration(rsh_batch b, int parity) {
    RSH_HIP_TRY(launch_batch_zero(b->d_probs, b->capacity, b->d_counters, kMaxClasses * kXcdLists, parity,
                                  b->stream));
    for (auto& kv : b->classes) {
        ClassState& cs = kv.second;
        if (cs.res)
            RSH_HIP_TRY(launch_grouped_iter_split_res(true, cs.form, cs.rb, cs.steps, cs.tile.waves, cs.grid,
                                                      b->d_probs, cs.d_items, cs.d_meta,
                                                      b->d_counters + cs.counter_index, parity, b->stream));
        else if (cs.form == RSH_FORM_FP64)
            RSH_HIP_TRY(launch_grouped_iter_f64(cs.rb, cs.tile.waves, cs.tile.chunk_steps, cs.grid, b->d_probs,
                                                cs.d_items, cs.d_n_items, b->d_counters + cs.counter_index, parity,
                                                b->stream));
        else
            RSH_HIP_TRY(launch_grouped_iter_split(cs.form, cs.rb, cs.grid, b->d_probs, cs.d_items, cs.d_n_items,
                                                  b->d_counters + cs.counter_index, parity, b->stream));
    }
    return RSH_OK;
}

int capture(rsh_batch b) {
    destroy_graph(b);
    RSH_HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_iteration(b, 0);
    if (rc == RSH_OK) rc = enqueue_iteration(b, 1);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc != RSH_OK) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    b->graph = g;
    RSH_HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    b->graph_valid = true;
    return RSH_OK;
}

// Rows allocated per source array (and FP64 state array): whole row blocks,
// for FP64 the iteration plan's padded row count. The grouped kernels read (and
// the FP64 one writes) the padding rows, which are zero, so that their residual
// term is 0 / 0 and ignored, instead of branching per row.
int64_t padded_rows(int form, int m) {
    return form == RSH_FORM_FP64 ? (int64_t)plan_f64(m, m, kGroupedF64MaxRb).rb_total * 16 : (m + 15) / 16 * 16;
}

int64_t arena_round(int64_t bytes) { return (bytes + kArenaAlign - 1) / kArenaAlign * kArenaAlign; }

// Device bytes of a slot's buffers as the arena carves them (each buffer
// rounded up to the arena alignment), so that a problem whose footprint fits the
// budget fits an empty arena.
int64_t footprint(int form, int64_t n, int m) {
    const int64_t rows = arena_round(padded_rows(form, m) * n * 8);  // FP64 B and I, padded rows
    if (form == RSH_FORM_FP64) return 3 * rows;                     // I ping-pong + B
    if (form == RSH_FORM_BF16X2 || form == RSH_FORM_FP16X2 || form == RSH_FORM_BF16X3)  // state ping-pong + FP32 B + scales
        return 2 * arena_round(split_state_units(form, 1, n, m) * 16) +
               arena_round(padded_rows(form, m) * ((n + 15) / 16 * 16) * 4) + arena_round(n * 4);
    return -1;
}

}  // namespace

extern "C" {

int rsh_batch_footprint(int32_t form, int64_t n_cells, int32_t m, int64_t* bytes) {
    if (!bytes || n_cells < 0 || m <= 0) return fail(RSH_E_ARG, "invalid footprint arguments");
    const int64_t f = footprint(form, n_cells, m);
    if (f < 0) return fail(RSH_E_UNSUPPORTED, "form not supported by batches");
    if (form == RSH_FORM_FP64 && n_cells >= kGroupedF64MaxCells)
        return fail(RSH_E_UNSUPPORTED, "FP64 batch problems hold fewer than 2^24 cells");
    *bytes = f;
    return RSH_OK;
}

int rsh_batch_create(int32_t device, int32_t capacity, int32_t use_graph, rsh_batch* out) {
    if (!out || capacity <= 0) return fail(RSH_E_ARG, "invalid batch arguments");
    *out = nullptr;
    int rc = check_device(device);
    if (rc) return rc;
    RSH_HIP_TRY(hipSetDevice(device));
    auto b = std::make_unique<rsh_batch_s>();
    b->device = device;
    b->capacity = capacity;
    b->use_graph = use_graph != 0;
    b->slots.resize(capacity);
    b->h_probs.assign(capacity, BatchProblem{});
    hipError_t e;
    if ((e = hipMalloc(&b->d_probs, capacity * sizeof(BatchProblem))) != hipSuccess ||
        (e = hipMalloc(&b->d_resid, 2 * (size_t)capacity * sizeof(unsigned long long))) != hipSuccess ||
        (e = hipMalloc(&b->d_counters, kMaxClasses * kXcdLists * sizeof(unsigned))) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) {
        hipFree(b->d_probs);
        hipFree(b->d_resid);
        hipFree(b->d_counters);
        return hip_fail(e, "batch allocation");
    }
    for (int i = 0; i < capacity; i++) b->h_probs[i].resid = b->d_resid + 2 * i;
    RSH_HIP_TRY(hipMemcpy(b->d_probs, b->h_probs.data(), capacity * sizeof(BatchProblem), hipMemcpyHostToDevice));
    *out = b.release();
    return RSH_OK;
}

int rsh_batch_destroy(rsh_batch b) {
    if (!b) return RSH_OK;
    hipSetDevice(b->device);
    hipStreamSynchronize(b->stream);
    destroy_graph(b);
    for (auto& s : b->slots) free_slot(b, s);
    hipFree(b->arena);
    hipFree(b->staging);
    for (auto& kv : b->classes) {
        hipFree(kv.second.d_items);
        hipFree(kv.second.d_n_items);
        if (kv.second.d_meta) hipFree(kv.second.d_meta);
    }
    hipFree(b->d_probs);
    hipFree(b->d_resid);
    hipFree(b->d_counters);
    hipStreamDestroy(b->stream);
    delete b;
    return RSH_OK;
}

}  // extern "C"

namespace {

// Adds a slot whose initial state and source term are either copied from the
// caller's device arrays (d_I0, d_B) or generated on the device from a seed
// (I0 ~ log-normal(0, 1) with seed 2 s + 1, B ~ U(0.5, 2) with seed 2 s + 2).
int add_impl(rsh_batch b, rsh_dirset ds, int64_t n_cells, double albedo, int form, const double* d_I0,
             const double* d_B, const uint64_t* seed, int32_t* slot) {
    if (!b || !ds || !slot || n_cells <= 0 || (!seed && (!d_I0 || !d_B)))
        return fail(RSH_E_ARG, "invalid batch_add arguments");
    if (footprint(form, 1, 1) < 0) return fail(RSH_E_ARG, "unknown form");
    if (ds->device != b->device) return fail(RSH_E_ARG, "direction set lives on another device");
    // the FP64 iteration kernel's 32-bit buffer offsets span one row block
    if (form == RSH_FORM_FP64 && n_cells >= kGroupedF64MaxCells)
        return fail(RSH_E_UNSUPPORTED, "FP64 batch problems hold fewer than 2^24 cells");
    int id = -1;
    for (int i = 0; i < b->capacity; i++)
        if (!b->slots[i].active) {
            id = i;
            break;
        }
    if (id < 0) return fail(RSH_E_NOMEM, "batch is full");
    RSH_HIP_TRY(hipSetDevice(b->device));
    if (!seed) {
        // The batch stream is non-blocking: order its copies of d_I0 / d_B after
        // all work already submitted to the null stream (where callers usually
        // made them).
        hipEvent_t ready;
        RSH_HIP_TRY(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
        hipError_t er = hipEventRecord(ready, nullptr);
        if (er == hipSuccess) er = hipStreamWaitEvent(b->stream, ready, 0);
        hipEventDestroy(ready);
        if (er != hipSuccess) return hip_fail(er, "ordering batch_add after the null stream");
    }
    Slot s;
    s.ds = ds;
    s.n = n_cells;
    s.m = ds->m;
    s.form = form;
    s.albedo = albedo;
    s.plan = plan_f64(ds->m, ds->m, kGroupedF64MaxRb);
    s.splan = plan_split(ds->m, ds->m);
    int split_kernel = b->split_kernel;
    rsh_tile tuned{};
    if (form != RSH_FORM_FP64 && split_kernel == 0 && grouped_tuned(b->device, form, ds->m, &tuned))
        split_kernel = tuned.waves == 2 || tuned.waves == 3 ? tuned.waves : 0;  // tuned, else auto (resident, 8 waves)
    if (form != RSH_FORM_FP64 && split_kernel != 2 && n_cells < (int64_t(1) << 24)) {
        // resident-panel grouped kernel where its plan fits and is specialised
        SplitPlan rplan{};
        int sets = 0;
        if (plan_split_resident(form, ds->m, ds->m, device_cus(b->device), &rplan, &sets, true) &&
            iter_split_res_supported(form, rplan.rb_per_pass, rplan.steps)) {
            s.splan = rplan;
            s.res = true;
            s.res_waves = split_kernel == 3 ? 4 : 8;
        } else if (split_kernel == 1 || split_kernel == 3) {
            return fail(RSH_E_UNSUPPORTED, "resident-panel iteration kernel: no specialised plan for this M");
        }
    }
    const size_t fbytes = (size_t)ds->m * n_cells * sizeof(double);
    const size_t sbytes = form == RSH_FORM_FP64 ? 0 : (size_t)split_state_units(form, 1, n_cells, ds->m) * 16;
    s.fbytes = fbytes;
    // The split forms keep the source term in FP32 (half the bytes the
    // iteration epilogue reads; its rounding, 2^-24 relative, is far below
    // their tolerance); the update (1 - a) B + a S is formed in FP64.
    const size_t bbytes = (size_t)ds->m * n_cells * sizeof(double);
    // split: FP32 in whole 16-cell tiles (MFMA output-tile order, padding written 0 by the conversion)
    s.abytes = form == RSH_FORM_FP64 ? (size_t)padded_rows(form, ds->m) * n_cells * sizeof(double)
                                     : (size_t)padded_rows(form, ds->m) * ((n_cells + 15) / 16 * 16) * sizeof(float);
    s.sbytes = sbytes;
    s.cbytes = form == RSH_FORM_FP64 ? 0 : (size_t)n_cells * sizeof(float);
    void* bsrc_alloc = nullptr;
    hipError_t e = dev_alloc(b, s.abytes, &bsrc_alloc);
    if (form == RSH_FORM_FP64)
        s.B = static_cast<double*>(bsrc_alloc);
    else
        s.Bs = static_cast<float*>(bsrc_alloc);
    if (e == hipSuccess && form == RSH_FORM_FP64 && s.abytes > bbytes)  // zero padding rows of B (read by the iteration kernels)
        e = hipMemsetAsync(static_cast<char*>(bsrc_alloc) + bbytes, 0, s.abytes - bbytes, b->stream);
    if (e == hipSuccess && form == RSH_FORM_FP64) {
        if ((e = dev_alloc(b, s.abytes, (void**)&s.I[0])) == hipSuccess) e = dev_alloc(b, s.abytes, (void**)&s.I[1]);
        for (double* a : {s.I[0], s.I[1]})  // zero padding rows (read, and written back as zeros)
            if (e == hipSuccess && s.abytes > fbytes)
                e = hipMemsetAsync(reinterpret_cast<char*>(a) + fbytes, 0, s.abytes - fbytes, b->stream);
    } else if (e == hipSuccess) {
        if ((e = dev_alloc(b, sbytes, &s.state[0])) == hipSuccess && (e = dev_alloc(b, sbytes, &s.state[1])) == hipSuccess)
            e = dev_alloc(b, s.cbytes, (void**)&s.cell_inv);
    }
    if (e != hipSuccess) {
        free_slot(b, s);
        if (e == hipErrorOutOfMemory && b->arena) return fail(RSH_E_NOMEM, "batch arena has no free block this large");
        return hip_fail(e, "batch slot allocation");
    }
    s.bytes = footprint(form, n_cells, ds->m);
    const int p = b->parity;  // the next iteration reads buffer p
    const int64_t count = (int64_t)ds->m * n_cells;
    const double* src_I0 = d_I0;
    const double* src_B = d_B;  // FP64 source term of a split slot (its cell scale, then rounded into Bs)
    if (seed) {
        const uint64_t sd = *seed;
        if (form == RSH_FORM_FP64) {
            e = launch_fill_uniform(s.B, count, 0.5, 2.0, sd * 2 + 2, b->stream);
            if (e == hipSuccess) e = launch_fill_lognormal(s.I[p], count, 0.0, 1.0, sd * 2 + 1, b->stream);
        } else {
            if (2 * (int64_t)fbytes > b->staging_bytes) {
                // grow-only; the old staging may still be read by queued packs
                RSH_HIP_TRY(hipStreamSynchronize(b->stream));
                hipFree(b->staging);
                b->staging = nullptr;
                b->staging_bytes = 0;
                if ((e = hipMalloc(&b->staging, 2 * fbytes)) == hipSuccess) b->staging_bytes = 2 * (int64_t)fbytes;
            }
            double* st_I0 = static_cast<double*>(b->staging);
            double* st_B = st_I0 + count;
            if (e == hipSuccess) e = launch_fill_uniform(st_B, count, 0.5, 2.0, sd * 2 + 2, b->stream);
            if (e == hipSuccess) e = launch_fill_lognormal(st_I0, count, 0.0, 1.0, sd * 2 + 1, b->stream);
            src_I0 = st_I0;
            src_B = st_B;
        }
    } else if (form == RSH_FORM_FP64) {
        if ((e = hipMemcpyAsync(s.B, d_B, fbytes, hipMemcpyDeviceToDevice, b->stream)) == hipSuccess)
            e = hipMemcpyAsync(s.I[p], d_I0, fbytes, hipMemcpyDeviceToDevice, b->stream);
    }
    if (e == hipSuccess && form != RSH_FORM_FP64) {
        // padding positions of the other buffer are never written by the epilogue
        if ((e = hipMemsetAsync(s.state[p ^ 1], 0, sbytes, b->stream)) == hipSuccess) {
            if (pack_fused_supported(ds->m))  // scale (I0 and B bound every iterate) and pack in one pass
                e = launch_pack_fused(form, src_I0, src_B, 1, n_cells, ds->m, s.cell_inv, s.state[p], b->stream);
            else if ((e = launch_cell_scale(src_I0, src_B, n_cells, ds->m, 1, s.cell_inv, b->stream)) == hipSuccess)
                e = launch_pack_state(form, src_I0, 1, n_cells, ds->m, s.cell_inv, s.state[p], b->stream);
        }
        if (e == hipSuccess)
            e = launch_source_tiles(src_B, ds->m, n_cells, (int)padded_rows(form, ds->m), s.Bs, b->stream);
    }
    if (e != hipSuccess) {
        free_slot(b, s);
        return hip_fail(e, "batch slot initialisation");
    }
    PanelRef pe;
    int rc = get_panel(ds, 0, ds->m, 1, form, b->stream, &pe, s.res ? &s.splan : nullptr, kGroupedF64MaxRb);
    if (rc) {
        free_slot(b, s);
        return rc;
    }
    // Copies from caller arrays complete before returning (the caller may free
    // them); seeded slots need no wait: everything above is ordered on the
    // batch stream before the next iteration.
    if (!seed) RSH_HIP_TRY(hipStreamSynchronize(b->stream));
    s.active = true;
    s.panel = pe;
    BatchProblem& P = b->h_probs[id];
    P = BatchProblem{};
    P.B = s.B;
    P.Bs = s.Bs;
    P.resid = b->d_resid + 2 * id;
    P.n = n_cells;
    P.m = ds->m;
    P.albedo = albedo;
    P.one_minus_albedo = 1.0 - albedo;
    if (form == RSH_FORM_FP64) {
        P.panel = static_cast<const double*>(pe->d_panel);
        P.I[0] = s.I[0];
        P.I[1] = s.I[1];
        P.steps = pe->plan.steps;
    } else {
        P.spanel = pe->d_panel;
        P.row_inv = pe->d_row_inv;
        P.state[0] = s.state[0];
        P.state[1] = s.state[1];
        P.cell_inv = s.cell_inv;
        P.steps = pe->splan.steps;
    }
    b->slots[id] = s;
    b->lists_dirty = true;
    *slot = id;
    return RSH_OK;
}

}  // namespace

extern "C" {

int rsh_batch_add(rsh_batch b, rsh_dirset ds, int64_t n_cells, const double* d_I0, const double* d_B, double albedo,
                  int32_t form, int32_t* slot) {
    return add_impl(b, ds, n_cells, albedo, form, d_I0, d_B, nullptr, slot);
}

int rsh_batch_add_seeded(rsh_batch b, rsh_dirset ds, int64_t n_cells, uint64_t seed, double albedo, int32_t form,
                         int32_t* slot) {
    return add_impl(b, ds, n_cells, albedo, form, nullptr, nullptr, &seed, slot);
}

int rsh_batch_set_f64_tile(rsh_batch b, const rsh_tile* tile) {
    if (!b) return fail(RSH_E_ARG, "null batch");
    rsh_tile t = tile ? *tile : rsh_tile{0, 0};
    if (t.waves != 0 && !((t.waves == 8 && t.chunk_steps == 2) || (t.waves == 4 && (t.chunk_steps == 2 ||
                                                                                         t.chunk_steps == 4))))
        return fail(RSH_E_UNSUPPORTED, "grouped FP64 tiles: (8,2), (4,2), (4,4); (0,0) = tuned or default");
    b->f64_tile = t;
    return RSH_OK;
}

int rsh_batch_slot_config(rsh_batch b, int32_t slot, rsh_tile* config) {
    if (!b || !config || slot < 0 || slot >= b->capacity || !b->slots[slot].active)
        return fail(RSH_E_ARG, "invalid slot");
    const Slot& s = b->slots[slot];
    if (s.form != RSH_FORM_FP64) {
        *config = rsh_tile{s.res ? (s.res_waves == 4 ? 3 : 1) : 2, 0};
        return RSH_OK;
    }
    RSH_HIP_TRY(hipSetDevice(b->device));
    if (b->lists_dirty) {
        int rc = rebuild(b);
        if (rc) return rc;
    }
    auto it = b->classes.find(std::make_pair(s.form, slot_class(s)));
    if (it == b->classes.end()) return fail(RSH_E_STATE, "slot has no class");
    *config = it->second.tile;
    return RSH_OK;
}

int rsh_batch_tuned(int32_t device, int32_t form, int32_t m, rsh_tile* choice) {
    if (!choice || m <= 0 || footprint(form, 1, 1) < 0) return fail(RSH_E_ARG, "invalid arguments");
    if (!grouped_tuned(device, form, tune_key(form, m), choice))
        *choice = form == RSH_FORM_FP64 ? rsh_tile{kGroupedF64Waves, kGroupedF64ChunkSteps} : rsh_tile{0, 0};
    return RSH_OK;
}

int rsh_batch_autotune(int32_t device, int32_t form, int32_t m, int64_t n_cells, int32_t iterations, rsh_tile* best,
                       float* best_ms) {
    if (!best || !best_ms || m <= 0 || n_cells <= 0 || iterations <= 0 || footprint(form, 1, 1) < 0)
        return fail(RSH_E_ARG, "invalid autotune arguments");
    int rc = check_device(device);
    if (rc) return rc;
    RSH_HIP_TRY(hipSetDevice(device));
    std::vector<double> xyz(3 * (size_t)m), w(m);
    double F = 0;
    rsh_fibonacci_directions(m, xyz.data(), w.data(), &F);
    rsh_dirset ds = nullptr;
    if ((rc = rsh_dirset_create(device, m, xyz.data(), w.data(), F, RSH_PHASE_HG, 0.6, 1, &ds))) return rc;
    std::vector<rsh_tile> cands;
    if (form == RSH_FORM_FP64)
        cands = {{4, 2}, {4, 4}, {8, 2}};
    else
        cands = {{1, 0}, {3, 0}, {2, 0}};  // resident-panel (8 waves, 4 waves), streaming-panel grouped kernel
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e;
    if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
        if (e0) hipEventDestroy(e0);
        rsh_dirset_destroy(ds);
        return hip_fail(e, "autotune events");
    }
    // One batch per candidate (a class keeps the tile it was created with), all
    // set up and warmed up first; then kAutotuneRounds rounds time every
    // candidate in turn and each keeps its fastest round, so that a clock
    // ramp or a neighbour's transient does not decide the choice.
    constexpr int kAutotuneRounds = 3;
    std::vector<rsh_batch> batches(cands.size(), nullptr);
    std::vector<float> best_of(cands.size(), -1.f);
    for (size_t c = 0; c < cands.size() && !rc; c++) {
        rsh_batch bb = nullptr;
        if ((rc = rsh_batch_create(device, 1, 1, &bb))) break;
        int slot = -1;
        if (form == RSH_FORM_FP64)
            rc = rsh_batch_set_f64_tile(bb, &cands[c]);
        else
            rc = rsh_batch_set_split_kernel(bb, cands[c].waves);
        if (!rc) rc = rsh_batch_add_seeded(bb, ds, n_cells, 0x7E57ull, 0.5, form, &slot);
        if (rc == RSH_E_UNSUPPORTED) {  // candidate does not apply (e.g. no resident plan for this M)
            rsh_batch_destroy(bb);
            rc = RSH_OK;
            continue;
        }
        batches[c] = bb;
        if (!rc) rc = rsh_batch_run(bb, 2);  // graph capture, first launches
        if (!rc) rc = rsh_batch_sync(bb);
    }
    for (int round = 0; round < kAutotuneRounds && !rc; round++) {
        for (size_t c = 0; c < cands.size() && !rc; c++) {
            rsh_batch bb = batches[c];
            if (!bb) continue;
            float ms = 0.f;
            if ((e = hipEventRecord(e0, bb->stream)) != hipSuccess) rc = hip_fail(e, "event");
            if (!rc) rc = rsh_batch_run(bb, iterations);
            if (!rc) {
                if ((e = hipEventRecord(e1, bb->stream)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess ||
                    (e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess)
                    rc = hip_fail(e, "event timing");
            }
            ms /= iterations;
            if (!rc && (best_of[c] < 0.f || ms < best_of[c])) best_of[c] = ms;
        }
    }
    for (rsh_batch bb : batches)
        if (bb) rsh_batch_destroy(bb);
    float bms = -1.f;
    rsh_tile bt = cands[0];
    for (size_t c = 0; c < cands.size(); c++)
        if (best_of[c] >= 0.f && (bms < 0.f || best_of[c] < bms)) {
            bms = best_of[c];
            bt = cands[c];
        }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    rsh_dirset_destroy(ds);
    if (rc) return rc;
    if (bms < 0.f) return fail(RSH_E_UNSUPPORTED, "no grouped kernel applies to this form and M");
    set_grouped_tuned(device, form, tune_key(form, m), bt);
    *best = bt;
    *best_ms = bms;
    return RSH_OK;
}

int rsh_batch_set_split_kernel(rsh_batch b, int32_t kind) {
    if (!b || kind < 0 || kind > 3)
        return fail(RSH_E_ARG, "split kernel: 0 (auto), 1 (resident panel), 2 (streaming), 3 (resident panel, 4 waves)");
    b->split_kernel = kind;
    return RSH_OK;
}

int rsh_batch_reserve(rsh_batch b, int64_t bytes) {
    if (!b || bytes < 0) return fail(RSH_E_ARG, "invalid reserve arguments");
    for (const auto& s : b->slots)
        if (s.active) return fail(RSH_E_STATE, "rsh_batch_reserve needs an empty batch");
    RSH_HIP_TRY(hipSetDevice(b->device));
    RSH_HIP_TRY(hipStreamSynchronize(b->stream));
    hipFree(b->arena);
    b->arena = nullptr;
    b->arena_size = 0;
    b->free_list.clear();
    if (bytes == 0) return RSH_OK;
    const int64_t size = (bytes + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
    RSH_HIP_TRY(hipMalloc(&b->arena, (size_t)size));
    b->arena_size = size;
    b->free_list[0] = size;
    return RSH_OK;
}

int rsh_batch_remove(rsh_batch b, int32_t slot) {
    if (!b || slot < 0 || slot >= b->capacity || !b->slots[slot].active) return fail(RSH_E_ARG, "invalid slot");
    RSH_HIP_TRY(hipSetDevice(b->device));
    RSH_HIP_TRY(hipStreamSynchronize(b->stream));
    Slot& s = b->slots[slot];
    free_slot(b, s);
    s = Slot{};
    BatchProblem& P = b->h_probs[slot];
    unsigned long long* r = P.resid;
    P = BatchProblem{};
    P.resid = r;
    b->lists_dirty = true;
    return RSH_OK;
}

int rsh_batch_run(rsh_batch b, int32_t iterations) {
    if (!b || iterations < 0) return fail(RSH_E_ARG, "invalid batch_run arguments");
    RSH_HIP_TRY(hipSetDevice(b->device));
    if (b->lists_dirty) {
        int rc = rebuild(b);
        if (rc) return rc;
    }
    if (b->classes.empty() || iterations == 0) return RSH_OK;
    int left = iterations;
    auto eager = [&]() -> int {
        int rc = enqueue_iteration(b, b->parity);
        b->parity ^= 1;
        left--;
        return rc;
    };
    if (b->parity == 1) {
        if (int rc = eager()) return rc;
    }
    if (b->use_graph && left >= 2) {
        if (!b->graph_valid) {
            if (int rc = capture(b)) return rc;
        }
        while (left >= 2) {
            RSH_HIP_TRY(hipGraphLaunch(b->exec, b->stream));
            left -= 2;
        }
    }
    while (left > 0) {
        if (int rc = eager()) return rc;
    }
    for (auto& s : b->slots)
        if (s.active) s.iterations += iterations;
    return RSH_OK;
}

int rsh_batch_sync(rsh_batch b) {
    if (!b) return fail(RSH_E_ARG, "null batch");
    RSH_HIP_TRY(hipStreamSynchronize(b->stream));
    return RSH_OK;
}

int rsh_batch_residuals(rsh_batch b, double* resid, int32_t* iterations) {
    if (!b || !resid || !iterations) return fail(RSH_E_ARG, "null argument");
    RSH_HIP_TRY(hipSetDevice(b->device));
    RSH_HIP_TRY(hipStreamSynchronize(b->stream));
    std::vector<unsigned long long> bits(2 * (size_t)b->capacity);
    RSH_HIP_TRY(hipMemcpy(bits.data(), b->d_resid, bits.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    // the last iteration read I[parity ^ 1] and wrote its residual to slot parity ^ 1
    const int last = b->parity ^ 1;
    for (int i = 0; i < b->capacity; i++) {
        const Slot& s = b->slots[i];
        iterations[i] = s.active ? s.iterations : -1;
        double v = 0.0;
        std::memcpy(&v, &bits[2 * (size_t)i + last], sizeof v);
        resid[i] = (s.active && s.iterations > 0) ? v : -1.0;
    }
    return RSH_OK;
}

int rsh_batch_get_state(rsh_batch b, int32_t slot, double* d_out) {
    if (!b || slot < 0 || slot >= b->capacity || !b->slots[slot].active || !d_out)
        return fail(RSH_E_ARG, "invalid slot or output");
    RSH_HIP_TRY(hipSetDevice(b->device));
    const Slot& s = b->slots[slot];
    if (s.form == RSH_FORM_FP64)
        RSH_HIP_TRY(hipMemcpyAsync(d_out, s.I[b->parity], (size_t)s.m * s.n * sizeof(double),
                                   hipMemcpyDeviceToDevice, b->stream));
    else
        RSH_HIP_TRY(launch_unpack_state(s.form, s.state[b->parity], s.n, s.m, s.cell_inv, d_out, b->stream));
    RSH_HIP_TRY(hipStreamSynchronize(b->stream));
    return RSH_OK;
}

int rsh_batch_stream(rsh_batch b, void** stream) {
    if (!b || !stream) return fail(RSH_E_ARG, "null argument");
    *stream = (void*)b->stream;
    return RSH_OK;
}

// ------------------------------------------------------------------ autotuner

int rsh_autotune_f64(int32_t device, int32_t m, int32_t faces, int64_t n_cells, rsh_tile* best, float* best_ms) {
    if (m <= 0 || n_cells <= 0 || (faces != 1 && faces != 4) || !best || !best_ms)
        return fail(RSH_E_ARG, "invalid autotune arguments");
    int rc = check_device(device);
    if (rc) return rc;
    RSH_HIP_TRY(hipSetDevice(device));
    std::vector<double> xyz(3 * (size_t)m), w(m);
    double F = 0;
    rsh_fibonacci_directions(m, xyz.data(), w.data(), &F);
    rsh_dirset ds = nullptr;
    if ((rc = rsh_dirset_create(device, m, xyz.data(), w.data(), F, RSH_PHASE_THOMSON, 0.0, 1, &ds))) return rc;
    double *illum = nullptr, *out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&](int status) {
        hipFree(illum);
        hipFree(out);
        if (st) hipStreamDestroy(st);
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
        rsh_dirset_destroy(ds);
        return status;
    };
    hipError_t e;
    if ((e = hipMalloc(&illum, (size_t)faces * m * n_cells * sizeof(double))) != hipSuccess ||
        (e = hipMalloc(&out, (size_t)m * n_cells * sizeof(double))) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess)
        return cleanup(hip_fail(e, "autotune allocation"));
    if ((e = launch_fill_lognormal(illum, (int64_t)faces * m * n_cells, 0.0, 1.0, 0xA070ull, st)) != hipSuccess)
        return cleanup(hip_fail(e, "autotune fill"));
    const rsh_tile cands[] = {{8, 2}, {4, 2}, {4, 4}};
    float bms = -1.f;
    rsh_tile bt{8, 2};
    for (const rsh_tile& t : cands) {
        rc = rsh_scatter_f64_dev(ds, illum, faces, n_cells, 0, m, out, n_cells, st, RSH_F64_MFMA, &t);
        if (rc) return cleanup(rc);
        if ((e = hipEventRecord(e0, st)) != hipSuccess) return cleanup(hip_fail(e, "event"));
        const int reps = 3;
        for (int r = 0; r < reps; r++) {
            rc = rsh_scatter_f64_dev(ds, illum, faces, n_cells, 0, m, out, n_cells, st, RSH_F64_MFMA, &t);
            if (rc) return cleanup(rc);
        }
        float ms = 0.f;
        if ((e = hipEventRecord(e1, st)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess ||
            (e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess)
            return cleanup(hip_fail(e, "event timing"));
        ms /= reps;
        if (bms < 0.f || ms < bms) {
            bms = ms;
            bt = t;
        }
    }
    set_tuned_tile(device, plan_f64(m, m).rb_per_pass, bt);
    *best = bt;
    *best_ms = bms;
    return cleanup(RSH_OK);
}

int rsh_tuned_tile_f64(int32_t device, int32_t m, rsh_tile* tile) {
    if (!tile || m <= 0) return fail(RSH_E_ARG, "invalid arguments");
    *tile = f64_tile_for(device, plan_f64(m, m).rb_per_pass);
    return RSH_OK;
}

}  // extern "C"

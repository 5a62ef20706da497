// Host side of the transport sweep: validates the reference-format sweep
// structures of each direction (sorted_id_cell, face states, vec_x0 records),
// derives the level of every cell, uploads everything once and launches the
// sweep kernel. Replaces ReadIllumGeometry's in-memory result
// (file_module/reader_bin.cpp:164-203) and the direction loop of
// MPI_CalculateIllumAsync (solve_module/illum/Illum_async.cpp:725-784).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "rsh_internal.hpp"
#include "rsh_scatter.h"

using namespace rsh;

struct rsh_sweep_s {
    int device = 0;
    int64_t n = 0;
    int m = 0;
    // class 0's source test per cell, bit 3 f + j for node j of face f (|x -
    // (0.5, 0, 0)| < 0.05, GetCurIllum's expression); empty without nodes
    std::vector<uint16_t> node_in_source;
    std::vector<int32_t> pairs;  // host copy: validation, levels, per-position copies
    struct Dir {
        void* block = nullptr;  // one device allocation per direction
        SweepDirDev dev{};
        int64_t records = 0;
        bool set = false;
    };
    std::vector<Dir> dirs;
    SweepDirDev* d_table = nullptr;
    bool table_dirty = true;
    std::mutex mu;
};

namespace {

bool valid_bound(int32_t t) { return t >= RSH_BOUND_INNER_SOURCE && t <= RSH_BOUND_FREE; }

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

extern "C" {

int rsh_sweep_create(int32_t device, int64_t n_cells, int32_t n_dirs, const int32_t* pairs, const double* nodes,
                     rsh_sweep* out) {
    if (!out || !pairs || n_cells <= 0 || n_dirs <= 0 || n_cells > INT32_MAX / 4)
        return fail(RSH_E_ARG, "invalid sweep size or null pointer");
    for (int64_t i = 0; i < n_cells * 4; i++) {
        const int32_t p = pairs[i];
        if (p >= 0 ? (p >= n_cells * 4 || p / 4 == i / 4 || pairs[p] != i) : !valid_bound(p))
            return fail(RSH_E_ARG, "pairs: entry " + std::to_string(i) +
                                       " is neither a symmetric neighbour face nor a boundary type -1..-4");
    }
    auto* s = new rsh_sweep_s();
    s->device = device;
    s->n = n_cells;
    s->m = n_dirs;
    s->pairs.assign(pairs, pairs + n_cells * 4);
    s->dirs.resize((size_t)n_dirs);
    auto cleanup = [&](int rc) {
        rsh_sweep_destroy(s);
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess) return cleanup(fail(RSH_E_HIP, "hipSetDevice failed"));
    if (hipMalloc(&s->d_table, (size_t)n_dirs * sizeof(SweepDirDev)) != hipSuccess)
        return cleanup(fail(RSH_E_NOMEM, "sweep: device allocation failed"));
    if (nodes) {
        s->node_in_source.assign((size_t)n_cells, 0);
        for (int64_t c = 0; c < n_cells; c++)
            for (int b = 0; b < 12; b++) {
                const double* x = nodes + (c * 12 + b) * 3;
                const double dx = x[0] - 0.5, dy = x[1] - 0.0, dz = x[2] - 0.0;
                if (std::sqrt(dx * dx + dy * dy + dz * dz) < 0.05) s->node_in_source[(size_t)c] |= (uint16_t)(1u << b);
            }
    }
    *out = s;
    return RSH_OK;
}

int rsh_sweep_destroy(rsh_sweep s) {
    if (!s) return RSH_OK;
    (void)hipSetDevice(s->device);
    for (auto& d : s->dirs)
        if (d.block) (void)hipFree(d.block);
    if (s->d_table) (void)hipFree(s->d_table);
    delete s;
    return RSH_OK;
}

int rsh_sweep_set_direction(rsh_sweep s, int32_t k, const int32_t* order, const int32_t* face_state,
                            const double* x0_s, const uint8_t* x0_face, int64_t n_records) {
    if (!s || k < 0 || k >= s->m || !order || !face_state || n_records < 0 || (n_records > 0 && (!x0_s || !x0_face)))
        return fail(RSH_E_ARG, "invalid direction, null pointer or record count");
    const int64_t n = s->n;
    std::vector<int32_t> pos((size_t)n, -1), level((size_t)n, 0);
    for (int64_t h = 0; h < n; h++) {
        const int32_t c = order[h];
        if (c < 0 || c >= n || pos[(size_t)c] >= 0) return fail(RSH_E_ARG, "order is not a permutation of the cells");
        pos[(size_t)c] = (int32_t)h;
    }
    std::vector<int64_t> x0_off((size_t)n);
    int64_t r = 0;
    int32_t levels = 0;
    for (int64_t h = 0; h < n; h++) {
        const int32_t c = order[h];
        const int st = face_state[c];
        if (st < 0 || st > 15) return fail(RSH_E_ARG, "face state outside 0..15");
        x0_off[(size_t)c] = r;
        int32_t lv = 0;
        for (int f = 0; f < 4; f++) {
            const int32_t p = s->pairs[(size_t)(c * 4 + f)];
            const bool in = (st >> f) & 1;
            if (!in) {
                for (int j = 0; j < 3; j++, r++) {
                    if (r >= n_records) return fail(RSH_E_ARG, "fewer x0 records than out-face nodes");
                    const int g = x0_face[r];
                    if (g > 3 || !((st >> g) & 1))
                        return fail(RSH_E_ARG, "x0 record " + std::to_string(r) + " does not name an in face");
                }
            }
            if (p < 0) continue;
            const int32_t u = p / 4;
            // Reads across an in face, and a face both cells treat as out
            // (written by both: the later one in the order wins), order u first.
            const bool dep = in || (!((face_state[u] >> (p % 4)) & 1) && pos[(size_t)u] < h);
            if (!dep) continue;
            if (pos[(size_t)u] > h)
                return fail(RSH_E_ARG, "order is not upwind: cell " + std::to_string(c) + " reads cell " +
                                           std::to_string(u) + " that comes later");
            lv = std::max(lv, level[(size_t)u] + 1);
        }
        level[(size_t)c] = lv;
        levels = std::max(levels, lv + 1);
    }
    if (r != n_records) return fail(RSH_E_ARG, "more x0 records than out-face nodes");
    // Cells by level, ascending cell number inside a level (cells of one level
    // are independent; neighbouring numbers share cache lines of the per-cell
    // arrays the pass reads and writes).
    std::vector<int32_t> level_ptr((size_t)levels + 1, 0), cells((size_t)n);
    for (int64_t c = 0; c < n; c++) level_ptr[(size_t)level[(size_t)c] + 1]++;
    for (int32_t l = 0; l < levels; l++) level_ptr[(size_t)l + 1] += level_ptr[(size_t)l];
    std::vector<int32_t> fill(level_ptr.begin(), level_ptr.end() - 1);
    for (int64_t c = 0; c < n; c++) cells[(size_t)fill[(size_t)level[(size_t)c]]++] = (int32_t)c;
    // Per position in the level order: cell, face states, write-back faces (out
    // on both sides with the neighbour earlier in the given order: the later
    // writer wins both entries, as in the sequential walk), the in-face ids of
    // the cell's records, its pairs; and the records' s values.
    std::vector<SweepPos> posv((size_t)n);
    std::vector<double> s9((size_t)n * kSweepRecStride, 0.0);
    for (int64_t i = 0; i < n; i++) {
        const int32_t c = cells[(size_t)i];
        const int st = face_state[c];
        uint32_t meta = (uint32_t)st;
        SweepPos& P = posv[(size_t)i];
        P.cell = c;
        P.ring[0] = P.ring[1] = 0;
        int q = 0;
        for (int f = 0; f < 4; f++) {
            const int32_t p = s->pairs[(size_t)(c * 4 + f)];
            P.pr[f] = p;
            if ((st >> f) & 1) continue;
            if (p >= 0) {
                const int32_t u = p / 4;
                if (!((face_state[u] >> (p % 4)) & 1) && pos[(size_t)u] < pos[(size_t)c]) meta |= 1u << (4 + f);
            }
            if (!s->node_in_source.empty())
                P.ring[f / 2] |= (uint32_t)((s->node_in_source[(size_t)c] >> (3 * f)) & 7u) << (16 * (f % 2));
            for (int j = 0; j < 3; j++, q++) {
                const int64_t r = x0_off[(size_t)c] + q;  // records of the cell in out-face order
                s9[(size_t)(i * kSweepRecStride + q)] = x0_s[r];
                meta |= (uint32_t)x0_face[r] << (8 + 2 * q);
            }
        }
        P.meta = meta;
    }
    std::vector<int32_t> pass;
    for (int32_t l = 0; l < levels; l++)
        for (int32_t b0 = level_ptr[(size_t)l]; b0 < level_ptr[(size_t)l + 1]; b0 += kSweepBlock) {
            pass.push_back(b0);
            pass.push_back(std::min(kSweepBlock, level_ptr[(size_t)l + 1] - b0));
        }
    const int32_t passes = (int32_t)(pass.size() / 2);
    // Ring codes: an in face whose upstream cell was swept 1..kSweepRingDelta
    // passes earlier reads that cell's values from the LDS ring. A pass that
    // rewrites the entry of a cell swept one pass earlier (a face out on both
    // sides) is flagged: the stores of the pass before it are completed ahead
    // of its barrier, so that the later write lands last.
    {
        std::vector<int32_t> pass_of((size_t)n), where((size_t)n);
        for (int32_t pi = 0; pi < passes; pi++) {
            const int32_t b0 = pass[(size_t)(2 * pi)], cnt = pass[(size_t)(2 * pi + 1)];
            for (int32_t j = 0; j < cnt; j++) pass_of[(size_t)(b0 + j)] = pi;
        }
        for (int64_t i = 0; i < n; i++) where[(size_t)cells[(size_t)i]] = (int32_t)i;
        for (int64_t i = 0; i < n; i++) {
            SweepPos& P = posv[(size_t)i];
            for (int f = 0; f < 4; f++) {
                const int32_t p = P.pr[f];
                if (p >= 0 && ((P.meta >> (4 + f)) & 1) &&
                    pass_of[(size_t)i] - pass_of[(size_t)where[(size_t)(p / 4)]] == 1)
                    pass[(size_t)(2 * pass_of[(size_t)i] + 1)] |= (int32_t)0x80000000u;
                if (p < 0 || !((P.meta >> f) & 1)) continue;
                const int32_t up = where[(size_t)(p / 4)];
                const int32_t delta = pass_of[(size_t)i] - pass_of[(size_t)up];
                if (delta < 1 || delta > kSweepRingDelta) continue;
                const uint32_t code = ((uint32_t)delta << 9) | (uint32_t)(up - pass[(size_t)(2 * pass_of[(size_t)up])]);
                P.ring[f / 2] |= code << (16 * (f % 2));
            }
        }
    }

    const size_t b_pos = align256((size_t)n * sizeof(SweepPos)), b_s9 = align256((size_t)n * kSweepRecStride * 8),
                 b_pass = align256(pass.size() * 4);
    std::lock_guard<std::mutex> lock(s->mu);
    RSH_HIP_TRY(hipSetDevice(s->device));
    auto& d = s->dirs[(size_t)k];
    if (d.block) RSH_HIP_TRY(hipFree(d.block));
    d.block = nullptr;
    d.set = false;
    if (hipMalloc(&d.block, b_pos + b_s9 + b_pass) != hipSuccess)
        return fail(RSH_E_NOMEM, "sweep: direction allocation failed");
    char* base = static_cast<char*>(d.block);
    RSH_HIP_TRY(hipMemcpy(base, posv.data(), (size_t)n * sizeof(SweepPos), hipMemcpyHostToDevice));
    RSH_HIP_TRY(hipMemcpy(base + b_pos, s9.data(), (size_t)n * kSweepRecStride * 8, hipMemcpyHostToDevice));
    RSH_HIP_TRY(hipMemcpy(base + b_pos + b_s9, pass.data(), pass.size() * 4, hipMemcpyHostToDevice));
    d.dev.pos = reinterpret_cast<const SweepPos*>(base);
    d.dev.s9 = reinterpret_cast<const double*>(base + b_pos);
    d.dev.pass = reinterpret_cast<const int32_t*>(base + b_pos + b_s9);
    d.dev.passes = passes;
    d.dev.levels = levels;
    d.records = n_records;
    d.set = true;
    s->table_dirty = true;
    return RSH_OK;
}

int rsh_sweep_levels(rsh_sweep s, int32_t k, int32_t* levels) {
    if (!s || k < 0 || k >= s->m || !levels) return fail(RSH_E_ARG, "invalid direction or null pointer");
    if (!s->dirs[(size_t)k].set) return fail(RSH_E_STATE, "direction not set");
    *levels = s->dirs[(size_t)k].dev.levels;
    return RSH_OK;
}

int rsh_sweep_run(rsh_sweep s, int32_t cls, const double* d_scattering, int64_t ld, const double* d_illum_old,
                  double* d_illum_new, double* d_norm, void* stream) {
    if (!s || !d_scattering || !d_illum_new || ld < s->n) return fail(RSH_E_ARG, "null pointer or ld < n_cells");
    if (cls != 0 && cls != 1 && cls != 2 && cls != 3 && cls != 10)
        return fail(RSH_E_UNSUPPORTED, "class_vtk must be 0, 1, 2, 3 or 10 (11 couples to the gas dynamics state)");
    if (cls == 0 && s->node_in_source.empty()) return fail(RSH_E_ARG, "class 0 needs the face nodes (rsh_sweep_create nodes)");
    if (d_illum_old && (d_illum_old == d_illum_new || !d_norm))
        return fail(RSH_E_ARG, "the residual needs distinct old and new arrays and a norm array");
    if ((reinterpret_cast<uintptr_t>(d_illum_new) | reinterpret_cast<uintptr_t>(d_illum_old)) & 15)
        return fail(RSH_E_ARG, "illum arrays must be 16-byte aligned");
    std::lock_guard<std::mutex> lock(s->mu);
    for (const auto& d : s->dirs)
        if (!d.set) return fail(RSH_E_STATE, "every direction needs rsh_sweep_set_direction first");
    RSH_HIP_TRY(hipSetDevice(s->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (s->table_dirty) {
        std::vector<SweepDirDev> table;
        for (const auto& d : s->dirs) table.push_back(d.dev);
        RSH_HIP_TRY(hipMemcpy(s->d_table, table.data(), table.size() * sizeof(SweepDirDev), hipMemcpyHostToDevice));
        s->table_dirty = false;
    }
    RSH_HIP_TRY(launch_sweep(cls, s->d_table, s->m, s->n, d_scattering, ld, d_illum_old, d_illum_new,
                             d_norm, st));
    return RSH_OK;
}

}  // extern "C"

// Serving loop for scattering-solve requests (C ABI: rsh_server_*).
//
// A request is one source-iteration problem (cells, directions, phase
// function, albedo, seed) that iterates until its FP64 residual falls below its
// tolerance. The server keeps a FIFO of pending requests and one batch of
// resident ones (rsh_batch). Each step:
//   1. admits pending requests, in arrival order, while a slot is free and the
//      HBM footprint of the admitted set stays within the budget (reserved once
//      as the batch's device arena, so admission allocates nothing); their
//      intensities and source terms are generated on the device from the seed;
//   2. runs `iters_per_step` iterations of every resident request (grouped
//      launches replayed from a captured hipGraph);
//   3. reads the residuals and retires every request that converged (or hit
//      its iteration cap), freeing its slot and memory.
// Requests therefore join and leave between graph launches while others keep
// iterating (continuous batching). Timestamps use a monotonic clock.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <random>
#include <thread>
#include <tuple>
#include <vector>

#include "rsh_objects.hpp"

using namespace rsh;

namespace {

double now_s() {
    using clk = std::chrono::steady_clock;
    static const clk::time_point t0 = clk::now();
    return std::chrono::duration<double>(clk::now() - t0).count();
}

struct Request {
    rsh_request_desc desc{};
    int status = RSH_REQ_PENDING;
    int slot = -1;
    int iterations = 0;
    double residual = -1.0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int64_t bytes = 0;
    double* d_result = nullptr;  // FP64 [M][N] when keep_result
};

}  // namespace

struct rsh_server_s {
    int device = 0;
    int64_t budget = 0;
    int max_slots = 0;
    int iters_per_step = 1;
    int keep_results = 0;
    rsh_batch batch = nullptr;
    std::map<std::tuple<int, int, double>, rsh_dirset> dirsets;  // (M, phase, g)
    std::vector<Request> reqs;
    std::deque<int> pending;
    std::vector<int> slot_req;  // slot -> request id
    int64_t used = 0;
    int64_t completed = 0;
    int64_t pair_iterations = 0;  // sum over retired requests of N * M * iterations
};

namespace {

int get_dirset(rsh_server srv, int m, int phase, double g, rsh_dirset* out) {
    auto key = std::make_tuple(m, phase, g);
    auto it = srv->dirsets.find(key);
    if (it != srv->dirsets.end()) {
        *out = it->second;
        return RSH_OK;
    }
    std::vector<double> xyz(3 * (size_t)m), w(m);
    double F = 0;
    int rc = rsh_fibonacci_directions(m, xyz.data(), w.data(), &F);
    if (rc) return rc;
    rsh_dirset ds = nullptr;
    rc = rsh_dirset_create(srv->device, m, xyz.data(), w.data(), F, phase, g, 1, &ds);
    if (rc) return rc;
    srv->dirsets[key] = ds;
    *out = ds;
    return RSH_OK;
}

// Admits a pending request into a batch slot; its data are generated on the
// device in the slot. RSH_E_NOMEM = no arena block fits right now (the request
// stays pending).
int admit(rsh_server srv, int id) {
    Request& r = srv->reqs[id];
    rsh_dirset ds = nullptr;
    int rc = get_dirset(srv, r.desc.n_dirs, r.desc.phase, r.desc.g, &ds);
    if (rc) return rc;
    int slot = -1;
    rc = rsh_batch_add_seeded(srv->batch, ds, r.desc.n_cells, r.desc.seed, r.desc.albedo, r.desc.form, &slot);
    if (rc) return rc;
    r.slot = slot;
    r.status = RSH_REQ_RUNNING;
    r.t_admit = now_s();
    srv->slot_req[slot] = id;
    srv->used += r.bytes;
    return RSH_OK;
}

int retire(rsh_server srv, int id, int status) {
    Request& r = srv->reqs[id];
    if (srv->keep_results) {
        RSH_HIP_TRY(hipMalloc(&r.d_result, (size_t)r.desc.n_dirs * r.desc.n_cells * sizeof(double)));
        int rc = rsh_batch_get_state(srv->batch, r.slot, r.d_result);
        if (rc) return rc;
    }
    int rc = rsh_batch_remove(srv->batch, r.slot);
    if (rc) return rc;
    srv->slot_req[r.slot] = -1;
    r.slot = -1;
    r.status = status;
    r.t_done = now_s();
    srv->used -= r.bytes;
    srv->completed++;
    srv->pair_iterations += (int64_t)r.desc.n_cells * r.desc.n_dirs * r.iterations;
    return RSH_OK;
}

}  // namespace

extern "C" {

int rsh_server_create(int32_t device, int64_t hbm_budget_bytes, int32_t max_slots, int32_t iters_per_step,
                      int32_t keep_results, rsh_server* out) {
    if (!out || hbm_budget_bytes <= 0 || max_slots <= 0 || iters_per_step <= 0)
        return fail(RSH_E_ARG, "invalid server arguments");
    *out = nullptr;
    auto srv = std::make_unique<rsh_server_s>();
    srv->device = device;
    srv->budget = hbm_budget_bytes;
    srv->max_slots = max_slots;
    srv->iters_per_step = iters_per_step;
    srv->keep_results = keep_results ? 1 : 0;
    srv->slot_req.assign(max_slots, -1);
    int rc = rsh_batch_create(device, max_slots, 1, &srv->batch);
    if (rc) return rc;
    // The budget is reserved once as the slots' arena: admission then costs no
    // device allocation and no device-wide synchronisation.
    rc = rsh_batch_reserve(srv->batch, hbm_budget_bytes);
    if (rc) {
        rsh_batch_destroy(srv->batch);
        return rc;
    }
    now_s();
    *out = srv.release();
    return RSH_OK;
}

int rsh_server_destroy(rsh_server srv) {
    if (!srv) return RSH_OK;
    hipSetDevice(srv->device);
    rsh_batch_destroy(srv->batch);
    for (auto& r : srv->reqs) hipFree(r.d_result);
    for (auto& kv : srv->dirsets) rsh_dirset_destroy(kv.second);
    delete srv;
    return RSH_OK;
}

int rsh_server_submit(rsh_server srv, const rsh_request_desc* d, int32_t* request_id) {
    if (!srv || !d || !request_id) return fail(RSH_E_ARG, "null argument");
    if (d->n_cells <= 0 || d->n_dirs <= 0 || !(d->tolerance > 0) || d->max_iterations <= 0)
        return fail(RSH_E_ARG, "invalid request");
    if (d->phase != RSH_PHASE_THOMSON && d->phase != RSH_PHASE_HG) return fail(RSH_E_ARG, "unknown phase function");
    // Requests the batch cannot hold (unknown form, FP64 with 2^24 cells or
    // more) or whose footprint (the arena's own rounding included) exceeds the
    // whole budget are refused here, so that none can wait for ever.
    int64_t bytes = 0;
    int rc = rsh_batch_footprint(d->form, d->n_cells, d->n_dirs, &bytes);
    if (rc) return rc;
    if (bytes > srv->budget) return fail(RSH_E_NOMEM, "request footprint exceeds the server's HBM budget");
    Request r;
    r.desc = *d;
    r.bytes = bytes;
    r.t_submit = now_s();
    srv->reqs.push_back(r);
    const int id = (int)srv->reqs.size() - 1;
    srv->pending.push_back(id);
    *request_id = id;
    return RSH_OK;
}

int rsh_server_cancel(rsh_server srv, int32_t request_id) {
    if (!srv || request_id < 0 || request_id >= (int)srv->reqs.size()) return fail(RSH_E_ARG, "invalid request id");
    Request& r = srv->reqs[request_id];
    if (r.status == RSH_REQ_PENDING) {
        srv->pending.erase(std::remove(srv->pending.begin(), srv->pending.end(), request_id), srv->pending.end());
        r.status = RSH_REQ_CANCELLED;
        r.t_done = now_s();
        return RSH_OK;
    }
    if (r.status == RSH_REQ_RUNNING) {
        RSH_HIP_TRY(hipSetDevice(srv->device));
        int rc = rsh_batch_sync(srv->batch);
        if (rc) return rc;
        rc = rsh_batch_remove(srv->batch, r.slot);
        if (rc) return rc;
        srv->slot_req[r.slot] = -1;
        srv->used -= r.bytes;
        r.slot = -1;
        r.status = RSH_REQ_CANCELLED;
        r.t_done = now_s();
        return RSH_OK;
    }
    return fail(RSH_E_STATE, "request already finished");
}

int rsh_server_step(rsh_server srv, int32_t* running, int32_t* finished_now) {
    if (!srv) return fail(RSH_E_ARG, "null server");
    RSH_HIP_TRY(hipSetDevice(srv->device));
    // 1. admission (FIFO, by footprint)
    auto count_running = [&]() {
        int k = 0;
        for (int s : srv->slot_req)
            if (s >= 0) k++;
        return k;
    };
    while (!srv->pending.empty()) {
        const int id = srv->pending.front();
        Request& r = srv->reqs[id];
        const bool slot_free = std::find(srv->slot_req.begin(), srv->slot_req.end(), -1) != srv->slot_req.end();
        if (!slot_free || srv->used + r.bytes > srv->budget) break;
        int rc = admit(srv, id);
        if (rc == RSH_E_NOMEM && count_running() > 0) break;  // arena fragmented: retry after retirements
        if (rc) {
            // Cannot be admitted even with nothing else resident (or the batch
            // refused it): the request fails instead of blocking the queue.
            r.status = RSH_REQ_FAILED;
            r.t_done = now_s();
            srv->pending.pop_front();
            continue;
        }
        srv->pending.pop_front();
    }
    const int nrun = count_running();
    if (running) *running = nrun;
    if (finished_now) *finished_now = 0;
    if (nrun == 0) return RSH_OK;
    // 2. iterate every resident request
    int rc = rsh_batch_run(srv->batch, srv->iters_per_step);
    if (rc) return rc;
    // 3. residuals -> retire
    std::vector<double> resid(srv->max_slots);
    std::vector<int32_t> iters(srv->max_slots);
    rc = rsh_batch_residuals(srv->batch, resid.data(), iters.data());
    if (rc) return rc;
    int done = 0;
    for (int s = 0; s < srv->max_slots; s++) {
        const int id = srv->slot_req[s];
        if (id < 0) continue;
        Request& r = srv->reqs[id];
        r.iterations = iters[s];
        r.residual = resid[s];
        int status = -1;
        if (resid[s] >= 0 && resid[s] < r.desc.tolerance) status = RSH_REQ_CONVERGED;
        else if (r.iterations >= r.desc.max_iterations) status = RSH_REQ_MAXED;
        if (status >= 0) {
            if ((rc = retire(srv, id, status))) return rc;
            done++;
        }
    }
    if (finished_now) *finished_now = done;
    return RSH_OK;
}

int rsh_server_poll(rsh_server srv, int32_t request_id, rsh_request_info* info) {
    if (!srv || !info || request_id < 0 || request_id >= (int)srv->reqs.size())
        return fail(RSH_E_ARG, "invalid request id");
    const Request& r = srv->reqs[request_id];
    info->status = r.status;
    info->iterations = r.iterations;
    info->residual = r.residual;
    info->t_submit = r.t_submit;
    info->t_admit = r.t_admit;
    info->t_done = r.t_done;
    info->footprint_bytes = r.bytes;
    return RSH_OK;
}

int rsh_server_result(rsh_server srv, int32_t request_id, double* d_out) {
    if (!srv || !d_out || request_id < 0 || request_id >= (int)srv->reqs.size())
        return fail(RSH_E_ARG, "invalid request id");
    const Request& r = srv->reqs[request_id];
    if (!r.d_result) return fail(RSH_E_STATE, "no stored result (request unfinished or keep_results = 0)");
    RSH_HIP_TRY(hipMemcpy(d_out, r.d_result, (size_t)r.desc.n_dirs * r.desc.n_cells * sizeof(double),
                          hipMemcpyDeviceToDevice));
    return RSH_OK;
}

int rsh_server_stats(rsh_server srv, int64_t* used_bytes, int32_t* pending, int32_t* running, int64_t* completed) {
    if (!srv) return fail(RSH_E_ARG, "null server");
    if (used_bytes) *used_bytes = srv->used;
    if (pending) *pending = (int32_t)srv->pending.size();
    if (running) {
        int n = 0;
        for (int s : srv->slot_req)
            if (s >= 0) n++;
        *running = n;
    }
    if (completed) *completed = srv->completed;
    return RSH_OK;
}

double rsh_server_clock(void) { return now_s(); }

// Real-time Poisson driver: submits requests[i] at t0 + arrival[i] (seconds,
// ascending) and steps the server until every request finished. Latency of a
// request = completion - arrival.
int rsh_server_run_trace(rsh_server srv, int32_t n, const rsh_request_desc* requests, const double* arrival,
                         double* latency, double* wall_seconds) {
    if (!srv || n < 0 || (n > 0 && (!requests || !arrival || !latency)) || !wall_seconds)
        return fail(RSH_E_ARG, "invalid trace arguments");
    std::vector<int> ids(n, -1);
    const double t0 = now_s();
    int next = 0, finished = 0;
    std::vector<double> arrive_abs(n);
    while (finished < n) {
        const double t = now_s() - t0;
        while (next < n && arrival[next] <= t) {
            int id = -1;
            int rc = rsh_server_submit(srv, &requests[next], &id);
            if (rc) return rc;
            ids[next] = id;
            arrive_abs[next] = t0 + arrival[next];
            next++;
        }
        int running = 0, done = 0;
        int rc = rsh_server_step(srv, &running, &done);
        if (rc) return rc;
        if (running == 0 && srv->pending.empty() && next < n) {
            // idle until the next arrival
            const double wait = t0 + arrival[next] - now_s();
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
        }
        finished = 0;
        for (int i = 0; i < next; i++) {
            const Request& r = srv->reqs[ids[i]];
            if (r.status != RSH_REQ_PENDING && r.status != RSH_REQ_RUNNING) finished++;
        }
    }
    for (int i = 0; i < n; i++) latency[i] = srv->reqs[ids[i]].t_done - arrive_abs[i];
    *wall_seconds = now_s() - t0;
    return RSH_OK;
}

}  // extern "C"

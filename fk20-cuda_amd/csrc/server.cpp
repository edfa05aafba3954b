// Continuous-batching request loop over a context (fk20_server_* in include/fk20.h).
//
// Requests are host polynomial blobs of one setup each.  A worker thread keeps up to
// `nslots` batches in flight, one per stream.  Whenever a slot is free it takes pending
// requests in arrival order while they fit the slot's share of the device-memory budget
// (per row: toeplitz_fft 16*n2 Fr + hext n2 G1 workspace, plus the polynomial in and h_fft
// out in the context's output format, fixed when the server is created), packs them per setup into one fk20_prove_batch_ws pass (copy in, prove, copy out on
// the slot's stream) and records an event.  Completed slots are reaped by event query, so
// requests that arrive while batches run are admitted as soon as a slot frees: batches on
// different streams overlap on the device.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/fk20.h"
#include "common.hip.h"

using fk20::status_from_hip;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t id;
    int setup;
    unsigned N, rows;
    const fk20_fr_t *poly;  // host
    fk20_g1p_t *out;        // host
    Clock::time_point t_submit, t_done;
    int status = 0;
    bool done = false;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done_ev = nullptr;
    char *dev = nullptr;  // slot arena: inputs, outputs, workspace
    size_t bytes = 0;
    bool busy = false;
    std::vector<uint64_t> reqs;
};

// device bytes of one row: workspace, polynomial in, h_fft out (in the context's format)
size_t row_bytes(unsigned N, size_t point_bytes) {
    const size_t n2 = N / 8;
    return fk20_rows_workspace_bytes(1, N) + (size_t)N * 32 + n2 * point_bytes;
}

}  // namespace

struct fk20_server {
    fk20_ctx_t *ctx = nullptr;
    int device = 0;
    size_t budget = 0;
    unsigned max_batch_rows = 0;
    size_t point_bytes = 144;  // output bytes per point, from the context's format at creation
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, Request> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    // statistics
    uint64_t batches = 0, rows_done = 0, max_inflight_rows = 0, inflight_rows = 0;
    int first_error = 0;
    void run();
    int launch(Slot &s, std::vector<uint64_t> &ids);
};

int fk20_server::launch(Slot &s, std::vector<uint64_t> &ids) {
    // output buffers were sized for the format fixed at creation
    int format = FK20_OUT_G1P;
    if (fk20_ctx_get_output(ctx, &format) != FK20_OK || fk20_output_point_bytes(format) != point_bytes)
        return FK20_EINVAL;
    // group the requests by setup: one batch item per setup, rows packed contiguously
    std::map<int, std::vector<uint64_t>> by_setup;
    for (uint64_t id : ids) by_setup[reqs[id].setup].push_back(id);
    std::vector<fk20_batch_item_t> items;
    char *p = s.dev;
    size_t ws_bytes = 0;
    struct Copy { fk20_g1p_t *dst; const char *src; size_t bytes; };
    std::vector<Copy> outs;
    for (auto &kv : by_setup) {
        unsigned rows = 0, N = reqs[kv.second[0]].N;
        for (uint64_t id : kv.second) rows += reqs[id].rows;
        ws_bytes += fk20_rows_workspace_bytes(rows, N);
        fk20_batch_item_t it;
        it.setup_id = kv.first;
        it.rows = rows;
        it.polynomial = reinterpret_cast<const fk20_fr_t *>(p);
        // copy the polynomials in
        for (uint64_t id : kv.second) {
            const Request &r = reqs[id];
            const size_t b = (size_t)r.rows * r.N * 32;
            hipError_t e = hipMemcpyAsync(p, r.poly, b, hipMemcpyHostToDevice, s.stream);
            if (e != hipSuccess) return status_from_hip(e);
            p += b;
        }
        it.h_fft = reinterpret_cast<fk20_g1p_t *>(p);
        for (uint64_t id : kv.second) {
            const Request &r = reqs[id];
            const size_t b = (size_t)r.rows * (r.N / 8) * point_bytes;
            outs.push_back({r.out, p, b});
            p += b;
        }
        items.push_back(it);
    }
    const size_t used = (size_t)(p - s.dev);
    int rc = fk20_prove_batch_ws(ctx, items.data(), (unsigned)items.size(), p, s.bytes - used, s.stream);
    if (rc != FK20_OK) return rc;
    for (const Copy &c : outs) {
        hipError_t e = hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, s.stream);
        if (e != hipSuccess) return status_from_hip(e);
    }
    hipError_t e = hipEventRecord(s.done_ev, s.stream);
    if (e != hipSuccess) return status_from_hip(e);
    (void)ws_bytes;
    return FK20_OK;
}

void fk20_server::run() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        bool progress = false;
        // reap finished slots
        for (Slot &s : slots) {
            if (!s.busy) continue;
            hipError_t q = hipEventQuery(s.done_ev);
            if (q == hipErrorNotReady) continue;
            const auto now = Clock::now();
            for (uint64_t id : s.reqs) {
                Request &r = reqs[id];
                r.t_done = now;
                r.done = true;
                r.status = q == hipSuccess ? FK20_OK : status_from_hip(q);
                rows_done += r.rows;
                inflight_rows -= r.rows;
            }
            if (q != hipSuccess && !first_error) first_error = status_from_hip(q);
            s.reqs.clear();
            s.busy = false;
            progress = true;
            cv_done.notify_all();
        }
        // admit pending requests into free slots
        for (Slot &s : slots) {
            if (s.busy || pending.empty()) continue;
            std::vector<uint64_t> take;
            size_t bytes = 0;
            unsigned rows = 0;
            std::map<int, int> setups;
            while (!pending.empty()) {
                const Request &r = reqs[pending.front()];
                const size_t b = (size_t)r.rows * row_bytes(r.N, point_bytes);
                const bool new_setup = !setups.count(r.setup);
                if (!take.empty() && (bytes + b > s.bytes || rows + r.rows > max_batch_rows ||
                                      (new_setup && setups.size() >= 16)))
                    break;
                take.push_back(pending.front());
                pending.pop_front();
                bytes += b;
                rows += r.rows;
                setups[r.setup] = 1;
            }
            if (take.empty()) continue;
            int rc = launch(s, take);
            if (rc != FK20_OK) {
                for (uint64_t id : take) {
                    Request &r = reqs[id];
                    r.status = rc;
                    r.done = true;
                    r.t_done = Clock::now();
                }
                if (!first_error) first_error = rc;
                cv_done.notify_all();
                continue;
            }
            s.reqs = take;
            s.busy = true;
            batches++;
            inflight_rows += rows;
            if (inflight_rows > max_inflight_rows) max_inflight_rows = inflight_rows;
            progress = true;
        }
        bool any_busy = false;
        for (Slot &s : slots) any_busy |= s.busy;
        if (stop && pending.empty() && !any_busy) break;
        if (!progress) {
            // poll completions every 20 us while work is in flight, else sleep until a submit
            if (any_busy) cv_work.wait_for(lk, std::chrono::microseconds(20));
            else cv_work.wait(lk, [&] { return stop || !pending.empty(); });
        }
    }
}

extern "C" {

int fk20_server_create(fk20_server_t **out, fk20_ctx_t *ctx, size_t budget_bytes, unsigned nslots,
                       unsigned max_batch_rows) {
    if (!out || !ctx || !nslots || nslots > 8 || !max_batch_rows) return FK20_EINVAL;
    int format = FK20_OUT_G1P;
    if (fk20_ctx_get_output(ctx, &format) != FK20_OK) return FK20_EINVAL;
    const size_t point_bytes = fk20_output_point_bytes(format);
    // each slot must hold at least one N=4096 row
    if (budget_bytes / nslots < row_bytes(4096, point_bytes)) return FK20_EINVAL;
    fk20_server *srv = new fk20_server();
    srv->ctx = ctx;
    srv->point_bytes = point_bytes;
    srv->budget = budget_bytes;
    srv->max_batch_rows = max_batch_rows;
    hipError_t e = hipGetDevice(&srv->device);
    srv->slots.resize(nslots);
    for (Slot &s : srv->slots) {
        if (e != hipSuccess) break;
        s.bytes = budget_bytes / nslots;
        e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done_ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&s.dev), s.bytes);
    }
    if (e != hipSuccess) {
        for (Slot &s : srv->slots) {
            if (s.dev) (void)hipFree(s.dev);
            if (s.done_ev) (void)hipEventDestroy(s.done_ev);
            if (s.stream) (void)hipStreamDestroy(s.stream);
        }
        delete srv;
        return status_from_hip(e);
    }
    srv->worker = std::thread([srv] { srv->run(); });
    *out = srv;
    return FK20_OK;
}

int fk20_server_submit(fk20_server_t *srv, int setup_id, const fk20_fr_t *host_poly, unsigned rows,
                       fk20_g1p_t *host_h_fft, uint64_t *request_id) {
    if (!srv || !host_poly || !host_h_fft || !rows || !request_id) return FK20_EINVAL;
    unsigned N = 0, win = 0;
    int rc = fk20_setup_info(srv->ctx, setup_id, &N, &win);
    if (rc != FK20_OK) return rc;
    if ((size_t)rows * row_bytes(N, srv->point_bytes) > srv->budget / srv->slots.size() || rows > srv->max_batch_rows)
        return FK20_ENOMEM;  // larger than one slot's share of the budget
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stop) return FK20_EINVAL;
    Request r;
    r.id = srv->next_id++;
    r.setup = setup_id;
    r.N = N;
    r.rows = rows;
    r.poly = host_poly;
    r.out = host_h_fft;
    r.t_submit = Clock::now();
    srv->reqs.emplace(r.id, r);
    srv->pending.push_back(r.id);
    *request_id = r.id;
    srv->cv_work.notify_one();
    return FK20_OK;
}

int fk20_server_wait(fk20_server_t *srv, uint64_t id, double *latency_s) {
    if (!srv) return FK20_EINVAL;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return FK20_EINVAL;
    srv->cv_done.wait(lk, [&] { return it->second.done; });
    if (latency_s)
        *latency_s = std::chrono::duration<double>(it->second.t_done - it->second.t_submit).count();
    const int status = it->second.status;
    srv->reqs.erase(it);
    return status;
}

int fk20_server_stats(fk20_server_t *srv, uint64_t *batches, uint64_t *rows_done, uint64_t *max_inflight_rows) {
    if (!srv) return FK20_EINVAL;
    std::lock_guard<std::mutex> lk(srv->mu);
    if (batches) *batches = srv->batches;
    if (rows_done) *rows_done = srv->rows_done;
    if (max_inflight_rows) *max_inflight_rows = srv->max_inflight_rows;
    return srv->first_error;
}

int fk20_server_destroy(fk20_server_t *srv) {
    if (!srv) return FK20_EINVAL;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
        srv->cv_work.notify_all();
    }
    srv->worker.join();
    for (Slot &s : srv->slots) {
        (void)hipStreamSynchronize(s.stream);
        (void)hipFree(s.dev);
        (void)hipEventDestroy(s.done_ev);
        (void)hipStreamDestroy(s.stream);
    }
    const int err = srv->first_error;
    delete srv;
    return err;
}

int fk20_host_alloc(void **ptr, size_t bytes) {
    if (!ptr) return FK20_EINVAL;
    return status_from_hip(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
}

int fk20_host_free(void *ptr) { return status_from_hip(hipHostFree(ptr)); }

}  // extern "C"

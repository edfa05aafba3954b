// Continuous-batching request loop over a context (fk20_server_* in include/fk20.h).
//
// Requests are host polynomial blobs of one setup each.  A worker thread keeps up to
// `nslots` batches in flight, one per stream.  Whenever a slot is free it takes pending
// requests in arrival order while they fit the free part of the device-memory budget (one
// arena of budget bytes shared by all slots, first-fit; per row: toeplitz_fft 16*n2 Fr + hext
// n2 G1 workspace, plus the polynomial in and h_fft out in the context's output format, fixed
// when the server is created), packs them per setup into one fk20_prove_batch_ws pass (copy
// in, prove, copy out on the slot's stream) and records an event.  The launch itself runs
// without the server lock, so submit and wait never block behind it.  Completed slots are
// reaped by event query and their arena block returned, so requests that arrive while
// batches run are admitted as soon as a slot and memory free up: batches on different streams
// overlap on the device.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <iterator>
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
    size_t off = 0, bytes = 0;  // arena block of the batch in flight: inputs, outputs, workspace
    bool busy = false;
    std::vector<uint64_t> reqs;
};

// What a launch needs of a request, copied out of the shared map under the lock.
struct Job {
    int setup;
    unsigned N, rows;
    const fk20_fr_t *poly;
    fk20_g1p_t *out;
};

// First-fit allocator over the arena offsets [0, size), blocks rounded up to 256 bytes.
class Arena {
  public:
    void reset(size_t size) {
        free_.clear();
        free_[0] = size;
    }
    size_t largest() const {
        size_t m = 0;
        for (const auto &kv : free_) m = kv.second > m ? kv.second : m;
        return m;
    }
    static size_t round(size_t b) { return (b + 255) & ~(size_t)255; }
    bool alloc(size_t bytes, size_t *off) {
        bytes = round(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes, at = it->first + bytes;
            free_.erase(it);
            if (rest) free_[at] = rest;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = round(bytes);
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {  // merge the next block
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {  // merge into the previous block
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }

  private:
    std::map<size_t, size_t> free_;  // offset -> size of each free block
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
    char *arena_dev = nullptr;  // budget bytes of device memory shared by the slots
    Arena arena;
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
    int launch(Slot &s, const std::vector<Job> &jobs) const;
};

// Enqueues one batch on the slot's stream in its arena block; touches no shared state.
int fk20_server::launch(Slot &s, const std::vector<Job> &jobs) const {
    // output buffers were sized for the format fixed at creation
    int format = FK20_OUT_G1P;
    if (fk20_ctx_get_output(ctx, &format) != FK20_OK || fk20_output_point_bytes(format) != point_bytes)
        return FK20_EINVAL;
    // group the requests by setup: one batch item per setup, rows packed contiguously
    std::map<int, std::vector<const Job *>> by_setup;
    for (const Job &j : jobs) by_setup[j.setup].push_back(&j);
    std::vector<fk20_batch_item_t> items;
    char *const base = arena_dev + s.off;
    char *p = base;
    struct Copy { fk20_g1p_t *dst; const char *src; size_t bytes; };
    std::vector<Copy> outs;
    for (auto &kv : by_setup) {
        unsigned rows = 0;
        for (const Job *j : kv.second) rows += j->rows;
        fk20_batch_item_t it;
        it.setup_id = kv.first;
        it.rows = rows;
        it.polynomial = reinterpret_cast<const fk20_fr_t *>(p);
        // copy the polynomials in
        for (const Job *j : kv.second) {
            const size_t b = (size_t)j->rows * j->N * 32;
            hipError_t e = hipMemcpyAsync(p, j->poly, b, hipMemcpyHostToDevice, s.stream);
            if (e != hipSuccess) return status_from_hip(e);
            p += b;
        }
        it.h_fft = reinterpret_cast<fk20_g1p_t *>(p);
        for (const Job *j : kv.second) {
            const size_t b = (size_t)j->rows * (j->N / 8) * point_bytes;
            outs.push_back({j->out, p, b});
            p += b;
        }
        items.push_back(it);
    }
    const size_t used = (size_t)(p - base);
    int rc = fk20_prove_batch_ws(ctx, items.data(), (unsigned)items.size(), p, s.bytes - used, s.stream);
    if (rc != FK20_OK) return rc;
    for (const Copy &c : outs) {
        hipError_t e = hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, s.stream);
        if (e != hipSuccess) return status_from_hip(e);
    }
    hipError_t e = hipEventRecord(s.done_ev, s.stream);
    if (e != hipSuccess) return status_from_hip(e);
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
            arena.release(s.off, s.bytes);
            progress = true;
            cv_done.notify_all();
        }
        // admit pending requests into free slots, within the free part of the budget
        for (Slot &s : slots) {
            if (s.busy || pending.empty()) continue;
            const size_t room = arena.largest();
            std::vector<uint64_t> take;
            std::vector<Job> jobs;
            size_t bytes = 0;
            unsigned rows = 0;
            std::map<int, int> setups;
            while (!pending.empty()) {
                const Request &r = reqs[pending.front()];
                const size_t b = (size_t)r.rows * row_bytes(r.N, point_bytes);
                const bool new_setup = !setups.count(r.setup);
                if (bytes + b > room || rows + r.rows > max_batch_rows || (new_setup && setups.size() >= 16))
                    break;
                take.push_back(pending.front());
                jobs.push_back({r.setup, r.N, r.rows, r.poly, r.out});
                pending.pop_front();
                bytes += b;
                rows += r.rows;
                setups[r.setup] = 1;
            }
            if (take.empty()) break;  // the oldest request waits for memory: keep arrival order
            if (!arena.alloc(bytes, &s.off)) {  // cannot happen: bytes <= the largest free block
                for (auto it = take.rbegin(); it != take.rend(); ++it) pending.push_front(*it);
                break;
            }
            s.bytes = Arena::round(bytes);
            s.busy = true;  // reserved while the launch runs unlocked
            lk.unlock();
            const int rc = launch(s, jobs);
            lk.lock();
            if (rc != FK20_OK) {
                for (uint64_t id : take) {
                    Request &r = reqs[id];
                    r.status = rc;
                    r.done = true;
                    r.t_done = Clock::now();
                }
                if (!first_error) first_error = rc;
                s.busy = false;
                (void)hipStreamSynchronize(s.stream);  // nothing may still use the block
                arena.release(s.off, s.bytes);
                cv_done.notify_all();
                continue;
            }
            s.reqs = take;
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
    // the budget must hold at least one N=4096 row
    if (budget_bytes < Arena::round(row_bytes(4096, point_bytes))) return FK20_EINVAL;
    fk20_server *srv = new fk20_server();
    srv->ctx = ctx;
    srv->point_bytes = point_bytes;
    srv->budget = budget_bytes;
    srv->max_batch_rows = max_batch_rows;
    hipError_t e = hipGetDevice(&srv->device);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&srv->arena_dev), budget_bytes);
    srv->arena.reset(budget_bytes);
    srv->slots.resize(nslots);
    for (Slot &s : srv->slots) {
        if (e != hipSuccess) break;
        e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done_ev, hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        for (Slot &s : srv->slots) {
            if (s.done_ev) (void)hipEventDestroy(s.done_ev);
            if (s.stream) (void)hipStreamDestroy(s.stream);
        }
        if (srv->arena_dev) (void)hipFree(srv->arena_dev);
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
    if (Arena::round((size_t)rows * row_bytes(N, srv->point_bytes)) > srv->budget || rows > srv->max_batch_rows)
        return FK20_ENOMEM;  // larger than the whole budget
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
        (void)hipEventDestroy(s.done_ev);
        (void)hipStreamDestroy(s.stream);
    }
    (void)hipFree(srv->arena_dev);
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

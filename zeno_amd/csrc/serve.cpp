// serve.cpp — the serving loop: continuous batching of circuit requests
// on one GPU.
//
// The reference runs one circuit per process invocation; admission there is
// mem::plan against RAM (src/mem.rs:79-130, exec.rs:261-304).  Here one HBM
// pool is allocated once when the server starts and each admitted request
// gets its state vector carved out of it, so admission is "does the pool
// have room for 2 * 2^n * s bytes right now" and no request allocates its
// state.  A fixed set of worker slots each own a backend (its own HIP
// stream, parameter arena and scratch; the scratch grows when a request
// first needs more, e.g. for a large shot count) that is re-bound to the request's
// slice of the pool; requests that arrive while others run are admitted as
// soon as a slot and pool space are free, and their kernels run
// concurrently with the others' on separate streams.  Admission is first
// fit in arrival order, except that a request that cannot fit is passed
// over at most kMaxSkips times before the server stops admitting behind
// it, so large requests are not starved by a stream of small ones, and that
// requests of 1 GiB of state and more run one at a time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "host.hpp"
#include "zeno_hip.h"

using namespace zeno;

namespace {

constexpr uint64_t kAlign = 1ull << 16;
constexpr int kMaxSkips = 16;
// Requests whose state is at least this large run one at a time: each one's
// sweeps stream the whole chip's HBM, so two side by side both finish later.
// Smaller requests keep the other slots meanwhile.
constexpr uint64_t kLargeBytes = 1ull << 30;

using Clock = std::chrono::steady_clock;

// First-fit allocator over [0, size) with coalescing free list.
class Pool {
  public:
    explicit Pool(uint64_t size) { free_[0] = size; }
    bool alloc(uint64_t bytes, uint64_t *off) {
        bytes = (bytes + kAlign - 1) & ~(kAlign - 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const uint64_t rest = it->second - bytes, start = it->first + bytes;
            free_.erase(it);
            if (rest) free_[start] = rest;
            used_[*off] = bytes;
            return true;
        }
        return false;
    }
    void release(uint64_t off) {
        auto u = used_.find(off);
        if (u == used_.end()) return;
        uint64_t start = off, len = u->second;
        used_.erase(u);
        auto next = free_.lower_bound(start);
        if (next != free_.end() && start + len == next->first) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                prev->second += len;
                return;
            }
        }
        free_[start] = len;
    }
    uint64_t largest_free() const {
        uint64_t m = 0;
        for (const auto &kv : free_) m = std::max(m, kv.second);
        return m;
    }

  private:
    std::map<uint64_t, uint64_t> free_;  // offset -> length
    std::map<uint64_t, uint64_t> used_;
};

struct Request {
    int64_t id = 0;
    zr_compiled circuit;
    int precision = 64;
    uint64_t shots = 0, seed = 0;
    bool noisy = false;
    zr_noise noise{};
    uint64_t bytes = 0, off = 0;
    int skips = 0;
    double t_submit = 0, t_start = 0, t_end = 0;
    int status = ZB_OK;
    std::string error;
    bool done = false;
    std::vector<std::pair<uint64_t, uint64_t>> counts;
};

}  // namespace

struct zs_server {
    char *pool_mem = nullptr;
    uint64_t pool_bytes = 0;
    std::unique_ptr<Pool> pool;
    int mode = 1;
    Clock::time_point epoch = Clock::now();
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending;
    std::unordered_map<int64_t, std::unique_ptr<Request>> requests;
    std::vector<std::thread> workers;
    int64_t next_id = 1;
    int running = 0, running_large = 0;
    bool stop = false;

    double now() const { return std::chrono::duration<double>(Clock::now() - epoch).count(); }

    // Next admissible request (caller holds mu), or nullptr.
    Request *admit() {
        for (size_t i = 0; i < pending.size(); ++i) {
            Request *r = pending[i];
            if (r->bytes >= kLargeBytes && running_large > 0) continue;  // waits for the running one, not starved
            uint64_t off = 0;
            if (pool->alloc(r->bytes, &off)) {
                r->off = off;
                pending.erase(pending.begin() + (std::ptrdiff_t)i);
                return r;
            }
            if (r->skips >= kMaxSkips) return nullptr;  // hold the line behind a starved request
            ++r->skips;
        }
        return nullptr;
    }

    void worker_loop() {
        zb_backend *h = nullptr;
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            Request *r = nullptr;
            cv_work.wait(lk, [&] { return stop || (r = admit()) != nullptr; });
            if (!r) break;
            const bool large = r->bytes >= kLargeBytes;
            ++running;
            running_large += large;
            r->t_start = now();
            lk.unlock();
            run(h, r);
            lk.lock();
            pool->release(r->off);
            --running;
            running_large -= large;
            r->t_end = now();
            r->done = true;
            cv_done.notify_all();
            cv_work.notify_all();
        }
        lk.unlock();
        if (h) zb_destroy(h);
    }

    void run(zb_backend *&h, Request *r) {
        const uint32_t n = r->circuit.c.n_qubits;
        const uint64_t half = r->bytes / 2;
        void *re = pool_mem + r->off, *im = pool_mem + r->off + half;
        int rc = h ? zb_rebind(h, n, r->precision, re, im) : zb_create_bound(n, r->precision, re, im, &h);
        if (rc == ZB_OK)
            rc = run_compiled_on(h, &r->circuit, r->shots, r->seed, r->noisy ? &r->noise : nullptr, mode, true,
                                 r->counts);
        if (rc == ZB_OK && h) rc = zb_finish(h);  // the slice is released after this returns
        if (rc != ZB_OK) {
            r->status = rc;
            r->error = last_error();
        }
    }
};

extern "C" {

int zs_create(uint64_t pool_bytes, int slots, int mode, zs_server **out) {
    *out = nullptr;
    if (slots < 1 || slots > 64) {
        set_error("zs_create: slots must be in 1..64");
        return ZB_ERR_ARG;
    }
    if (mode != 0 && mode != 1) {
        set_error("zs_create: mode must be 0 (exact) or 1 (tiled)");
        return ZB_ERR_ARG;
    }
    if (zb_device_count() == 0) {
        set_error("no HIP device is available (the hip backend needs an AMD GPU)");
        return ZB_ERR_UNSUPPORTED;
    }
    uint64_t free_b = 0, total_b = 0;
    if (const int rc = zb_mem_info(&free_b, &total_b); rc != ZB_OK) return rc;
    if (pool_bytes == 0) pool_bytes = free_b / 10 * 8;  // default: 80% of free HBM
    pool_bytes &= ~(kAlign - 1);
    if (pool_bytes < kAlign || pool_bytes + (256ull << 20) > free_b) {
        set_error("zs_create: a pool of " + std::to_string(pool_bytes) + " B does not fit in " +
                  std::to_string(free_b) + " B of free device memory");
        return ZB_ERR_MEMORY;
    }
    auto *s = new zs_server();
    if (hipMalloc((void **)&s->pool_mem, pool_bytes) != hipSuccess) {
        delete s;
        set_error("zs_create: hipMalloc of the state pool failed");
        return ZB_ERR_MEMORY;
    }
    s->pool_bytes = pool_bytes;
    s->pool = std::make_unique<Pool>(pool_bytes);
    s->mode = mode;
    for (int i = 0; i < slots; ++i) s->workers.emplace_back([s] { s->worker_loop(); });
    *out = s;
    return ZB_OK;
}

int64_t zs_submit(zs_server *s, const zr_compiled *c, int precision, uint64_t shots, uint64_t seed,
                  const zr_noise *noise) {
    if (precision != 32 && precision != 64) {
        set_error("precision must be 32 or 64");
        return ZB_ERR_ARG;
    }
    const uint32_t n = c->c.n_qubits;
    if (n == 0 || n > 35) {
        set_error("the HIP backend supports 1..35 qubits");
        return ZB_ERR_UNSUPPORTED;
    }
    const uint64_t bytes = 2 * ((1ull << n) * (precision == 64 ? 8 : 4));
    if (((bytes + kAlign - 1) & ~(kAlign - 1)) > s->pool_bytes) {  // mem::plan (src/mem.rs:114-130)
        set_error("state vector for " + std::to_string(n) + " qubits at f" + std::to_string(precision) + " needs " +
                  std::to_string(bytes) + " B but the serving pool holds " + std::to_string(s->pool_bytes) + " B");
        return ZB_ERR_MEMORY;
    }
    auto r = std::make_unique<Request>();
    r->circuit.c = c->c;  // the request keeps its own copy of the op stream
    r->precision = precision;
    r->shots = shots;
    r->seed = seed;
    if (noise) {
        r->noisy = true;
        r->noise = *noise;
    }
    r->bytes = bytes;
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->next_id++;
    r->t_submit = s->now();
    const int64_t id = r->id;
    s->pending.push_back(r.get());
    s->requests[id] = std::move(r);
    s->cv_work.notify_one();
    return id;
}

int64_t zs_wait(zs_server *s, int64_t id, uint64_t *keys, uint64_t *counts, int64_t cap, double *times) {
    std::unique_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->requests.find(id);
        if (it == s->requests.end()) {
            set_error("zs_wait: unknown request id " + std::to_string(id));
            return ZB_ERR_ARG;
        }
        Request *p = it->second.get();
        s->cv_done.wait(lk, [&] { return p->done; });
        r = std::move(it->second);
        s->requests.erase(it);
    }
    if (times) times[0] = r->t_submit, times[1] = r->t_start, times[2] = r->t_end;
    if (r->status != ZB_OK) {
        set_error(r->error);
        return r->status;
    }
    if ((int64_t)r->counts.size() > cap) {
        set_error("key buffer too small");
        return ZB_ERR_ARG;
    }
    for (size_t i = 0; i < r->counts.size(); ++i) keys[i] = r->counts[i].first, counts[i] = r->counts[i].second;
    return (int64_t)r->counts.size();
}

void zs_status(zs_server *s, int64_t *out) {
    std::lock_guard<std::mutex> lk(s->mu);
    out[0] = (int64_t)s->pending.size();
    out[1] = s->running;
    out[2] = (int64_t)s->pool->largest_free();
    out[3] = (int64_t)s->pool_bytes;
}

void zs_destroy(zs_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (Request *r : s->pending) {  // never admitted: fail them
            r->status = ZB_ERR_ARG;
            r->error = "the server was destroyed before the request ran";
            r->done = true;
        }
        s->pending.clear();
        s->cv_work.notify_all();
        s->cv_done.notify_all();
    }
    for (auto &t : s->workers) t.join();  // running requests finish first
    if (s->pool_mem) (void)hipFree(s->pool_mem);
    delete s;
}

}  // extern "C"

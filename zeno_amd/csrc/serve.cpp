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
// requests of 1 GiB of state and more run one at a time.  A noisy or
// dynamic request whose shots run as batched trajectories is charged that
// workspace and runs in its slice, on its slot's stream.
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
// A noisy or dynamic request whose shots run as batched trajectories keeps
// its concurrent trajectory states in its pool slice: up to two per CU, but
// no more than this many bytes of them, so that one noisy request does not
// take the pool from the others.
constexpr uint64_t kTrajServeBytes = 2ull << 30;
// Small static requests are served in batches by the batcher thread: all
// of them that have arrived, up to this many, share one persistent launch
// in which each keeps its state on chip.  The batch workspace (op tables,
// running sums, sampled indices) is allocated once, with the server.
constexpr int kDefaultMaxBatch = 1024;
constexpr uint64_t kBatchWorkspaceBytes = 64ull << 20;

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
    uint64_t state_half = 0;  // bytes of one state array (re or im) at the start of the slice
    int skips = 0;
    double t_submit = 0, t_start = 0, t_end = 0;
    int status = ZB_OK;
    std::string error;
    bool done = false;
    std::vector<std::pair<uint64_t, uint64_t>> counts;
    // the state (not a trajectory workspace) decides the one-at-a-time rule
    bool large() const { return 2 * state_half >= kLargeBytes; }
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
    // the batcher: small static requests, one launch per batch
    std::condition_variable cv_batch;
    std::deque<Request *> pending_small;
    std::thread batcher;
    BatchRunner *runner = nullptr;
    int max_batch = kDefaultMaxBatch;  // 0: no batching, every request takes a slot
    int64_t batches = 0, batched = 0;

    double now() const { return std::chrono::duration<double>(Clock::now() - epoch).count(); }

    // Next admissible request (caller holds mu), or nullptr.  A request that
    // does not fit counts one skip each time a later request is admitted
    // past it; after kMaxSkips nothing more is admitted behind it.
    Request *admit() {
        std::vector<Request *> passed;
        for (size_t i = 0; i < pending.size(); ++i) {
            Request *r = pending[i];
            if (r->large() && running_large > 0) continue;  // waits for the running one, not starved
            uint64_t off = 0;
            if (pool->alloc(r->bytes, &off)) {
                r->off = off;
                for (Request *p : passed) ++p->skips;
                pending.erase(pending.begin() + (std::ptrdiff_t)i);
                return r;
            }
            if (r->skips >= kMaxSkips) return nullptr;  // hold the line behind a starved request
            passed.push_back(r);
        }
        return nullptr;
    }

    void worker_loop() {
        zb_backend *h = nullptr;
        // the slot's trajectory launch block: fixed, so the slot's captured
        // trajectory launch replays for every later request of its shape
        TrajWorkspace tw;
        if (hipMalloc(&tw.launch_block, kTrajLaunchBlockBytes) != hipSuccess) {
            (void)hipGetLastError();
            tw.launch_block = nullptr;  // the launch block then lives in each request's slice
        }
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            Request *r = nullptr;
            cv_work.wait(lk, [&] { return stop || (r = admit()) != nullptr; });
            if (!r) break;
            const bool large = r->large();
            ++running;
            running_large += large;
            r->t_start = now();
            lk.unlock();
            run(h, tw, r);
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
        traj_graph_free(tw.graph);
        if (tw.launch_block) (void)hipFree(tw.launch_block);
        if (h) zb_destroy(h);
    }

    // Continuous batching of small static requests: every request that has
    // arrived while the previous batch ran joins the next one (up to
    // max_batch, and while their tables fit the workspace).
    void batcher_loop() {
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            cv_batch.wait(lk, [&] { return stop || !pending_small.empty(); });
            if (pending_small.empty()) break;  // stopping, nothing left
            std::vector<Request *> take;
            uint64_t bytes = 1 << 16;
            while (!pending_small.empty() && (int)take.size() < std::max(1, max_batch)) {
                Request *r = pending_small.front();
                const uint64_t b = batch_item_bytes(r->circuit.c, r->precision, r->shots);
                if (!take.empty() && bytes + b > batch_runner_capacity(runner)) break;
                take.push_back(r);
                pending_small.pop_front();
                bytes += b;
            }
            const double t0 = now();
            for (Request *r : take) r->t_start = t0;
            running += (int)take.size();
            lk.unlock();
            std::vector<BatchItem> items(take.size());
            for (size_t i = 0; i < take.size(); ++i) {
                items[i].c = &take[i]->circuit.c;
                items[i].precision = take[i]->precision;
                items[i].shots = take[i]->shots;
                items[i].seed = take[i]->seed;
            }
            const int rc = batch_run(runner, items);
            const std::string err = rc == ZB_OK ? std::string() : std::string(last_error());
            for (size_t i = 0; i < take.size(); ++i) {
                Request *r = take[i];
                if (rc != ZB_OK) {
                    r->status = rc;
                    r->error = err;
                    continue;
                }
                std::vector<uint64_t> words(items[i].indices.size());
                batch_words(r->circuit.c, items[i].indices, words.data());
                std::unordered_map<uint64_t, uint64_t> raw;
                for (uint64_t w : words) ++raw[w];
                r->counts.assign(raw.begin(), raw.end());
                std::sort(r->counts.begin(), r->counts.end());
            }
            lk.lock();
            running -= (int)take.size();
            ++batches;
            batched += (int64_t)take.size();
            const double t1 = now();
            for (Request *r : take) {
                r->t_end = t1;
                r->done = true;
            }
            cv_done.notify_all();
        }
    }

    void run(zb_backend *&h, TrajWorkspace &tw, Request *r) {
        const uint32_t n = r->circuit.c.n_qubits;
        void *re = pool_mem + r->off, *im = pool_mem + r->off + r->state_half;
        int rc = h ? zb_rebind(h, n, r->precision, re, im) : zb_create_bound(n, r->precision, re, im, &h);
        if (rc == ZB_OK) {
            // batched trajectories take the whole slice (the bound state is
            // not used on that path) and the slot's stream
            tw.mem = pool_mem + r->off;
            tw.bytes = r->bytes;
            tw.stream = zb_stream(h);
            rc = run_compiled_on(h, &r->circuit, r->shots, r->seed, r->noisy ? &r->noise : nullptr, mode, true,
                                 r->counts, &tw);
        }
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
    if (mode == 1) s->runner = batch_runner_new(kBatchWorkspaceBytes);
    if (!s->runner) s->max_batch = 0;  // exact mode, or no workspace: every request takes a slot
    for (int i = 0; i < slots; ++i) s->workers.emplace_back([s] { s->worker_loop(); });
    if (s->runner) s->batcher = std::thread([s] { s->batcher_loop(); });
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
    const uint64_t half = (1ull << n) * (precision == 64 ? 8 : 4);
    uint64_t bytes = 2 * half;
    if (((bytes + kAlign - 1) & ~(kAlign - 1)) > s->pool_bytes) {  // mem::plan (src/mem.rs:114-130)
        set_error("state vector for " + std::to_string(n) + " qubits at f" + std::to_string(precision) + " needs " +
                  std::to_string(bytes) + " B but the serving pool holds " + std::to_string(s->pool_bytes) + " B");
        return ZB_ERR_MEMORY;
    }
    // Requests that will run as batched trajectories are charged their
    // workspace: as many concurrent trajectory states as fit kTrajServeBytes
    // and the pool (at least one), plus the op tables and per-shot words.
    const Compiled &cc = c->c;
    const bool noisy_run = noise && !(noise->depolarizing_1q == 0.0 && noise->depolarizing_2q == 0.0 &&
                                      noise->bit_flip == 0.0 && noise->phase_flip == 0.0 &&
                                      noise->amplitude_damping == 0.0 && noise->readout_flip_0to1 == 0.0 &&
                                      noise->readout_flip_1to0 == 0.0);
    if ((noisy_run || cc.dynamic) && (cc.dynamic || !cc.final_measures.empty()) && s->mode != 0 &&
        trajectories_supported(cc) && zb_device_count() > 0) {
        uint64_t grid = std::min<uint64_t>(traj_max_grid(shots), std::max<uint64_t>(1, kTrajServeBytes / bytes));
        while (grid > 1 && traj_workspace_bytes(cc, precision, shots, grid) > s->pool_bytes) grid /= 2;
        const uint64_t ws = traj_workspace_bytes(cc, precision, shots, grid);
        // a workspace the batched launch would not use (traj_preferred) is
        // not charged: the request then runs the per-shot loop in its state
        if (ws <= s->pool_bytes && traj_preferred(cc, precision, shots, ws)) bytes = std::max(bytes, ws);
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
    r->state_half = half;
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->next_id++;
    r->t_submit = s->now();
    const int64_t id = r->id;
    if (s->max_batch > 0 && !r->noisy && batch_supported(r->circuit.c, precision, shots)) {
        s->pending_small.push_back(r.get());  // state on chip: no pool space
        s->requests[id] = std::move(r);
        s->cv_batch.notify_one();
        return id;
    }
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

int zs_set_batch(zs_server *s, int max_batch) {
    if (max_batch < 0) {
        set_error("zs_set_batch: max_batch must be >= 0");
        return ZB_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (max_batch > 0 && !s->runner) {
        set_error("zs_set_batch: this server does not batch (exact mode)");
        return ZB_ERR_UNSUPPORTED;
    }
    s->max_batch = max_batch;
    return ZB_OK;
}

void zs_counters(zs_server *s, int64_t *out) {
    std::lock_guard<std::mutex> lk(s->mu);
    out[0] = s->batches;
    out[1] = s->batched;
    out[2] = (int64_t)batch_runner_replays(s->runner);
    out[3] = s->max_batch;
}

void zs_status(zs_server *s, int64_t *out) {
    std::lock_guard<std::mutex> lk(s->mu);
    out[0] = (int64_t)(s->pending.size() + s->pending_small.size());
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
        for (Request *r : s->pending_small) {
            r->status = ZB_ERR_ARG;
            r->error = "the server was destroyed before the request ran";
            r->done = true;
        }
        s->pending_small.clear();
        s->cv_work.notify_all();
        s->cv_batch.notify_all();
        s->cv_done.notify_all();
    }
    for (auto &t : s->workers) t.join();  // running requests finish first
    if (s->batcher.joinable()) s->batcher.join();
    batch_runner_free(s->runner);
    if (s->pool_mem) (void)hipFree(s->pool_mem);
    delete s;
}

}  // extern "C"

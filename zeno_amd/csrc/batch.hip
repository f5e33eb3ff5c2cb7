// batch.hip — host side of the small-circuit batch (batch.hpp): packs many
// static circuits into one persistent launch, captured in a hipGraph that
// is replayed whenever a batch of the same precision and state size comes,
// and returns each instance's sampled indices.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "batch.hpp"
#include "host.hpp"
#include "zeno_hip.h"

using namespace zeno;

namespace {

struct BatchError {
    int code;
    std::string msg;
};

void bcheck(hipError_t e, const char *what) {
    if (e != hipSuccess) throw BatchError{ZB_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e)};
}

constexpr uint64_t kAlign = 256;
uint64_t up(uint64_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

uint32_t max_qubits_of(int precision) { return precision == 64 ? 13 : 14; }

uint64_t cu_count() {
    static const uint64_t cus = [] {
        int d = 0, n = 0;
        if (hipGetDevice(&d) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, d) !=
                                                   hipSuccess || n <= 0)
            n = 256;
        return (uint64_t)n;
    }();
    return cus;
}

// Device bytes of one instance besides the shared tables: its running sums
// and its shots' indices.
uint64_t inst_scratch_bytes(uint32_t n, uint64_t shots) { return up((1ull << n) * 8) + up(shots * 8); }

template <typename T>
uint64_t inst_table_bytes(const Compiled &c) {
    uint64_t ops = 0, data = 0;
    for (const Op &o : c.ops) {
        if (o.kind == kBarrier) continue;
        ++ops;
        if (o.kind == kUnitary || o.kind == kDiagonal) data += 2 * o.data.size();
    }
    return ops * sizeof(dev::BatchOp) + data * sizeof(T);
}

}  // namespace

namespace zeno {

bool batch_supported(const Compiled &c, int precision, uint64_t shots) {
    if (precision != 32 && precision != 64) return false;
    if (c.dynamic || c.n_qubits < 1 || c.n_qubits > max_qubits_of(precision)) return false;
    if (shots > kBatchMaxShots) return false;
    for (const Op &o : c.ops) {
        const size_t dense = precision == 64 ? dev::kBatchMaxDense<double> : dev::kBatchMaxDense<float>;
        if (o.kind == kUnitary && o.qs.size() > dense) return false;
        if (o.kind == kMeasure || o.kind == kReset || o.is_if) return false;
    }
    return true;
}

uint64_t batch_item_bytes(const Compiled &c, int precision, uint64_t shots) {
    const uint64_t t = precision == 64 ? inst_table_bytes<double>(c) : inst_table_bytes<float>(c);
    return t + sizeof(dev::BatchInst) + inst_scratch_bytes(c.n_qubits, shots) + 3 * kAlign;
}

struct BatchRunner {
    hipStream_t stream = nullptr;
    char *dev = nullptr, *host = nullptr;  // [inst | ops | data | sums | indices] and its pinned staging
    uint64_t cap = 0;
    void *block = nullptr;  // the fixed launch block
    struct Graph {
        hipGraph_t g = nullptr;
        hipGraphExec_t e = nullptr;
    };
    std::map<std::tuple<int, uint64_t, uint64_t>, Graph> graphs;  // (precision, LDS bytes, grid)
    uint64_t launches = 0, replays = 0;
};

BatchRunner *batch_runner_new(uint64_t bytes) {
    auto *r = new BatchRunner();
    try {
        bcheck(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking), "hipStreamCreate");
        bcheck(hipMalloc((void **)&r->dev, bytes), "hipMalloc batch workspace");
        bcheck(hipHostMalloc((void **)&r->host, bytes, hipHostMallocDefault), "hipHostMalloc batch staging");
        bcheck(hipMalloc(&r->block, kAlign), "hipMalloc launch block");
        r->cap = bytes;
    } catch (const BatchError &e) {
        set_error(e.msg);
        batch_runner_free(r);
        return nullptr;
    }
    return r;
}

void batch_runner_free(BatchRunner *r) {
    if (!r) return;
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    for (auto &kv : r->graphs) {
        if (kv.second.e) (void)hipGraphExecDestroy(kv.second.e);
        if (kv.second.g) (void)hipGraphDestroy(kv.second.g);
    }
    if (r->block) (void)hipFree(r->block);
    if (r->dev) (void)hipFree(r->dev);
    if (r->host) (void)hipHostFree(r->host);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
}

uint64_t batch_runner_capacity(const BatchRunner *r) { return r ? r->cap : 0; }
uint64_t batch_runner_replays(const BatchRunner *r) { return r ? r->replays : 0; }

}  // namespace zeno

namespace {

template <typename T>
void run_T(BatchRunner &R, const std::vector<BatchItem *> &items) {
    using dev::BatchInst;
    using dev::BatchOp;
    const int precision = sizeof(T) == 8 ? 64 : 32;
    const size_t m = items.size();
    // largest first, so the long instances start while the short ones fill in
    std::vector<size_t> order(m);
    std::iota(order.begin(), order.end(), 0);
    auto cost = [&](size_t i) {
        const Compiled &c = *items[i]->c;
        return (double)(c.ops.size() + 1) * (double)(1ull << c.n_qubits) + (double)items[i]->shots * c.n_qubits;
    };
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost(a) > cost(b); });
    // layout: [inst][ops][data][sums][indices]
    uint64_t nops = 0, ndata = 0, nsum = 0, nout = 0;
    uint32_t maxn = 1;
    for (BatchItem *it : items) {
        for (const Op &o : it->c->ops) {
            if (o.kind == kBarrier) continue;
            ++nops;
            if (o.kind == kUnitary || o.kind == kDiagonal) ndata += 2 * o.data.size();
        }
        nsum += 1ull << it->c->n_qubits;
        nout += it->shots;
        maxn = std::max(maxn, it->c->n_qubits);
    }
    const uint64_t o_inst = 0, o_ops = up(o_inst + m * sizeof(BatchInst)), o_data = up(o_ops + nops * sizeof(BatchOp)),
                   o_sum = up(o_data + std::max<uint64_t>(ndata, 1) * sizeof(T)), o_out = up(o_sum + nsum * 8),
                   end = up(o_out + std::max<uint64_t>(nout, 1) * 8);
    if (end > R.cap)
        throw BatchError{ZB_ERR_MEMORY, "a batch of " + std::to_string(m) + " circuits needs " + std::to_string(end) +
                                            " B of workspace, more than its " + std::to_string(R.cap) + " B"};
    auto *hinst = reinterpret_cast<BatchInst *>(R.host + o_inst);
    auto *hops = reinterpret_cast<BatchOp *>(R.host + o_ops);
    auto *hdata = reinterpret_cast<T *>(R.host + o_data);
    uint64_t po = 0, pd = 0, ps = 0, pq = 0;
    std::vector<uint64_t> out_at(m);
    for (size_t slot = 0; slot < m; ++slot) {
        const size_t i = order[slot];
        const BatchItem &it = *items[i];
        const Compiled &c = *it.c;
        BatchInst in{};
        in.ops_off = po;
        in.data_off = pd;
        in.x_off = ps;
        in.out_off = pq;
        in.seed = splitmix64(it.seed);  // run_sampled hands splitmix64(seed) to sample (exec.rs:436)
        in.n = c.n_qubits;
        in.shots = (uint32_t)(c.final_measures.empty() ? 0 : it.shots);
        uint32_t k = 0;
        for (const Op &o : c.ops) {
            if (o.kind == kBarrier) continue;
            BatchOp b{};
            b.data_off = (uint32_t)(pd - in.data_off);
            if (o.kind == kUnitary || o.kind == kDiagonal) {
                b.kind = o.kind == kDiagonal ? 2 : (o.qs.size() == 1 ? 0 : 1);
                b.k = (uint8_t)o.qs.size();
                for (size_t j = 0; j < o.qs.size(); ++j) b.q[j] = (uint8_t)o.qs[j];
                for (const cplx &z : o.data) hdata[pd++] = (T)z.re;  // T::fr
                for (const cplx &z : o.data) hdata[pd++] = (T)z.im;
            } else {  // cx
                b.kind = 3;
                b.k = 2;
                b.q[0] = (uint8_t)o.a;
                b.q[1] = (uint8_t)o.b;
            }
            hops[po++] = b;
            ++k;
        }
        in.nops = k;
        hinst[slot] = in;
        out_at[i] = pq;
        ps += 1ull << c.n_qubits;
        pq += it.shots;
    }
    dev::BatchLaunch<T> L{};
    L.inst = reinterpret_cast<const BatchInst *>(R.dev + o_inst);
    L.ops = reinterpret_cast<const BatchOp *>(R.dev + o_ops);
    L.data = reinterpret_cast<const T *>(R.dev + o_data);
    L.xs = reinterpret_cast<double *>(R.dev + o_sum);
    L.out = reinterpret_cast<uint64_t *>(R.dev + o_out);
    L.count = (uint32_t)m;
    L.next = 0;
    // The grid depends on the state size only (workgroups that find no
    // instance left exit at once), so batches of any count with the same
    // largest state replay one captured launch.
    const uint64_t lds = 2 * (1ull << maxn) * sizeof(T);
    const uint64_t per_cu = std::max<uint64_t>(1, std::min<uint64_t>(8, (160ull << 10) / (lds + 256)));
    const uint64_t grid = per_cu * cu_count();
    static const bool lds_ok = [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&dev::k_batch<double>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 << 10);
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&dev::k_batch<float>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 << 10);
        (void)hipGetLastError();
        return true;
    }();
    (void)lds_ok;
    auto key = std::make_tuple(precision, lds, grid);
    auto it = R.graphs.find(key);
    if (it == R.graphs.end()) {
        BatchRunner::Graph g;
        bcheck(hipStreamBeginCapture(R.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        hipLaunchKernelGGL(dev::k_batch<T>, dim3((unsigned)grid), dim3(dev::kBatchBlock), (unsigned)lds, R.stream,
                           reinterpret_cast<dev::BatchLaunch<T> *>(R.block));
        const hipError_t le = hipGetLastError();
        const hipError_t ce = hipStreamEndCapture(R.stream, &g.g);
        bcheck(le, "k_batch");
        bcheck(ce, "hipStreamEndCapture");
        const hipError_t ie = hipGraphInstantiate(&g.e, g.g, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            (void)hipGraphDestroy(g.g);
            bcheck(ie, "hipGraphInstantiate");
        }
        it = R.graphs.emplace(key, g).first;
    } else {
        ++R.replays;
    }
    // the tables, then the launch block, on the batch stream (pinned: no
    // staging copy); the graph launch follows them in stream order
    bcheck(hipMemcpyAsync(R.dev, R.host, o_sum, hipMemcpyHostToDevice, R.stream), "batch tables");
    bcheck(hipMemcpyAsync(R.block, &L, sizeof(L), hipMemcpyHostToDevice, R.stream), "batch launch block");
    ++R.launches;
    bcheck(hipGraphLaunch(it->second.e, R.stream), "hipGraphLaunch");
    if (nout) bcheck(hipMemcpyAsync(R.host + o_out, R.dev + o_out, nout * 8, hipMemcpyDeviceToHost, R.stream), "indices");
    bcheck(hipStreamSynchronize(R.stream), "hipStreamSynchronize");
    const uint64_t *hout = reinterpret_cast<const uint64_t *>(R.host + o_out);
    for (size_t i = 0; i < m; ++i) {
        BatchItem &b = *items[i];
        b.indices.assign(hout + out_at[i], hout + out_at[i] + (b.c->final_measures.empty() ? 0 : b.shots));
    }
}

}  // namespace

namespace zeno {

int batch_run(BatchRunner *R, std::vector<BatchItem> &items) {
    try {
        for (const BatchItem &it : items)
            if (!batch_supported(*it.c, it.precision, it.shots))
                throw BatchError{ZB_ERR_UNSUPPORTED, "circuit not batchable (static, n <= 13 in f64 / 14 in f32, "
                                                     "dense ops on <= 3 / 4 qubits, <= 2^17 shots)"};
        for (int prec : {64, 32}) {
            std::vector<BatchItem *> part;
            for (BatchItem &it : items)
                if (it.precision == prec) part.push_back(&it);
            if (part.empty()) continue;
            if (prec == 64) run_T<double>(*R, part);
            else run_T<float>(*R, part);
        }
    } catch (const BatchError &e) {
        set_error(e.msg);
        return e.code;
    }
    return ZB_OK;
}

// Sampled indices -> clbit words through the final measures (run_sampled,
// src/exec.rs:437-447).
void batch_words(const Compiled &c, const std::vector<uint64_t> &idx, uint64_t *words) {
    for (size_t s = 0; s < idx.size(); ++s) {
        uint64_t cl = 0;
        for (const auto &fm : c.final_measures) {
            const uint64_t bit = (idx[s] >> fm.first) & 1;
            cl = (cl & ~(1ull << fm.second)) | (bit << fm.second);
        }
        words[s] = cl;
    }
}

}  // namespace zeno

extern "C" {

int zr_batch_supported(const zr_compiled *c, int precision, uint64_t shots) {
    return batch_supported(c->c, precision, shots) ? 1 : 0;
}

int zr_run_batch(const zr_compiled *const *circuits, const int32_t *precisions, const uint64_t *shots,
                 const uint64_t *seeds, int64_t count, uint64_t *words, double *seconds) {
    if (count < 0) {
        set_error("zr_run_batch: negative count");
        return ZB_ERR_ARG;
    }
    if (zb_device_count() == 0) {
        set_error("no HIP device is available (the hip backend needs an AMD GPU)");
        return ZB_ERR_UNSUPPORTED;
    }
    std::vector<BatchItem> items((size_t)count);
    uint64_t bytes = 1 << 20;
    for (int64_t i = 0; i < count; ++i) {
        items[i].c = &circuits[i]->c;
        items[i].precision = precisions[i];
        items[i].shots = shots[i];
        items[i].seed = seeds[i];
        if (!batch_supported(circuits[i]->c, precisions[i], shots[i])) {
            set_error("zr_run_batch: circuit " + std::to_string(i) + " is not batchable");
            return ZB_ERR_UNSUPPORTED;
        }
        bytes += batch_item_bytes(circuits[i]->c, precisions[i], shots[i]);
    }
    BatchRunner *R = batch_runner_new(bytes);
    if (!R) return ZB_ERR_MEMORY;
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = batch_run(R, items);
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    batch_runner_free(R);
    if (rc != ZB_OK) return rc;
    uint64_t at = 0;
    for (int64_t i = 0; i < count; ++i) {
        batch_words(items[i].c[0], items[i].indices, words + at);
        at += shots[i];
    }
    return ZB_OK;
}

}  // extern "C"

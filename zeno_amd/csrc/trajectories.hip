// trajectories.hip — host side of the batched trajectory kernel: packs the
// compiled ops, runs every shot in one persistent launch captured in a
// hipGraph, and returns each shot's clbit word.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "host.hpp"
#include "trajectories.hpp"
#include "zeno_hip.h"

using namespace zeno;

namespace {

struct TrajError {
    int code;
    std::string msg;
};

void hcheck(hipError_t e, const char *what) {
    if (e != hipSuccess) throw TrajError{ZB_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e)};
}

// Device buffers freed on every exit path.
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    template <typename U>
    U *get() {
        return static_cast<U *>(p);
    }
};

bool supported(const Compiled &c) {
    if (c.n_qubits > 20) return false;
    for (const Op &o : c.ops) {
        if ((o.kind == kUnitary || o.kind == kDiagonal) && o.qs.size() > (size_t)dev::kTrajMaxK) return false;
    }
    return true;
}

template <typename T>
void run_T(const Compiled &c, uint64_t shots, uint64_t seed, const zr_noise *noise, uint64_t *out, double *secs) {
    std::vector<dev::TrajOp> ops;
    std::vector<T> mats;
    for (const Op &o : c.ops) {
        if (o.kind == kBarrier) continue;
        dev::TrajOp t{};
        t.is_if = o.is_if;
        t.creg_off = o.creg_off;
        t.creg_len = o.creg_len;
        t.value = o.value;
        if (o.kind == kUnitary || o.kind == kDiagonal) {
            t.kind = o.kind == kUnitary ? 0 : 1;
            t.k = (uint32_t)o.qs.size();
            for (size_t b = 0; b < o.qs.size(); ++b) t.q[b] = o.qs[b];
            t.data_off = (uint32_t)mats.size();
            for (const cplx &z : o.data) mats.push_back((T)z.re);  // T::fr
            for (const cplx &z : o.data) mats.push_back((T)z.im);
        } else if (o.kind == kCx) {
            t.kind = 2;
            t.q[0] = o.a, t.q[1] = o.b;
        } else if (o.kind == kMeasure) {
            t.kind = 3;
            t.q[0] = o.a;
            t.clbit = o.b;
        } else if (o.kind == kReset) {
            t.kind = 4;
            t.q[0] = o.a;
        }
        ops.push_back(t);
    }
    std::vector<uint32_t> fm;
    for (const auto &f : c.final_measures) fm.push_back(f.first), fm.push_back(f.second);
    const size_t nops = ops.size(), nfm = c.final_measures.size();
    // keep every upload non-empty; the kernel reads only nops / nfm entries
    if (mats.empty()) mats.push_back(T(0));
    if (ops.empty()) ops.push_back(dev::TrajOp{});
    if (fm.empty()) fm.push_back(0);

    int dev_id = 0, cus = 0;
    hcheck(hipGetDevice(&dev_id), "hipGetDevice");
    hcheck(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_id), "cu count");
    const uint64_t len = 1ull << c.n_qubits;
    const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>(shots, (uint64_t)cus * 2));
    size_t free_b = 0, total_b = 0;
    hcheck(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const size_t ws_bytes = (size_t)grid * 2 * len * sizeof(T);
    if (ws_bytes + shots * 8 + (64ull << 20) > free_b)
        throw TrajError{ZB_ERR_MEMORY, "trajectory workspace of " + std::to_string(ws_bytes) + " B does not fit"};
    DevBuf ws, dops, dmats, dfm, dout;
    hcheck(hipMalloc(&ws.p, ws_bytes), "hipMalloc workspace");
    hcheck(hipMalloc(&dops.p, ops.size() * sizeof(dev::TrajOp)), "hipMalloc ops");
    hcheck(hipMalloc(&dmats.p, mats.size() * sizeof(T)), "hipMalloc mats");
    hcheck(hipMalloc(&dfm.p, fm.size() * 4), "hipMalloc fm");
    hcheck(hipMalloc(&dout.p, std::max<uint64_t>(shots, 1) * 8), "hipMalloc out");
    dev::TrajNoise nz{};
    int noisy = 0;
    if (noise) {
        nz = dev::TrajNoise{noise->depolarizing_1q, noise->depolarizing_2q, noise->bit_flip, noise->phase_flip,
                            noise->amplitude_damping, noise->readout_flip_0to1, noise->readout_flip_1to0};
        noisy = 1;
    }
    hipStream_t stream;
    hcheck(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    try {
        // Uploads on the launch stream, waited for: a plain hipMemcpy from
        // pageable memory may return before its DMA lands, and a
        // non-blocking stream would not wait for it.
        hcheck(hipMemcpyAsync(dops.p, ops.data(), ops.size() * sizeof(dev::TrajOp), hipMemcpyHostToDevice, stream),
               "ops");
        hcheck(hipMemcpyAsync(dmats.p, mats.data(), mats.size() * sizeof(T), hipMemcpyHostToDevice, stream), "mats");
        hcheck(hipMemcpyAsync(dfm.p, fm.data(), fm.size() * 4, hipMemcpyHostToDevice, stream), "fm");
        hcheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
        // One persistent launch, captured in a graph and replayed once.
        hcheck(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        hipLaunchKernelGGL(dev::k_trajectories<T>, dim3((unsigned)grid), dim3(dev::kBlock), 0, stream, ws.get<T>(),
                           c.n_qubits, dops.get<dev::TrajOp>(), (uint32_t)nops, dfm.get<uint32_t>(), (uint32_t)nfm,
                           dmats.get<T>(), nz, noisy, seed, shots, dout.get<uint64_t>());
        hcheck(hipGetLastError(), "k_trajectories");
        hcheck(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
        hcheck(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        hcheck(hipGraphLaunch(exec, stream), "hipGraphLaunch");
        hcheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
    } catch (...) {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipStreamDestroy(stream);
        throw;
    }
    if (secs) *secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    (void)hipStreamDestroy(stream);
    if (shots) hcheck(hipMemcpy(out, dout.p, shots * 8, hipMemcpyDeviceToHost), "out copy");
}

}  // namespace

extern "C" int zr_trajectories_supported(const zr_compiled *c) { return supported(c->c) ? 1 : 0; }

extern "C" int zr_run_trajectories(const zr_compiled *c, int precision, uint64_t shots, uint64_t seed,
                                   const zr_noise *noise, uint64_t *out, double *kernel_seconds) {
    if (!supported(c->c)) {
        set_error("batched trajectories need <= 20 qubits and unitary/diagonal ops on <= 3 qubits");
        return ZB_ERR_UNSUPPORTED;
    }
    if (precision != 32 && precision != 64) {
        set_error("precision must be 32 or 64");
        return ZB_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("no HIP device is available (the hip backend needs an AMD GPU)");
        return ZB_ERR_UNSUPPORTED;
    }
    try {
        if (precision == 64) run_T<double>(c->c, shots, seed, noise, out, kernel_seconds);
        else run_T<float>(c->c, shots, seed, noise, out, kernel_seconds);
    } catch (const TrajError &e) {
        set_error(e.msg);
        return e.code;
    }
    return ZB_OK;
}

// backend.hip — the HIP implementation of the reference's Backend trait
// (src/exec.rs:134-162) behind the C ABI of include/zeno_hip.h.
//
// One handle owns one split-complex state in HBM and one HIP stream.  Ops
// are enqueued asynchronously; reads synchronise.  Matrices and diagonal
// tables travel through a pinned-host -> device parameter arena that is
// only reused after a stream synchronisation, so the caller's buffers are
// borrowed for the call only (as src/metal.rs:365-367 requires).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <tuple>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "host.hpp"
#include "kernels.hpp"
#include "zeno_hip.h"

namespace zeno {
// sort.hip: the sampler's draws sorted and routed to chunks on the device
size_t draw_sort_temp_bytes(uint64_t n);
size_t draw_jump_bytes(uint64_t shots);
hipError_t draws_on_device(const uint64_t *state, uint64_t shots, double total, void *jump_dev,
                           std::vector<uint64_t> &host, double *us, hipStream_t stream);
hipError_t draw_sort_and_route(void *temp, size_t temp_bytes, const double *us, double *sorted, uint64_t n,
                               const double *cum, uint64_t m, uint64_t *starts, hipStream_t stream);
}  // namespace zeno

using namespace zeno;

namespace {

constexpr size_t kArenaBytes = 64ull << 20;
// Sampler pass 1 walks 8 chunks per workgroup (k_chunk_sums_multi) from
// this many 2^16-value chunks up (28 qubits: 2.13 -> 1.66 ms, 30 qubits
// f64: 7.67 -> 4.14 ms); below, one chunk per workgroup is faster.
constexpr uint64_t kChunkMultiDefault = 4096;

// Scratch of the allocation-free sampler (zb_sample_noalloc): one copy of
// this array per GPU, held until the process ends.  Host code never uses its
// name as an address: each device's copy is looked up once, with that device
// current, when the first handle on it is set up, and kept in g_noalloc_ptr.
// Every output slot the sampler reads back is written by the same call, so
// nothing in it has to start at zero.
constexpr size_t kNoallocScratchBytes = 64ull << 20;
static __device__ __align__(256) char scratch[kNoallocScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kMaxDevices] = {};

struct HipError {
    std::string msg;
};

inline void check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw HipError{std::string(what) + ": " + hipGetErrorString(e)};
}

unsigned grid_for(uint64_t work) {
    const uint64_t blocks = (work + dev::kBlock - 1) / dev::kBlock;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(blocks, 65536));
}

}  // namespace

struct zb_backend {
    int device = 0;  // the HIP device current when the handle was created
    uint32_t n = 0;
    int prec = 64;
    uint64_t len = 0;
    size_t elem = 8;
    void *re = nullptr, *im = nullptr;
    bool owns_state = true;  // false when bound to caller-owned device memory
    // qubit layout: circuit qubit q is bit phys[q] of the state index
    uint8_t phys[64];
    bool ident = true;
    int layout_kind = -1;  // -1: by size and precision; 0 identity; 1 spread; >= 2 seeded shuffle
    uint64_t *lut_dev = nullptr;
    zb_backend() {
        for (int i = 0; i < 64; ++i) phys[i] = (uint8_t)i;
    }
    dev::Layout layout() const { return dev::Layout{ident ? nullptr : lut_dev}; }
    hipStream_t stream = nullptr;
    int mode = 1;
    bool prefetch = true;  // tiled kernel variant: register prefetch of the next tile
    int slots = 0;          // tiled stage width: 3 (512-lane workgroups) or 4 (256 lanes); 0 = 3
    uint64_t chunk_multi = kChunkMultiDefault;  // sampler pass 1: fewest chunks for the multi-chain walk
    unsigned win_grid = 16384;  // workgroups per window launch (at most one per tile)
    uint32_t page_bits = 6;  // most tile qubits at or above the page bit in one window
    uint32_t tile_bits = 0;  // tile width of the tiled kernel: 12 or 13 (0 = 12)
    std::string name;
    // parameter arena
    char *arena_dev = nullptr, *arena_host = nullptr;
    size_t arena_off = 0;
    // reduction scratch (device) + pinned result slot
    double *red_dev = nullptr;
    size_t red_cap = 0;
    double *red_host = nullptr;
    // grow-only scratch for the sampler
    char *scratch = nullptr;
    size_t scratch_cap = 0;
    // pending tiled window (mode 1): ops with global qubits, data in T
    struct POp {
        uint32_t kind, k;
        uint32_t q[dev::kMaxDiagK];
        uint32_t data_off;
    };
    uint64_t win_mask = 0;
    std::vector<POp> win_ops;
    // CX ops held back (physical control, target; pairwise disjoint): a CX is a
    // pure permutation, so it commutes bit-exactly with every op on other
    // qubits and can wait for a window that already holds its target
    std::vector<std::pair<uint32_t, uint32_t>> deferred;  // circuit qubits (control, target)
    bool defer_cx = true;
    // Relabelling: when the caller has announced the coming ops (their
    // tile needs as masks of circuit qubits), a window that closes because
    // an op does not fit also permutes its tile's qubits on write-back, so
    // that the low state bits every window holds go to the tile's qubits
    // needed soonest.  The layout LUT then lags behind phys[] until the
    // next reduction or copy refreshes it.
    bool relabel = true;
    const uint64_t *la_need = nullptr;
    size_t la_count = 0, la_pos = 0;
    std::vector<std::pair<uint32_t, uint32_t>> wb_swaps;  // state-bit swaps of the window being flushed
    bool lut_dirty = false;
    bool mfma = false;  // dense 4..6-qubit ops on the matrix cores (within tolerance)
    // dense 2..6-qubit ops through LDS tiles (k_kq's results): 0 never,
    // 1 where the per-lane gather is slow, 2 always
    int kq_tiled = 1;
    bool fma = false;   // fused multiply-adds in the sweep arithmetic (within tolerance)
    std::vector<unsigned char> win_data;    // 1q / dense matrices (copied into LDS)
    std::vector<unsigned char> win_tables;  // diagonal tables (read from global)
    // launch profiling: HIP events around every state-sweep launch
    struct Rec {
        int kernel;
        double bytes;
        hipEvent_t a, b;
        int32_t detail[4];  // windows: ops, stages, diagonal ops, CX ops (else zero)
    };
    bool prof = false;
    std::vector<Rec> recs, last_recs;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_next = 0;
    hipEvent_t next_event() {
        if (ev_next == ev_pool.size()) {
            hipEvent_t e;
            check(hipEventCreate(&e), "hipEventCreate");
            ev_pool.push_back(e);
        }
        return ev_pool[ev_next++];
    }
    // Brackets one launch with events when profiling (kernel ids in zeno_hip.h).
    int32_t next_detail[4] = {0, 0, 0, 0};  // detail of the next timed launch
    template <class F>
    void timed(int kernel, double bytes, F &&launch) {
        if (!prof) return launch();
        Rec r{kernel, bytes, next_event(), next_event(), {next_detail[0], next_detail[1], next_detail[2], next_detail[3]}};
        for (int32_t &d : next_detail) d = 0;
        check(hipEventRecord(r.a, stream), "hipEventRecord");
        launch();
        check(hipEventRecord(r.b, stream), "hipEventRecord");
        recs.push_back(r);
    }
    double sweep_bytes() const { return 4.0 * (double)len * (double)elem; }

    void sync() { check(hipStreamSynchronize(stream), "hipStreamSynchronize"); }

    // Copy `bytes` of host data into the arena; returns the device address.
    void *upload(const void *src, size_t bytes) {
        const size_t need = (bytes + 255) & ~size_t(255);
        if (arena_off + need > kArenaBytes) {
            sync();
            arena_off = 0;
        }
        std::memcpy(arena_host + arena_off, src, bytes);
        void *d = arena_dev + arena_off;
        check(hipMemcpyAsync(d, arena_host + arena_off, bytes, hipMemcpyHostToDevice, stream), "param upload");
        arena_off += need;
        return d;
    }

    void ensure_red(size_t n_doubles) {
        if (n_doubles <= red_cap) return;
        sync();
        if (red_dev) check(hipFree(red_dev), "hipFree");
        red_cap = std::max<size_t>(n_doubles, 1 << 16);
        check(hipMalloc(&red_dev, red_cap * sizeof(double)), "hipMalloc reduce");
    }

    char *scratch_get(size_t bytes) {
        if (bytes > scratch_cap) {
            sync();
            if (scratch) check(hipFree(scratch), "hipFree");
            scratch = nullptr;
            scratch_cap = std::max<size_t>(bytes, 1 << 20);
            check(hipMalloc((void **)&scratch, scratch_cap), "hipMalloc scratch");
        }
        return scratch;
    }

    template <typename T>
    T *R() {
        return static_cast<T *>(re);
    }
    template <typename T>
    T *I() {
        return static_cast<T *>(im);
    }
};

namespace {

void flush(zb_backend *h);
void flush_window(zb_backend *h);

template <typename T>
void launch_zero(zb_backend *h) {
    h->timed(ZB_K_ZERO, h->sweep_bytes() / 2, [&] {
        hipLaunchKernelGGL(dev::k_zero<T>, dim3(grid_for(h->len)), dim3(dev::kBlock), 0, h->stream, h->R<T>(),
                           h->I<T>(), h->len);
    });
    check(hipGetLastError(), "k_zero");
}

// Page-selecting bit: bits at or above it pick a different 2 MiB page of a
// state array (2^18 doubles, 2^19 floats).
uint32_t page_bit(const zb_backend *h) { return h->prec == 64 ? 18 : 19; }

// Qubit layout.  Windows are runs of consecutive ops, and circuits touch
// qubits in runs of consecutive indices, so with the identity layout a
// window on the top qubits gathers a tile from 2^8 pages per array and the
// per-CU translation cache (about 16 entries) misses on almost every
// access.  The spread layout keeps the lowest `low` qubits in place (they
// carry the coalesced 128 B rows) and deals the rest alternately onto the
// bits below the page bit and the page-selecting bits, in proportion and
// each in ascending order, so any run of qubits holds few page bits.
// Results do not depend on the layout: every op acts on the same
// amplitudes with the same arithmetic, and every reduction or copy walks
// logical indices.  Set only while the state is |0...0> (index 0 is fixed).
void upload_layout(zb_backend *h);

void set_layout(zb_backend *h, int kind) {
    const uint32_t n = h->n, low = std::min<uint32_t>(h->prec == 64 ? 4 : 5, n), pb = page_bit(h);
    if (kind < 0) kind = n >= pb + 4 ? 1 : 0;
    for (int i = 0; i < 64; ++i) h->phys[i] = (uint8_t)i;
    if (kind == 1 && n > pb) {
        const uint32_t m = n - low, np = n - pb;  // qubits dealt, page bits among them
        uint32_t next_n = low, next_p = pb, taken_p = 0;
        for (uint32_t i = 0; i < m; ++i) {
            const uint32_t want_p = (uint32_t)(((uint64_t)(i + 1) * np) / m);
            if (taken_p < want_p) h->phys[low + i] = (uint8_t)next_p++, ++taken_p;
            else h->phys[low + i] = (uint8_t)next_n++;
        }
    } else if (kind >= 2 && n > low) {  // seeded shuffle of the bits above `low` (tests)
        Rng rng((uint64_t)kind);
        for (uint32_t i = n - 1; i > low; --i) {
            const uint32_t j = low + (uint32_t)rng.range(0, i - low + 1);
            std::swap(h->phys[i], h->phys[j]);
        }
    }
    h->layout_kind = kind;
    upload_layout(h);
}

// Uploads the LUT of phys[] (index -> state index); returns when it landed.
void upload_layout(zb_backend *h) {
    const uint32_t n = h->n;
    h->lut_dirty = false;
    h->ident = true;
    for (uint32_t q = 0; q < n; ++q) h->ident = h->ident && h->phys[q] == q;
    if (h->ident) return;
    std::vector<uint64_t> lut(5 * 256, 0);
    for (int c = 0; c < 5; ++c)
        for (int v = 0; v < 256; ++v) {
            uint64_t o = 0;
            for (int b = 0; b < 8; ++b) {
                const uint32_t q = (uint32_t)(8 * c + b);
                if (((v >> b) & 1) && q < n) o |= 1ull << h->phys[q];
            }
            lut[(size_t)c * 256 + v] = o;
        }
    if (!h->lut_dev) check(hipMalloc((void **)&h->lut_dev, lut.size() * 8), "hipMalloc layout");
    // On the backend's own stream and waited for: a plain hipMemcpy from
    // pageable memory may return before its DMA lands, and the backend's
    // non-blocking stream would not wait for it.
    check(hipMemcpyAsync(h->lut_dev, lut.data(), lut.size() * 8, hipMemcpyHostToDevice, h->stream), "layout upload");
    h->sync();
}

template <typename T>
void to_T(const double *src, size_t n, std::vector<T> &re, std::vector<T> &im) {
    re.resize(n);
    im.resize(n);
    for (size_t i = 0; i < n; ++i) {
        re[i] = (T)src[2 * i];  // T::fr: round to nearest (src/state.rs:41-50)
        im[i] = (T)src[2 * i + 1];
    }
}

// State bits of an op's qubits in the op's own order, plus sorted.
dev::Qubits make_qubits(const uint32_t *qs, uint32_t k) {
    dev::Qubits q{};
    q.k = (int)k;
    for (uint32_t b = 0; b < k; ++b) q.q[b] = q.s[b] = qs[b];
    std::sort(q.s, q.s + k);
    return q;
}

template <typename T, int K>
void launch_kq(zb_backend *h, const T *dmat, const dev::Qubits &qs) {
    const uint64_t groups = h->len >> K;
    h->timed(ZB_K_KQ + K, h->sweep_bytes(), [&] {
        if (h->fma)
            hipLaunchKernelGGL((dev::k_kq<T, K, true>), dim3(grid_for(groups)), dim3(dev::kBlock), 0, h->stream,
                               h->R<T>(), h->I<T>(), dmat, qs, groups);
        else
            hipLaunchKernelGGL((dev::k_kq<T, K, false>), dim3(grid_for(groups)), dim3(dev::kBlock), 0, h->stream,
                               h->R<T>(), h->I<T>(), dmat, qs, groups);
    });
    check(hipGetLastError(), "k_kq");
}

template <typename T, int K>
void launch_kq_mfma(zb_backend *h, const std::vector<T> &mr, const std::vector<T> &mi, const uint32_t *pq) {
    constexpr int D = 1 << K, D2 = 2 * D;
    std::vector<T> w((size_t)D2 * D2);
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) {
            const T a = mr[r * D + c], b = mi[r * D + c];
            w[r * D2 + c] = a;
            w[r * D2 + D + c] = -b;
            w[(D + r) * D2 + c] = b;
            w[(D + r) * D2 + D + c] = a;
        }
    const T *dw = static_cast<const T *>(h->upload(w.data(), w.size() * sizeof(T)));
    const dev::KqTile t = kq_tile(h->prec == 64 ? 4 : 5, pq, K);
    const uint64_t tiles = h->len >> dev::kKqTileBits;
    const unsigned grid = (unsigned)std::min<uint64_t>(tiles, 1024);
    h->timed(ZB_K_KQ_MFMA + K, h->sweep_bytes(), [&] {
        hipLaunchKernelGGL((dev::k_kq_mfma<T, K>), dim3(grid), dim3(dev::kKqThreads), 0, h->stream, h->R<T>(),
                           h->I<T>(), dw, t, tiles);
    });
    check(hipGetLastError(), "k_kq_mfma");
}

// The exact dense kernel on LDS tiles: the matrix packed per wave and
// column, [wave][column][row of the wave's 2^K/8][re, im].
template <typename T, int K>
void launch_kq_tiled(zb_backend *h, const std::vector<T> &mr, const std::vector<T> &mi, const uint32_t *pq) {
    const std::vector<T> pk = kq_tiled_pack<T>(K, mr.data(), mi.data());
    const T *dp = static_cast<const T *>(h->upload(pk.data(), pk.size() * sizeof(T)));
    dev::KqTile t = kq_tile(h->prec == 64 ? 4 : 5, pq, K);
    kq_swizzle(t);
    const uint64_t tiles = h->len >> dev::kKqTileBits;
    const unsigned grid = (unsigned)std::min<uint64_t>(tiles, 1024);
    h->timed(ZB_K_KQ + K, h->sweep_bytes(), [&] {
        if (h->fma)
            hipLaunchKernelGGL((dev::k_kq_tiled<T, K, true>), dim3(grid), dim3(dev::kKqThreads), 0, h->stream,
                               h->R<T>(), h->I<T>(), dp, t, tiles);
        else
            hipLaunchKernelGGL((dev::k_kq_tiled<T, K, false>), dim3(grid), dim3(dev::kKqThreads), 0, h->stream,
                               h->R<T>(), h->I<T>(), dp, t, tiles);
    });
    check(hipGetLastError(), "k_kq_tiled");
}

// Whether a dense op on state bits pq runs on LDS tiles.  The per-lane
// gather of k_kq is up to 10% faster while at most two op bits lie inside
// a 128-byte run (its loads stay close to coalesced); with three it is
// 1.4x (5 qubits) to 2.6x (3 qubits) slower, and with five 3.6x; a
// 2-qubit op is tiled when both its bits lie in the run (1.7x).  In f64 a
// 6-qubit group (256 registers of amplitudes) does not fit one lane well,
// and tiles are 3.6x faster.
bool use_kq_tiled(const zb_backend *h, const uint32_t *pq, uint32_t k) {
    if (h->kq_tiled == 0 || k < 2 || h->n < (uint32_t)dev::kKqTileBits) return false;
    if (h->kq_tiled == 2) return true;
    const uint32_t run = h->prec == 64 ? 4 : 5;
    uint32_t low = 0;
    for (uint32_t b = 0; b < k; ++b) low += pq[b] < run;
    return low >= std::min(k, 3u) || (h->prec == 64 && k == 6);
}

template <typename T>
void exact_unitary_T(zb_backend *h, const uint32_t *qs, uint32_t k, const double *mat) {
    const size_t dim = (size_t)1 << k;
    std::vector<T> mr, mi;
    to_T<T>(mat, dim * dim, mr, mi);
    if (h->mfma && k >= 4 && h->n >= (uint32_t)dev::kKqTileBits) {
        switch (k) {
            case 4: return launch_kq_mfma<T, 4>(h, mr, mi, qs);
            case 5: return launch_kq_mfma<T, 5>(h, mr, mi, qs);
            case 6: return launch_kq_mfma<T, 6>(h, mr, mi, qs);
        }
    }
    if (use_kq_tiled(h, qs, k)) {
        switch (k) {
            case 2: return launch_kq_tiled<T, 2>(h, mr, mi, qs);
            case 3: return launch_kq_tiled<T, 3>(h, mr, mi, qs);
            case 4: return launch_kq_tiled<T, 4>(h, mr, mi, qs);
            case 5: return launch_kq_tiled<T, 5>(h, mr, mi, qs);
            case 6: return launch_kq_tiled<T, 6>(h, mr, mi, qs);
        }
    }
    if (k == 1) {
        dev::Mat2<T> m;
        for (int i = 0; i < 4; ++i) m.r[i] = mr[i], m.i[i] = mi[i];
        const uint64_t half = h->len >> 1;
        h->timed(ZB_K_1Q, h->sweep_bytes(), [&] {
            if (h->fma)
                hipLaunchKernelGGL((dev::k_1q<T, true>), dim3(grid_for(half)), dim3(dev::kBlock), 0, h->stream,
                                   h->R<T>(), h->I<T>(), qs[0], half, m);
            else
                hipLaunchKernelGGL((dev::k_1q<T, false>), dim3(grid_for(half)), dim3(dev::kBlock), 0, h->stream,
                                   h->R<T>(), h->I<T>(), qs[0], half, m);
        });
        check(hipGetLastError(), "k_1q");
        return;
    }
    std::vector<T> packed(2 * dim * dim);
    std::copy(mr.begin(), mr.end(), packed.begin());
    std::copy(mi.begin(), mi.end(), packed.begin() + dim * dim);
    const T *d = static_cast<const T *>(h->upload(packed.data(), packed.size() * sizeof(T)));
    const dev::Qubits q = make_qubits(qs, k);
    switch (k) {
        case 2: launch_kq<T, 2>(h, d, q); break;
        case 3: launch_kq<T, 3>(h, d, q); break;
        case 4: launch_kq<T, 4>(h, d, q); break;
        case 5: launch_kq<T, 5>(h, d, q); break;
        case 6: launch_kq<T, 6>(h, d, q); break;
    }
}

template <typename T>
void exact_diag_T(zb_backend *h, const uint32_t *qs, uint32_t k, const double *d) {
    const size_t dim = (size_t)1 << k;
    std::vector<T> dr, di;
    to_T<T>(d, dim, dr, di);
    if (k == 1) {
        const uint64_t half = h->len >> 1;
        h->timed(ZB_K_DIAG1, h->sweep_bytes(), [&] {
            if (h->fma)
                hipLaunchKernelGGL((dev::k_diag1<T, true>), dim3(grid_for(half)), dim3(dev::kBlock), 0, h->stream,
                                   h->R<T>(), h->I<T>(), qs[0], half, dr[0], di[0], dr[1], di[1]);
            else
                hipLaunchKernelGGL((dev::k_diag1<T, false>), dim3(grid_for(half)), dim3(dev::kBlock), 0, h->stream,
                                   h->R<T>(), h->I<T>(), qs[0], half, dr[0], di[0], dr[1], di[1]);
        });
        check(hipGetLastError(), "k_diag1");
        return;
    }
    std::vector<T> packed(2 * dim);
    std::copy(dr.begin(), dr.end(), packed.begin());
    std::copy(di.begin(), di.end(), packed.begin() + dim);
    const T *tab = static_cast<const T *>(h->upload(packed.data(), packed.size() * sizeof(T)));
    const dev::Qubits q = make_qubits(qs, k);
    const size_t lds = 2 * (size_t)dim * sizeof(T);
    h->timed(ZB_K_DIAGK, h->sweep_bytes(), [&] {
        if (h->fma)
            hipLaunchKernelGGL((dev::k_diagk<T, true>), dim3(grid_for(h->len)), dim3(dev::kBlock), lds, h->stream,
                               h->R<T>(), h->I<T>(), tab, q, h->len);
        else
            hipLaunchKernelGGL((dev::k_diagk<T, false>), dim3(grid_for(h->len)), dim3(dev::kBlock), lds, h->stream,
                               h->R<T>(), h->I<T>(), tab, q, h->len);
    });
    check(hipGetLastError(), "k_diagk");
}

template <typename T>
void exact_cx_T(zb_backend *h, uint32_t c, uint32_t t) {
    const uint64_t groups = h->len >> 2;
    h->timed(ZB_K_CX, h->sweep_bytes() / 2, [&] {
        hipLaunchKernelGGL(dev::k_cx<T>, dim3(grid_for(groups)), dim3(dev::kBlock), 0, h->stream, h->R<T>(),
                           h->I<T>(), std::min(c, t), std::max(c, t), 1ull << c, 1ull << t, groups);
    });
    check(hipGetLastError(), "k_cx");
}

// ---------------------------------------------------------------- windows
// Tiled mode: ops are queued into the open window while the window's qubit
// set Q (the lowest `low` qubits plus every qubit an op needs inside the
// tile) stays within the tile width; otherwise the window is flushed as one
// k_window pass.  A window holding a single op runs the per-op kernel, and
// so do dense ops on 5 or 6 qubits and every op of a state under 12 qubits.
uint32_t tile_width(const zb_backend *h) { return h->tile_bits ? h->tile_bits : 12; }
uint32_t tile_bits(const zb_backend *h) { return std::min<uint32_t>(h->n, tile_width(h)); }
// Tiled windows need n >= 12 (full tiles); lane offsets and rows are 64-bit
// state indices, so every size the backend holds (to 35 qubits) is tiled.
bool tiled(const zb_backend *h) { return h->mode == 1 && h->n >= tile_width(h); }

uint64_t low_mask(const zb_backend *h) {
    const uint32_t low = std::min<uint32_t>(h->prec == 64 ? 4 : 5, h->n);
    return (1ull << low) - 1;
}

// A tile touches 2^(its page-selecting bits) pages per array; a window
// keeps at most `page_bits` of them (the layout above makes this rarely
// bind).
uint32_t page_bits_of(const zb_backend *h, uint64_t mask) {
    const uint32_t pb = page_bit(h);
    return pb >= 64 ? 0 : (uint32_t)__builtin_popcountll(mask >> pb);
}

// Bytes of the descriptor blob for a window of `nops` ops and `data`
// bytes of matrices (stages <= ops).
size_t desc_bytes(size_t nops, size_t data) {
    auto pad = [](size_t x) { return (x + 15) & ~size_t(15); };
    return pad(sizeof(dev::WinHeader)) + pad(nops * sizeof(dev::WinStage)) + pad(nops * sizeof(dev::WinOp)) + data;
}

int nslots(const zb_backend *h) { return h->slots ? h->slots : 3; }

template <typename T, int TB, int NS>
void launch_window_shape(zb_backend *h, bool dense, unsigned grid, const uint32_t *blob, const T *tab,
                         uint64_t tiles) {
    const dim3 g(grid), b(dev::WinShape<T, TB, NS>::kLanes);
    // the fused-multiply-add variants exist with register prefetch only
    if (h->fma && dense)
        hipLaunchKernelGGL((dev::k_window<T, TB, NS, true, true, true>), g, b, 0, h->stream, h->R<T>(), h->I<T>(),
                           blob, tab, tiles);
    else if (h->fma)
        hipLaunchKernelGGL((dev::k_window<T, TB, NS, false, true, true>), g, b, 0, h->stream, h->R<T>(), h->I<T>(),
                           blob, tab, tiles);
    else if (dense && h->prefetch)
        hipLaunchKernelGGL((dev::k_window<T, TB, NS, true, true, false>), g, b, 0, h->stream, h->R<T>(), h->I<T>(),
                           blob, tab, tiles);
    else if (dense)
        hipLaunchKernelGGL((dev::k_window<T, TB, NS, true, false, false>), g, b, 0, h->stream, h->R<T>(), h->I<T>(),
                           blob, tab, tiles);
    else if (h->prefetch)
        hipLaunchKernelGGL((dev::k_window<T, TB, NS, false, true, false>), g, b, 0, h->stream, h->R<T>(), h->I<T>(),
                           blob, tab, tiles);
    else
        hipLaunchKernelGGL((dev::k_window<T, TB, NS, false, false, false>), g, b, 0, h->stream, h->R<T>(),
                           h->I<T>(), blob, tab, tiles);
}

template <typename T>
void launch_window(zb_backend *h, int ns, bool dense, unsigned grid, const uint32_t *blob, const T *tab,
                   uint64_t tiles) {
    if (tile_width(h) == 13) {
        if (ns == 3) launch_window_shape<T, 13, 3>(h, dense, grid, blob, tab, tiles);
        else launch_window_shape<T, 13, 4>(h, dense, grid, blob, tab, tiles);
    } else {
        if (ns == 3) launch_window_shape<T, 12, 3>(h, dense, grid, blob, tab, tiles);
        else launch_window_shape<T, 12, 4>(h, dense, grid, blob, tab, tiles);
    }
}

void close_window(zb_backend *h);
template <typename T>
void exact_unitary_T(zb_backend *h, const uint32_t *qs, uint32_t k, const double *mat);

// Queues an op on circuit qubits lq into the open window.  Its state bits
// are taken from phys[] now, and again if closing the window relabelled.
template <typename T>
void enqueue(zb_backend *h, uint32_t kind, const uint32_t *lq, uint32_t k, const double *src, size_t n_complex) {
    const uint64_t m = tile_bits(h);
    const bool table = kind == dev::kWDiag;
    const size_t add = table ? 0 : 2 * n_complex * sizeof(T);
    uint32_t qs[dev::kMaxDiagK];
    uint64_t need = 0;
    auto map = [&] {
        need = 0;
        for (uint32_t b = 0; b < k; ++b) qs[b] = h->phys[lq[b]];
        if (kind == dev::kW1q || kind == dev::kWDense)
            for (uint32_t b = 0; b < k; ++b) need |= 1ull << qs[b];
        else if (kind == dev::kWCx)
            need = 1ull << qs[0];  // the target; the control may stay outside the tile
    };
    map();
    if (!h->win_ops.empty() && ((uint64_t)__builtin_popcountll(h->win_mask | need) > m ||
                                page_bits_of(h, h->win_mask | need) > h->page_bits ||
                                desc_bytes(h->win_ops.size() + 1, h->win_data.size() + add) > dev::kDescBytes)) {
        close_window(h);
        map();
    }
    if (h->win_ops.empty()) h->win_mask = low_mask(h);
    h->win_mask |= need;
    zb_backend::POp op{};
    op.kind = kind;
    op.k = k;
    for (uint32_t b = 0; b < k; ++b) op.q[b] = qs[b];
    std::vector<unsigned char> &buf = table ? h->win_tables : h->win_data;
    op.data_off = (uint32_t)(buf.size() / sizeof(T));
    const size_t at = buf.size();
    buf.resize(at + 2 * n_complex * sizeof(T));
    T *dst = reinterpret_cast<T *>(buf.data() + at);
    for (size_t i = 0; i < n_complex; ++i) {
        dst[i] = (T)src[2 * i];  // T::fr
        dst[n_complex + i] = (T)src[2 * i + 1];
    }
    h->win_ops.push_back(op);
}

template <typename T>
void flush_T(zb_backend *h) {
    // one op: the per-op kernel is the cheaper pass, except for a dense block
    // on 3 or 4 qubits, whose per-op kernel is bound by its arithmetic
    if (h->win_ops.size() == 1 && !(h->win_ops[0].kind == dev::kWDense && h->win_ops[0].k >= 3)) {
        h->wb_swaps.clear();
        const zb_backend::POp op = h->win_ops[0];
        std::vector<double> inter;
        if (op.kind != dev::kWCx) {  // a CX carries no matrix
            const T *d =
                reinterpret_cast<const T *>(op.kind == dev::kWDiag ? h->win_tables.data() : h->win_data.data()) +
                op.data_off;
            const size_t len = op.kind == dev::kWDiag ? (1u << op.k) : (1u << (2 * op.k));
            inter.resize(2 * len);
            for (size_t i = 0; i < len; ++i) inter[2 * i] = (double)d[i], inter[2 * i + 1] = (double)d[len + i];
        }
        h->win_ops.clear();
        h->win_data.clear();
        h->win_tables.clear();
        if (op.kind == dev::kWCx) exact_cx_T<T>(h, op.q[1], op.q[0]);
        else if (op.kind == dev::kWDiag) exact_diag_T<T>(h, op.q, op.k, inter.data());
        else exact_unitary_T<T>(h, op.q, op.k, inter.data());
        return;
    }
    const uint32_t m = tile_bits(h);
    uint64_t mask = h->win_mask;
    for (uint32_t q = 0; q < h->n && (uint32_t)__builtin_popcountll(mask) < m; ++q) mask |= 1ull << q;
    dev::WinHeader hdr{};
    hdr.m = m;
    int local[64];
    for (int i = 0; i < 64; ++i) local[i] = -1;
    uint32_t idx = 0;
    for (uint32_t q = 0; q < h->n; ++q)
        if ((mask >> q) & 1) hdr.q[idx] = q, local[q] = (int)idx++;
    // fwd[i]: the local bit that local bit i moves to; the kernel reads
    // through the inverse
    uint32_t fwd[dev::kMaxTileBits];
    for (uint32_t i = 0; i < m; ++i) fwd[i] = i;
    for (const auto &sw : h->wb_swaps) std::swap(fwd[local[sw.first]], fwd[local[sw.second]]);
    for (uint32_t i = 0; i < m; ++i) hdr.wb[fwd[i]] = i;
    hdr.relabel = h->wb_swaps.empty() ? 0 : 1;
    h->wb_swaps.clear();
    // Stages: greedy over the ops in order; a stage's slots must hold every
    // local bit its ops act on (CX: the target only; diagonals: none).
    auto needs = [&](const zb_backend::POp &p) {
        uint32_t nb = 0;
        if (p.kind == dev::kW1q || p.kind == dev::kWDense)
            for (uint32_t b = 0; b < p.k; ++b) nb |= 1u << local[p.q[b]];
        else if (p.kind == dev::kWCx)
            nb |= 1u << local[p.q[0]];
        return nb;
    };
    std::vector<dev::WinStage> stages;
    std::vector<uint32_t> stage_slots;  // bitmask of local bits per stage
    std::vector<std::array<int, 64>> stage_index;  // local bit -> register bit, per stage
    // A dense op's matrix bit b must sit on a lower register bit than its
    // bit b + 1 (the register kernels walk a block's bits in that order), so
    // a stage's slots are ordered: before[a] holds the local bits that must
    // precede local bit a.  Windows with a 4-qubit block run 4 slots.
    int ns = nslots(h);
    for (const auto &p : h->win_ops)
        if (p.kind == dev::kWDense) ns = std::max(ns, (int)p.k);
    uint32_t cur = 0, before[dev::kMaxTileBits] = {};
    size_t begin = 0;
    bool dense = false;
    // a topological order of the stage's slots (ties: lower local bit first)
    // or an empty vector if the constraints have a cycle
    auto order_of = [&](uint32_t sl, const uint32_t *bef) {
        std::vector<int> ord;
        uint32_t done = 0;
        while (done != sl) {
            int pick = -1;
            for (uint32_t b = 0; b < m && pick < 0; ++b)
                if (((sl & ~done) >> b) & 1 && (bef[b] & sl & ~done) == 0) pick = (int)b;
            if (pick < 0) return std::vector<int>{};
            ord.push_back(pick);
            done |= 1u << pick;
        }
        return ord;
    };
    auto close = [&](size_t end) {
        uint32_t sl = cur;
        for (uint32_t b = 0; b < m && __builtin_popcount(sl) < ns; ++b) sl |= 1u << b;
        const std::vector<int> ord = order_of(sl, before);
        dev::WinStage st{};
        std::array<int, 64> index;
        index.fill(-1);
        for (int i = 0; i < (int)ord.size(); ++i) st.slot[i] = (uint32_t)ord[i], index[ord[i]] = i;
        int j = 0;
        for (uint32_t b = 0; b < m; ++b)
            if ((sl >> b) & 1) st.ins[j++] = b;
        st.op_begin = (uint32_t)begin;
        st.op_end = (uint32_t)end;
        stages.push_back(st);
        stage_slots.push_back(sl);
        stage_index.push_back(index);
    };
    for (size_t i = 0; i < h->win_ops.size(); ++i) {
        const zb_backend::POp &p = h->win_ops[i];
        const uint32_t nb = needs(p);
        uint32_t bef[dev::kMaxTileBits];
        std::copy(before, before + m, bef);
        if (p.kind == dev::kWDense)
            for (uint32_t b = 1; b < p.k; ++b)
                for (uint32_t a = 0; a < b; ++a) bef[local[p.q[b]]] |= 1u << local[p.q[a]];
        if (__builtin_popcount(cur | nb) > ns || order_of(cur | nb, bef).empty()) {
            close(i);
            begin = i;
            cur = 0;
            std::fill(before, before + m, 0u);
            if (p.kind == dev::kWDense)
                for (uint32_t b = 1; b < p.k; ++b)
                    for (uint32_t a = 0; a < b; ++a) before[local[p.q[b]]] |= 1u << local[p.q[a]];
        } else {
            std::copy(bef, bef + m, before);
        }
        cur |= nb;
    }
    close(h->win_ops.size());
    hdr.nstages = (uint32_t)stages.size();
    hdr.nslots = (uint32_t)ns;
    std::vector<dev::WinOp> wops(h->win_ops.size());
    for (size_t si = 0; si < stages.size(); ++si) {
        const uint32_t sl = stage_slots[si];
        auto slot_of = [&](int lb) { return (uint32_t)stage_index[si][lb]; };
        for (uint32_t oi = stages[si].op_begin; oi < stages[si].op_end; ++oi) {
            const zb_backend::POp &p = h->win_ops[oi];
            dev::WinOp &w = wops[oi];
            w = dev::WinOp{};
            w.kind = p.kind;
            w.k = p.k;
            w.data_off = p.data_off;
            if (p.kind == dev::kW1q) {
                w.sel = slot_of(local[p.q[0]]);
            } else if (p.kind == dev::kWDense) {
                for (uint32_t b = 0; b < p.k; ++b) w.sel |= 1u << slot_of(local[p.q[b]]);
                dense = true;
            } else if (p.kind == dev::kWCx) {
                const uint32_t tq = p.q[0], cq = p.q[1];
                w.sel = slot_of(local[tq]);
                if (local[cq] >= 0 && ((sl >> local[cq]) & 1)) {
                    w.sel |= slot_of(local[cq]) << 2 | dev::kCtrlSlot << 4;
                } else if (local[cq] >= 0) {
                    w.sel |= dev::kCtrlLocal << 4;
                    w.aux = (uint32_t)local[cq];
                } else {
                    w.sel |= dev::kCtrlFixed << 4;
                    w.aux = cq;
                }
            } else {
                for (uint32_t b = 0; b < p.k; ++b)
                    w.bits[b] = local[p.q[b]] >= 0 ? (uint32_t)local[p.q[b]] : (dev::kFixedBit | p.q[b]);
            }
        }
    }
    auto pad = [](size_t x) { return (x + 15) & ~size_t(15); };
    const size_t hb = pad(sizeof(hdr)), sb = pad(stages.size() * sizeof(dev::WinStage)),
                 ob = pad(wops.size() * sizeof(dev::WinOp));
    const size_t total = hb + sb + ob + pad(h->win_data.size());
    hdr.nops = (uint32_t)wops.size();
    hdr.desc_words = (uint32_t)(total / 4);
    hdr.stage_off = (uint32_t)(hb / 4);
    hdr.op_off = (uint32_t)((hb + sb) / 4);
    hdr.data_off = (uint32_t)((hb + sb + ob) / 4);
    if (total > (size_t)dev::kDescBytes) throw std::runtime_error("window descriptor exceeds its size limit");
    std::vector<unsigned char> blob(total, 0);
    std::memcpy(blob.data(), &hdr, sizeof(hdr));
    std::memcpy(blob.data() + hb, stages.data(), stages.size() * sizeof(dev::WinStage));
    std::memcpy(blob.data() + hb + sb, wops.data(), wops.size() * sizeof(dev::WinOp));
    std::memcpy(blob.data() + hb + sb + ob, h->win_data.data(), h->win_data.size());
    const uint32_t *dblob = static_cast<const uint32_t *>(h->upload(blob.data(), blob.size()));
    const T *dtab = h->win_tables.empty()
                        ? nullptr
                        : static_cast<const T *>(h->upload(h->win_tables.data(), h->win_tables.size()));
    const uint64_t tiles = h->len >> m;
    const unsigned grid = (unsigned)std::min<uint64_t>(tiles, h->win_grid);
    if (h->prof) {
        int32_t nd = 0, nc = 0;
        for (const auto &p : h->win_ops) nd += p.kind == dev::kWDiag, nc += p.kind == dev::kWCx;
        h->next_detail[0] = (int32_t)h->win_ops.size();
        h->next_detail[1] = (int32_t)stages.size();
        h->next_detail[2] = nd;
        h->next_detail[3] = nc;
    }
    h->timed(ZB_K_WINDOW, h->sweep_bytes(), [&] {
        launch_window<T>(h, ns, dense, grid, dblob, dtab, tiles);
    });
    check(hipGetLastError(), "k_window");
    h->win_ops.clear();
    h->win_data.clear();
    h->win_tables.clear();
}

// Launches the open window (if any) without touching held-back CX ops.
void flush_window(zb_backend *h) {
    if (h->win_ops.empty()) return;
    if (h->prec == 64) flush_T<double>(h);
    else flush_T<float>(h);
}

// Closes the open window because the next op does not fit.  With the
// coming ops known, the window's write-back also moves onto the low state
// bits (which every window holds) those of its tile's qubits that the
// coming ops need soonest, swapping them with the low bits' present
// occupants; phys[] follows.  A relabelling is only a permutation of the
// state index bits, applied to every amplitude of every tile: results do
// not change.
void close_window(zb_backend *h) {
    h->wb_swaps.clear();
    std::vector<std::pair<uint32_t, uint32_t>> moves;  // circuit qubits (into low, out of low)
    if (h->relabel && h->la_need && h->win_ops.size() > 1) {
        const uint32_t low = (uint32_t)__builtin_popcountll(low_mask(h));
        uint64_t in_tile = 0;  // circuit qubits of the tile
        for (uint32_t q = 0; q < h->n; ++q)
            if ((h->win_mask >> h->phys[q]) & 1) in_tile |= 1ull << q;
        std::vector<std::pair<size_t, uint32_t>> next;  // (next use, qubit)
        uint64_t open = in_tile;
        for (size_t i = h->la_pos; i < h->la_count && open; ++i) {
            uint64_t hit = h->la_need[i] & open;
            open &= ~hit;
            for (; hit; hit &= hit - 1) next.emplace_back(i, (uint32_t)__builtin_ctzll(hit));
        }
        for (uint64_t r = open; r; r &= r - 1) next.emplace_back(SIZE_MAX, (uint32_t)__builtin_ctzll(r));
        // soonest first; among qubits no coming op uses, circuit qubits
        // 0..low-1 first (circuit-order reads of the finished state then
        // walk whole 128-byte runs); on other ties the present low
        // occupants stay
        std::stable_sort(next.begin(), next.end(), [&](const auto &a, const auto &b) {
            if (a.first != b.first) return a.first < b.first;
            if (a.first == SIZE_MAX && (a.second < low) != (b.second < low)) return a.second < low;
            return (h->phys[a.second] < low) > (h->phys[b.second] < low);
        });
        uint64_t keep = 0;
        for (size_t i = 0; i < next.size() && i < low; ++i) keep |= 1ull << next[i].second;
        std::vector<uint32_t> in, out;
        for (uint32_t q = 0; q < h->n; ++q) {
            const bool is_low = h->phys[q] < low, kept = (keep >> q) & 1;
            if (kept && !is_low) in.push_back(q);
            if (!kept && is_low) out.push_back(q);
        }
        for (size_t i = 0; i < in.size() && i < out.size(); ++i) {
            moves.emplace_back(in[i], out[i]);
            h->wb_swaps.emplace_back(h->phys[in[i]], h->phys[out[i]]);
        }
    }
    flush_window(h);
    for (const auto &mv : moves) std::swap(h->phys[mv.first], h->phys[mv.second]);
    if (!moves.empty()) h->lut_dirty = true;
}

// Brings the device copy of the layout up to date with phys[] before a
// kernel that walks circuit-order indices.
void refresh_layout(zb_backend *h) {
    if (!h->lut_dirty) return;
    h->sync();
    upload_layout(h);
}

// CX on circuit qubits (control c, target t).
template <typename T>
void enqueue_cx(zb_backend *h, uint32_t c, uint32_t t) {
    const uint32_t lq[2] = {t, c};  // bits[0] = target, bits[1] = control
    enqueue<T>(h, dev::kWCx, lq, 2, nullptr, 0);
}

// Queues, in order, the held-back CX ops that share a qubit with `qubits`
// (a mask of circuit qubits).
template <typename T>
void settle(zb_backend *h, uint64_t qubits) {
    if (h->deferred.empty()) return;
    std::vector<std::pair<uint32_t, uint32_t>> keep, emit;
    for (const auto &cx : h->deferred)
        (((qubits >> cx.first) & 1) || ((qubits >> cx.second) & 1) ? emit : keep).push_back(cx);
    h->deferred.swap(keep);
    for (const auto &cx : emit) enqueue_cx<T>(h, cx.first, cx.second);
}

// Everything queued reaches the device: held-back CX ops, then the window.
void flush(zb_backend *h) {
    if (h->prec == 64) settle<double>(h, ~0ull);
    else settle<float>(h, ~0ull);
    flush_window(h);
}

// The state is being overwritten: queued work is dropped.
void discard_pending(zb_backend *h) {
    h->win_ops.clear();
    h->win_data.clear();
    h->win_tables.clear();
    h->deferred.clear();
}

// The apply_* entry points take circuit qubits and hand state bits to the
// kernels.  A dense op whose qubits land on state bits in a different order
// keeps its matrix as written: the per-op kernel gathers in op order, and a
// window stage orders its register slots by the matrix's bits.  Dense ops on
// up to 4 qubits join windows (a window holding a 4-qubit block runs with 4
// slots); larger ones, and 4-qubit ones under the knob "mfma", run the per-op
// kernels.
template <typename T>
void apply_unitary_T(zb_backend *h, const uint32_t *qs, uint32_t k, const double *mat) {
    uint32_t pq[dev::kMaxDiagK];
    uint64_t qubits = 0;
    for (uint32_t b = 0; b < k; ++b) qubits |= 1ull << qs[b];
    const bool on_mfma = h->mfma && k >= 4 && h->n >= (uint32_t)dev::kKqTileBits;  // knob "mfma"
    if (!tiled(h) || k > (uint32_t)dev::kMaxSlots || on_mfma) {
        flush(h);  // held-back CX ops may close a relabelling window: map again
        for (uint32_t b = 0; b < k; ++b) pq[b] = h->phys[qs[b]];
        return exact_unitary_T<T>(h, pq, k, mat);
    }
    settle<T>(h, qubits);
    enqueue<T>(h, k == 1 ? dev::kW1q : dev::kWDense, qs, k, mat, (size_t)1 << (2 * k));
}

template <typename T>
void apply_diag_T(zb_backend *h, const uint32_t *qs, uint32_t k, const double *d) {
    if (!tiled(h)) {
        uint32_t pq[dev::kMaxDiagK];
        for (uint32_t b = 0; b < k; ++b) pq[b] = h->phys[qs[b]];
        return exact_diag_T<T>(h, pq, k, d);
    }
    uint64_t qubits = 0;
    for (uint32_t b = 0; b < k; ++b) qubits |= 1ull << qs[b];
    settle<T>(h, qubits);
    enqueue<T>(h, dev::kWDiag, qs, k, d, (size_t)1 << k);
}

template <typename T>
void apply_cx_T(zb_backend *h, uint32_t c, uint32_t t) {
    if (!tiled(h)) return exact_cx_T<T>(h, h->phys[c], h->phys[t]);
    settle<T>(h, (1ull << c) | (1ull << t));
    const uint64_t mask = h->win_ops.empty() ? low_mask(h) : h->win_mask;
    if (h->defer_cx && !((mask >> h->phys[t]) & 1)) {  // wait for a window that holds the target
        if (h->deferred.size() >= 64) {
            const auto oldest = h->deferred.front();
            h->deferred.erase(h->deferred.begin());
            enqueue_cx<T>(h, oldest.first, oldest.second);
        }
        h->deferred.emplace_back(c, t);
        return;
    }
    enqueue_cx<T>(h, c, t);
}

// Fixed-order reduction (mode 0: norm, mode 1: prob_one of q).
template <typename T>
double reduce_T(zb_backend *h, int mode, uint32_t q) {
    flush(h);
    refresh_layout(h);
    uint64_t n = mode ? h->len >> 1 : h->len;
    uint64_t nb = (n + dev::kReduceSpan - 1) / dev::kReduceSpan;
    h->ensure_red(2 * nb + 2);
    double *a = h->red_dev, *b = h->red_dev + nb + 1;
    hipLaunchKernelGGL(dev::k_reduce_state<T>, dim3((unsigned)nb), dim3(dev::kBlock), 0, h->stream, h->R<T>(),
                       h->I<T>(), h->layout(), n, mode, q, a);
    check(hipGetLastError(), "k_reduce_state");
    while (nb > 1) {
        const uint64_t nn = (nb + dev::kReduceSpan - 1) / dev::kReduceSpan;
        hipLaunchKernelGGL(dev::k_reduce_f64, dim3((unsigned)nn), dim3(dev::kBlock), 0, h->stream, a, nb, b);
        check(hipGetLastError(), "k_reduce_f64");
        std::swap(a, b);
        nb = nn;
    }
    check(hipMemcpyAsync(h->red_host, a, sizeof(double), hipMemcpyDeviceToHost, h->stream), "reduce copy");
    h->sync();
    return *h->red_host;
}

template <typename T>
int measure_T(zb_backend *h, uint32_t q, double u) {
    flush(h);
    const double p1 = reduce_T<T>(h, 1, q);
    const int outcome = u < p1;
    const double p = outcome ? p1 : 1.0 - p1;
    const T scale = (T)(1.0 / std::sqrt(p));
    const uint64_t half = h->len >> 1;
    hipLaunchKernelGGL(dev::k_collapse<T>, dim3(grid_for(half)), dim3(dev::kBlock), 0, h->stream, h->R<T>(),
                       h->I<T>(), (uint32_t)h->phys[q], half, outcome, scale);
    check(hipGetLastError(), "k_collapse");
    return outcome;
}

// The sampler's chunk walks are latency-bound serial chains: from 8 walks
// per CU on, the narrow shape lets more of them share a CU.
bool samp_narrow(uint64_t walks) {
    static const uint64_t cus = [] {
        int d = 0, c = 0;
        if (hipGetDevice(&d) != hipSuccess || hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, d) !=
                                                   hipSuccess || c <= 0)
            c = 256;
        return (uint64_t)c;
    }();
    return walks >= 8 * cus;
}

template <typename T, typename S>
void launch_rescan(zb_backend *h, uint64_t blocks, uint64_t chunk, const uint32_t *touched, const double *cum,
                   const uint64_t *starts, const double *us, uint64_t *out) {
    hipLaunchKernelGGL((dev::k_rescan<T, S>), dim3((unsigned)blocks), dim3(S::kBlock), 0, h->stream, h->R<T>(),
                       h->I<T>(), h->layout(), h->len, chunk, touched, cum, starts, us, out);
    check(hipGetLastError(), "k_rescan");
}

template <typename T>
void rescan_T(zb_backend *h, uint64_t blocks, uint64_t chunk, const uint32_t *touched, const double *cum,
              const uint64_t *starts, const double *us, uint64_t *out) {
    if (samp_narrow(blocks)) launch_rescan<T, dev::SampNarrow>(h, blocks, chunk, touched, cum, starts, us, out);
    else launch_rescan<T, dev::SampWide>(h, blocks, chunk, touched, cum, starts, us, out);
}

template <typename T, int G, int SLAB>
void launch_chunk_sums_multi(zb_backend *h, double *dsums, uint64_t nch) {
    hipLaunchKernelGGL((dev::k_chunk_sums_multi<T, G, SLAB>), dim3((unsigned)((nch + G - 1) / G)), dim3(256), 0,
                       h->stream, h->R<T>(), h->I<T>(), h->layout(), nch, dsums);
}

template <typename T>
void chunk_sums_T(zb_backend *h, double *dsums, uint64_t chunk, uint64_t nch) {
    // many full chunks: 8 serial chains per workgroup (measured at 26-30
    // qubits against 4..64 chains and 1024..4096-value slabs)
    if (chunk == (1ull << 16) && h->chunk_multi != 0 && nch >= h->chunk_multi) {
        launch_chunk_sums_multi<T, 8, 2048>(h, dsums, nch);
    } else if (samp_narrow(nch))
        hipLaunchKernelGGL((dev::k_chunk_sums<T, dev::SampNarrow>), dim3((unsigned)nch),
                           dim3(dev::SampNarrow::kBlock), 0, h->stream, h->R<T>(), h->I<T>(), h->layout(), h->len,
                           chunk, dsums);
    else
        hipLaunchKernelGGL((dev::k_chunk_sums<T, dev::SampWide>), dim3((unsigned)nch), dim3(dev::SampWide::kBlock),
                           0, h->stream, h->R<T>(), h->I<T>(), h->layout(), h->len, chunk, dsums);
    check(hipGetLastError(), "k_chunk_sums");
}

// sample_indices (src/sample.rs:29-111).  Passes 1 and 5-6 run on the
// device; the serial prefix and the serial draws (passes 2-3) run on the
// host, in the reference's order.  The sort and the routing (passes 3-4)
// run on the host below kDeviceSortShots shots and on the device (sort.hip)
// from there, where the host sort would cost more than the launches.
constexpr uint64_t kDeviceSortShots = 1ull << 15;
constexpr uint64_t kDeviceSortMaxShots = 1ull << 30;  // the radix sort counts items in an int

// Device scratch the sampler needs for `shots` draws from a state of `len`
// amplitudes: the largest of its three layouts (chunk sums; the device sort
// and routing; the host-routed rescan with at most one touched entry per
// chunk), each as laid out below.
size_t sample_scratch_bytes(uint64_t len, uint64_t shots) {
    const uint64_t chunk = std::min<uint64_t>(len, 1ull << 16);
    const uint64_t nch = (len + chunk - 1) / chunk;
    size_t need = nch * sizeof(double);
    if (shots >= kDeviceSortShots && shots <= kDeviceSortMaxShots) {
        const size_t tb = (draw_sort_temp_bytes(shots) + 255) & ~size_t(255);
        const size_t nb = ((nch + 1) * 8 + 255) & ~size_t(255), sb = (shots * 8 + 255) & ~size_t(255);
        const size_t jb = (draw_jump_bytes(shots) + 255) & ~size_t(255);
        need = std::max(need, tb + 2 * nb + 3 * sb + jb + 256);
    } else {
        need = std::max<size_t>(need, std::min(nch, shots) * 4 + (nch + 1) * 8 * 2 + shots * 8 * 2 + 64);
    }
    return need;
}

// The sampler's device scratch comes from the handle's grow-only buffer, or,
// when `fixed` is given, from that region (which the caller has checked
// holds sample_scratch_bytes); nothing is then allocated.
template <typename T>
void sample_T(zb_backend *h, uint64_t shots, uint64_t seed, uint64_t *out, char *fixed = nullptr) {
    flush(h);
    refresh_layout(h);
    if (shots == 0) return;
    auto get_scratch = [&](size_t bytes) { return fixed ? fixed : h->scratch_get(bytes); };
    const uint64_t chunk = std::min<uint64_t>(h->len, 1ull << 16);
    const uint64_t nch = (h->len + chunk - 1) / chunk;
    double *dsums = (double *)get_scratch(nch * sizeof(double));
    chunk_sums_T<T>(h, dsums, chunk, nch);
    std::vector<double> sums(nch);
    check(hipMemcpyAsync(sums.data(), dsums, nch * sizeof(double), hipMemcpyDeviceToHost, h->stream), "sums copy");
    h->sync();
    std::vector<double> cum(nch + 1, 0.0);
    for (uint64_t c = 0; c < nch; ++c) cum[c + 1] = cum[c] + sums[c];
    const double total = cum[nch];
    Rng rng(seed);
    if (shots >= kDeviceSortShots && shots <= kDeviceSortMaxShots) {
        // draw, sort, route and rescan on the device; every chunk gets a block
        const size_t tb = (draw_sort_temp_bytes(shots) + 255) & ~size_t(255);
        const size_t nb = ((nch + 1) * 8 + 255) & ~size_t(255), sb = (shots * 8 + 255) & ~size_t(255);
        const size_t jb = (draw_jump_bytes(shots) + 255) & ~size_t(255);
        char *p = get_scratch(tb + 2 * nb + 3 * sb + jb + 256);
        p = (char *)(((uintptr_t)p + 255) & ~(uintptr_t)255);
        void *temp = p;
        double *dcum = (double *)(p + tb);
        uint64_t *dstarts = (uint64_t *)(p + tb + nb);
        double *dus = (double *)(p + tb + 2 * nb);
        double *dsorted = (double *)(p + tb + 2 * nb + sb);
        uint64_t *dout = (uint64_t *)(p + tb + 2 * nb + 2 * sb);
        void *djump = p + tb + 2 * nb + 3 * sb;
        std::vector<uint64_t> staging;  // read by the copies below until the final sync
        check(hipMemcpyAsync(dcum, cum.data(), (nch + 1) * 8, hipMemcpyHostToDevice, h->stream), "cum");
        check(draws_on_device(rng.state(), shots, total, djump, staging, dus, h->stream), "k_draws");
        check(draw_sort_and_route(temp, tb, dus, dsorted, shots, dcum, nch + 1, dstarts, h->stream), "draw sort");
        rescan_T<T>(h, nch, chunk, nullptr, dcum, dstarts, dsorted, dout);
        check(hipMemcpyAsync(out, dout, shots * 8, hipMemcpyDeviceToHost, h->stream), "out copy");
        h->sync();
        return;
    }
    std::vector<double> us(shots);
    for (uint64_t i = 0; i < shots; ++i) us[i] = rng.f64() * total;
    std::sort(us.begin(), us.end());
    std::vector<uint64_t> starts(nch + 1);
    for (uint64_t c = 0; c <= nch; ++c)
        starts[c] = (uint64_t)(std::lower_bound(us.begin(), us.end(), cum[c]) - us.begin());
    std::vector<uint32_t> touched;
    for (uint64_t c = 0; c < nch; ++c)
        if (starts[c] < starts[c + 1]) touched.push_back((uint32_t)c);
    if (!touched.empty()) {
        const size_t bytes = touched.size() * 4 + (nch + 1) * 8 * 2 + shots * 8 * 2;
        char *p = get_scratch(bytes + 64);
        double *dcum = (double *)p;
        p += (nch + 1) * 8;
        uint64_t *dstarts = (uint64_t *)p;
        p += (nch + 1) * 8;
        double *dus = (double *)p;
        p += shots * 8;
        uint64_t *dout = (uint64_t *)p;
        p += shots * 8;
        uint32_t *dtouched = (uint32_t *)p;
        check(hipMemcpyAsync(dcum, cum.data(), (nch + 1) * 8, hipMemcpyHostToDevice, h->stream), "cum");
        check(hipMemcpyAsync(dstarts, starts.data(), (nch + 1) * 8, hipMemcpyHostToDevice, h->stream), "starts");
        check(hipMemcpyAsync(dus, us.data(), shots * 8, hipMemcpyHostToDevice, h->stream), "us");
        check(hipMemcpyAsync(dtouched, touched.data(), touched.size() * 4, hipMemcpyHostToDevice, h->stream),
              "touched");
        rescan_T<T>(h, touched.size(), chunk, dtouched, dcum, dstarts, dus, dout);
        check(hipMemcpyAsync(out, dout, shots * 8, hipMemcpyDeviceToHost, h->stream), "out copy");
    }
    h->sync();
}

// Host copies of the state in logical index order: straight copies under
// the identity layout, else chunk by chunk through a device gather/scatter.
constexpr uint64_t kCopyChunk = 1ull << 22;

template <typename T>
void read_state_T(zb_backend *h, T *re, T *im) {
    flush(h);
    refresh_layout(h);
    if (h->ident) {
        check(hipMemcpyAsync(re, h->re, h->len * sizeof(T), hipMemcpyDeviceToHost, h->stream), "copy re");
        check(hipMemcpyAsync(im, h->im, h->len * sizeof(T), hipMemcpyDeviceToHost, h->stream), "copy im");
        h->sync();
        return;
    }
    const uint64_t chunk = std::min<uint64_t>(h->len, kCopyChunk);
    T *buf = reinterpret_cast<T *>(h->scratch_get(2 * chunk * sizeof(T)));
    for (uint64_t lo = 0; lo < h->len; lo += chunk) {
        const uint64_t cnt = std::min<uint64_t>(chunk, h->len - lo);
        hipLaunchKernelGGL(dev::k_gather<T>, dim3(grid_for(cnt)), dim3(dev::kBlock), 0, h->stream, h->R<T>(),
                           h->I<T>(), h->layout(), lo, cnt, buf, buf + chunk);
        check(hipGetLastError(), "k_gather");
        check(hipMemcpyAsync(re + lo, buf, cnt * sizeof(T), hipMemcpyDeviceToHost, h->stream), "copy re");
        check(hipMemcpyAsync(im + lo, buf + chunk, cnt * sizeof(T), hipMemcpyDeviceToHost, h->stream), "copy im");
        h->sync();
    }
}

template <typename T>
void write_state_T(zb_backend *h, const T *re, const T *im) {
    discard_pending(h);  // the state is being overwritten
    refresh_layout(h);
    if (h->ident) {
        check(hipMemcpyAsync(h->re, re, h->len * sizeof(T), hipMemcpyHostToDevice, h->stream), "copy re");
        check(hipMemcpyAsync(h->im, im, h->len * sizeof(T), hipMemcpyHostToDevice, h->stream), "copy im");
        h->sync();
        return;
    }
    const uint64_t chunk = std::min<uint64_t>(h->len, kCopyChunk);
    T *buf = reinterpret_cast<T *>(h->scratch_get(2 * chunk * sizeof(T)));
    for (uint64_t lo = 0; lo < h->len; lo += chunk) {
        const uint64_t cnt = std::min<uint64_t>(chunk, h->len - lo);
        check(hipMemcpyAsync(buf, re + lo, cnt * sizeof(T), hipMemcpyHostToDevice, h->stream), "copy re");
        check(hipMemcpyAsync(buf + chunk, im + lo, cnt * sizeof(T), hipMemcpyHostToDevice, h->stream), "copy im");
        hipLaunchKernelGGL(dev::k_scatter<T>, dim3(grid_for(cnt)), dim3(dev::kBlock), 0, h->stream, h->R<T>(),
                           h->I<T>(), h->layout(), lo, cnt, buf, buf + chunk);
        check(hipGetLastError(), "k_scatter");
        h->sync();
    }
}

template <typename F>
int guarded(F &&f) {
    try {
        f();
        return ZB_OK;
    } catch (const HipError &e) {
        set_error(e.msg);
        return ZB_ERR_DEVICE;
    } catch (const std::exception &e) {
        set_error(e.what());
        return ZB_ERR_ARG;
    }
}

bool sorted_distinct(const uint32_t *qs, uint32_t k, uint32_t n) {
    for (uint32_t b = 0; b < k; ++b) {
        if (qs[b] >= n) return false;
        if (b && qs[b] <= qs[b - 1]) return false;
    }
    return true;
}

}  // namespace

namespace zeno {

void backend_lookahead(zb_backend *h, const uint64_t *need, size_t count) {
    h->la_need = need;
    h->la_count = need ? count : 0;
    h->la_pos = 0;
}

void backend_lookahead_pos(zb_backend *h, size_t pos) { h->la_pos = pos; }

// End of a run with lookahead: the held-back CX ops are queued and the last
// window closes as one that no coming op uses, which moves circuit qubits
// 0..low-1 of its tile back onto the low state bits.
int backend_end_run(zb_backend *h) {
    return guarded([&] {
        h->la_pos = h->la_count;
        if (h->prec == 64) settle<double>(h, ~0ull);
        else settle<float>(h, ~0ull);
        // The tile is filled up to its width anyway: its spare bits go to
        // circuit qubits 0..low-1 that are off the low bits, so that the
        // write-back can home them too.
        if (h->relabel && h->la_need && h->win_ops.size() > 1) {
            const uint32_t m = tile_bits(h), low = (uint32_t)__builtin_popcountll(low_mask(h));
            for (uint32_t q = 0; q < low; ++q) {
                const uint64_t next = h->win_mask | (1ull << h->phys[q]);
                if ((uint32_t)__builtin_popcountll(next) > m) break;
                if (page_bits_of(h, next) <= h->page_bits) h->win_mask = next;
            }
        }
        close_window(h);
    });
}

}  // namespace zeno

extern "C" {

const char *zb_last_error(void) { return last_error(); }


namespace {

// Process-wide knob defaults per (qubits, precision), applied to every
// backend of that shape when it is created or re-bound (the autotuner
// installs its choices here).
std::mutex g_tune_mu;
std::map<std::tuple<uint32_t, int, std::string>, int64_t> g_tune_defaults;

int apply_knob(zb_backend *h, const std::string &k, int64_t value);

void apply_defaults(zb_backend *h) {
    std::vector<std::pair<std::string, int64_t>> kv;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        for (const auto &e : g_tune_defaults)
            if (std::get<0>(e.first) == h->n && std::get<1>(e.first) == h->prec)
                kv.emplace_back(std::get<2>(e.first), e.second);
    }
    for (const auto &p : kv)
        if (p.first != "layout") apply_knob(h, p.first, p.second);
}

int64_t default_layout(const zb_backend *h) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_defaults.find({h->n, h->prec, "layout"});
    return it == g_tune_defaults.end() ? -1 : it->second;
}

int check_shape(uint32_t n_qubits, int precision) {
    if (precision != 32 && precision != 64) {
        set_error("precision must be 32 or 64");
        return ZB_ERR_ARG;
    }
    if (n_qubits == 0 || n_qubits > 35) {
        set_error("the HIP backend supports 1..35 qubits (got " + std::to_string(n_qubits) + ")");
        return ZB_ERR_UNSUPPORTED;
    }
    if (zb_device_count() == 0) {
        set_error("no HIP device is available (the hip backend needs an AMD GPU)");
        return ZB_ERR_UNSUPPORTED;
    }
    return ZB_OK;
}

// Every tuning knob back to the library default (a fresh handle's value),
// so a re-bound handle carries nothing over from the shape it ran before.
void reset_knobs(zb_backend *h) {
    const zb_backend d;
    h->prefetch = d.prefetch;
    h->slots = d.slots;
    h->win_grid = d.win_grid;
    h->page_bits = d.page_bits;
    h->tile_bits = d.tile_bits;
    h->defer_cx = d.defer_cx;
    h->relabel = d.relabel;
    h->mfma = d.mfma;
    h->kq_tiled = d.kq_tiled;
    h->fma = d.fma;
    h->chunk_multi = d.chunk_multi;
}

void set_shape(zb_backend *h, uint32_t n_qubits, int precision) {
    h->n = n_qubits;
    h->prec = precision;
    h->len = 1ull << n_qubits;
    h->elem = precision == 64 ? 8 : 4;
    h->name = precision == 64 ? "hip-f64" : "hip-f32";
}

// The device's copy of the allocation-free sampler's scratch array, looked
// up once per device with that device current.  A failed lookup is kept as
// an error message for zb_sample_noalloc (every other entry point works
// without the array).
std::string g_noalloc_err[kMaxDevices];

void noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        g_noalloc_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return;
    }
    g_noalloc_ptr[device] = p;
}

void *noalloc_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    return g_noalloc_ptr[device];
}

// Stream, parameter arena, pinned result slot and layout table: everything
// but the state (the layout table too, so that no read of the state has to
// allocate later).
void create_resources(zb_backend *h) {
    check(hipGetDevice(&h->device), "hipGetDevice");
    noalloc_setup(h->device);
    check(hipMalloc((void **)&h->lut_dev, 5 * 256 * sizeof(uint64_t)), "hipMalloc layout");
    check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
    check(hipMalloc((void **)&h->arena_dev, kArenaBytes), "hipMalloc arena");
    check(hipHostMalloc((void **)&h->arena_host, kArenaBytes, hipHostMallocDefault), "hipHostMalloc arena");
    check(hipHostMalloc((void **)&h->red_host, sizeof(double), hipHostMallocDefault), "hipHostMalloc red");
}

void zero_state(zb_backend *h) {
    if (h->prec == 64) launch_zero<double>(h);
    else launch_zero<float>(h);
    h->sync();
}

// One tuning knob (none but "mfma" and "fma" changes results, and those
// only within tolerance); ZB_ERR_ARG for an unknown key or value.  "layout"
// resets the state to |0...0>.
int apply_knob(zb_backend *h, const std::string &k, int64_t value) {
    if (k == "prefetch") h->prefetch = value != 0;
    else if (k == "win_grid" && value >= 1 && value <= (1 << 20)) h->win_grid = (unsigned)value;
    else if (k == "slots" && (value == 0 || value == 3 || value == 4)) h->slots = (int)value;
    else if (k == "page_bits" && value >= 1 && value <= 12) h->page_bits = (uint32_t)value;
    else if (k == "tile_bits" && (value == 0 || value == 12 || value == 13)) h->tile_bits = (uint32_t)value;
    else if (k == "defer_cx") h->defer_cx = value != 0;
    else if (k == "mfma") h->mfma = value != 0;
    else if (k == "kq_tiled" && value >= 0 && value <= 2) h->kq_tiled = (int)value;
    else if (k == "fma") h->fma = value != 0;
    else if (k == "relabel") h->relabel = value != 0;
    else if (k == "chunk_multi" && value >= 0) h->chunk_multi = (uint64_t)value;
    else if (k == "layout" && value >= -1 && value <= (1 << 30)) {
        h->sync();
        set_layout(h, (int)value);
        zero_state(h);
    } else return ZB_ERR_ARG;
    return ZB_OK;
}

}  // namespace

int zb_create(uint32_t n_qubits, int precision, zb_backend **out) {
    *out = nullptr;
    if (const int rc = check_shape(n_qubits, precision); rc != ZB_OK) return rc;
    zb_backend *h = new zb_backend();
    set_shape(h, n_qubits, precision);
    const int rc = guarded([&] {
        size_t free_b = 0, total_b = 0;
        check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        const size_t need = 2 * h->len * h->elem;
        if (need + kArenaBytes + (64ull << 20) > free_b)
            throw std::runtime_error("state vector for " + std::to_string(n_qubits) + " qubits at f" +
                                     std::to_string(precision) + " needs " + std::to_string(need) +
                                     " B but only " + std::to_string(free_b) + " B of device memory are free");
        create_resources(h);
        check(hipMalloc(&h->re, h->len * h->elem), "hipMalloc re");
        check(hipMalloc(&h->im, h->len * h->elem), "hipMalloc im");
        set_layout(h, (int)default_layout(h));
        apply_defaults(h);
        zero_state(h);
    });
    if (rc != ZB_OK) {
        zb_destroy(h);
        return std::string(last_error()).find("needs") != std::string::npos ? ZB_ERR_MEMORY : rc;
    }
    *out = h;
    return ZB_OK;
}

int zb_create_bound(uint32_t n_qubits, int precision, void *re, void *im, zb_backend **out) {
    *out = nullptr;
    if (const int rc = check_shape(n_qubits, precision); rc != ZB_OK) return rc;
    if (!re || !im) {
        set_error("zb_create_bound: state pointers must be non-null device memory");
        return ZB_ERR_ARG;
    }
    zb_backend *h = new zb_backend();
    set_shape(h, n_qubits, precision);
    h->owns_state = false;
    h->re = re;
    h->im = im;
    const int rc = guarded([&] {
        create_resources(h);
        set_layout(h, (int)default_layout(h));
        apply_defaults(h);
        zero_state(h);
    });
    if (rc != ZB_OK) {
        zb_destroy(h);
        return rc;
    }
    *out = h;
    return ZB_OK;
}

int zb_rebind(zb_backend *h, uint32_t n_qubits, int precision, void *re, void *im) {
    if (const int rc = check_shape(n_qubits, precision); rc != ZB_OK) return rc;
    if (!re || !im) {
        set_error("zb_rebind: state pointers must be non-null device memory");
        return ZB_ERR_ARG;
    }
    return guarded([&] {
        h->sync();  // work queued on the old state completes first
        discard_pending(h);
        if (h->owns_state) {
            if (h->re) check(hipFree(h->re), "hipFree re");
            if (h->im) check(hipFree(h->im), "hipFree im");
        }
        h->owns_state = false;
        h->re = re;
        h->im = im;
        set_shape(h, n_qubits, precision);
        reset_knobs(h);
        set_layout(h, (int)default_layout(h));
        apply_defaults(h);
        zero_state(h);
    });
}

void zb_destroy(zb_backend *h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->owns_state && h->re) (void)hipFree(h->re);
    if (h->owns_state && h->im) (void)hipFree(h->im);
    if (h->arena_dev) (void)hipFree(h->arena_dev);
    if (h->arena_host) (void)hipHostFree(h->arena_host);
    if (h->red_dev) (void)hipFree(h->red_dev);
    if (h->red_host) (void)hipHostFree(h->red_host);
    if (h->scratch) (void)hipFree(h->scratch);
    if (h->lut_dev) (void)hipFree(h->lut_dev);
    for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char *zb_name(const zb_backend *h) { return h->name.c_str(); }
uint32_t zb_n_qubits(const zb_backend *h) { return h->n; }
int zb_precision(const zb_backend *h) { return h->prec; }
void *zb_stream(zb_backend *h) { return (void *)h->stream; }

int zb_set_mode(zb_backend *h, int mode) {
    if (mode != 0 && mode != 1) {
        set_error("mode must be 0 (exact) or 1 (tiled)");
        return ZB_ERR_ARG;
    }
    const int rc = guarded([&] { flush(h); });
    if (rc != ZB_OK) return rc;
    h->mode = mode;
    return ZB_OK;
}

int zb_apply_unitary(zb_backend *h, const uint32_t *qs, uint32_t k, const double *mat) {
    if (k < 1 || k > (uint32_t)dev::kMaxFusedK || !sorted_distinct(qs, k, h->n)) {
        set_error("apply_unitary: qubits must be sorted, distinct, < n, and 1 <= k <= 6");
        return ZB_ERR_ARG;
    }
    return guarded([&] {
        if (h->prec == 64) apply_unitary_T<double>(h, qs, k, mat);
        else apply_unitary_T<float>(h, qs, k, mat);
    });
}

int zb_apply_diagonal(zb_backend *h, const uint32_t *qs, uint32_t k, const double *diag) {
    if (k < 1 || k > (uint32_t)dev::kMaxDiagK || !sorted_distinct(qs, k, h->n)) {
        set_error("apply_diagonal: qubits must be sorted, distinct, < n, and 1 <= k <= 12");
        return ZB_ERR_ARG;
    }
    return guarded([&] {
        if (h->prec == 64) apply_diag_T<double>(h, qs, k, diag);
        else apply_diag_T<float>(h, qs, k, diag);
    });
}

int zb_apply_cx(zb_backend *h, uint32_t control, uint32_t target) {
    if (control >= h->n || target >= h->n || control == target) {
        set_error("apply_cx: control and target must be distinct qubits < n");
        return ZB_ERR_ARG;
    }
    return guarded([&] {
        if (h->prec == 64) apply_cx_T<double>(h, control, target);
        else apply_cx_T<float>(h, control, target);
    });
}

double zb_prob_one(zb_backend *h, uint32_t q) {
    if (q >= h->n) {
        set_error("prob_one: qubit out of range");
        return std::numeric_limits<double>::quiet_NaN();
    }
    double r = std::numeric_limits<double>::quiet_NaN();
    guarded([&] { r = h->prec == 64 ? reduce_T<double>(h, 1, q) : reduce_T<float>(h, 1, q); });
    return r;
}

int zb_measure(zb_backend *h, uint32_t q, double u) {
    if (q >= h->n) {
        set_error("measure: qubit out of range");
        return ZB_ERR_ARG;
    }
    int out = 0;
    const int rc = guarded([&] { out = h->prec == 64 ? measure_T<double>(h, q, u) : measure_T<float>(h, q, u); });
    return rc == ZB_OK ? out : rc;
}

int zb_reset_all(zb_backend *h) {
    return guarded([&] {
        discard_pending(h);  // pending ops act on a state that is being overwritten
        if (h->prec == 64) launch_zero<double>(h);
        else launch_zero<float>(h);
        h->sync();
    });
}

int zb_sample(zb_backend *h, uint64_t shots, uint64_t seed, uint64_t *out) {
    return guarded([&] {
        if (h->prec == 64) sample_T<double>(h, shots, seed, out);
        else sample_T<float>(h, shots, seed, out);
    });
}

int zb_sample_noalloc(zb_backend *h, uint64_t shots, uint64_t seed, uint64_t *out) {
    const size_t need = sample_scratch_bytes(h->len, shots);
    if (need > kNoallocScratchBytes) {
        set_error("zb_sample_noalloc: " + std::to_string(shots) + " shots from " + std::to_string(h->n) +
                  " qubits need " + std::to_string(need) + " B of scratch, more than its " +
                  std::to_string(kNoallocScratchBytes) + " B (use zb_sample)");
        return ZB_ERR_ARG;
    }
    char *p = static_cast<char *>(noalloc_ptr(h->device));
    if (!p) {
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        set_error("zb_sample_noalloc: no scratch address for device " + std::to_string(h->device) + ": " +
                  (h->device >= 0 && h->device < kMaxDevices ? g_noalloc_err[h->device] : std::string("out of range")));
        return ZB_ERR_DEVICE;
    }
    return guarded([&] {
        if (h->prec == 64) sample_T<double>(h, shots, seed, out, p);
        else sample_T<float>(h, shots, seed, out, p);
    });
}

int zb_chunk_sums(zb_backend *h, double *out) {
    return guarded([&] {
        flush(h);
        refresh_layout(h);
        const uint64_t chunk = std::min<uint64_t>(h->len, 1ull << 16);
        const uint64_t nch = (h->len + chunk - 1) / chunk;
        double *d = (double *)h->scratch_get(nch * 8);
        if (h->prec == 64) chunk_sums_T<double>(h, d, chunk, nch);
        else chunk_sums_T<float>(h, d, chunk, nch);
        check(hipMemcpyAsync(out, d, nch * 8, hipMemcpyDeviceToHost, h->stream), "copy");
        h->sync();
    });
}

int zb_state_raw(zb_backend *h, void *re, void *im) {
    return guarded([&] {
        if (h->prec == 64) read_state_T<double>(h, (double *)re, (double *)im);
        else read_state_T<float>(h, (float *)re, (float *)im);
    });
}

int zb_state_set_raw(zb_backend *h, const void *re, const void *im) {
    return guarded([&] {
        if (h->prec == 64) write_state_T<double>(h, (const double *)re, (const double *)im);
        else write_state_T<float>(h, (const float *)re, (const float *)im);
    });
}

int zb_layout(const zb_backend *h, uint32_t *phys) {
    for (uint32_t q = 0; q < h->n; ++q) phys[q] = h->phys[q];
    return ZB_OK;
}

int zb_state_device_ptrs(zb_backend *h, void **re, void **im) {
    const int rc = guarded([&] {
        flush(h);
        h->sync();
    });
    if (rc != ZB_OK) return rc;
    *re = h->re;
    *im = h->im;
    return ZB_OK;
}

int zb_statevector(zb_backend *h, double *out) {
    return guarded([&] {
        const uint64_t n = h->len;
        if (h->prec == 64) {
            std::vector<double> re(n), im(n);
            read_state_T<double>(h, re.data(), im.data());
            for (uint64_t i = 0; i < n; ++i) out[2 * i] = re[i], out[2 * i + 1] = im[i];
        } else {
            std::vector<float> re(n), im(n);
            read_state_T<float>(h, re.data(), im.data());
            for (uint64_t i = 0; i < n; ++i) out[2 * i] = (double)re[i], out[2 * i + 1] = (double)im[i];
        }
    });
}

double zb_norm_sqr(zb_backend *h) {
    double r = std::numeric_limits<double>::quiet_NaN();
    guarded([&] { r = h->prec == 64 ? reduce_T<double>(h, 0, 0) : reduce_T<float>(h, 0, 0); });
    return r;
}

int zb_finish(zb_backend *h) {
    return guarded([&] {
        flush(h);
        h->sync();
    });
}

int zb_device_count(void) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) return 0;
    return ndev;
}

int zb_mem_info(uint64_t *free_bytes, uint64_t *total_bytes) {
    if (zb_device_count() == 0) {
        set_error("no HIP device is available (the hip backend needs an AMD GPU)");
        return ZB_ERR_UNSUPPORTED;
    }
    return guarded([&] {
        size_t f = 0, t = 0;
        check(hipMemGetInfo(&f, &t), "hipMemGetInfo");
        *free_bytes = f;
        *total_bytes = t;
    });
}

int zb_calibration_copy(uint64_t n_doubles, int reps, double *ms) {
    if (zb_device_count() == 0) {
        set_error("no HIP device is available (the hip backend needs an AMD GPU)");
        return ZB_ERR_UNSUPPORTED;
    }
    double *src = nullptr, *dst = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const int rc = guarded([&] {
        check(hipMalloc(&src, n_doubles * 8), "hipMalloc");
        check(hipMalloc(&dst, n_doubles * 8), "hipMalloc");
        check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
        check(hipEventCreate(&e0), "hipEventCreate");
        check(hipEventCreate(&e1), "hipEventCreate");
        const unsigned grid = (unsigned)std::min<uint64_t>((n_doubles + dev::kBlock - 1) / dev::kBlock, 1u << 16);
        hipLaunchKernelGGL(dev::k_fill, dim3(grid), dim3(dev::kBlock), 0, st, src, n_doubles, 1.0);
        check(hipEventRecord(e0, st), "hipEventRecord");
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(dev::k_copy, dim3(grid), dim3(dev::kBlock), 0, st, dst, src, n_doubles);
        check(hipGetLastError(), "k_copy");
        check(hipEventRecord(e1, st), "hipEventRecord");
        check(hipStreamSynchronize(st), "hipStreamSynchronize");
        float f = 0.f;
        check(hipEventElapsedTime(&f, e0, e1), "hipEventElapsedTime");
        if (ms) *ms = reps > 0 ? (double)f / reps : 0.0;
    });
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    return rc;
}

int zb_copy_probe(uint64_t n_doubles, int reps, int kind, double *ms) {
    if (kind < 0 || kind > 7) {
        set_error("zb_copy_probe: kind must be 0..7");
        return ZB_ERR_ARG;
    }
    if (kind == 0) return zb_calibration_copy(n_doubles, reps, ms);
    const bool hashed = kind >= 4;  // 4..7: kinds 0..3 on hashed, not constant, data
    if (hashed) kind -= 4;
    if (zb_device_count() == 0) {
        set_error("no HIP device is available (the hip backend needs an AMD GPU)");
        return ZB_ERR_UNSUPPORTED;
    }
    n_doubles &= ~uint64_t(1);  // whole 16-byte elements
    double *src = nullptr, *dst = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const int rc = guarded([&] {
        check(hipMalloc(&src, n_doubles * 8), "hipMalloc");
        check(hipMalloc(&dst, n_doubles * 8), "hipMalloc");
        check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
        check(hipEventCreate(&e0), "hipEventCreate");
        check(hipEventCreate(&e1), "hipEventCreate");
        const unsigned fill_grid = (unsigned)std::min<uint64_t>((n_doubles + dev::kBlock - 1) / dev::kBlock, 1u << 16);
        if (hashed) hipLaunchKernelGGL(dev::k_fill_hash, dim3(fill_grid), dim3(dev::kBlock), 0, st, src, n_doubles);
        else hipLaunchKernelGGL(dev::k_fill, dim3(fill_grid), dim3(dev::kBlock), 0, st, src, n_doubles, 1.0);
        const uint64_t n2 = n_doubles / 2;
        const int u = kind == 1 ? 1 : 4;
        const unsigned grid =
            (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n2 / u + dev::kBlock - 1) / dev::kBlock, 1u << 16));
        auto launch = [&] {
            double2 *d = reinterpret_cast<double2 *>(dst);
            const double2 *sv = reinterpret_cast<const double2 *>(src);
            if (kind == 0)
                hipLaunchKernelGGL(dev::k_copy, dim3(fill_grid), dim3(dev::kBlock), 0, st, dst, src, n_doubles);
            else if (kind == 1) hipLaunchKernelGGL((dev::k_copy_wide<1, false>), dim3(grid), dim3(dev::kBlock), 0, st, d, sv, n2);
            else if (kind == 2) hipLaunchKernelGGL((dev::k_copy_wide<4, false>), dim3(grid), dim3(dev::kBlock), 0, st, d, sv, n2);
            else hipLaunchKernelGGL((dev::k_copy_wide<4, true>), dim3(grid), dim3(dev::kBlock), 0, st, d, sv, n2);
        };
        launch();  // warm
        check(hipEventRecord(e0, st), "hipEventRecord");
        for (int r = 0; r < reps; ++r) launch();
        check(hipGetLastError(), "k_copy_wide");
        check(hipEventRecord(e1, st), "hipEventRecord");
        check(hipStreamSynchronize(st), "hipStreamSynchronize");
        float f = 0.f;
        check(hipEventElapsedTime(&f, e0, e1), "hipEventElapsedTime");
        if (ms) *ms = reps > 0 ? (double)f / reps : 0.0;
    });
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    return rc;
}

int zb_tune(zb_backend *h, const char *key, int64_t value) {
    const std::string k(key ? key : "");
    return guarded([&] {
        flush(h);
        if (apply_knob(h, k, value) != ZB_OK) throw std::runtime_error("zb_tune: unknown key or bad value: " + k);
    });
}

int zb_tune_default(uint32_t n_qubits, int precision, const char *key, int64_t value) {
    static const char *keys[] = {"prefetch", "win_grid", "slots", "page_bits", "tile_bits",
                                 "defer_cx", "layout",   "mfma",  "fma",       "relabel", "chunk_multi", "kq_tiled"};
    const std::string k(key ? key : "");
    if (std::find(std::begin(keys), std::end(keys), k) == std::end(keys) || (precision != 32 && precision != 64)) {
        set_error("zb_tune_default: unknown key or precision: " + k);
        return ZB_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (value == INT64_MIN) g_tune_defaults.erase({n_qubits, precision, k});
    else g_tune_defaults[{n_qubits, precision, k}] = value;
    return ZB_OK;
}

int zb_profile_begin(zb_backend *h) {
    return guarded([&] {
        flush(h);
        h->sync();
        h->recs.clear();
        h->ev_next = 0;
        h->prof = true;
    });
}

int64_t zb_profile_detail(const zb_backend *h, int32_t *detail, int64_t cap) {
    int64_t n = 0;
    for (const auto &r : h->last_recs) {
        if (n >= cap) break;
        for (int i = 0; i < 4; ++i) detail[4 * n + i] = r.detail[i];
        ++n;
    }
    return n;
}

int64_t zb_profile_end(zb_backend *h, int32_t *kernel, double *ms, double *bytes, int64_t cap) {
    int64_t n = 0;
    const int rc = guarded([&] {
        flush(h);
        h->sync();
        h->prof = false;
        for (const auto &r : h->recs) {
            if (n >= cap) break;
            float t = 0.f;
            check(hipEventElapsedTime(&t, r.a, r.b), "hipEventElapsedTime");
            kernel[n] = r.kernel;
            ms[n] = (double)t;
            bytes[n] = r.bytes;
            ++n;
        }
        h->last_recs.assign(h->recs.begin(), h->recs.begin() + n);
        h->recs.clear();
        h->ev_next = 0;
    });
    return rc == ZB_OK ? n : rc;
}

int zb_max_abs_diff(zb_backend *a, zb_backend *b, double *out) {
    if (a->n != b->n) {
        set_error("max_abs_diff: states have different qubit counts");
        return ZB_ERR_ARG;
    }
    return guarded([&] {
        flush(a);
        flush(b);
        refresh_layout(a);
        refresh_layout(b);
        b->sync();
        unsigned long long *d = reinterpret_cast<unsigned long long *>(a->scratch_get(8));
        check(hipMemsetAsync(d, 0, 8, a->stream), "hipMemsetAsync");
        const dim3 g(grid_for(a->len)), blk(dev::kBlock);
        if (a->prec == 64 && b->prec == 64)
            hipLaunchKernelGGL((dev::k_max_abs_diff<double, double>), g, blk, 0, a->stream, a->R<double>(),
                               a->I<double>(), a->layout(), b->R<double>(), b->I<double>(), b->layout(), a->len, d);
        else if (a->prec == 64)
            hipLaunchKernelGGL((dev::k_max_abs_diff<double, float>), g, blk, 0, a->stream, a->R<double>(),
                               a->I<double>(), a->layout(), b->R<float>(), b->I<float>(), b->layout(), a->len, d);
        else if (b->prec == 64)
            hipLaunchKernelGGL((dev::k_max_abs_diff<float, double>), g, blk, 0, a->stream, a->R<float>(),
                               a->I<float>(), a->layout(), b->R<double>(), b->I<double>(), b->layout(), a->len, d);
        else
            hipLaunchKernelGGL((dev::k_max_abs_diff<float, float>), g, blk, 0, a->stream, a->R<float>(),
                               a->I<float>(), a->layout(), b->R<float>(), b->I<float>(), b->layout(), a->len, d);
        check(hipGetLastError(), "k_max_abs_diff");
        check(hipMemcpyAsync(a->red_host, d, 8, hipMemcpyDeviceToHost, a->stream), "copy");
        a->sync();
        *out = *a->red_host;
    });
}

}  // extern "C"

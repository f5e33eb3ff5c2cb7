// executor.cpp — the host side above the Backend boundary: compile() and
// run_compiled() (src/compiler.rs:118-200, src/exec.rs:334-619,
// src/noise.rs:91-192) driving the HIP backend only through the zb_* C ABI,
// exactly as the reference's executor drives `dyn Backend`.
#include <algorithm>
#include <iterator>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>

#include "host.hpp"
#include "zeno_hip.h"

using namespace zeno;

namespace {

struct BackendDeleter {
    void operator()(zb_backend *h) const { zb_destroy(h); }
};
using BackendPtr = std::unique_ptr<zb_backend, BackendDeleter>;

struct RunError {
    int code;
};

void ok(int rc) {
    if (rc < 0) throw RunError{rc};
}

const double kX[8] = {0, 0, 1, 0, 1, 0, 0, 0};
const double kY[8] = {0, 0, 0, -1, 0, 1, 0, 0};
const double kZ[4] = {1, 0, -1, 0};

// noise.rs:91-100
void apply_pauli(zb_backend *h, uint32_t q, uint64_t code) {
    if (code == 1) ok(zb_apply_unitary(h, &q, 1, kX));
    else if (code == 2) ok(zb_apply_unitary(h, &q, 1, kY));
    else if (code == 3) ok(zb_apply_diagonal(h, &q, 1, kZ));
}

int measure(zb_backend *h, uint32_t q, double u) {
    const int r = zb_measure(h, q, u);
    ok(r);
    return r;
}

// noise.rs:114-127: jump = collapse to 1 then X; no jump = the renormalised
// diagonal diag(1, sqrt(1-g)) / sqrt(1 - P(jump)) in one sweep.
void amplitude_damp(zb_backend *h, uint32_t q, double gamma, double u) {
    const double p1 = zb_prob_one(h, q);
    if (std::isnan(p1)) throw RunError{ZB_ERR_DEVICE};
    const double pj = gamma * p1;
    if (u < pj) {
        measure(h, q, 0.0);
        apply_pauli(h, q, 1);
    } else {
        const double s = 1.0 / std::sqrt(1.0 - pj);
        const double d[4] = {s, 0.0, s * std::sqrt(1.0 - gamma), 0.0};
        ok(zb_apply_diagonal(h, &q, 1, d));
    }
}

// noise.rs:145-175, frozen draw order.
void gate_noise(zb_backend *h, const zr_noise &m, const std::vector<uint32_t> &qs, Rng &rng) {
    const int k = (int)qs.size();
    const double pd = k == 1 ? m.depolarizing_1q : m.depolarizing_2q;
    if (pd > 0.0 && rng.f64() < pd) {
        const uint64_t idx = rng.range(1, 1ull << (2 * k));
        for (int b = 0; b < k; ++b) apply_pauli(h, qs[b], (idx >> (2 * b)) & 3);
    }
    for (uint32_t q : qs) {
        if (m.bit_flip > 0.0 && rng.f64() < m.bit_flip) apply_pauli(h, q, 1);
        if (m.phase_flip > 0.0 && rng.f64() < m.phase_flip) apply_pauli(h, q, 3);
        if (m.amplitude_damping > 0.0) {
            const double u = rng.f64();
            amplitude_damp(h, q, m.amplitude_damping, u);
        }
    }
}

// noise.rs:181-192
int readout(const zr_noise &m, int bit, Rng &rng) {
    const double p = bit ? m.readout_flip_1to0 : m.readout_flip_0to1;
    if (p > 0.0 && rng.f64() < p) return !bit;
    return bit;
}

void apply_op(zb_backend *h, const Op &o) {  // exec.rs:398-405
    if (o.kind == kUnitary)
        ok(zb_apply_unitary(h, o.qs.data(), (uint32_t)o.qs.size(), reinterpret_cast<const double *>(o.data.data())));
    else if (o.kind == kDiagonal)
        ok(zb_apply_diagonal(h, o.qs.data(), (uint32_t)o.qs.size(), reinterpret_cast<const double *>(o.data.data())));
    else if (o.kind == kCx)
        ok(zb_apply_cx(h, o.a, o.b));
}

// Applies ops [lo, hi) in order, with the backend told which circuit
// qubits each coming op needs in a tile (its window scheduling uses it).
void apply_ops_ahead(zb_backend *h, const std::vector<Op> &ops, size_t lo, size_t hi) {
    std::vector<uint64_t> need(hi > lo ? hi - lo : 0, 0);
    for (size_t i = lo; i < hi; ++i) {
        const Op &o = ops[i];
        if (o.kind == kUnitary)
            for (uint32_t q : o.qs) need[i - lo] |= 1ull << q;
        else if (o.kind == kCx)
            need[i - lo] = 1ull << o.b;
    }
    backend_lookahead(h, need.data(), need.size());
    try {
        for (size_t i = lo; i < hi; ++i) {
            backend_lookahead_pos(h, i - lo);
            apply_op(h, ops[i]);
        }
        ok(backend_end_run(h));
    } catch (...) {
        backend_lookahead(h, nullptr, 0);
        throw;
    }
    backend_lookahead(h, nullptr, 0);
}

// exec.rs:556-619
void exec_shot_op(zb_backend *h, const Op &o, const zr_noise *nz, Rng &rng, uint64_t &clbits) {
    if (o.is_if) {
        const uint64_t mask = o.creg_len >= 64 ? ~0ull : ((1ull << o.creg_len) - 1);
        if (((clbits >> o.creg_off) & mask) != o.value) return;
    }
    switch (o.kind) {
        case kUnitary:
        case kDiagonal:
            apply_op(h, o);
            if (nz) gate_noise(h, *nz, o.qs, rng);
            break;
        case kCx:
            apply_op(h, o);
            if (nz) gate_noise(h, *nz, {std::min(o.a, o.b), std::max(o.a, o.b)}, rng);
            break;
        case kMeasure: {
            int bit = measure(h, o.a, rng.f64());
            if (nz) bit = readout(*nz, bit, rng);
            clbits = (clbits & ~(1ull << o.b)) | ((uint64_t)bit << o.b);
            break;
        }
        case kReset:
            if (measure(h, o.a, rng.f64())) ok(zb_apply_unitary(h, &o.a, 1, kX));
            break;
        default:
            break;
    }
}

bool trivial(const zr_noise *m) {
    return !m || (m->depolarizing_1q == 0.0 && m->depolarizing_2q == 0.0 && m->bit_flip == 0.0 &&
                  m->phase_flip == 0.0 && m->amplitude_damping == 0.0 && m->readout_flip_0to1 == 0.0 &&
                  m->readout_flip_1to0 == 0.0);
}

// The sampled path's host buffers (the draws' indices and the counts) are
// kept per thread across runs: a fresh multi-megabyte vector costs a page
// fault per 4 KiB on first touch, about 10 ms for 1M shots, more than the
// counting itself.  A buffer past kKeepBytes is released after its run.
constexpr size_t kKeepBytes = 64ull << 20;

template <typename V>
void trim(V &v) {
    if (v.capacity() * sizeof(typename V::value_type) > kKeepBytes) V().swap(v);
}

bool valid(const zr_noise *m) {  // noise.rs:78-87
    if (!m) return true;
    for (double v : {m->depolarizing_1q, m->depolarizing_2q, m->bit_flip, m->phase_flip, m->amplitude_damping,
                     m->readout_flip_0to1, m->readout_flip_1to0})
        if (!(v >= 0.0 && v <= 1.0)) return false;
    return true;
}

}  // namespace

namespace zeno {

int run_compiled_on(zb_backend *hb, const zr_compiled *cc, uint64_t shots, uint64_t seed, const zr_noise *noise,
                    int mode, bool state_is_zero, std::vector<std::pair<uint64_t, uint64_t>> &out,
                    TrajWorkspace *traj) {
    const Compiled &c = cc->c;
    if (!valid(noise)) {
        set_error("invalid noise model: every probability must be in [0, 1]");
        return ZB_ERR_ARG;
    }
    if (zb_n_qubits(hb) != c.n_qubits) {
        set_error("backend and circuit qubit counts differ");
        return ZB_ERR_ARG;
    }
    const int precision = zb_precision(hb);
    const zr_noise *nz = trivial(noise) ? nullptr : noise;
    std::unordered_map<uint64_t, uint64_t> raw;
    // run_sampled's counts, ascending by word (a per-thread buffer: see kKeepBytes)
    static thread_local std::vector<std::pair<uint64_t, uint64_t>> sampled;
    trim(sampled);  // an earlier run that failed may have left a large buffer
    sampled.clear();
    try {
        zb_backend *h = hb;
        ok(zb_set_mode(h, mode == 2 ? 1 : mode));  // 2: tiled, per-shot loop on the host
        if (!state_is_zero) ok(zb_reset_all(h));
        if (!nz && !c.dynamic) {  // run_sampled (exec.rs:407-452)
            apply_ops_ahead(h, c.ops, 0, c.ops.size());
            ok(zb_finish(h));
            if (!c.final_measures.empty() && shots > 0) {
                static thread_local std::vector<uint64_t> smp_buf;
                trim(smp_buf);
                if (smp_buf.size() < shots) smp_buf.resize(shots);
                const uint64_t *smp = smp_buf.data();
                ok(zb_sample(h, shots, splitmix64(seed), smp_buf.data()));
                // Counts per basis index (exec.rs:437-441): the samples come
                // back ascending, so equal indices are runs.  Then each unique
                // index's clbit word (exec.rs:442-447); when the words come
                // out strictly ascending (every qubit measured, clbits in
                // qubit order) they are the result as they are, else they are
                // sorted and equal words merged.
                // The word is a bit gather (the last measure into a clbit
                // decides it, clbits never measured stay 0), so it is the OR
                // of one table entry per byte of the index.
                int src[64];
                std::fill(std::begin(src), std::end(src), -1);
                for (const auto &fm : c.final_measures) src[fm.second] = (int)fm.first;
                const uint32_t nbytes = (c.n_qubits + 7) / 8;
                std::vector<uint64_t> lut((size_t)nbytes * 256, 0);
                for (int cb = 0; cb < 64; ++cb)
                    if (src[cb] >= 0)
                        for (uint32_t v = 0; v < 256; ++v)
                            if ((v >> (src[cb] & 7)) & 1) lut[(size_t)(src[cb] >> 3) * 256 + v] |= 1ull << cb;
                bool ascending = true;
                sampled.reserve(shots);
                for (size_t i = 0; i < shots;) {
                    size_t j = i + 1;
                    while (j < shots && smp[j] == smp[i]) ++j;
                    uint64_t cl = 0;
                    for (uint32_t b = 0; b < nbytes; ++b) cl |= lut[(size_t)b * 256 + ((smp[i] >> (8 * b)) & 255)];
                    ascending = ascending && (sampled.empty() || sampled.back().first < cl);
                    sampled.emplace_back(cl, (uint64_t)(j - i));
                    i = j;
                }
                if (!ascending) {
                    std::sort(sampled.begin(), sampled.end());
                    size_t m = 0;
                    for (size_t i = 0; i < sampled.size(); ++i) {
                        if (m > 0 && sampled[m - 1].first == sampled[i].first) sampled[m - 1].second += sampled[i].second;
                        else sampled[m++] = sampled[i];
                    }
                    sampled.resize(m);
                }
                trim(smp_buf);
            }
        } else {  // run_per_shot (exec.rs:459-549)
            bool batched = false;
            uint64_t avail = 0;  // the workspace the batched launch would get
            if (traj) {
                avail = traj->bytes;
            } else {
                uint64_t fr = 0, tot = 0;
                if (zb_mem_info(&fr, &tot) == ZB_OK) avail = fr > kTrajMargin ? fr - kTrajMargin : 0;
            }
            if ((c.dynamic || !c.final_measures.empty()) && mode == 1 &&
                traj_preferred(c, precision, shots, avail)) {
                // every shot in one batched trajectory launch; when no
                // workspace fits, the shots run one by one below
                std::vector<uint64_t> words(shots);
                const int rc = traj ? run_trajectories_in(c, precision, shots, seed, nz, words.data(), nullptr, *traj)
                                    : zr_run_trajectories(cc, precision, shots, seed, nz, words.data(), nullptr);
                if (rc != ZB_ERR_MEMORY) ok(rc);
                if (rc == ZB_OK) {
                    batched = true;
                    for (uint64_t w : words) ++raw[w];
                }
            }
            if (!batched && (c.dynamic || !c.final_measures.empty())) {
                for (uint64_t shot = 0; shot < shots; ++shot) {
                    ok(zb_reset_all(h));
                    Rng rng(splitmix64(seed ^ shot));
                    uint64_t cl = 0;
                    for (const Op &o : c.ops) exec_shot_op(h, o, nz, rng, cl);
                    for (const auto &fm : c.final_measures) {
                        int bit = measure(h, fm.first, rng.f64());
                        if (nz) bit = readout(*nz, bit, rng);
                        cl = (cl & ~(1ull << fm.second)) | ((uint64_t)bit << fm.second);
                    }
                    ++raw[cl];
                }
            }
        }
    } catch (const RunError &e) {
        return e.code;
    }
    if (!nz && !c.dynamic) {
        out.swap(sampled);
        trim(sampled);
        return ZB_OK;
    }
    out.assign(raw.begin(), raw.end());
    std::sort(out.begin(), out.end());
    return ZB_OK;
}

}  // namespace zeno

extern "C" {

zr_compiled *zr_compile(const zr_instr *prog, int64_t n_instr, uint32_t n_qubits, const uint32_t *creg_sizes,
                        uint32_t n_cregs, int fusion_max, int diag_max) {
    try {
        auto *c = new zr_compiled();
        c->c = compile(prog, n_instr, n_qubits, creg_sizes, n_cregs, fusion_max, diag_max);
        return c;
    } catch (const std::exception &e) {
        set_error(e.what());
        return nullptr;
    }
}

void zr_compiled_free(zr_compiled *c) { delete c; }

int64_t zr_compiled_nops(const zr_compiled *c) { return (int64_t)c->c.ops.size(); }

int zr_compiled_op(const zr_compiled *c, int64_t i, int32_t *kind, int32_t *k, uint32_t *qs, double *data,
                   int32_t *is_if, uint64_t *cond) {
    if (i < 0 || i >= (int64_t)c->c.ops.size()) return ZB_ERR_ARG;
    const Op &o = c->c.ops[(size_t)i];
    *kind = o.kind;
    *is_if = o.is_if;
    cond[0] = o.creg_off, cond[1] = o.creg_len, cond[2] = o.value;
    if (o.kind == kUnitary || o.kind == kDiagonal) {
        *k = (int32_t)o.qs.size();
        std::copy(o.qs.begin(), o.qs.end(), qs);
        if (data) std::memcpy(data, o.data.data(), o.data.size() * sizeof(cplx));
    } else {
        *k = 2;
        qs[0] = o.a, qs[1] = o.b;
    }
    return ZB_OK;
}

void zr_compiled_stats(const zr_compiled *c, int64_t *s) {
    s[0] = c->c.stats.input_gates;
    s[1] = c->c.stats.cancelled;
    s[2] = c->c.stats.output_ops;
    s[3] = c->c.stats.max_fused;
    s[4] = c->c.dynamic;
    s[5] = (int64_t)c->c.final_measures.size();
}

int64_t zr_compiled_final_measures(const zr_compiled *c, uint32_t *pairs) {
    for (size_t i = 0; i < c->c.final_measures.size(); ++i)
        pairs[2 * i] = c->c.final_measures[i].first, pairs[2 * i + 1] = c->c.final_measures[i].second;
    return (int64_t)c->c.final_measures.size();
}

int zr_apply_ops(zb_backend *h, const zr_compiled *c, int64_t lo, int64_t hi) {
    hi = std::min<int64_t>(hi, (int64_t)c->c.ops.size());
    try {
        const int64_t from = std::max<int64_t>(lo, 0);
        if (from < hi) apply_ops_ahead(h, c->c.ops, (size_t)from, (size_t)hi);
    } catch (const RunError &e) {
        return e.code;
    }
    return ZB_OK;
}

}  // extern "C"

namespace {

// mem::state_bytes (src/mem.rs:79-81), saturated at 2^64 - 1.
uint64_t state_bytes(uint32_t n, int precision) {
    const unsigned __int128 b = ((unsigned __int128)1 << n) * (unsigned __int128)(precision == 64 ? 16 : 8);
    return b > (unsigned __int128)UINT64_MAX ? UINT64_MAX : (uint64_t)b;
}

// mem::max_qubits (src/mem.rs:84-93).
uint32_t max_qubits(uint64_t budget, int precision) {
    uint32_t n = 0;
    while (state_bytes(n + 1, precision) <= budget) {
        ++n;
        if (n >= 60) break;
    }
    return n;
}

// lib.rs human_bytes: B below 1 KiB, else one decimal and the largest unit.
std::string human_bytes(uint64_t b) {
    static const char *units[] = {"B", "KiB", "MiB", "GiB", "TiB"};
    double v = (double)b;
    int u = 0;
    while (v >= 1024.0 && u + 1 < 5) v /= 1024.0, ++u;
    char buf[64];
    if (u == 0) std::snprintf(buf, sizeof buf, "%llu B", (unsigned long long)b);
    else std::snprintf(buf, sizeof buf, "%.1f %s", v, units[u]);
    return buf;
}

}  // namespace

extern "C" {

uint64_t zr_budget_bytes(uint64_t mem_limit, double fraction) {
    if (mem_limit > 0) return mem_limit;
    uint64_t physical = 0;
    if (const char *env = std::getenv("ZENO_MEM_BYTES")) {
        char *end = nullptr;
        const unsigned long long v = std::strtoull(env, &end, 10);
        if (end && end != env && *end == '\0') physical = v;
    }
    if (physical == 0) {
        uint64_t free_b = 0, total_b = 0;
        if (zb_device_count() == 0 || zb_mem_info(&free_b, &total_b) != ZB_OK) return 0;
        physical = total_b;
    }
    return (uint64_t)((double)physical * fraction);
}

int zr_plan(uint32_t n_qubits, int precision, uint64_t budget, zr_mem_plan *plan) {
    if (precision != 32 && precision != 64) {
        set_error("precision must be 32 or 64");
        return ZB_ERR_ARG;
    }
    zr_mem_plan p{};
    p.n_qubits = n_qubits;
    p.precision = precision;
    p.needed_bytes = state_bytes(n_qubits, precision);
    p.budget_bytes = budget;
    p.max_qubits = max_qubits(budget, precision);
    if (plan) *plan = p;
    if (p.needed_bytes > budget) {  // Error::Memory's text (src/lib.rs:82-88)
        set_error("state vector for " + std::to_string(n_qubits) + " qubits at f" + std::to_string(precision) +
                  " needs " + human_bytes(p.needed_bytes) + " but the budget is " + human_bytes(budget) + " (max " +
                  std::to_string(p.max_qubits) +
                  " qubits at this precision; raise --mem-limit or lower qubits)");
        return ZB_ERR_MEMORY;
    }
    return ZB_OK;
}

int zr_resolve_precision(uint32_t n_qubits, int requested, uint64_t budget, int *precision, zr_mem_plan *plan,
                         char *notice, size_t notice_cap) {
    if (notice && notice_cap) notice[0] = '\0';
    if (requested != 0 && requested != 32 && requested != 64) {
        set_error("requested precision must be 0 (auto), 32 or 64");
        return ZB_ERR_ARG;
    }
    int chosen = requested;
    if (requested == 0) {
        chosen = zr_plan(n_qubits, 64, budget, nullptr) == ZB_OK ? 64 : 32;
        if (chosen == 32 && zr_plan(n_qubits, 32, budget, nullptr) == ZB_OK && notice && notice_cap) {
            const std::string t = "auto-precision: " + std::to_string(n_qubits) +
                                  " qubits exceed the f64 budget, using f32 (8 B/amplitude)";
            std::snprintf(notice, notice_cap, "%s", t.c_str());
        }
    }
    if (precision) *precision = chosen;
    return zr_plan(n_qubits, chosen, budget, plan);
}

int64_t zr_run(const zr_compiled *cc, int precision, uint64_t shots, uint64_t seed, const zr_noise *noise, int mode,
               uint64_t *keys, uint64_t *counts, int64_t cap, double *sim_seconds) {
    if (!valid(noise)) {
        set_error("invalid noise model: every probability must be in [0, 1]");
        return ZB_ERR_ARG;
    }
    static thread_local std::vector<std::pair<uint64_t, uint64_t>> v;  // see kKeepBytes
    trim(v);
    v.clear();
    const auto t0 = std::chrono::steady_clock::now();
    zb_backend *hp = nullptr;
    if (const int rc = zb_create(cc->c.n_qubits, precision, &hp); rc < 0) return rc;
    BackendPtr h(hp);
    if (const int rc = run_compiled_on(h.get(), cc, shots, seed, noise, mode, true, v); rc < 0) return rc;
    if (sim_seconds) *sim_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if ((int64_t)v.size() > cap) {
        set_error("key buffer too small");
        return ZB_ERR_ARG;
    }
    for (size_t i = 0; i < v.size(); ++i) keys[i] = v[i].first, counts[i] = v[i].second;
    const int64_t m = (int64_t)v.size();
    trim(v);
    return m;
}

}  // extern "C"

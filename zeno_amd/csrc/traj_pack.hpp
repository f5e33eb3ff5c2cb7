// traj_pack.hpp — a compiled circuit's ops in the batched trajectory
// kernel's format (host side; trajectories.hpp holds the kernel).
#pragma once

#include <vector>

#include "host.hpp"
#include "trajectories.hpp"

namespace zeno {

// Circuits the batched kernel runs: up to kTrajMaxQubits qubits, dense ops
// on up to 6 qubits, diagonals on up to 12.
inline bool traj_fits(const Compiled &c) {
    if (c.n_qubits > (uint32_t)dev::kTrajMaxQubits) return false;
    for (const Op &o : c.ops) {
        if (o.kind == kUnitary && o.qs.size() > (size_t)dev::kTrajWideDense) return false;
        if (o.kind == kDiagonal && o.qs.size() > (size_t)dev::kTrajOpQubits) return false;
    }
    return true;
}

// Whether the circuit needs the wide kernel instance (an op on more than 3
// qubits).
inline bool traj_wide(const Compiled &c) {
    for (const Op &o : c.ops)
        if ((o.kind == kUnitary || o.kind == kDiagonal) && o.qs.size() > (size_t)dev::kTrajMaxK) return true;
    return false;
}

// Bytes of one workgroup's slot: re and im of 2^n values of `elem` bytes,
// then, when the probability sums have more than 256 blocks, the block
// sums and the next level's totals (f64), rounded up to 256 bytes.
inline uint64_t traj_slot_bytes(uint32_t n, uint64_t elem) {
    const uint64_t len = 1ull << n, nb = (len >> 1) / dev::kReduceSpan;
    uint64_t b = 2 * len * elem;
    if (nb > (uint64_t)dev::kTrajLdsSumBlocks) b += (nb + (nb + dev::kReduceSpan - 1) / dev::kReduceSpan + 1) * 8;
    return (b + 255) & ~uint64_t(255);
}

// The op tables of one run in the kernel's format: ops, matrices in T,
// final-measure pairs.
template <typename T>
struct Tables {
    std::vector<dev::TrajOp> ops;
    std::vector<T> mats;
    std::vector<uint32_t> fm;
    size_t nops = 0, nfm = 0;
};

template <typename T>
inline Tables<T> pack(const Compiled &c) {
    Tables<T> t;
    for (const Op &o : c.ops) {
        if (o.kind == kBarrier) continue;
        dev::TrajOp d{};
        d.is_if = o.is_if;
        d.creg_off = o.creg_off;
        d.creg_len = o.creg_len;
        d.value = o.value;
        if (o.kind == kUnitary || o.kind == kDiagonal) {
            d.kind = o.kind == kUnitary ? 0 : 1;
            d.k = (uint32_t)o.qs.size();
            for (size_t b = 0; b < o.qs.size(); ++b) d.q[b] = o.qs[b];
            d.data_off = (uint32_t)t.mats.size();
            for (const cplx &z : o.data) t.mats.push_back((T)z.re);  // T::fr
            for (const cplx &z : o.data) t.mats.push_back((T)z.im);
        } else if (o.kind == kCx) {
            d.kind = 2;
            d.q[0] = o.a, d.q[1] = o.b;
        } else if (o.kind == kMeasure) {
            d.kind = 3;
            d.q[0] = o.a;
            d.clbit = o.b;
        } else if (o.kind == kReset) {
            d.kind = 4;
            d.q[0] = o.a;
        }
        t.ops.push_back(d);
    }
    for (const auto &f : c.final_measures) t.fm.push_back(f.first), t.fm.push_back(f.second);
    t.nops = t.ops.size();
    t.nfm = c.final_measures.size();
    // keep every region non-empty; the kernel reads only nops / nfm entries
    if (t.mats.empty()) t.mats.push_back(T(0));
    if (t.ops.empty()) t.ops.push_back(dev::TrajOp{});
    if (t.fm.empty()) t.fm.push_back(0);
    return t;
}

}  // namespace zeno

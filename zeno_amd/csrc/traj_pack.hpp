// traj_pack.hpp — a compiled circuit's ops in the batched trajectory
// kernel's format (host side; trajectories.hpp holds the kernel).
#pragma once

#include <vector>

#include "host.hpp"
#include "trajectories.hpp"

namespace zeno {

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

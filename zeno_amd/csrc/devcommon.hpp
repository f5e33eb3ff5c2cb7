// devcommon.hpp — the few definitions the state kernels (kernels.hpp,
// kq_tiles.hpp) and the batched trajectory kernel (trajectories.hpp) share:
// the workgroup size of the sweeps and fixed-order reductions, the reduction
// block, the insertion of a zero bit into an index, and the sweeps'
// multiply-accumulate steps.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace zeno {
namespace dev {

constexpr int kBlock = 256;        // 4 wave64 per workgroup
constexpr int kReduceSpan = 4096;  // elements per reduction block (16 per lane)

__device__ __forceinline__ uint64_t insert_zero(uint64_t x, uint32_t q) {
    return ((x >> q) << (q + 1)) | (x & ((1ull << q) - 1));
}

// Multiply-accumulate steps of the sweep arithmetic, in the CPU path's
// expression order.  F = false is the CPU path's own rounding: a separately
// rounded product and sum (the library is built with -ffp-contract=off, so
// acc + a*b is never fused).  F = true (knob "fma") rounds once per step
// with a fused multiply-add: fewer VALU operations, not bit-identical.
template <bool F, typename T>
__device__ __forceinline__ T mac(T acc, T a, T b) {
    if constexpr (!F) return acc + a * b;
    else if constexpr (sizeof(T) == 4) return __builtin_fmaf(a, b, acc);
    else return __builtin_fma(a, b, acc);
}
template <bool F, typename T>
__device__ __forceinline__ T msc(T acc, T a, T b) {
    if constexpr (!F) return acc - a * b;
    else if constexpr (sizeof(T) == 4) return __builtin_fmaf(-a, b, acc);
    else return __builtin_fma(-a, b, acc);
}

}  // namespace dev
}  // namespace zeno

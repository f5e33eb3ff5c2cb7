// devcommon.hpp — the few definitions the state kernels (kernels.hpp) and the
// batched trajectory kernel (trajectories.hpp) share: the workgroup size of
// the sweeps and fixed-order reductions, the reduction block, and the
// insertion of a zero bit into an index.
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

}  // namespace dev
}  // namespace zeno

// Test infrastructure: host stand-ins for the HIP built-ins that the batched
// trajectory kernel (zeno_amd/csrc/trajectories.hpp) and the exact tiled
// dense kernel (zeno_amd/csrc/kq_tiles.hpp) use, so that the kernels' own
// source runs on the CPU under sanitizers.  A workgroup's lanes are
// cooperative contexts on one thread (tests/emu/lanes.hpp):
// __syncthreads hands control back to the scheduler, which resumes every
// lane to its next barrier, in an order of the test's choosing, before any
// passes it.  __shared__ arrays are function statics (one workgroup runs at
// a time).  ZENO_EMU_OUTLINE builds every __forceinline__ helper as an
// out-of-line call instead.  Nothing here is part of the library.
#pragma once

#include <cmath>
#include <cstdint>

struct dim3 {
    unsigned x = 1, y = 1, z = 1;
};

namespace emu {
inline dim3 thread_idx, block_idx, grid_dim, block_dim;
void barrier();  // tests/emu/lanes.hpp
}  // namespace emu

#define threadIdx (emu::thread_idx)
#define blockIdx (emu::block_idx)
#define gridDim (emu::grid_dim)
#define blockDim (emu::block_dim)

#define __device__
#define __host__
#define __global__
#define __launch_bounds__(...)
#define __shared__ static
#ifdef ZENO_EMU_OUTLINE
#define __forceinline__ __attribute__((noinline))
#else
#define __forceinline__ inline __attribute__((always_inline))
#endif

inline void __syncthreads() { emu::barrier(); }

// A value the kernel already holds uniform across its wave: every lane's own.
#define __builtin_amdgcn_readfirstlane(x) (x)

inline uint64_t __umul64hi(uint64_t a, uint64_t b) {
    return (uint64_t)(((unsigned __int128)a * (unsigned __int128)b) >> 64);
}

using std::sqrt;

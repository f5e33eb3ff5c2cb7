// sort.hip — the sampler's draws on the device for large shot counts: the
// draws themselves (one xoshiro256++ stream, as the host would draw it),
// hipCUB's radix sort of the f64 draws, and the routing of the sorted draws
// to chunks by lower bound (sample_indices steps 2-4, src/sample.rs:52-66).
// A sorted sequence of doubles is unique, and the lower bound here is
// std::lower_bound's (the first draw not less than the chunk's start), so
// the result is the host path's.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>
#include <vector>

#include "draws.hpp"
#include "host.hpp"

namespace zeno {
namespace {

// starts[c] = first i with us[i] >= cum[c], for c < m (us sorted ascending).
__global__ __launch_bounds__(256) void k_route(const double *__restrict__ us, uint64_t n,
                                               const double *__restrict__ cum, uint64_t m,
                                               uint64_t *__restrict__ starts) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m) return;
    const double v = cum[c];
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (us[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    starts[c] = lo;
}

// ------------------------------------------------------------ the draws
// Thread k of k_draws starts draws::kDrawBlock * k steps into the stream
// (draws.hpp) and writes those draws in stream order: draw i is
// f64(output i) * total, the product the host path forms.
using draws::kDrawBlock;

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

__global__ __launch_bounds__(64) void k_draws(const uint64_t *__restrict__ s0, const uint64_t *__restrict__ jump,
                                              int levels, uint64_t shots, double total, double *__restrict__ us) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t s[4] = {s0[0], s0[1], s0[2], s0[3]};
    // every lane runs every level (the map's columns are uniform loads) and
    // keeps the result where its bit of k is set
    for (int i = 0; i < levels; ++i) {
        const uint64_t *col = jump + (size_t)i * 256 * 4;
        uint64_t r[4] = {0, 0, 0, 0};
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const uint64_t x = s[w];
#pragma unroll 8
            for (int b = 0; b < 64; ++b) {
                const uint64_t m = 0 - ((x >> b) & 1);
                const uint64_t *c = col + (size_t)(w * 64 + b) * 4;
#pragma unroll
                for (int u = 0; u < 4; ++u) r[u] ^= c[u] & m;
            }
        }
        const bool take = (k >> i) & 1;
#pragma unroll
        for (int u = 0; u < 4; ++u) s[u] = take ? r[u] : s[u];
    }
    const uint64_t start = k * kDrawBlock;
    if (start >= shots) return;
    const uint64_t end = start + kDrawBlock < shots ? start + kDrawBlock : shots;
    for (uint64_t i = start; i < end; ++i) {
        const uint64_t out = rotl64(s[0] + s[3], 23) + s[0];
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl64(s[3], 45);
        us[i] = (double)(out >> 11) * 0x1.0p-53 * total;
    }
}

}  // namespace

size_t draw_jump_bytes(uint64_t shots) { return 32 + (size_t)draws::draw_levels(shots) * sizeof(draws::Gf2Map); }

// `host` stages the generator state and the jump maps; the caller keeps it
// alive until the stream has been synchronised.
hipError_t draws_on_device(const uint64_t *state, uint64_t shots, double total, void *jump_dev,
                           std::vector<uint64_t> &host, double *us, hipStream_t stream) {
    const int levels = draws::draw_levels(shots);
    host.resize(4 + (size_t)levels * 1024);
    for (int w = 0; w < 4; ++w) host[w] = state[w];
    draws::copy_jump_maps(levels, host.data() + 4);
    uint64_t *d = static_cast<uint64_t *>(jump_dev);
    const hipError_t e = hipMemcpyAsync(d, host.data(), host.size() * 8, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return e;
    const uint64_t blocks = (shots + kDrawBlock - 1) / kDrawBlock;
    hipLaunchKernelGGL(k_draws, dim3((unsigned)((blocks + 63) / 64)), dim3(64), 0, stream, d, d + 4, levels, shots,
                       total, us);
    return hipGetLastError();
}

size_t draw_sort_temp_bytes(uint64_t n) {
    size_t bytes = 0;
    const hipError_t e =
        hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, (const double *)nullptr, (double *)nullptr, (int)n);
    return e == hipSuccess ? bytes : 0;
}

hipError_t draw_sort_and_route(void *temp, size_t temp_bytes, const double *us, double *sorted, uint64_t n,
                               const double *cum, uint64_t m, uint64_t *starts, hipStream_t stream) {
    size_t bytes = temp_bytes;
    hipError_t e = hipcub::DeviceRadixSort::SortKeys(temp, bytes, us, sorted, (int)n, 0, 64, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_route, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, stream, sorted, n, cum, m, starts);
    return hipGetLastError();
}

}  // namespace zeno

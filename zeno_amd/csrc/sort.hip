// sort.hip — the sampler's draws on the device for large shot counts:
// hipCUB's radix sort of the f64 draws, and the routing of the sorted draws
// to chunks by lower bound (sample_indices steps 3-4, src/sample.rs:60-66).
// A sorted sequence of doubles is unique, and the lower bound here is
// std::lower_bound's (the first draw not less than the chunk's start), so
// the result is the host path's.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

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

}  // namespace

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

// kernels.hpp — CDNA4 (gfx950) HIP kernels of the zeno state-vector backend.
//
// The state is split complex (SoA): re[2^n], im[2^n] of T = float or double
// in HBM.  Every kernel below reproduces the reference CPU path's
// expression tree for its op (src/state.rs), and the library is compiled
// with -ffp-contract=off, so no a*b+c is contracted to an FMA: results are
// bit-identical to the scalar CPU path for the same inputs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace zeno {
namespace dev {

constexpr int kBlock = 256;        // 4 wave64 per workgroup
constexpr int kReduceSpan = 4096;  // elements per reduction block (16 per lane)
constexpr int kMaxDiagK = 12;
constexpr int kMaxFusedK = 6;

template <typename T>
struct Mat2 {
    T r[4], i[4];
};

struct Qubits {
    uint32_t q[kMaxDiagK];
    int k;
};

__device__ __forceinline__ uint64_t insert_zero(uint64_t x, uint32_t q) {
    return ((x >> q) << (q + 1)) | (x & ((1ull << q) - 1));
}

__device__ __forceinline__ uint64_t gtid() { return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint64_t gstride() { return (uint64_t)gridDim.x * blockDim.x; }

// ------------------------------------------------------------- zero state
// StateVec::zero_state / reset_zero (src/state.rs:109-132).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_zero(T *__restrict__ re, T *__restrict__ im, uint64_t len) {
    for (uint64_t i = gtid(); i < len; i += gstride()) {
        re[i] = i == 0 ? T(1) : T(0);
        im[i] = T(0);
    }
}

// -------------------------------------------------------------- 1q dense
// kern_1q_scalar (src/state.rs:232-260): pair (a, a + 2^q) per half index.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_1q(T *__restrict__ re, T *__restrict__ im, uint32_t q, uint64_t half,
                                               Mat2<T> m) {
    const uint64_t s = 1ull << q;
    for (uint64_t h = gtid(); h < half; h += gstride()) {
        const uint64_t a = insert_zero(h, q), b = a + s;
        const T ar = re[a], ai = im[a], br = re[b], bi = im[b];
        re[a] = m.r[0] * ar - m.i[0] * ai + m.r[1] * br - m.i[1] * bi;
        im[a] = m.r[0] * ai + m.i[0] * ar + m.r[1] * bi + m.i[1] * br;
        re[b] = m.r[2] * ar - m.i[2] * ai + m.r[3] * br - m.i[3] * bi;
        im[b] = m.r[2] * ai + m.i[2] * ar + m.r[3] * bi + m.i[3] * br;
    }
}

// -------------------------------------------------------------- kq dense
// kern_kq_scalar (src/state.rs:355-392): gather 2^K amplitudes of a group,
// row-major matvec with ar = (ar + wr*xr) - wi*xi, ai = (ai + wr*xi) + wi*xr
// over ascending columns, scatter.  The matrix ([re | im], 2*4^K of T) is
// staged in LDS; every lane reads the same element (broadcast).
template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_kq(T *__restrict__ re, T *__restrict__ im, const T *__restrict__ mat,
                                               Qubits qs, uint64_t groups) {
    constexpr int D = 1 << K;
    __shared__ T sm[2 * D * D];
    for (int i = threadIdx.x; i < 2 * D * D; i += blockDim.x) sm[i] = mat[i];
    __syncthreads();
    uint64_t bit[K];
#pragma unroll
    for (int b = 0; b < K; ++b) bit[b] = 1ull << qs.q[b];
    for (uint64_t g = gtid(); g < groups; g += gstride()) {
        uint64_t base = g;
#pragma unroll
        for (int b = 0; b < K; ++b) base = insert_zero(base, qs.q[b]);
        T xr[D], xi[D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            uint64_t o = base;
#pragma unroll
            for (int b = 0; b < K; ++b)
                if ((j >> b) & 1) o |= bit[b];
            xr[j] = re[o];
            xi[j] = im[o];
        }
        constexpr int kRowUnroll = K <= 4 ? D : 1;
#pragma unroll kRowUnroll
        for (int r = 0; r < D; ++r) {
            T ar = T(0), ai = T(0);
            const T *wr = sm + r * D, *wi = sm + D * D + r * D;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                ar = ar + wr[c] * xr[c] - wi[c] * xi[c];
                ai = ai + wr[c] * xi[c] + wi[c] * xr[c];
            }
            uint64_t o = base;
#pragma unroll
            for (int b = 0; b < K; ++b)
                if ((r >> b) & 1) o |= bit[b];
            re[o] = ar;
            im[o] = ai;
        }
    }
}

// -------------------------------------------------------------------- CX
// apply_cx (src/state.rs:394-417): swap the target pair inside the
// control=1 half; no arithmetic.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cx(T *__restrict__ re, T *__restrict__ im, uint32_t lo, uint32_t hi,
                                               uint64_t coff, uint64_t toff, uint64_t groups) {
    for (uint64_t g = gtid(); g < groups; g += gstride()) {
        const uint64_t a = insert_zero(insert_zero(g, lo), hi) | coff, b = a | toff;
        const T r0 = re[a], r1 = re[b], i0 = im[a], i1 = im[b];
        re[a] = r1;
        re[b] = r0;
        im[a] = i1;
        im[b] = i0;
    }
}

// -------------------------------------------------------------- diagonal
// kern_diag1_scalar (src/state.rs:498-527): d[0..1] on the a side, d[2..3]
// on the b side, re = wr*ar - wi*ai, im = wr*ai + wi*ar.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_diag1(T *__restrict__ re, T *__restrict__ im, uint32_t q, uint64_t half,
                                                  T d0r, T d0i, T d1r, T d1i) {
    const uint64_t s = 1ull << q;
    for (uint64_t h = gtid(); h < half; h += gstride()) {
        const uint64_t a = insert_zero(h, q), b = a + s;
        const T ar = re[a], ai = im[a], br = re[b], bi = im[b];
        re[a] = d0r * ar - d0i * ai;
        im[a] = d0r * ai + d0i * ar;
        re[b] = d1r * br - d1i * bi;
        im[b] = d1r * bi + d1i * br;
    }
}

// kern_diagk_scalar (src/state.rs:577-596): table index j gathered from the
// bits of i at qs; table ([re | im], 2*2^k of T) staged in LDS.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_diagk(T *__restrict__ re, T *__restrict__ im, const T *__restrict__ tab,
                                                  Qubits qs, uint64_t len) {
    __shared__ T st[2 << kMaxDiagK];
    const int dim = 1 << qs.k;
    for (int i = threadIdx.x; i < dim; i += blockDim.x) {
        st[i] = tab[i];
        st[(1 << kMaxDiagK) + i] = tab[dim + i];
    }
    __syncthreads();
    for (uint64_t i = gtid(); i < len; i += gstride()) {
        uint32_t j = 0;
        for (int b = 0; b < qs.k; ++b) j |= (uint32_t)((i >> qs.q[b]) & 1) << b;
        const T wr = st[j], wi = st[(1 << kMaxDiagK) + j];
        const T ar = re[i], ai = im[i];
        re[i] = wr * ar - wi * ai;
        im[i] = wr * ai + wi * ar;
    }
}

// -------------------------------------------------------------- collapse
// collapse (src/state.rs:628-661): keep side scaled, other side zeroed.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_collapse(T *__restrict__ re, T *__restrict__ im, uint32_t q,
                                                     uint64_t half, int outcome, T scale) {
    const uint64_t s = 1ull << q;
    for (uint64_t h = gtid(); h < half; h += gstride()) {
        const uint64_t a = insert_zero(h, q);
        const uint64_t keep = outcome ? a + s : a, zero = outcome ? a : a + s;
        re[keep] = re[keep] * scale;
        im[keep] = im[keep] * scale;
        re[zero] = T(0);
        im[zero] = T(0);
    }
}

// ------------------------------------------------------------- reductions
// Fixed-order f64 reduction shared with the CPU oracle: block b owns
// e[b*4096 .. b*4096+4095]; lane t adds e[b*4096 + j*256 + t], j = 0..15,
// serially from 0.0; then lanes combine pairwise with strides 128..1.
__device__ __forceinline__ double block_tree(double a, double *s) {
    const int t = threadIdx.x;
    s[t] = a;
    __syncthreads();
#pragma unroll
    for (int st = kBlock / 2; st >= 1; st >>= 1) {
        if (t < st) s[t] = s[t] + s[t + st];
        __syncthreads();
    }
    return s[0];
}

// mode 0: |amp|^2 over all indices (norm_sqr); mode 1: over the q=1 half
// indexed by half index (prob_one, src/state.rs:598-626).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_reduce_state(const T *__restrict__ re, const T *__restrict__ im,
                                                         uint64_t n_elems, int mode, uint32_t q,
                                                         double *__restrict__ out) {
    __shared__ double s[kBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kReduceSpan;
    const uint64_t qs = 1ull << q;
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < kReduceSpan / kBlock; ++j) {
        const uint64_t e = base + (uint64_t)j * kBlock + threadIdx.x;
        if (e < n_elems) {
            const uint64_t idx = mode ? insert_zero(e, q) + qs : e;
            const double r = (double)re[idx], m = (double)im[idx];
            a = a + (r * r + m * m);
        }
    }
    const double tot = block_tree(a, s);
    if (threadIdx.x == 0) out[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBlock) void k_reduce_f64(const double *__restrict__ in, uint64_t n,
                                                       double *__restrict__ out) {
    __shared__ double s[kBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kReduceSpan;
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < kReduceSpan / kBlock; ++j) {
        const uint64_t e = base + (uint64_t)j * kBlock + threadIdx.x;
        if (e < n) a = a + in[e];
    }
    const double tot = block_tree(a, s);
    if (threadIdx.x == 0) out[blockIdx.x] = tot;
}

// max_i |a_i - b_i| (complex modulus, f64) between two states of possibly
// different precision; non-negative doubles order like their bit patterns,
// so the grid combines block maxima with an integer atomicMax.
template <typename A, typename B>
__global__ __launch_bounds__(kBlock) void k_max_abs_diff(const A *__restrict__ ar, const A *__restrict__ ai,
                                                         const B *__restrict__ br, const B *__restrict__ bi,
                                                         uint64_t len, unsigned long long *__restrict__ out) {
    __shared__ double s[kBlock];
    double m = 0.0;
    for (uint64_t i = gtid(); i < len; i += gstride()) {
        const double dr = (double)ar[i] - (double)br[i], di = (double)ai[i] - (double)bi[i];
        m = fmax(m, sqrt(dr * dr + di * di));
    }
    s[threadIdx.x] = m;
    __syncthreads();
    for (int st = kBlock / 2; st >= 1; st >>= 1) {
        if (threadIdx.x < st) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(s[0]));
}

// --------------------------------------------------------------- sampler
// sample_indices pass 1 (src/sample.rs:38-50): per chunk a strictly serial
// f64 sum acc += r*r + i*i in index order.  The 256 lanes load a 4096
// element slab coalesced and square it into LDS; lane 0 then accumulates
// the slab serially, so the sum order is exactly the reference's.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_chunk_sums(const T *__restrict__ re, const T *__restrict__ im,
                                                       uint64_t len, uint64_t chunk, double *__restrict__ sums) {
    __shared__ double tv[kReduceSpan];
    const uint64_t lo = (uint64_t)blockIdx.x * chunk;
    const uint64_t hi = lo + chunk < len ? lo + chunk : len;
    double acc = 0.0;
    for (uint64_t slab = lo; slab < hi; slab += kReduceSpan) {
        const int cnt = (int)(hi - slab < (uint64_t)kReduceSpan ? hi - slab : (uint64_t)kReduceSpan);
        for (int j = threadIdx.x; j < cnt; j += kBlock) {
            const double r = (double)re[slab + j], m = (double)im[slab + j];
            tv[j] = r * r + m * m;
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int j = 0; j < cnt; ++j) acc += tv[j];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = acc;
}

// sample_indices pass 5 + 6 (src/sample.rs:68-107): rescan one touched
// chunk from cum_before[c], emitting the sorted draws us[s0..s1) that fall
// below the running sum; leftovers land on the chunk's last index with
// nonzero probability.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_rescan(const T *__restrict__ re, const T *__restrict__ im, uint64_t len,
                                                   uint64_t chunk, const uint32_t *__restrict__ touched,
                                                   const double *__restrict__ cum, const uint64_t *__restrict__ starts,
                                                   const double *__restrict__ us, uint64_t *__restrict__ out) {
    __shared__ double tv[kReduceSpan];
    __shared__ uint64_t sptr;
    __shared__ int sdone;
    __shared__ uint64_t slast[kBlock];
    const uint32_t c = touched[blockIdx.x];
    const uint64_t lo = (uint64_t)c * chunk;
    const uint64_t hi = lo + chunk < len ? lo + chunk : len;
    const uint64_t s1 = starts[c + 1];
    double acc = cum[c];
    if (threadIdx.x == 0) {
        sptr = starts[c];
        sdone = 0;
    }
    __syncthreads();
    for (uint64_t slab = lo; slab < hi; slab += kReduceSpan) {
        const int cnt = (int)(hi - slab < (uint64_t)kReduceSpan ? hi - slab : (uint64_t)kReduceSpan);
        for (int j = threadIdx.x; j < cnt; j += kBlock) {
            const double r = (double)re[slab + j], m = (double)im[slab + j];
            tv[j] = r * r + m * m;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t ptr = sptr;
            for (int j = 0; j < cnt; ++j) {
                acc += tv[j];
                while (ptr < s1 && us[ptr] < acc) out[ptr++] = slab + (uint64_t)j;
                if (ptr == s1) break;
            }
            sptr = ptr;
            sdone = ptr == s1;
        }
        __syncthreads();
        if (sdone) break;
    }
    if (sdone) return;
    // Fallback: last index of the chunk with nonzero probability (else hi-1).
    uint64_t best = 0;
    int found = 0;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const double r = (double)re[i], m = (double)im[i];
        if (r * r + m * m > 0.0) best = i, found = 1;
    }
    slast[threadIdx.x] = found ? best + 1 : 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t m = 0;
        for (int t = 0; t < kBlock; ++t) m = slast[t] > m ? slast[t] : m;
        const uint64_t last = m ? m - 1 : hi - 1;
        for (uint64_t p = sptr; p < s1; ++p) out[p] = last;
    }
}

}  // namespace dev
}  // namespace zeno

// batch.hpp — many small static circuits as uneven instances of one
// persistent launch (each instance reproduces run_sampled,
// src/exec.rs:407-452: the op loop, then sample_indices, src/sample.rs:29-111).
//
// A workgroup claims instances from a work counter, largest first, and runs
// each one whole: the state (2^n re + 2^n im of T) lives in LDS for the
// instance's life, so its sweeps never touch HBM; only the op tables come in
// and the sampled indices go out.  Every op uses the per-op kernels'
// expression tree (k_1q, k_kq, k_diag1 / k_diagk, k_cx), so the final state
// is bit-identical to the CPU path's, and the sampler reproduces the
// reference's indices: below 17 qubits the state is one sampler chunk, whose
// serial f64 running sum (acc += r*r + i*i in index order) gives both the
// total mass and every rescan threshold, and a draw u = f64 * total lands on
// the first index whose running sum exceeds it (the rescan's emit rule, with
// its fallback to the last index of nonzero probability).  One lane builds
// the running sums while a lane of another wave draws the xoshiro256++
// stream; then every lane places its draws by binary search.  The indices
// come out in draw order (the reference returns them sorted; the counts
// are the same multiset).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"
#include "trajectories.hpp"

namespace zeno {
namespace dev {

constexpr int kBatchBlock = 256;
// Dense ops on up to 3 qubits in f64 and 4 in f32: 2^K complex inputs per
// lane stay in registers (2^K * 2 * sizeof(T) / 4 VGPRs) next to a matrix row.
template <typename T>
constexpr int kBatchMaxDense = sizeof(T) == 8 ? 3 : 4;

struct BatchOp {
    uint32_t data_off;       // T offset (from the instance's data) of the matrix [re | im] or table [re | im]
    uint8_t kind;            // 0 1q, 1 dense k = 2..4, 2 diagonal k = 1..12, 3 cx
    uint8_t k;
    uint8_t q[kMaxDiagK];    // sorted qubits; cx: control, target
    uint8_t pad[2];
};

struct BatchInst {
    uint64_t ops_off;   // first op of the instance in the op table
    uint64_t data_off;  // first T of the instance's matrices and tables
    uint64_t x_off;     // first double of its running sums (2^n)
    uint64_t out_off;   // first index of its shots (raw draws staged there first)
    uint64_t seed;      // the sampler's seed_from_u64 input (splitmix64 of the run seed)
    uint32_t n, nops;
    uint32_t shots, pad;
};

// Everything a launch reads, at a fixed device address: a captured graph of
// the launch (whose only argument is this block's address) replays for any
// batch written there.
template <typename T>
struct BatchLaunch {
    const BatchInst *inst;
    const BatchOp *ops;
    const T *data;
    double *xs;
    uint64_t *out;
    uint32_t count;  // instances
    uint32_t next;   // work counter: uploaded as 0 with the block
};

// A value every lane loads from the same address, moved to scalar
// registers word by word: the compiler cannot prove the tables unwritten
// (the kernel stores to other global memory), so it loads them per lane,
// and the op dispatch would then branch as if divergent.
template <typename S>
__device__ __forceinline__ S uniform(const S &v) {
    static_assert(sizeof(S) % 4 == 0, "whole words");
    S r;
    const uint32_t *w = reinterpret_cast<const uint32_t *>(&v);
    uint32_t *o = reinterpret_cast<uint32_t *>(&r);
#pragma unroll
    for (unsigned i = 0; i < sizeof(S) / 4; ++i) o[i] = __builtin_amdgcn_readfirstlane(w[i]);
    return r;
}

// Indices inside an on-chip state fit 32 bits (n <= 14).
__device__ __forceinline__ uint32_t ins0(uint32_t x, uint32_t q) { return ((x >> q) << (q + 1)) | (x & ((1u << q) - 1)); }

template <typename T>
__device__ __forceinline__ void b_1q(T *re, T *im, uint32_t len, uint32_t q, const T *__restrict__ d) {
    const T m0r = d[0], m1r = d[1], m2r = d[2], m3r = d[3];
    const T m0i = d[4], m1i = d[5], m2i = d[6], m3i = d[7];
    const uint32_t s = 1u << q, half = len >> 1;
    for (uint32_t h = threadIdx.x; h < half; h += kBatchBlock) {
        const uint32_t a = ins0(h, q), b = a + s;
        const T ar = re[a], ai = im[a], br = re[b], bi = im[b];
        re[a] = msc<false>(mac<false>(msc<false>(m0r * ar, m0i, ai), m1r, br), m1i, bi);
        im[a] = mac<false>(mac<false>(mac<false>(m0r * ai, m0i, ar), m1r, bi), m1i, br);
        re[b] = msc<false>(mac<false>(msc<false>(m2r * ar, m2i, ai), m3r, br), m3i, bi);
        im[b] = mac<false>(mac<false>(mac<false>(m2r * ai, m2i, ar), m3r, bi), m3i, br);
    }
}

template <typename T, int K>
__device__ __forceinline__ void b_kq(T *re, T *im, uint32_t len, const uint8_t *qs, const T *__restrict__ d) {
    constexpr int D = 1 << K;
    uint32_t q[K];
#pragma unroll
    for (int b = 0; b < K; ++b) q[b] = qs[b];
    for (uint32_t g = threadIdx.x; g < (len >> K); g += kBatchBlock) {
        uint32_t base = g;
#pragma unroll
        for (int b = 0; b < K; ++b) base = ins0(base, q[b]);
        T xr[D], xi[D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            uint32_t o = base;
#pragma unroll
            for (int b = 0; b < K; ++b)
                if ((j >> b) & 1) o |= 1u << q[b];
            xr[j] = re[o];
            xi[j] = im[o];
        }
        // each group belongs to one lane, whose inputs are all in registers:
        // its outputs can be stored row by row.  Rows stay a runtime loop, so
        // the matrix is re-read per row instead of hoisted into registers.
#pragma unroll 1
        for (int r = 0; r < D; ++r) {
            T ar = T(0), ai = T(0);
            const T *wr = d + r * D, *wi = d + D * D + r * D;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                ar = msc<false>(mac<false>(ar, wr[c], xr[c]), wi[c], xi[c]);
                ai = mac<false>(mac<false>(ai, wr[c], xi[c]), wi[c], xr[c]);
            }
            uint32_t o = base;
#pragma unroll
            for (int b = 0; b < K; ++b)
                if ((r >> b) & 1) o |= 1u << q[b];
            re[o] = ar;
            im[o] = ai;
        }
    }
}

// kern_diag1_scalar / kern_diagk_scalar: re = wr*ar - wi*ai, im = wr*ai + wi*ar
// with the table index gathered from the bits at qs.
template <typename T>
__device__ __forceinline__ void b_diag(T *re, T *im, uint32_t len, int k, const uint8_t *qs,
                                       const T *__restrict__ d) {
    const int D = 1 << k;
    for (uint32_t i = threadIdx.x; i < len; i += kBatchBlock) {
        uint32_t j = 0;
#pragma unroll
        for (int b = 0; b < kMaxDiagK; ++b)  // static indices keep the op's qubits in registers
            if (b < k) j |= (uint32_t)((i >> qs[b]) & 1) << b;
        const T wr = d[j], wi = d[D + j];
        const T ar = re[i], ai = im[i];
        re[i] = msc<false>(wr * ar, wi, ai);
        im[i] = mac<false>(wr * ai, wi, ar);
    }
}

template <typename T>
__device__ __forceinline__ void b_cx(T *re, T *im, uint32_t len, uint32_t c, uint32_t t) {
    const uint32_t lo = c < t ? c : t, hi = c < t ? t : c;
    const uint32_t cb = 1u << c, tb = 1u << t;
    for (uint32_t g = threadIdx.x; g < (len >> 2); g += kBatchBlock) {
        const uint32_t a = ins0(ins0(g, lo), hi) | cb, b = a | tb;
        const T r0 = re[a], r1 = re[b], i0 = im[a], i1 = im[b];
        re[a] = r1;
        re[b] = r0;
        im[a] = i1;
        im[b] = i0;
    }
}

template <typename T>
__global__ __launch_bounds__(kBatchBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_batch(BatchLaunch<T> *__restrict__ L) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t s_inst;
    __shared__ double s_total;
    T *const sre = reinterpret_cast<T *>(smem);
    const uint32_t count = L->count;
    while (true) {
        if (threadIdx.x == 0) s_inst = atomicAdd(&L->next, 1u);
        __syncthreads();
        // the claim is the workgroup's: readfirstlane keeps it (and every
        // address and branch derived from it) in scalar registers
        const uint32_t ii = __builtin_amdgcn_readfirstlane(s_inst);
        __syncthreads();  // every lane has read s_inst before lane 0 claims again
        if (ii >= count) break;  // uniform: every wave of the workgroup leaves together
        const BatchInst I = uniform(L->inst[ii]);
        const uint32_t len = 1u << I.n;
        T *re = sre, *im = sre + len;
        for (uint32_t i = threadIdx.x; i < len; i += kBatchBlock) {
            re[i] = i == 0 ? T(1) : T(0);
            im[i] = T(0);
        }
        __syncthreads();
        const BatchOp *ops = L->ops + I.ops_off;
        const T *data = L->data + I.data_off;
        for (uint32_t oi = 0; oi < I.nops; ++oi) {
            const BatchOp op = uniform(ops[oi]);
            const T *d = data + op.data_off;
            switch (op.kind) {
                case 0: b_1q(re, im, len, op.q[0], d); break;
                case 1:
                    if (op.k == 2) b_kq<T, 2>(re, im, len, op.q, d);
                    else if (kBatchMaxDense<T> == 3 || op.k == 3) b_kq<T, 3>(re, im, len, op.q, d);
                    else b_kq<T, kBatchMaxDense<T>>(re, im, len, op.q, d);
                    break;
                case 2: b_diag(re, im, len, op.k, op.q, d); break;
                default: b_cx(re, im, len, op.q[0], op.q[1]); break;
            }
            __syncthreads();
        }
        if (I.shots > 0) {
            double *X = L->xs + I.x_off;
            uint64_t *out = L->out + I.out_off;
            if (threadIdx.x == 0) {
                // the chunk's serial running sum (sample.rs:38-50 and :68-77)
                double acc = 0.0;
                for (uint32_t i = 0; i < len; i += 8) {
                    double p[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        if (i + u < len) {
                            const double a = (double)re[i + u], b = (double)im[i + u];
                            p[u] = a * a + b * b;
                        }
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        if (i + u < len) {
                            acc += p[u];
                            X[i + u] = acc;
                        }
                    }
                }
                s_total = acc;
            } else if (threadIdx.x == 64) {
                // the draws' unit uniforms, in stream order (sample.rs:60)
                DevRng rng;
                rng.seed(I.seed);
                for (uint32_t j = 0; j < I.shots; ++j) out[j] = __double_as_longlong(rng.f64());
            }
            __syncthreads();
            const double total = s_total;
            for (uint32_t j = threadIdx.x; j < I.shots; j += kBatchBlock) {
                const double u = __longlong_as_double((long long)out[j]) * total;
                // first index whose running sum exceeds u (the rescan's emit rule)
                uint32_t lo = 0, hi = len;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (u < X[mid]) hi = mid;
                    else lo = mid + 1;
                }
                if (lo == len) {  // unresolved draw: the last index with nonzero probability
                    lo = len - 1;
                    for (uint32_t i = len; i-- > 0;) {
                        const double a = (double)re[i], b = (double)im[i];
                        if (a * a + b * b > 0.0) {
                            lo = i;
                            break;
                        }
                    }
                }
                out[j] = lo;
            }
        }
        __syncthreads();  // the state's LDS is free for the next instance
    }
}

}  // namespace dev
}  // namespace zeno

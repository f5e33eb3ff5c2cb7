// trajectories.hpp — batched noisy trajectories (src/exec.rs:459-619,
// src/noise.rs:91-192) in one persistent launch.
//
// Each workgroup runs whole trajectories, one after another (shot =
// blockIdx.x, += gridDim.x), on its own split-complex state in HBM.  The
// per-shot RNG is Xoshiro256++ seeded with splitmix64(seed ^ shot), drawn in
// the reference's frozen order (docs/NOISE.md:150-166); it is uniform across
// the workgroup, so every lane computes the same draws.  Sweeps use exactly
// the per-op kernels' arithmetic and prob_one uses the fixed reduction order
// of the rest of the backend, so every measurement decision — hence every
// shot's clbit word — equals the CPU path's for the same seed.
//
// Every helper is force-inlined: the kernel then makes no device-function
// calls and needs no scratch.  No-jump damping steps are folded into the
// next sweep (Pend below), which removes about half of the state passes.
//
// Two instances: KM = 3 takes unitaries and diagonals on <= 3 qubits (every
// noisy circuit: the reference compiles those without fusion, src/lib.rs:
// 226-233); KM = 6 also takes fused dense blocks on up to 6 qubits and
// diagonals on up to 12, with a larger LDS matrix stage.  States of more
// than 21 qubits keep their probability sums' block totals in a global
// scratch area after the state (more than 256 blocks).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "devcommon.hpp"

namespace zeno {
namespace dev {

constexpr int kTrajMaxK = 3;        // damping scales held at once; the narrow instance's op width
constexpr int kTrajWideDense = 6;   // the wide instance: dense blocks on up to 6 qubits
constexpr int kTrajOpQubits = 12;   // ... and diagonals on up to 12
constexpr int kTrajMaxQubits = 26;  // largest trajectory state
constexpr int kTrajLdsSumBlocks = 256;  // block totals kept in LDS (n <= 21)

struct TrajOp {
    uint32_t kind;  // 0 unitary, 1 diagonal, 2 cx, 3 measure, 4 reset
    uint32_t k;     // qubits of a unitary/diagonal
    uint32_t q[kTrajOpQubits];  // sorted qubits; cx: control, target; measure/reset: qubit
    uint32_t clbit;
    uint32_t is_if, creg_off, creg_len;
    uint32_t data_off;  // matrix (2*4^k T) or diagonal (2*2^k T) in the data block
    uint64_t value;
};

struct TrajNoise {
    double dep1, dep2, bit_flip, phase_flip, damping, ro01, ro10;
};

struct DevRng {
    uint64_t s0, s1, s2, s3;
    __device__ static uint64_t mix(uint64_t x) {
        uint64_t z = x + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    __device__ void seed(uint64_t x) {  // rand_xoshiro seed_from_u64
        s0 = mix(x);
        s1 = mix(x + 0x9E3779B97F4A7C15ull);
        s2 = mix(x + 2 * 0x9E3779B97F4A7C15ull);
        s3 = mix(x + 3 * 0x9E3779B97F4A7C15ull);
    }
    __device__ uint64_t next() {
        const uint64_t r = ((s0 + s3) << 23 | (s0 + s3) >> 41) + s0;
        const uint64_t t = s1 << 17;
        s2 ^= s0;
        s3 ^= s1;
        s1 ^= s2;
        s0 ^= s3;
        s2 ^= t;
        s3 = s3 << 45 | s3 >> 19;
        return r;
    }
    __device__ double f64() { return (double)(next() >> 11) * 0x1.0p-53; }
    __device__ uint64_t range(uint64_t lo, uint64_t hi) {  // rand 0.8 UniformInt<u64>
        const uint64_t r = hi - lo;
        const uint64_t zone = (r << __builtin_clzll(r)) - 1;
        while (true) {
            const uint64_t v = next();
            const uint64_t mlo = v * r, mhi = __umul64hi(v, r);
            if (mlo <= zone) return lo + mhi;
        }
    }
};

template <typename T>
struct TrajCtx {
    T *re, *im;
    uint64_t len;
    double *red;     // LDS: 8 x 256 lane partials + 256 block sums
    const T *pauli;  // LDS: X then Y as [re 4 | im 4]
    double *gsum;    // global: the block sums when there are more than 256, then the next level's
};

// Where block b's sum goes: LDS while the blocks fit it, else global.
template <typename T>
__device__ __forceinline__ double *block_sums(const TrajCtx<T> &c) {
    return c.gsum ? c.gsum : c.red + 8 * kBlock;
}

// Pending damping scales.  A no-jump amplitude-damping step is the
// diagonal diag(sc, sc*sqrt(1-gamma)) on one qubit (noise.rs:114-127).
// Instead of a sweep of its own it is held here (uniform across the
// workgroup) and applied to every amplitude by the next sweep as that sweep
// loads it, with kern_diag1_scalar's expressions (imaginary parts T(0),
// signed zeros included) and in the order the steps were taken; prob_one
// applies it on the fly without writing.  The state therefore goes through
// exactly the values of the one-sweep-per-step path in fewer passes.
template <typename T>
struct Pend {
    int n = 0;
    uint32_t q[kTrajMaxK];
    T d0[kTrajMaxK], d1[kTrajMaxK];
};

template <typename T>
__device__ __forceinline__ void pend_apply(const Pend<T> &p, uint64_t idx, T &ar, T &ai) {
#pragma unroll
    for (int e = 0; e < kTrajMaxK; ++e) {
        if (e >= p.n) break;
        const T d = ((idx >> p.q[e]) & 1) ? p.d1[e] : p.d0[e], z = T(0);
        const T r = d * ar - z * ai, i = d * ai + z * ar;
        ar = r;
        ai = i;
    }
}

// Applies the pending scales with a sweep of their own.
template <typename T>
__device__ __forceinline__ void t_pend_flush(const TrajCtx<T> c, Pend<T> &p) {
    for (uint64_t i = threadIdx.x; i < c.len; i += blockDim.x) {
        T ar = c.re[i], ai = c.im[i];
        pend_apply(p, i, ar, ai);
        c.re[i] = ar;
        c.im[i] = ai;
    }
    p.n = 0;
    __syncthreads();
}

// prob_one in the backend's fixed order: 4096-element blocks (16 serial
// adds per lane, pairwise tree over 256 lanes), then the block sums the
// same way, level by level (k_reduce_f64's order).  Needs blockDim.x ==
// 256.  Pending damping scales are applied to each loaded amplitude, not
// stored.
//
// t_prob_blocks: the lane partials of blocks b0 .. b0+nchunk-1 (nchunk <= 8)
// are in red[bi * 256 + lane]; each block's pairwise tree goes to bsum.
template <typename T>
__device__ __forceinline__ void t_prob_blocks(const TrajCtx<T> c, int b0, int nchunk) {
    double *s = c.red;
    double *bsum = block_sums(c);
    const int t = threadIdx.x;
    __syncthreads();
    for (int st = kBlock / 2; st >= 1; st >>= 1) {
        if (t < st)
            for (int bi = 0; bi < nchunk; ++bi) s[bi * kBlock + t] = s[bi * kBlock + t] + s[bi * kBlock + t + st];
        __syncthreads();
    }
    if (t < nchunk) bsum[b0 + t] = s[t * kBlock];
    __syncthreads();
}

// The block sums of all nb blocks, combined by the same tree.
template <typename T>
__device__ __forceinline__ double t_prob_final(const TrajCtx<T> c, int nb) {
    double *s = c.red;
    double *bsum = block_sums(c);
    const int t = threadIdx.x;
    if (nb > kTrajLdsSumBlocks) {
        // k_reduce_f64 level by level: blocks of 4096 sums, lane t adding
        // elements t + 256 j serially from 0.0, then the pairwise tree; the
        // level's totals go after its inputs until one remains
        double *in = c.gsum, *out = c.gsum + nb;
        uint64_t m = (uint64_t)nb;
        while (m > 1) {
            const uint64_t nn = (m + kReduceSpan - 1) / kReduceSpan;
            for (uint64_t b = 0; b < nn; ++b) {
                double a = 0.0;
#pragma unroll
                for (int j = 0; j < kReduceSpan / kBlock; ++j) {
                    const uint64_t e = b * kReduceSpan + (uint64_t)j * kBlock + t;
                    if (e < m) a = a + in[e];
                }
                s[t] = a;
                __syncthreads();
                for (int st = kBlock / 2; st >= 1; st >>= 1) {
                    if (t < st) s[t] = s[t] + s[t + st];
                    __syncthreads();
                }
                if (t == 0) out[b] = s[0];
                __syncthreads();
            }
            double *nx = in;
            in = out;
            out = nx;
            m = nn;
        }
        const double r = in[0];
        __syncthreads();
        return r;
    }
    if (nb > 1) {
        double a = 0.0;
        if (t < nb) a = a + bsum[t];  // the 4096-wide pass: only j = 0 is populated
        s[t] = a;
        __syncthreads();
        for (int st = kBlock / 2; st >= 1; st >>= 1) {
            if (t < st) s[t] = s[t] + s[t + st];
            __syncthreads();
        }
        const double r = s[0];
        __syncthreads();
        return r;
    }
    const double r = bsum[0];
    __syncthreads();
    return r;
}

template <typename T>
__device__ __forceinline__ double t_prob_one(const TrajCtx<T> c, const Pend<T> &p, uint32_t q) {
    const uint64_t half = c.len >> 1, qs = 1ull << q;
    const int t = threadIdx.x;
    const int nb = (int)((half + kReduceSpan - 1) / kReduceSpan);
    double *s = c.red;  // [8][256] lane partials
    for (int b0 = 0; b0 < nb; b0 += 8) {
        const int nchunk = nb - b0 < 8 ? nb - b0 : 8;
        for (int bi = 0; bi < nchunk; ++bi) {
            double a = 0.0;
            const uint64_t base = (uint64_t)(b0 + bi) * kReduceSpan;
#pragma unroll
            for (int j = 0; j < kReduceSpan / kBlock; ++j) {
                const uint64_t e = base + (uint64_t)j * kBlock + t;
                if (e < half) {
                    const uint64_t idx = insert_zero(e, q) + qs;
                    T vr = c.re[idx], vi = c.im[idx];
                    pend_apply(p, idx, vr, vi);
                    const double r = (double)vr, m = (double)vi;
                    a = a + (r * r + m * m);
                }
            }
            s[bi * kBlock + t] = a;
        }
        t_prob_blocks(c, b0, nchunk);
    }
    return t_prob_final(c, nb);
}

// The state of a trajectory lives in L2 / the last-level cache; a sweep is
// bound by the latency of those loads.  Each lane therefore loads a batch
// of kTrajBatch<T> independent pairs before computing and storing any of
// them (the loop order does not change any pair's arithmetic).
template <typename T>
constexpr int kTrajBatch = sizeof(T) == 4 ? 8 : 4;

// With fuse (13 <= n <= 20: the b-side half is whole 4096-element blocks),
// the sweep also forms t_prob_one(q) of the new state: lane t visits half
// indices t + 256 m in increasing m, which is exactly the lane's serial order
// inside each block there, so it adds the same terms in the same order into
// the same partials; every lane reaches each group of 8 blocks at the same
// iteration, where the group's trees run as in t_prob_one.  Returns that
// probability, or -1 without fuse.
template <typename T>
__device__ __forceinline__ double t_1q(const TrajCtx<T> c, Pend<T> &p, uint32_t q, const T *m, bool fuse = false) {
    const T m00r = m[0], m01r = m[1], m10r = m[2], m11r = m[3];
    const T m00i = m[4], m01i = m[5], m10i = m[6], m11i = m[7];
    constexpr int U = kTrajBatch<T>;
    static_assert(kReduceSpan / kBlock % U == 0, "a batch stays inside one reduction block");
    const uint64_t s = 1ull << q, half = c.len >> 1, step = (uint64_t)blockDim.x;
    double acc = 0.0;
    int held = 0, blk = 0, b0 = 0;
    for (uint64_t h0 = threadIdx.x; h0 < half; h0 += U * step) {
        T ar[U], ai[U], br[U], bi[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t h = h0 + u * step;
            if (h < half) {
                const uint64_t a = insert_zero(h, q), b = a + s;
                ar[u] = c.re[a], ai[u] = c.im[a], br[u] = c.re[b], bi[u] = c.im[b];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t h = h0 + u * step;
            if (h < half) {
                const uint64_t a = insert_zero(h, q), b = a + s;
                pend_apply(p, a, ar[u], ai[u]);
                pend_apply(p, b, br[u], bi[u]);
                c.re[a] = m00r * ar[u] - m00i * ai[u] + m01r * br[u] - m01i * bi[u];
                c.im[a] = m00r * ai[u] + m00i * ar[u] + m01r * bi[u] + m01i * br[u];
                const T nr = m10r * ar[u] - m10i * ai[u] + m11r * br[u] - m11i * bi[u];
                const T ni = m10r * ai[u] + m10i * ar[u] + m11r * bi[u] + m11i * br[u];
                c.re[b] = nr;
                c.im[b] = ni;
                if (fuse) {
                    const double r = (double)nr, mi = (double)ni;
                    acc = acc + (r * r + mi * mi);
                }
            }
        }
        if (fuse) {
            held += U;
            if (held == kReduceSpan / kBlock) {
                c.red[blk * kBlock + threadIdx.x] = acc;
                acc = 0.0;
                held = 0;
                if (++blk == 8) {
                    t_prob_blocks(c, b0, 8);
                    b0 += 8;
                    blk = 0;
                }
            }
        }
    }
    p.n = 0;
    __syncthreads();
    if (!fuse) return -1.0;
    if (blk > 0) t_prob_blocks(c, b0, blk);
    return t_prob_final(c, (int)(half / kReduceSpan));
}

template <typename T, int K>
__device__ __forceinline__ void t_kq(const TrajCtx<T> c, Pend<T> &p, const uint32_t *qs, const T *m) {
    constexpr int D = 1 << K;
    for (uint64_t g = threadIdx.x; g < (c.len >> K); g += blockDim.x) {
        uint64_t base = g;
#pragma unroll
        for (int b = 0; b < K; ++b) base = insert_zero(base, qs[b]);
        T xr[D], xi[D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            uint64_t o = base;
#pragma unroll
            for (int b = 0; b < K; ++b)
                if ((j >> b) & 1) o |= 1ull << qs[b];
            xr[j] = c.re[o];
            xi[j] = c.im[o];
            pend_apply(p, o, xr[j], xi[j]);
        }
        // Rows stay a runtime loop so the LDS matrix is re-read per group
        // instead of being hoisted into registers.
#pragma unroll 1
        for (int r = 0; r < D; ++r) {
            T ar = T(0), ai = T(0);
#pragma unroll
            for (int cc = 0; cc < D; ++cc) {
                const T wr = m[r * D + cc], wi = m[D * D + r * D + cc];
                ar = ar + wr * xr[cc] - wi * xi[cc];
                ai = ai + wr * xi[cc] + wi * xr[cc];
            }
            uint64_t o = base;
#pragma unroll
            for (int b = 0; b < K; ++b)
                if ((r >> b) & 1) o |= 1ull << qs[b];
            c.re[o] = ar;
            c.im[o] = ai;
        }
    }
    p.n = 0;
    __syncthreads();
}

// The same op for K = 4..6 without holding a group's 2^K amplitudes in
// registers: lane t computes row t % D of group t / D of each batch of
// 256 / D groups (a group's lanes share a wave, so its loads are
// broadcasts), in kern_kq_scalar's column order, and every lane writes its
// row only after the whole workgroup has read the batch.
template <typename T, int K>
__device__ __forceinline__ void t_kq_rows(const TrajCtx<T> c, Pend<T> &p, const uint32_t *qs, const T *m) {
    constexpr int D = 1 << K, G = kBlock / D;
    const int t = threadIdx.x, r = t % D;
    const uint64_t groups = c.len >> K;
    for (uint64_t g0 = 0; g0 < groups; g0 += G) {
        const uint64_t g = g0 + (uint64_t)(t / D);
        uint64_t base = 0, o = 0;
        T ar = T(0), ai = T(0);
        if (g < groups) {
            base = g;
#pragma unroll
            for (int b = 0; b < K; ++b) base = insert_zero(base, qs[b]);
            for (int cc = 0; cc < D; ++cc) {
                uint64_t src = base;
#pragma unroll
                for (int b = 0; b < K; ++b)
                    if ((cc >> b) & 1) src |= 1ull << qs[b];
                T xr = c.re[src], xi = c.im[src];
                pend_apply(p, src, xr, xi);
                const T wr = m[r * D + cc], wi = m[D * D + r * D + cc];
                ar = ar + wr * xr - wi * xi;
                ai = ai + wr * xi + wi * xr;
            }
            o = base;
#pragma unroll
            for (int b = 0; b < K; ++b)
                if ((r >> b) & 1) o |= 1ull << qs[b];
        }
        __syncthreads();
        if (g < groups) {
            c.re[o] = ar;
            c.im[o] = ai;
        }
        __syncthreads();
    }
    p.n = 0;
}

// kern_diag1_scalar (src/state.rs:498-527): d0 on the a side, d1 on the b side.
template <typename T>
__device__ __forceinline__ void t_diag1(const TrajCtx<T> c, Pend<T> &p, uint32_t q, T d0r, T d1r, T d0i, T d1i) {
    constexpr int U = kTrajBatch<T>;
    const uint64_t s = 1ull << q, half = c.len >> 1, step = (uint64_t)blockDim.x;
    for (uint64_t h0 = threadIdx.x; h0 < half; h0 += U * step) {
        T ar[U], ai[U], br[U], bi[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t h = h0 + u * step;
            if (h < half) {
                const uint64_t a = insert_zero(h, q), b = a + s;
                ar[u] = c.re[a], ai[u] = c.im[a], br[u] = c.re[b], bi[u] = c.im[b];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t h = h0 + u * step;
            if (h < half) {
                const uint64_t a = insert_zero(h, q), b = a + s;
                pend_apply(p, a, ar[u], ai[u]);
                pend_apply(p, b, br[u], bi[u]);
                c.re[a] = d0r * ar[u] - d0i * ai[u];
                c.im[a] = d0r * ai[u] + d0i * ar[u];
                c.re[b] = d1r * br[u] - d1i * bi[u];
                c.im[b] = d1r * bi[u] + d1i * br[u];
            }
        }
    }
    p.n = 0;
    __syncthreads();
}

// Diagonal op on k <= KQ qubits; d = [re D | im D] (k == 1: d0r, d1r, d0i, d1i).
template <typename T, int KQ>
__device__ __forceinline__ void t_diag(const TrajCtx<T> c, Pend<T> &p, const uint32_t *qs, int k, const T *d) {
    const int D = 1 << k;
    if (k == 1) {
        t_diag1(c, p, qs[0], d[0], d[1], d[2], d[3]);
        return;
    } else {
        constexpr int U = kTrajBatch<T>;
        const uint64_t step = (uint64_t)blockDim.x;
        for (uint64_t i0 = threadIdx.x; i0 < c.len; i0 += U * step) {
            T ar[U], ai[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t i = i0 + u * step;
                if (i < c.len) ar[u] = c.re[i], ai[u] = c.im[i];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t i = i0 + u * step;
                if (i < c.len) {
                    uint32_t j = 0;
#pragma unroll
                    for (int b = 0; b < KQ; ++b)
                        if (b < k) j |= (uint32_t)((i >> qs[b]) & 1) << b;
                    pend_apply(p, i, ar[u], ai[u]);
                    const T wr = d[j], wi = d[D + j];
                    c.re[i] = wr * ar[u] - wi * ai[u];
                    c.im[i] = wr * ai[u] + wi * ar[u];
                }
            }
        }
    }
    p.n = 0;
    __syncthreads();
}

// apply_cx (src/state.rs:394-417).  With damping scales pending, every
// amplitude needs them, so the sweep then covers both halves.
template <typename T>
__device__ __forceinline__ void t_cx(const TrajCtx<T> c, Pend<T> &p, uint32_t ctl, uint32_t tgt) {
    const uint32_t lo = ctl < tgt ? ctl : tgt, hi = ctl < tgt ? tgt : ctl;
    constexpr int U = kTrajBatch<T>;
    const uint64_t quarter = c.len >> 2, step = (uint64_t)blockDim.x;
    const uint64_t cb = 1ull << ctl, tb = 1ull << tgt;
    if (p.n == 0) {
        for (uint64_t g0 = threadIdx.x; g0 < quarter; g0 += U * step) {
            T r0[U], r1[U], i0[U], i1[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t g = g0 + u * step;
                if (g < quarter) {
                    const uint64_t a = insert_zero(insert_zero(g, lo), hi) | cb, b = a | tb;
                    r0[u] = c.re[a], r1[u] = c.re[b], i0[u] = c.im[a], i1[u] = c.im[b];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t g = g0 + u * step;
                if (g < quarter) {
                    const uint64_t a = insert_zero(insert_zero(g, lo), hi) | cb, b = a | tb;
                    c.re[a] = r1[u];
                    c.re[b] = r0[u];
                    c.im[a] = i1[u];
                    c.im[b] = i0[u];
                }
            }
        }
    } else {
        constexpr int V = U / 2 > 0 ? U / 2 : 1;
        for (uint64_t g0 = threadIdx.x; g0 < quarter; g0 += V * step) {
            T xr[V][4], xi[V][4];
#pragma unroll
            for (int u = 0; u < V; ++u) {
                const uint64_t g = g0 + u * step;
                if (g < quarter) {
                    const uint64_t e = insert_zero(insert_zero(g, lo), hi);
                    const uint64_t idx[4] = {e, e | tb, e | cb, e | cb | tb};
#pragma unroll
                    for (int v = 0; v < 4; ++v) xr[u][v] = c.re[idx[v]], xi[u][v] = c.im[idx[v]];
                }
            }
#pragma unroll
            for (int u = 0; u < V; ++u) {
                const uint64_t g = g0 + u * step;
                if (g < quarter) {
                    const uint64_t e = insert_zero(insert_zero(g, lo), hi);
                    const uint64_t idx[4] = {e, e | tb, e | cb, e | cb | tb};
#pragma unroll
                    for (int v = 0; v < 4; ++v) pend_apply(p, idx[v], xr[u][v], xi[u][v]);
                    c.re[idx[0]] = xr[u][0];
                    c.im[idx[0]] = xi[u][0];
                    c.re[idx[1]] = xr[u][1];
                    c.im[idx[1]] = xi[u][1];
                    c.re[idx[2]] = xr[u][3];  // the control = 1 pair swaps
                    c.im[idx[2]] = xi[u][3];
                    c.re[idx[3]] = xr[u][2];
                    c.im[idx[3]] = xi[u][2];
                }
            }
        }
        p.n = 0;
    }
    __syncthreads();
}

template <typename T>
__device__ __forceinline__ void t_collapse(const TrajCtx<T> c, Pend<T> &p, uint32_t q, int outcome, T scale) {
    constexpr int U = kTrajBatch<T>;
    const uint64_t s = 1ull << q, half = c.len >> 1, step = (uint64_t)blockDim.x;
    for (uint64_t h0 = threadIdx.x; h0 < half; h0 += U * step) {
        T kr[U], ki[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t h = h0 + u * step;
            if (h < half) {
                const uint64_t keep = insert_zero(h, q) + (outcome ? s : 0);
                kr[u] = c.re[keep], ki[u] = c.im[keep];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t h = h0 + u * step;
            if (h < half) {
                const uint64_t a = insert_zero(h, q);
                const uint64_t keep = outcome ? a + s : a, zero = outcome ? a : a + s;
                pend_apply(p, keep, kr[u], ki[u]);
                c.re[keep] = kr[u] * scale;
                c.im[keep] = ki[u] * scale;
                c.re[zero] = T(0);
                c.im[zero] = T(0);
            }
        }
    }
    p.n = 0;
    __syncthreads();
}

// CpuBackend::measure (src/exec.rs:201-207)
template <typename T>
__device__ __forceinline__ int t_measure(const TrajCtx<T> c, Pend<T> &p, uint32_t q, double u) {
    const double p1 = t_prob_one(c, p, q);
    const int outcome = u < p1;
    const double pr = outcome ? p1 : 1.0 - p1;
    t_collapse(c, p, q, outcome, (T)(1.0 / sqrt(pr)));
    return outcome;
}

// noise.rs:91-100, through the same unitary / diagonal arithmetic the CPU
// path uses (signed zeros included).
template <typename T>
__device__ __forceinline__ void t_pauli(const TrajCtx<T> c, Pend<T> &p, uint32_t q, uint64_t code) {
    if (code == 1) t_1q(c, p, q, c.pauli + 0);
    else if (code == 2) t_1q(c, p, q, c.pauli + 8);
    else if (code == 3) t_diag1(c, p, q, T(1), T(-1), T(0), T(0));
}

// p1_first: prob_one of qs[0] formed by the gate's own sweep (or < 0); it
// stands until any Pauli changes the state.  k <= KQ.  Up to 3 qubits the
// loops over the op's qubits unroll (qs in registers); above, they stay
// loops over qs in the op record (uniform loads), one copy of their body.
template <typename T, int KQ>
__device__ __forceinline__ void t_gate_noise(const TrajCtx<T> c, Pend<T> &p, const TrajNoise &nz, const uint32_t *qs,
                                             int k, DevRng &rng, double p1_first = -1.0) {
    const double pd = k == 1 ? nz.dep1 : nz.dep2;
    if (pd > 0.0 && rng.f64() < pd) {
        const uint64_t idx = rng.range(1, 1ull << (2 * k));
        if constexpr (KQ <= kTrajMaxK) {
#pragma unroll
            for (int b = 0; b < KQ; ++b)
                if (b < k) t_pauli(c, p, qs[b], (idx >> (2 * b)) & 3);
        } else {
#pragma unroll 1
            for (int b = 0; b < k; ++b) t_pauli(c, p, qs[b], (idx >> (2 * b)) & 3);
        }
        p1_first = -1.0;
    }
    auto channels = [&](const int b) __attribute__((always_inline)) {
        const uint32_t q = qs[b];
        if (nz.bit_flip > 0.0 && rng.f64() < nz.bit_flip) t_pauli(c, p, q, 1), p1_first = -1.0;
        if (nz.phase_flip > 0.0 && rng.f64() < nz.phase_flip) t_pauli(c, p, q, 3), p1_first = -1.0;
        if (nz.damping > 0.0) {  // noise.rs:114-127
            const double u = rng.f64();
            const double pj = nz.damping * (b == 0 && p1_first >= 0.0 ? p1_first : t_prob_one(c, p, q));
            if (u < pj) {
                t_measure(c, p, q, 0.0);
                t_pauli(c, p, q, 1);
            } else {  // diag(sc, sc*sqrt(1-gamma)) on q, held for the next sweep
                const double sc = 1.0 / sqrt(1.0 - pj);
                if (p.n == kTrajMaxK) t_pend_flush(c, p);  // a gate on more than 3 qubits
                const T d0 = (T)sc, d1 = (T)(sc * sqrt(1.0 - nz.damping));
#pragma unroll
                for (int e = 0; e < kTrajMaxK; ++e)  // static indices keep Pend in registers
                    if (e == p.n) p.q[e] = q, p.d0[e] = d0, p.d1[e] = d1;
                ++p.n;
            }
        }
    };
    if constexpr (KQ <= kTrajMaxK) {
#pragma unroll
        for (int b = 0; b < KQ; ++b) {
            if (b >= k) break;
            channels(b);
        }
    } else {
#pragma unroll 1
        for (int b = 0; b < k; ++b) channels(b);
    }
}

template <typename T>
__device__ __forceinline__ int t_readout(const TrajNoise &nz, int bit, DevRng &rng) {
    const double p = bit ? nz.ro10 : nz.ro01;
    if (p > 0.0 && rng.f64() < p) return !bit;
    return bit;
}

// Everything one launch needs, in device memory, so that a captured graph
// of the launch (its only argument is this block's address) replays
// unchanged for any circuit whose launch block is written to the same place
// and that runs at the same grid.
template <typename T>
struct TrajLaunch {
    T *states;           // gridDim.x slots of `slot` values: re, im (2^n each), block sums
    const TrajOp *ops;
    const uint32_t *fm;  // final measures: (qubit, clbit) pairs
    const T *mats;
    uint64_t *out;       // per-shot clbit words
    uint64_t seed, shots;
    uint64_t slot;       // values of T per workgroup slot
    TrajNoise nz;
    uint32_t n, nops, nfm;
    int32_t noisy;
};

template <typename T, int KM>
__global__ __launch_bounds__(kBlock) void k_trajectories(const TrajLaunch<T> *__restrict__ L) {
    // ops on up to KQ qubits; the LDS stage holds a 2^KM dense matrix or a
    // 2^KQ diagonal (the same 2 << 2KM values in the wide instance)
    constexpr int KQ = KM == kTrajMaxK ? kTrajMaxK : kTrajOpQubits;
    constexpr int kStage = (2 << (2 * KM)) > (2 << KQ) ? (2 << (2 * KM)) : (2 << KQ);
    T *__restrict__ ws = L->states;
    const uint32_t n = L->n, nops = L->nops, nfm = L->nfm;
    const TrajOp *__restrict__ ops = L->ops;
    const uint32_t *__restrict__ fm = L->fm;
    const T *__restrict__ mats = L->mats;
    uint64_t *__restrict__ out = L->out;
    const TrajNoise nz = L->nz;
    const int noisy = L->noisy;
    const uint64_t seed = L->seed, shots = L->shots, slot = L->slot;
    __shared__ double red[9 * kBlock];
    __shared__ T smat[kStage];
    __shared__ T spauli[16];
    if (threadIdx.x < 16) {
        // X = [[0,1],[1,0]], Y = [[0,-i],[i,0]], each as [re 4 | im 4]
        const int t = threadIdx.x;
        spauli[t] = (t == 1 || t == 2 || t == 14) ? T(1) : t == 13 ? T(-1) : T(0);
    }
    __syncthreads();
    const uint64_t len = 1ull << n;
    // a 1q gate's sweep also forms the damping probability of its qubit
    const bool fuse_p1 = noisy && nz.damping > 0.0 && n >= 13;
    // this workgroup's slot: re, im, then (above 21 qubits) the block sums
    T *const mine = ws + (uint64_t)blockIdx.x * slot;
    double *const gsum = (len >> 1) / kReduceSpan > (uint64_t)kTrajLdsSumBlocks
                             ? reinterpret_cast<double *>(mine + 2 * len)
                             : nullptr;
    TrajCtx<T> c{mine, mine + len, len, red, spauli, gsum};
    for (uint64_t shot = blockIdx.x; shot < shots; shot += gridDim.x) {
        for (uint64_t i = threadIdx.x; i < len; i += blockDim.x) {
            c.re[i] = i == 0 ? T(1) : T(0);
            c.im[i] = T(0);
        }
        __syncthreads();
        DevRng rng;
        rng.seed(DevRng::mix(seed ^ shot));
        Pend<T> pend;
        uint64_t clbits = 0;
        for (uint32_t oi = 0; oi < nops; ++oi) {
            const TrajOp &op = ops[oi];
            if (op.is_if) {
                const uint64_t mask = op.creg_len >= 64 ? ~0ull : ((1ull << op.creg_len) - 1);
                if (((clbits >> op.creg_off) & mask) != op.value) continue;
            }
            const uint32_t kind = op.kind, k = op.k;
            // the narrow instance holds the op's qubits in registers; the
            // wide one reads them from the op record
            uint32_t qreg[kTrajMaxK];
#pragma unroll
            for (int b = 0; b < kTrajMaxK; ++b) qreg[b] = op.q[b];
            const uint32_t *qs = KM == kTrajMaxK ? qreg : op.q;
            if (kind <= 1) {  // stage the matrix / diagonal in LDS
                const uint32_t cnt = kind == 0 ? 2u << (2 * k) : 2u << k;
                for (uint32_t i = threadIdx.x; i < cnt; i += blockDim.x) smat[i] = mats[op.data_off + i];
                __syncthreads();
            }
            switch (kind) {
                case 0: {
                    double p1 = -1.0;
                    if (k == 1) p1 = t_1q(c, pend, qs[0], smat, fuse_p1);
                    else if (k == 2) t_kq<T, 2>(c, pend, qs, smat);
                    else if (k == 3) t_kq<T, 3>(c, pend, qs, smat);
                    else if constexpr (KM > kTrajMaxK) {
                        if (k == 4) t_kq_rows<T, 4>(c, pend, qs, smat);
                        else if (k == 5) t_kq_rows<T, 5>(c, pend, qs, smat);
                        else t_kq_rows<T, 6>(c, pend, qs, smat);
                    }
                    if (noisy) t_gate_noise<T, KQ>(c, pend, nz, qs, (int)k, rng, p1);
                    break;
                }
                case 1:
                    t_diag<T, KQ>(c, pend, qs, (int)k, smat);
                    if (noisy) t_gate_noise<T, KQ>(c, pend, nz, qs, (int)k, rng);
                    break;
                case 2: {
                    t_cx(c, pend, qs[0], qs[1]);
                    if (noisy) {
                        const uint32_t sq[2] = {qs[0] < qs[1] ? qs[0] : qs[1], qs[0] < qs[1] ? qs[1] : qs[0]};
                        t_gate_noise<T, 2>(c, pend, nz, sq, 2, rng);
                    }
                    break;
                }
                case 3: {
                    int bit = t_measure(c, pend, qs[0], rng.f64());
                    if (noisy) bit = t_readout<T>(nz, bit, rng);
                    clbits = (clbits & ~(1ull << op.clbit)) | ((uint64_t)bit << op.clbit);
                    break;
                }
                case 4:
                    if (t_measure(c, pend, qs[0], rng.f64())) t_pauli(c, pend, qs[0], 1);
                    break;
            }
        }
        for (uint32_t f = 0; f < nfm; ++f) {
            int bit = t_measure(c, pend, fm[2 * f], rng.f64());
            if (noisy) bit = t_readout<T>(nz, bit, rng);
            clbits = (clbits & ~(1ull << fm[2 * f + 1])) | ((uint64_t)bit << fm[2 * f + 1]);
        }
        if (threadIdx.x == 0) out[shot] = clbits;
        __syncthreads();
    }
}

}  // namespace dev
}  // namespace zeno

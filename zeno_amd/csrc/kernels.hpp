// kernels.hpp — CDNA4 (gfx950) HIP kernels of the zeno state-vector backend.
//
// The state is split complex (SoA): re[2^n], im[2^n] of T = float or double
// in HBM.  Every kernel below reproduces the reference CPU path's
// expression tree for its op (src/state.rs), and the library is compiled
// with -ffp-contract=off, so no a*b+c is contracted to an FMA: results are
// bit-identical to the scalar CPU path for the same inputs.  Two opt-in
// knobs trade that for speed within tolerance: "fma" (fused multiply-adds
// in the same expression order) and "mfma" (dense 4..6-qubit ops on the
// matrix cores).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "devcommon.hpp"
#include "kq_tiles.hpp"

namespace zeno {
namespace dev {

constexpr int kMaxDiagK = 12;
constexpr int kMaxFusedK = 6;

template <typename T>
struct Mat2 {
    T r[4], i[4];
};

// Qubits of one op: q[b] is the state bit of the op's b-th qubit (the
// order the matrix or table is written in); s[] holds the same bits sorted
// ascending, for inserting the zero bits of a group's base index.
struct Qubits {
    uint32_t q[kMaxDiagK];
    uint32_t s[kMaxDiagK];
    int k;
};

// Qubit layout: qubit q of the circuit is bit phys[q] of the state index in
// HBM.  lut holds 5 tables of 256 entries, one per byte of a logical index,
// whose OR is the physical index (null = identity).  Reductions, sampling
// and state copies walk logical indices in the reference's order and read
// through it; op kernels take physical bits directly.
struct Layout {
    const uint64_t *lut;
};

__device__ __forceinline__ uint64_t phys_index(const Layout &L, uint64_t l) {
    if (!L.lut) return l;
    return L.lut[l & 255] | L.lut[256 + ((l >> 8) & 255)] | L.lut[512 + ((l >> 16) & 255)] |
           L.lut[768 + ((l >> 24) & 255)] | L.lut[1024 + ((l >> 32) & 255)];
}
// The layout permutes index bits, so it distributes over disjoint bit sets:
// phys_index(a | b) = phys_index(a) | phys_index(b).  Walks over logical
// indices split each index into a lane part (looked up once per lane) and a
// part shared by the workgroup (the same entries for every lane).

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
template <typename T, bool F = false>
__global__ __launch_bounds__(kBlock) void k_1q(T *__restrict__ re, T *__restrict__ im, uint32_t q, uint64_t half,
                                               Mat2<T> m) {
    const uint64_t s = 1ull << q;
    for (uint64_t h = gtid(); h < half; h += gstride()) {
        const uint64_t a = insert_zero(h, q), b = a + s;
        const T ar = re[a], ai = im[a], br = re[b], bi = im[b];
        re[a] = msc<F>(mac<F>(msc<F>(m.r[0] * ar, m.i[0], ai), m.r[1], br), m.i[1], bi);
        im[a] = mac<F>(mac<F>(mac<F>(m.r[0] * ai, m.i[0], ar), m.r[1], bi), m.i[1], br);
        re[b] = msc<F>(mac<F>(msc<F>(m.r[2] * ar, m.i[2], ai), m.r[3], br), m.i[3], bi);
        im[b] = mac<F>(mac<F>(mac<F>(m.r[2] * ai, m.i[2], ar), m.r[3], bi), m.i[3], br);
    }
}

// -------------------------------------------------------------- kq dense
// kern_kq_scalar (src/state.rs:355-392): gather 2^K amplitudes of a group,
// row-major matvec with ar = (ar + wr*xr) - wi*xi, ai = (ai + wr*xi) + wi*xr
// over ascending columns, scatter.  The matrix ([re | im], 2*4^K of T) is
// the same for every lane: its elements are read with scalar loads into
// scalar registers and enter the arithmetic from there.
template <typename T, int K, bool F = false>
__global__ __launch_bounds__(kBlock) void k_kq(T *__restrict__ re, T *__restrict__ im, const T *__restrict__ mat,
                                               Qubits qs, uint64_t groups) {
    constexpr int D = 1 << K;
    const T *__restrict__ sm = mat;
    uint64_t bit[K];
#pragma unroll
    for (int b = 0; b < K; ++b) bit[b] = 1ull << qs.q[b];
    for (uint64_t g = gtid(); g < groups; g += gstride()) {
        uint64_t base = g;
#pragma unroll
        for (int b = 0; b < K; ++b) base = insert_zero(base, qs.s[b]);
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
        constexpr int kRowUnroll = K <= 3 ? D : 1;
#pragma unroll kRowUnroll
        for (int r = 0; r < D; ++r) {
            T ar = T(0), ai = T(0);
            const T *wr = sm + r * D, *wi = sm + D * D + r * D;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                ar = msc<F>(mac<F>(ar, wr[c], xr[c]), wi[c], xi[c]);
                ai = mac<F>(mac<F>(ai, wr[c], xi[c]), wi[c], xr[c]);
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

// ------------------------------------------------------ kq on matrix cores
// The same op as k_kq for K = 4..6, as a real GEMM on the matrix cores
// (opt-in: MFMA accumulation does not round like kern_kq_scalar's sum, so
// results agree with the CPU path within tolerance, not bit for bit).
// With D = 2^K, y = M x is [yr; yi] = [[Mr, -Mi], [Mi, Mr]] [xr; xi], a
// 2D x 2D real matrix W.  A workgroup loads a tile of 2^12 amplitudes that
// holds the op's K state bits, so the tile is G = 2^(12-K) whole groups,
// and stores it in LDS as X[2D][G] (row j < D: re of amplitude j of every
// group, row D + j: its im).  Y = W X is 32 output blocks of 16 x 16
// (v_mfma_*_16x16x4: A[i][k] in lane i + 16k, B[k][j] in lane j + 16k; D[i][j]
// in lane j + 16(i/4), element i%4 for f32, lane j + 16(i%4), element i/4 for
// f64); each of the 8 waves computes 4 blocks of one row band, holding that
// band of W in registers for all tiles.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct MfmaOps;
template <>
struct MfmaOps<float> {
    using Acc = f32x4;
    static __device__ __forceinline__ Acc step(float a, float b, Acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // row of a 16 x 16 output block held by element v of lane ln
    static __device__ __forceinline__ int row(int ln, int v) { return 4 * (ln >> 4) + v; }
};
template <>
struct MfmaOps<double> {
    using Acc = f64x4;
    static __device__ __forceinline__ Acc step(double a, double b, Acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int ln, int v) { return (ln >> 4) + 4 * v; }
};

template <typename T, int K>
__global__ __launch_bounds__(kKqThreads) void k_kq_mfma(T *__restrict__ re, T *__restrict__ im,
                                                          const T *__restrict__ w, KqTile mt, uint64_t tiles) {
    using Ops = MfmaOps<T>;
    using Acc = typename Ops::Acc;
    constexpr int TB = kKqTileBits;
    constexpr int D = 1 << K, D2 = 2 * D, G = 1 << (TB - K), SP = G + 16;
    constexpr int CB = G / 16, KS = D2 / 4;
    static_assert((32 * 16 * 16) == D2 * G, "tile shape");
    __shared__ T sx[D2 * SP];
    const int t = threadIdx.x, wv = t >> 6, ln = t & 63;
    const int rb = (wv * 4) / CB, cb0 = (wv * 4) % CB;
    T a[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) a[ks] = w[(16 * rb + (ln & 15)) * D2 + 4 * ks + (ln >> 4)];
    kq_tile_loop<T, K, SP>(re, im, mt, tiles, sx, [&] {
        Acc acc[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = Acc{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const T b = sx[(4 * ks + (ln >> 4)) * SP + 16 * (cb0 + i) + (ln & 15)];
                acc[i] = Ops::step(a[ks], b, acc[i]);
            }
        }
        __syncthreads();  // every wave has read X
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                sx[(16 * rb + Ops::row(ln, v)) * SP + 16 * (cb0 + i) + (ln & 15)] = acc[i][v];
    });
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
template <typename T, bool F = false>
__global__ __launch_bounds__(kBlock) void k_diag1(T *__restrict__ re, T *__restrict__ im, uint32_t q, uint64_t half,
                                                  T d0r, T d0i, T d1r, T d1i) {
    const uint64_t s = 1ull << q;
    for (uint64_t h = gtid(); h < half; h += gstride()) {
        const uint64_t a = insert_zero(h, q), b = a + s;
        const T ar = re[a], ai = im[a], br = re[b], bi = im[b];
        re[a] = msc<F>(d0r * ar, d0i, ai);
        im[a] = mac<F>(d0r * ai, d0i, ar);
        re[b] = msc<F>(d1r * br, d1i, bi);
        im[b] = mac<F>(d1r * bi, d1i, br);
    }
}

// kern_diagk_scalar (src/state.rs:577-596): table index j gathered from the
// bits of i at qs; table ([re | im], 2*2^k of T) staged in LDS.
template <typename T, bool F = false>
__global__ __launch_bounds__(kBlock) void k_diagk(T *__restrict__ re, T *__restrict__ im, const T *__restrict__ tab,
                                                  Qubits qs, uint64_t len) {
    // the table ([re | im], 2 * 2^k values) in dynamic LDS of its own size,
    // so small tables leave room for more workgroups per CU
    extern __shared__ __align__(16) unsigned char diag_lds[];
    T *const st = reinterpret_cast<T *>(diag_lds);
    const int dim = 1 << qs.k;
    for (int i = threadIdx.x; i < 2 * dim; i += blockDim.x) st[i] = tab[i];
    __syncthreads();
    // Element i = row + t: a workgroup's rows are multiples of kBlock, so the
    // table bits on qubits below log2(kBlock) come from the lane (once) and
    // the others from the row (uniform per iteration).
    static_assert(kBlock == 256, "lane bits are index bits 0..7");
    const uint32_t t = threadIdx.x;
    uint32_t jl = 0;
    for (int b = 0; b < qs.k; ++b)
        if (qs.q[b] < 8) jl |= ((t >> qs.q[b]) & 1) << b;
    for (uint64_t row = (uint64_t)blockIdx.x * kBlock; row < len; row += gstride()) {
        uint32_t jr = 0;
        for (int b = 0; b < qs.k; ++b)
            if (qs.q[b] >= 8) jr |= (uint32_t)((row >> qs.q[b]) & 1) << b;
        const uint64_t i = row + t;
        if (i >= len) break;
        const uint32_t j = jl | jr;
        const T wr = st[j], wi = st[dim + j];
        const T ar = re[i], ai = im[i];
        re[i] = msc<F>(wr * ar, wi, ai);
        im[i] = mac<F>(wr * ai, wi, ar);
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
// indexed by half index (prob_one, src/state.rs:598-626).  e and q are
// logical; the amplitude is read through the layout.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_reduce_state(const T *__restrict__ re, const T *__restrict__ im,
                                                         Layout L, uint64_t n_elems, int mode, uint32_t q,
                                                         double *__restrict__ out) {
    __shared__ double s[kBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kReduceSpan;
    // logical index insert_zero(e, q) | 2^q (mode 1) or e, where e = base +
    // j*kBlock + lane: insert_zero maps every bit on its own, so the lane's
    // bits, the block's, row j's (bits 8..11 of e, at most bit 12 after the
    // insertion: one byte-1 entry of the layout) and the q bit are disjoint
    static_assert(kBlock == 256 && kReduceSpan == 16 * kBlock, "row j holds bits 8..11 of e");
    const uint64_t t = threadIdx.x;
    const uint64_t lane = phys_index(L, mode ? insert_zero(t, q) : t) | (mode ? phys_index(L, 1ull << q) : 0);
    const uint64_t blk = phys_index(L, mode ? insert_zero(base, q) : base);
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < kReduceSpan / kBlock; ++j) {
        const uint64_t jr = (uint64_t)j * kBlock;
        if (base + jr + t < n_elems) {
            const uint64_t jl = mode ? insert_zero(jr, q) : jr;
            const uint64_t idx = lane | blk | (L.lut ? L.lut[256 + (jl >> 8)] : jl);
            const double r = (double)re[idx], m = (double)im[idx];
            a = a + (r * r + m * m);
        }
    }
    const double tot = block_tree(a, s);
    if (threadIdx.x == 0) out[blockIdx.x] = tot;
}

static __global__ __launch_bounds__(kBlock) void k_reduce_f64(const double *__restrict__ in, uint64_t n,
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

// ------------------------------------------------------------ tiled window
// One HBM pass applies a window of consecutive ops.  A tile is the 2^12
// amplitudes sharing every index bit outside the window's qubit set Q
// (|Q| = 12, Q sorted; local bit i <-> qubit Q[i]; Q holds the lowest
// qubits, so a tile's HBM accesses are 128-byte runs).  The tile is staged
// in LDS as (re, im) pairs with one pad pair per 16 (bank spread).  The window is cut into
// stages: a stage fixes NS local "slot" bits; each lane gathers the 2^NS
// amplitudes that differ only in those bits into registers, applies the
// stage's ops there with exactly the per-op kernels' arithmetic, and
// scatters them back.  A workgroup has 2^(TB-NS) lanes.  Ops admitted to a
// window: 1q dense and dense k <= NS whose qubits are slots of one stage;
// diagonals on any qubits (elementwise: bits outside Q are fixed per tile);
// CX whose target is a slot (its control may be a slot, another local bit,
// or fixed per tile).  Window descriptors are uniform (scalar loads).
constexpr int kMaxTileBits = 13;
constexpr int kMaxSlots = 4;
constexpr uint32_t kFixedBit = 0x80000000u;  // table bit outside Q: qubit in the low bits
constexpr int kDescBytes = 65536;           // window descriptor blob limit

enum WinKind : uint32_t { kW1q = 0, kWDense = 1, kWDiag = 2, kWCx = 3 };
enum CxCtrl : uint32_t { kCtrlSlot = 0, kCtrlLocal = 1, kCtrlFixed = 2 };

struct WinOp {
    uint32_t kind;
    uint32_t sel;       // 1q: slot; dense: slot mask; cx: target slot | ctrl slot << 2 | mode << 4
    uint32_t aux;       // cx: control local bit (mode 1) or qubit (mode 2)
    uint32_t k;         // diag: table bits
    uint32_t data_off;  // offset of the matrix (blob, in T) or diagonal table (tables, in T)
    uint32_t pad[3];
    uint32_t bits[kMaxDiagK];  // diag: local bit of table bit b, or kFixedBit | qubit
};

struct WinStage {
    // register bit i of a lane's index <-> local bit slot[i] (first NS
    // used, in any order: a dense op's matrix bits take increasing register
    // bits); ins[] holds the same bits ascending, where the lane index
    // takes zero bits
    uint32_t slot[kMaxSlots];
    uint32_t op_begin, op_end;
    uint32_t ins[kMaxSlots];
};

struct WinHeader {
    uint32_t m, nstages, nops, desc_words;
    uint32_t stage_off, op_off, data_off, nslots;  // word offsets inside the blob
    uint32_t q[kMaxTileBits];
    // Relabelling of the tile's qubits on write-back (relabel != 0): the
    // amplitude written back to local index u is read from the LDS position
    // whose bit wb[i] is bit i of u, so the tile leaves in the permuted
    // layout.  wb is a bit permutation (the inverse of the relabelling's).
    uint32_t relabel;
    uint32_t wb[kMaxTileBits];
};

template <typename T>
struct Cplx;
template <>
struct Cplx<double> {
    using type = double2;
};
template <>
struct Cplx<float> {
    using type = float2;
};

__device__ __forceinline__ int padded(int l) { return l + (l >> 4); }

// The window's ops run in a loop over a runtime op list, so the lane's
// registers are loop-carried and every op case must leave each value in the
// register it came in: a case whose results land elsewhere makes the
// register allocator put a full set of copies on the loop's common latch,
// paid by every op.  Hence (1) the arithmetic ops finish every input read
// before the first result is written (a result can then take its input's
// register), and (2) pure permutations swap in place through tied asm
// operands.

// Swap two register values in place (v_swap_b32 per 32-bit half).
template <typename T>
__device__ __forceinline__ void rswap(T &a, T &b) {
    if constexpr (sizeof(T) == 8) {
        uint32_t *pa = reinterpret_cast<uint32_t *>(&a), *pb = reinterpret_cast<uint32_t *>(&b);
        asm volatile("v_swap_b32 %0, %1\n\tv_swap_b32 %2, %3" : "+v"(pa[0]), "+v"(pb[0]), "+v"(pa[1]), "+v"(pb[1]));
    } else {
        asm volatile("v_swap_b32 %0, %1" : "+v"(a), "+v"(b));
    }
}

template <typename T, int NS, int S, bool F>
__device__ __forceinline__ void reg_1q(T *xr, T *xi, const T *__restrict__ d) {
    const T m00r = d[0], m01r = d[1], m10r = d[2], m11r = d[3];
    const T m00i = d[4], m01i = d[5], m10i = d[6], m11i = d[7];
#pragma unroll
    for (int r = 0; r < (1 << NS); ++r) {
        if ((r >> S) & 1) continue;
        const int a = r, c = r | (1 << S);
        const T ar = xr[a], ai = xi[a], br = xr[c], bi = xi[c];
        if constexpr (F) {
            xr[a] = msc<F>(mac<F>(msc<F>(m00r * ar, m00i, ai), m01r, br), m01i, bi);
            xi[a] = mac<F>(mac<F>(mac<F>(m00r * ai, m00i, ar), m01r, bi), m01i, br);
            xr[c] = msc<F>(mac<F>(msc<F>(m10r * ar, m10i, ai), m11r, br), m11i, bi);
            xi[c] = mac<F>(mac<F>(mac<F>(m10r * ai, m10i, ar), m11r, bi), m11i, br);
        } else {
            // the same roundings as ((m00r*ar - m00i*ai) + m01r*br) - m01i*bi,
            // with each last product formed before any result is written
            const T p0 = mac<F>(msc<F>(m00r * ar, m00i, ai), m01r, br);
            const T p1 = mac<F>(mac<F>(m00r * ai, m00i, ar), m01r, bi);
            const T p2 = mac<F>(msc<F>(m10r * ar, m10i, ai), m11r, br);
            const T p3 = mac<F>(mac<F>(m10r * ai, m10i, ar), m11r, bi);
            const T q0 = m01i * bi, q1 = m01i * br, q2 = m11i * bi, q3 = m11i * br;
            xr[a] = p0 - q0;
            xi[a] = p1 + q1;
            xr[c] = p2 - q2;
            xi[c] = p3 + q3;
        }
    }
}

// Dense op on the slots of mask M (matrix local bit b <-> b-th set bit).
template <typename T, int NS, int M, bool F>
__device__ __forceinline__ void reg_dense(T *xr, T *xi, const T *__restrict__ d) {
    constexpr int K = __builtin_popcount(M);
    constexpr int D = 1 << K;
    int pos[K > 0 ? K : 1];
    {
        int b = 0;
#pragma unroll
        for (int s = 0; s < NS; ++s)
            if ((M >> s) & 1) pos[b++] = s;
    }
#pragma unroll
    for (int g = 0; g < (1 << NS); ++g) {
        if (g & M) continue;  // g enumerates the register bits outside M
        T yr[D], yi[D], vr[D], vi[D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            int r = g;
#pragma unroll
            for (int b = 0; b < K; ++b)
                if ((j >> b) & 1) r |= 1 << pos[b];
            vr[j] = xr[r];
            vi[j] = xi[r];
        }
#pragma unroll
        for (int row = 0; row < D; ++row) {
            T ar = T(0), ai = T(0);
#pragma unroll
            for (int c = 0; c < D; ++c) {
                const T wr = d[row * D + c], wi = d[D * D + row * D + c];
                ar = msc<F>(mac<F>(ar, wr, vr[c]), wi, vi[c]);
                ai = mac<F>(mac<F>(ai, wr, vi[c]), wi, vr[c]);
            }
            yr[row] = ar;
            yi[row] = ai;
        }
#pragma unroll
        for (int j = 0; j < D; ++j) {
            int r = g;
#pragma unroll
            for (int b = 0; b < K; ++b)
                if ((j >> b) & 1) r |= 1 << pos[b];
            xr[r] = yr[j];
            xi[r] = yi[j];
        }
    }
}

template <typename T, int NS, int ST, int SC>
__device__ __forceinline__ void reg_cx_slot(T *xr, T *xi) {
#pragma unroll
    for (int r = 0; r < (1 << NS); ++r) {
        if (((r >> ST) & 1) || !((r >> SC) & 1)) continue;
        const int c = r | (1 << ST);
        rswap(xr[r], xr[c]);
        rswap(xi[r], xi[c]);
    }
}

// CX whose control is fixed for the whole tile and set: swap every pair.
template <typename T, int NS, int ST>
__device__ __forceinline__ void reg_cx_swap(T *xr, T *xi) {
#pragma unroll
    for (int r = 0; r < (1 << NS); ++r) {
        if ((r >> ST) & 1) continue;
        const int c = r | (1 << ST);
        rswap(xr[r], xr[c]);
        rswap(xi[r], xi[c]);
    }
}

// Swap the re and im values of two registers in the lanes of mask m only:
// exec is narrowed to m around plain in-place swaps and restored in the same
// asm block (every lane of a window workgroup is active here).
template <typename T>
__device__ __forceinline__ void mswap(T &ar, T &ai, T &br, T &bi, uint64_t m) {
    uint64_t sv;
    if constexpr (sizeof(T) == 8) {
        uint32_t *a0 = reinterpret_cast<uint32_t *>(&ar), *a1 = reinterpret_cast<uint32_t *>(&ai);
        uint32_t *b0 = reinterpret_cast<uint32_t *>(&br), *b1 = reinterpret_cast<uint32_t *>(&bi);
        asm volatile(
            "s_and_saveexec_b64 %[sv], %[m]\n\t"
            "v_swap_b32 %0, %1\n\tv_swap_b32 %2, %3\n\tv_swap_b32 %4, %5\n\tv_swap_b32 %6, %7\n\t"
            "s_mov_b64 exec, %[sv]"
            : "+v"(a0[0]), "+v"(b0[0]), "+v"(a0[1]), "+v"(b0[1]), "+v"(a1[0]), "+v"(b1[0]), "+v"(a1[1]),
              "+v"(b1[1]), [sv] "=&s"(sv)
            : [m] "s"(m)
            : "scc");
    } else {
        asm volatile(
            "s_and_saveexec_b64 %[sv], %[m]\n\tv_swap_b32 %0, %1\n\tv_swap_b32 %2, %3\n\ts_mov_b64 exec, %[sv]"
            : "+v"(ar), "+v"(br), "+v"(ai), "+v"(bi), [sv] "=&s"(sv)
            : [m] "s"(m)
            : "scc");
    }
}

// CX whose control is a lane bit: swap along the target slot in the lanes
// whose control bit is set (m: their lane mask).
template <typename T, int NS, int ST>
__device__ __forceinline__ void reg_cx_all(T *xr, T *xi, uint64_t m) {
#pragma unroll
    for (int r = 0; r < (1 << NS); ++r) {
        if ((r >> ST) & 1) continue;
        const int c = r | (1 << ST);
        mswap(xr[r], xi[r], xr[c], xi[c], m);
    }
}

template <typename T, int NS, int M, bool F>
__device__ __forceinline__ void reg_dense_if(T *xr, T *xi, const T *__restrict__ d) {
    if constexpr (M < (1 << NS)) reg_dense<T, NS, M, F>(xr, xi, d);
}

template <typename T, int NS, bool F>
__device__ __forceinline__ void reg_dense_dispatch(T *xr, T *xi, uint32_t mask, const T *__restrict__ d) {
    switch (mask) {
        case 0x3: reg_dense_if<T, NS, 0x3, F>(xr, xi, d); break;
        case 0x5: reg_dense_if<T, NS, 0x5, F>(xr, xi, d); break;
        case 0x6: reg_dense_if<T, NS, 0x6, F>(xr, xi, d); break;
        case 0x7: reg_dense_if<T, NS, 0x7, F>(xr, xi, d); break;
        case 0x9: reg_dense_if<T, NS, 0x9, F>(xr, xi, d); break;
        case 0xA: reg_dense_if<T, NS, 0xA, F>(xr, xi, d); break;
        case 0xC: reg_dense_if<T, NS, 0xC, F>(xr, xi, d); break;
        case 0xB: reg_dense_if<T, NS, 0xB, F>(xr, xi, d); break;
        case 0xD: reg_dense_if<T, NS, 0xD, F>(xr, xi, d); break;
        case 0xE: reg_dense_if<T, NS, 0xE, F>(xr, xi, d); break;
        case 0xF: reg_dense_if<T, NS, 0xF, F>(xr, xi, d); break;
    }
}

// Shape of the tiled kernel: a tile of 2^TB amplitudes, NS register
// slots per lane, hence 2^(TB-NS) lanes; the padded tile lives in LDS and
// fixes how many workgroups share a CU (160 KiB of LDS): the 12-bit f64
// tile takes 68 KiB (two per CU), the 13-bit one 136 KiB (one).
template <typename T, int TB, int NS>
struct WinShape {
    static constexpr int kTile = 1 << TB;
    static constexpr int kTilePad = kTile + (kTile >> 4);
    static constexpr int kLanes = kTile >> NS;
    static constexpr int kRegs = 1 << NS;
    static constexpr int kLdsBytes = 2 * kTilePad * (int)sizeof(T);
    static constexpr int kPerCu = (160 * 1024) / kLdsBytes < 1 ? 1 : (160 * 1024) / kLdsBytes;
    // occupancy asked of the compiler: two workgroups per CU for 12-bit
    // tiles, one for 13-bit ones, at most 4 waves per SIMD (128 VGPRs) ...
    static constexpr int kWant = (TB == 12 ? 2 : 1) * kLanes / 256;
    static constexpr int kWavesPerEu = kWant < 1 ? 1 : (kWant > 4 ? 4 : kWant);
    // ... or, for the 3-slot windows without dense ops (which fit without
    // spilling), as many workgroups per CU as the LDS holds, at most 8 waves
    // per SIMD (f32: 4 workgroups, 64 VGPRs)
    static constexpr int kWantLds = kPerCu * (kLanes / 64) / 4;
    static constexpr int kWavesPerEuLight = kWantLds < kWavesPerEu ? kWavesPerEu : (kWantLds > 8 ? 8 : kWantLds);
};

// kDense: the window holds dense k >= 2 ops; kPrefetch: the next tile's HBM
// loads go to registers while the current tile computes (separate instances
// keep the register budget of the common 1q/CX/diagonal windows low).
// Global addressing: row j of a tile starts at the uniform element offset
// base | hi(j); a lane adds its offset lo (its low local bits).
template <typename T, int TB, int NS, bool kDense, bool kPrefetch, bool F>
__global__ __attribute__((amdgpu_flat_work_group_size(1, WinShape<T, TB, NS>::kLanes)))
__attribute__((amdgpu_waves_per_eu(NS == 3 && !kDense ? WinShape<T, TB, NS>::kWavesPerEuLight
                                                       : WinShape<T, TB, NS>::kWavesPerEu))) void k_window(
    T *__restrict__ re, T *__restrict__ im, const uint32_t *__restrict__ blob, const T *__restrict__ tables,
    uint64_t tiles) {
    using Shape = WinShape<T, TB, NS>;
    constexpr int B = Shape::kLanes, R = Shape::kRegs;
    constexpr int LB = TB - NS;  // local bits carried by the lane index
    // re and im of an amplitude side by side: one 16-byte (f64) or 8-byte
    // (f32) LDS access per amplitude (same-box A/B against two arrays on
    // the 30q circuit: f64 windows 3% and f32 windows 1.5% faster)
    using T2 = typename Cplx<T>::type;
    __shared__ T2 sx[Shape::kTilePad];
    auto lds_put = [&](int i, T a, T b) { sx[i] = T2{a, b}; };
    auto lds_get = [&](int i, T &a, T &b) {
        const T2 v = sx[i];
        a = v.x, b = v.y;
    };
    const int t = threadIdx.x;
    const WinHeader *hdr = reinterpret_cast<const WinHeader *>(blob);
    const WinStage *stages = reinterpret_cast<const WinStage *>(blob + hdr->stage_off);
    const WinOp *ops = reinterpret_cast<const WinOp *>(blob + hdr->op_off);
    const T *mats = reinterpret_cast<const T *>(blob + hdr->data_off);
    const int nst = (int)hdr->nstages;
    uint64_t lo = 0;  // the lane's element offset within a row
#pragma unroll
    for (int i = 0; i < LB; ++i)
        if ((t >> i) & 1) lo |= 1ull << hdr->q[i];
    const int lt = padded(t);  // LDS slot of local index t (+ (B + B/16)*j for row j)
    // Write-back read positions of local index t + B*j: wlt + woff[j].  A
    // relabelling window reads through its bit map (bit i of the index
    // selects LDS bit wb[i]); the map is a bit permutation, so the lane part
    // and the row part stay disjoint and padded() splits over them.  woff
    // is uniform: computed once per launch.
    int wlt = lt, woff[R];
#pragma unroll
    for (int j = 0; j < R; ++j) woff[j] = (B + B / 16) * j;
    if (hdr->relabel != 0) {
        int v = 0;
#pragma unroll
        for (int i = 0; i < LB; ++i)
            if ((t >> i) & 1) v |= 1 << hdr->wb[i];
        wlt = padded(v);
#pragma unroll
        for (int j = 0; j < R; ++j) {
            int u = 0;
#pragma unroll
            for (int i = LB; i < TB; ++i)
                if ((j >> (i - LB)) & 1) u |= 1 << hdr->wb[i];
            woff[j] = padded(u);
        }
    }
    // Workgroups are dealt to the 8 XCDs in turn, so workgroups w and w + 8
    // share an XCD (and its L2) and start together: within each 16 tiles
    // they take tiles 2i and 2i + 1, which differ in the lowest state bit
    // outside the tile and may share cache lines.
    const bool pair = tiles % 16 == 0;
    auto tile_base = [&](uint64_t tile) {
        uint64_t b = pair ? (tile & ~15ull) | ((tile & 7) << 1) | ((tile >> 3) & 1) : tile;
#pragma unroll
        for (int i = 0; i < TB; ++i) b = insert_zero(b, hdr->q[i]);
        return b;
    };
    auto row = [&](uint64_t base, int j) {
        uint64_t h = base;
#pragma unroll
        for (int i = LB; i < TB; ++i)
            if ((j >> (i - LB)) & 1) h |= 1ull << hdr->q[i];
        return h;
    };
    T pr[R], pi[R];
    // The lane's offset lo and a row's uniform index have disjoint bits, so
    // element row | lo is at (array + lo) + row: one uniform add per access.
    T *__restrict__ const rel = re + lo;
    T *__restrict__ const iml = im + lo;
    uint64_t tile = blockIdx.x;
    uint64_t base = tile < tiles ? tile_base(tile) : 0;
    if (kPrefetch && tile < tiles) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const uint64_t g = row(base, j);
            pr[j] = rel[g];
            pi[j] = iml[g];
        }
    }
    // Write-back registers of the previous tile.  Its global stores are
    // issued only after this tile's loads have landed in LDS: the vector
    // memory counter retires in issue order, so a store issued before the
    // wait on a prefetched load would be waited on with it.
    T wr[R], wi[R];
    uint64_t wbase = 0;
    bool pending = false;
    while (tile < tiles) {
        if constexpr (!kPrefetch) {
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const uint64_t g = row(base, j);
                pr[j] = rel[g];
                pi[j] = iml[g];
            }
        }
#pragma unroll
        for (int j = 0; j < R; ++j) {
            lds_put(lt + (B + B / 16) * j, pr[j], pi[j]);
        }
        if (pending) {
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const uint64_t g = row(wbase, j);
                rel[g] = wr[j];
                iml[g] = wi[j];
            }
        }
        __syncthreads();
        const uint64_t next = tile + gridDim.x;
        const uint64_t nbase = next < tiles ? tile_base(next) : 0;
        if constexpr (kPrefetch) {
            if (next < tiles) {
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    const uint64_t g = row(nbase, j);
                    pr[j] = rel[g];
                    pi[j] = iml[g];
                }
            }
        }
        for (int s = 0; s < nst; ++s) {
            const WinStage *st = stages + s;
            int sl[NS];
#pragma unroll
            for (int i = 0; i < NS; ++i) sl[i] = (int)st->slot[i];
            int lb = t;
#pragma unroll
            for (int i = 0; i < NS; ++i) lb += lb & -(1 << st->ins[i]);  // insert a zero bit at ins[i]
            const int plb = padded(lb);
            T xr[R], xi[R];
            // lb and the register bits are disjoint, so padded(lb | c) =
            // padded(lb) + padded(c): a lane base plus a uniform offset
#pragma unroll
            for (int r = 0; r < R; ++r) {
                int c = 0;
#pragma unroll
                for (int i = 0; i < NS; ++i) c |= ((r >> i) & 1) << sl[i];
                lds_get(plb + padded(c), xr[r], xi[r]);
            }
            const int oe = (int)st->op_end;
            for (int o = (int)st->op_begin; o < oe; ++o) {
                const WinOp *op = ops + o;
                const uint32_t kind = op->kind, sel = op->sel;
                // one flat dispatch over (kind, sel): nested switches merge the
                // register state twice and the merges cost register copies
                constexpr uint32_t k1 = kW1q << 8, kc = kWCx << 8, ks = kCtrlSlot << 4, kl = kCtrlLocal << 4,
                                   kf = kCtrlFixed << 4;
                const uint64_t on = __builtin_amdgcn_ballot_w64((lb >> op->aux) & 1);  // kCtrlLocal lanes
                const bool fixed_on = (base >> op->aux) & 1;       // kCtrlFixed: uniform
                switch (kind << 8 | sel) {
                    case k1 | 0: reg_1q<T, NS, 0, F>(xr, xi, mats + op->data_off); break;
                    case k1 | 1: reg_1q<T, NS, 1, F>(xr, xi, mats + op->data_off); break;
                    case k1 | 2: if constexpr (NS > 2) reg_1q<T, NS, 2, F>(xr, xi, mats + op->data_off); break;
                    case k1 | 3: if constexpr (NS > 3) reg_1q<T, NS, 3, F>(xr, xi, mats + op->data_off); break;
                    case kc | ks | 0 | 1 << 2: reg_cx_slot<T, NS, 0, 1>(xr, xi); break;
                    case kc | ks | 0 | 2 << 2: if constexpr (NS > 2) reg_cx_slot<T, NS, 0, 2>(xr, xi); break;
                    case kc | ks | 0 | 3 << 2: if constexpr (NS > 3) reg_cx_slot<T, NS, 0, 3>(xr, xi); break;
                    case kc | ks | 1 | 0 << 2: reg_cx_slot<T, NS, 1, 0>(xr, xi); break;
                    case kc | ks | 1 | 2 << 2: if constexpr (NS > 2) reg_cx_slot<T, NS, 1, 2>(xr, xi); break;
                    case kc | ks | 1 | 3 << 2: if constexpr (NS > 3) reg_cx_slot<T, NS, 1, 3>(xr, xi); break;
                    case kc | ks | 2 | 0 << 2: if constexpr (NS > 2) reg_cx_slot<T, NS, 2, 0>(xr, xi); break;
                    case kc | ks | 2 | 1 << 2: if constexpr (NS > 2) reg_cx_slot<T, NS, 2, 1>(xr, xi); break;
                    case kc | ks | 2 | 3 << 2: if constexpr (NS > 3) reg_cx_slot<T, NS, 2, 3>(xr, xi); break;
                    case kc | ks | 3 | 0 << 2: if constexpr (NS > 3) reg_cx_slot<T, NS, 3, 0>(xr, xi); break;
                    case kc | ks | 3 | 1 << 2: if constexpr (NS > 3) reg_cx_slot<T, NS, 3, 1>(xr, xi); break;
                    case kc | ks | 3 | 2 << 2: if constexpr (NS > 3) reg_cx_slot<T, NS, 3, 2>(xr, xi); break;
                    // control on a lane bit: swaps in the lanes whose bit is set;
                    // fixed for the tile: a uniform branch around plain swaps
                    case kc | kl | 0: reg_cx_all<T, NS, 0>(xr, xi, on); break;
                    case kc | kl | 1: reg_cx_all<T, NS, 1>(xr, xi, on); break;
                    case kc | kl | 2: if constexpr (NS > 2) reg_cx_all<T, NS, 2>(xr, xi, on); break;
                    case kc | kl | 3: if constexpr (NS > 3) reg_cx_all<T, NS, 3>(xr, xi, on); break;
                    case kc | kf | 0: if (fixed_on) reg_cx_swap<T, NS, 0>(xr, xi); break;
                    case kc | kf | 1: if (fixed_on) reg_cx_swap<T, NS, 1>(xr, xi); break;
                    case kc | kf | 2: if constexpr (NS > 2) if (fixed_on) reg_cx_swap<T, NS, 2>(xr, xi); break;
                    case kc | kf | 3: if constexpr (NS > 3) if (fixed_on) reg_cx_swap<T, NS, 3>(xr, xi); break;
                    default:
                        if (kind == kWDense) {
                            if constexpr (kDense) reg_dense_dispatch<T, NS, F>(xr, xi, sel, mats + op->data_off);
                        } else if (kind == kWDiag) {
                            // table index: bits fixed by the tile or by the lane, plus the
                            // bits that vary across the lane's registers
                            const T *d = tables + op->data_off;
                            const int k = (int)op->k, D = 1 << k;
                            uint32_t jt = 0, js[NS];
#pragma unroll
                            for (int i = 0; i < NS; ++i) js[i] = 0;
                            for (int b = 0; b < k; ++b) {
                                const uint32_t bb = op->bits[b];
                                if (bb & kFixedBit) {
                                    jt |= (uint32_t)((base >> (bb & ~kFixedBit)) & 1) << b;
                                    continue;
                                }
                                bool slot = false;
#pragma unroll
                                for (int i = 0; i < NS; ++i)
                                    if ((int)bb == sl[i]) js[i] = 1u << b, slot = true;
                                if (!slot) jt |= (uint32_t)((lb >> bb) & 1) << b;
                            }
#pragma unroll
                            for (int r = 0; r < R; ++r) {
                                uint32_t j = jt;
#pragma unroll
                                for (int i = 0; i < NS; ++i) j |= ((r >> i) & 1) ? js[i] : 0u;
                                const T wr = d[j], wi = d[D + j];
                                const T ar = xr[r], ai = xi[r];
                                if constexpr (F) {
                                    xr[r] = msc<F>(wr * ar, wi, ai);
                                    xi[r] = mac<F>(wr * ai, wi, ar);
                                } else {  // same roundings, inputs read first
                                    const T p0 = wr * ar, q0 = wi * ai, p1 = wr * ai, q1 = wi * ar;
                                    xr[r] = p0 - q0;
                                    xi[r] = p1 + q1;
                                }
                            }
                        }
                        break;
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                int c = 0;
#pragma unroll
                for (int i = 0; i < NS; ++i) c |= ((r >> i) & 1) << sl[i];
                lds_put(plb + padded(c), xr[r], xi[r]);
            }
            __syncthreads();
        }
        // write-back through its own registers: pr/pi may already hold the
        // prefetched next tile.  The opaque copy of wlt makes the R row
        // addresses per-tile adds instead of loop-invariant registers.
        int w = wlt;
        asm volatile("" : "+v"(w));
#pragma unroll
        for (int j = 0; j < R; ++j) {
            lds_get(w + woff[j], wr[j], wi[j]);
        }
        wbase = base;
        pending = true;
        __syncthreads();
        tile = next;
        base = nbase;
    }
    if (pending) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const uint64_t g = row(wbase, j);
            rel[g] = wr[j];
            iml[g] = wi[j];
        }
    }
}

// Plain streaming copy of known size (8 B read + 8 B written per element):
// the reference kernel that scales hardware byte counters.
static __global__ __launch_bounds__(kBlock) void k_copy(double *__restrict__ dst, const double *__restrict__ src,
                                                        uint64_t n) {
    for (uint64_t i = gtid(); i < n; i += gstride()) dst[i] = src[i];
}

// Best-effort streaming copies, to measure what a read + write pass can
// reach: 16-byte accesses, U of them in flight per lane before the stores,
// and with NT the stores marked non-temporal (streamed past the caches).
// n2 counts 16-byte elements; the grid covers them U at a time per lane.
template <int U, bool NT>
__global__ __launch_bounds__(kBlock) void k_copy_wide(double2 *__restrict__ dst, const double2 *__restrict__ src,
                                                      uint64_t n2) {
    const uint64_t stride = gstride();
    uint64_t i = gtid();
    for (; i + (U - 1) * stride < n2; i += U * stride) {
        double2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = src[i + u * stride];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if constexpr (NT) {
                __builtin_nontemporal_store(v[u].x, &dst[i + u * stride].x);
                __builtin_nontemporal_store(v[u].y, &dst[i + u * stride].y);
            } else {
                dst[i + u * stride] = v[u];
            }
        }
    }
    for (; i < n2; i += stride) dst[i] = src[i];
}

static __global__ __launch_bounds__(kBlock) void k_fill(double *__restrict__ dst, uint64_t n, double v) {
    for (uint64_t i = gtid(); i < n; i += gstride()) dst[i] = v;
}

// dst[i] = a uniform double in [-1, 1) hashed from i (splitmix64 finaliser):
// copy probes on data whose bits change from element to element, as a
// state's do.
static __global__ __launch_bounds__(kBlock) void k_fill_hash(double *__restrict__ dst, uint64_t n) {
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        uint64_t z = i + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        dst[i] = (double)(z >> 11) * 0x1.0p-52 - 1.0;
    }
}

// max_i |a_i - b_i| (complex modulus, f64) between two states of possibly
// different precision; non-negative doubles order like their bit patterns,
// so the grid combines block maxima with an integer atomicMax.
template <typename A, typename B>
__global__ __launch_bounds__(kBlock) void k_max_abs_diff(const A *__restrict__ ar, const A *__restrict__ ai,
                                                         Layout la, const B *__restrict__ br,
                                                         const B *__restrict__ bi, Layout lb, uint64_t len,
                                                         unsigned long long *__restrict__ out) {
    __shared__ double s[kBlock];
    double m = 0.0;
    for (uint64_t i = gtid(); i < len; i += gstride()) {
        const uint64_t pa = phys_index(la, i), pb = phys_index(lb, i);
        const double dr = (double)ar[pa] - (double)br[pb], di = (double)ai[pa] - (double)bi[pb];
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
// Lane 0's strictly ordered walk over a slab of squared moduli in LDS, in
// register groups of G values: two groups alternate, the loads of one
// issued before the other's adds, so the serial f64 chain, not the LDS
// latency, sets the pace.  group(g, v) takes the G values from index g and
// returns false to stop; one(j, v) takes the tail one value at a time.
template <int G>
__device__ __forceinline__ void load_group(double *v, const double *tv) {
#pragma unroll
    for (int u = 0; u < G; ++u) v[u] = tv[u];
}

template <int G, typename Group, typename One>
__device__ __forceinline__ void serial_walk(const double *tv, int cnt, Group &&group, One &&one) {
    const int full = cnt / G * G;
    double a[G], b[G];
    if (full > 0) load_group<G>(a, tv);
    for (int g = 0; g < full;) {
        const bool more_b = g + G < full;
        if (more_b) load_group<G>(b, tv + g + G);
        if (!group(g, a)) return;
        g += G;
        if (!more_b) break;
        const bool more_a = g + G < full;
        if (more_a) load_group<G>(a, tv + g + G);
        if (!group(g, b)) return;
        g += G;
        if (!more_a) break;
    }
    for (int j = full; j < cnt; ++j)
        if (!one(j, tv[j])) return;
}

// Shapes of the sampler's chunk walks (one workgroup per chunk; lane 0's
// serial chain sets the time): with few chunks, 4 waves load 4096-value
// slabs just before squaring them and lane 0 walks 16-value groups, the
// shortest walk per chunk at a register count that lets several share a CU;
// with many chunks, one wave per chunk with 512-value slabs, loaded while
// lane 0 walks the previous one, and 8-value groups keeps registers and LDS
// small, so many chains share a CU.
template <int BS, int SPAN, int G, bool PF>
struct SampShape {
    static constexpr int kBlock = BS, kSpan = SPAN, kGroup = G;
    static constexpr bool kPrefetch = PF;
    static_assert(SPAN % BS == 0 && SPAN % G == 0, "whole lanes and groups per slab");
};
using SampWide = SampShape<256, 4096, 16, false>;
using SampNarrow = SampShape<64, 512, 8, true>;

// A workgroup's slab of amplitudes in circuit order: each lane holds
// S::kSpan / S::kBlock of them in registers (with S::kPrefetch, loaded while
// lane 0 still walks the previous slab) and squares them into LDS (r*r + i*i
// in f64).
template <typename T, typename S>
struct SampSlab {
    static constexpr int kPerLane = S::kSpan / S::kBlock;
    T r[kPerLane], m[kPerLane];
    // Slabs start at multiples of S::kSpan (or 0 for a state shorter than
    // one), so slab, the lane and u * S::kBlock are disjoint index parts.
    __device__ __forceinline__ void load(const T *__restrict__ re, const T *__restrict__ im, const Layout &L,
                                         uint64_t slab, int cnt) {
        const uint64_t lane = phys_index(L, threadIdx.x) | phys_index(L, slab);
#pragma unroll
        for (int u = 0; u < kPerLane; ++u) {
            const int j = threadIdx.x + u * S::kBlock;
            if (j < cnt) {
                const uint64_t p = lane | phys_index(L, (uint64_t)u * S::kBlock);
                r[u] = re[p];
                m[u] = im[p];
            }
        }
    }
    // Without prefetch: load, square and store one value at a time.
    __device__ __forceinline__ static void fill(const T *__restrict__ re, const T *__restrict__ im, const Layout &L,
                                                uint64_t slab, int cnt, double *tv) {
        const uint64_t lane = phys_index(L, threadIdx.x) | phys_index(L, slab);
        for (int j = threadIdx.x; j < cnt; j += S::kBlock) {
            const uint64_t p = lane | phys_index(L, (uint64_t)(j - (int)threadIdx.x));
            const double a = (double)re[p], b = (double)im[p];
            tv[j] = a * a + b * b;
        }
    }
    __device__ __forceinline__ void store(double *tv, int cnt) const {
#pragma unroll
        for (int u = 0; u < kPerLane; ++u) {
            const int j = threadIdx.x + u * S::kBlock;
            if (j < cnt) {
                const double a = (double)r[u], b = (double)m[u];
                tv[j] = a * a + b * b;
            }
        }
    }
};

// sample_indices pass 1 (src/sample.rs:38-50): per chunk a strictly serial
// f64 sum acc += r*r + i*i in index order.  The lanes load a slab coalesced
// and square it into LDS; lane 0 then accumulates the slab serially, so the
// sum order is exactly the reference's (with S::kPrefetch while the lanes'
// loads of the next slab are in flight).
template <typename T, typename S>
__global__ __launch_bounds__(S::kBlock) void k_chunk_sums(const T *__restrict__ re, const T *__restrict__ im, Layout L,
                                                       uint64_t len, uint64_t chunk, double *__restrict__ sums) {
    __shared__ double tv[S::kSpan];
    const uint64_t lo = (uint64_t)blockIdx.x * chunk;
    const uint64_t hi = lo + chunk < len ? lo + chunk : len;
    auto count = [&](uint64_t slab) {
        return (int)(hi - slab < (uint64_t)S::kSpan ? hi - slab : (uint64_t)S::kSpan);
    };
    double acc = 0.0;
    SampSlab<T, S> w;
    if (S::kPrefetch && lo < hi) w.load(re, im, L, lo, count(lo));
    for (uint64_t slab = lo; slab < hi; slab += S::kSpan) {
        const int cnt = count(slab);
        if (S::kPrefetch) w.store(tv, cnt);
        else SampSlab<T, S>::fill(re, im, L, slab, cnt, tv);
        __syncthreads();
        if (S::kPrefetch && slab + S::kSpan < hi) w.load(re, im, L, slab + S::kSpan, count(slab + S::kSpan));
        if (threadIdx.x == 0)
            serial_walk<S::kGroup>(
                tv, cnt,
                [&](int, const double *v) {
#pragma unroll
                    for (int u = 0; u < S::kGroup; ++u) acc += v[u];
                    return true;
                },
                [&](int, double v) {
                    acc += v;
                    return true;
                });
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = acc;
}

// sample_indices pass 1 for many chunks (src/sample.rs:38-50): a workgroup
// owns G consecutive full chunks of 2^16 values and lane g (g < G) walks
// chunk g, so G serial chains share one wave instead of one chain per
// workgroup.  Each slab holds S = SLAB / G values of every chunk: the 256
// lanes load it coalesced along each chunk (registers, issued while wave 0
// walks the previous slab), square it and store it transposed in LDS (value
// i of chunk g at i * (G + 1) + g: the chain reads are consecutive across
// lanes, the stores stride G + 1 doubles), double-buffered.  Every chain
// adds r*r + i*i (in f64) in index order, exactly the reference's order.
template <typename T, int G, int SLAB>
__global__ __launch_bounds__(256) void k_chunk_sums_multi(const T *__restrict__ re, const T *__restrict__ im,
                                                          Layout L, uint64_t nch, double *__restrict__ sums) {
    constexpr int kLanes = 256, kChunk = 1 << 16, S = SLAB / G, kPer = S * G / kLanes, kRow = G + 1;
    static_assert(kPer * kLanes == S * G && kChunk % S == 0, "whole slabs");
    __shared__ double tv[2][S * kRow];
    __shared__ uint64_t hp[G];  // physical bits of each chunk's base (bits 16 and up)
    const uint64_t c0 = (uint64_t)blockIdx.x * G;
    const int gc = nch - c0 < (uint64_t)G ? (int)(nch - c0) : G;
    const int t = threadIdx.x;
    if (t < G) {
        const uint64_t b = (c0 + t) << 16;
        hp[t] = L.lut ? L.lut[512 + ((b >> 16) & 255)] | L.lut[768 + ((b >> 24) & 255)] |
                            L.lut[1024 + ((b >> 32) & 255)]
                      : b;
    }
    __syncthreads();
    // value u of the lane: chunk g = e / S, offset i = e % S within the slab
    T r[kPer], m[kPer];
    auto load = [&](int off) {
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int e = t + u * kLanes, g = e / S, x = off + e % S;
            if (g < gc) {
                const uint64_t lo = L.lut ? L.lut[x & 255] | L.lut[256 + (x >> 8)] : (uint64_t)x;
                const uint64_t p = hp[g] | lo;
                r[u] = re[p];
                m[u] = im[p];
            }
        }
    };
    auto store = [&](double *buf) {
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int e = t + u * kLanes, g = e / S, i = e % S;
            if (g < gc) {
                const double a = (double)r[u], b = (double)m[u];
                buf[i * kRow + g] = a * a + b * b;
            }
        }
    };
    double acc = 0.0;
    load(0);
    store(tv[0]);
    __syncthreads();
    for (int off = 0, k = 0; off < kChunk; off += S, k ^= 1) {
        const bool more = off + S < kChunk;
        if (more) load(off + S);
        if (t < gc) {
            const double *col = tv[k] + t;
#pragma unroll 16
            for (int i = 0; i < S; ++i) acc += col[i * kRow];
        }
        if (more) store(tv[k ^ 1]);
        __syncthreads();
    }
    if (t < gc) sums[c0 + t] = acc;
}

// sample_indices pass 5 + 6 (src/sample.rs:68-107): rescan one touched
// chunk from cum_before[c], emitting the sorted draws us[s0..s1) that fall
// below the running sum; leftovers land on the chunk's last index with
// nonzero probability.
//
// Per slab, lane 0 forms the running sums in index order (the reference's
// serial chain, so the same bits) and keeps the one at the end of each
// group of S::kGroup values in LDS.  The reference emits draw u at the first
// index whose running sum exceeds u; the running sums never decrease (each
// step adds r*r + i*i >= 0 under monotonic rounding), so the slab takes
// exactly the draws below its last running sum, and every lane places its
// share of them: a binary search over the group ends finds the group, and
// the group's sums, recomputed from the previous group's end with the same
// adds in the same order, give the index.
template <typename T, typename S>
__global__ __launch_bounds__(S::kBlock) void k_rescan(const T *__restrict__ re, const T *__restrict__ im, Layout L,
                                                   uint64_t len,
                                                   uint64_t chunk, const uint32_t *__restrict__ touched,
                                                   const double *__restrict__ cum, const uint64_t *__restrict__ starts,
                                                   const double *__restrict__ us, uint64_t *__restrict__ out) {
    constexpr int kGroups = S::kSpan / S::kGroup;
    __shared__ double tv[S::kSpan];
    __shared__ double gend[kGroups];  // running sum at the end of each group
    __shared__ uint64_t sptr;
    __shared__ int sdone;
    __shared__ double sstart;         // running sum before the slab
    __shared__ int srun;              // values walked in the slab
    __shared__ uint64_t slast[S::kBlock];
    // touched lists the chunks with draws; without it block c is chunk c,
    // and a chunk without draws has nothing to do
    const uint32_t c = touched ? touched[blockIdx.x] : blockIdx.x;
    const uint64_t lo = (uint64_t)c * chunk;
    const uint64_t hi = lo + chunk < len ? lo + chunk : len;
    const uint64_t s1 = starts[c + 1];
    if (starts[c] == s1) return;
    double acc = cum[c];
    if (threadIdx.x == 0) {
        sptr = starts[c];
        sdone = 0;
    }
    __syncthreads();
    auto count = [&](uint64_t slab) {
        return (int)(hi - slab < (uint64_t)S::kSpan ? hi - slab : (uint64_t)S::kSpan);
    };
    SampSlab<T, S> w;
    if (S::kPrefetch) w.load(re, im, L, lo, count(lo));
    for (uint64_t slab = lo; slab < hi; slab += S::kSpan) {
        const int cnt = count(slab);
        if (S::kPrefetch) w.store(tv, cnt);
        else SampSlab<T, S>::fill(re, im, L, slab, cnt, tv);
        __syncthreads();
        if (S::kPrefetch && slab + S::kSpan < hi) w.load(re, im, L, slab + S::kSpan, count(slab + S::kSpan));
        if (threadIdx.x == 0) {
            // the walk stops once the running sum passes the chunk's last draw
            const double ulast = us[s1 - 1];
            int run = cnt;
            sstart = acc;
            serial_walk<S::kGroup>(
                tv, cnt,
                [&](int g, const double *v) {
                    double x = acc;
#pragma unroll
                    for (int u = 0; u < S::kGroup; ++u) x = x + v[u];
                    acc = x;
                    gend[g / S::kGroup] = x;
                    if (x <= ulast) return true;
                    run = g + S::kGroup;
                    return false;
                },
                [&](int j, double v) {
                    acc = acc + v;
                    gend[j / S::kGroup] = acc;  // the partial last group ends at its last value
                    if (acc <= ulast) return true;
                    run = j + 1;
                    return false;
                });
            srun = run;
        }
        __syncthreads();
        const int run = srun, ngroups = (run + S::kGroup - 1) / S::kGroup;
        const double top = gend[ngroups - 1];  // the running sum after the last value walked
        uint64_t first_out = s1;               // the lane's first draw at or above top
        for (uint64_t i = sptr + threadIdx.x; i < s1; i += S::kBlock) {
            const double u = us[i];
            if (!(u < top)) {
                first_out = i;
                break;
            }
            int a = 0, b = ngroups - 1;  // gend[ngroups - 1] = top > u
            while (a < b) {
                const int mid = (a + b) >> 1;
                if (gend[mid] > u) b = mid;
                else a = mid + 1;
            }
            double x = a > 0 ? gend[a - 1] : sstart;
            int j = a * S::kGroup;
            const int jend = j + S::kGroup < run ? j + S::kGroup : run;
            for (; j < jend - 1; ++j) {
                x = x + tv[j];
                if (x > u) break;
            }
            out[i] = slab + (uint64_t)j;
        }
        slast[threadIdx.x] = first_out;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t m = s1;
            for (int t = 0; t < S::kBlock; ++t) m = slast[t] < m ? slast[t] : m;
            sptr = m;
            sdone = m == s1;
        }
        __syncthreads();
        if (sdone) break;
    }
    if (sdone) return;
    // Fallback: last index of the chunk with nonzero probability (else hi-1).
    uint64_t best = 0;
    int found = 0;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += S::kBlock) {
        const uint64_t p = phys_index(L, i);
        const double r = (double)re[p], m = (double)im[p];
        if (r * r + m * m > 0.0) best = i, found = 1;
    }
    slast[threadIdx.x] = found ? best + 1 : 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t m = 0;
        for (int t = 0; t < S::kBlock; ++t) m = slast[t] > m ? slast[t] : m;
        const uint64_t last = m ? m - 1 : hi - 1;
        for (uint64_t p = sptr; p < s1; ++p) out[p] = last;
    }
}

// State copies in logical order through the layout: out[j] = state at
// logical index lo + j (gather), or the reverse (scatter).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather(const T *__restrict__ re, const T *__restrict__ im, Layout L,
                                                   uint64_t lo, uint64_t count, T *__restrict__ ore,
                                                   T *__restrict__ oim) {
    for (uint64_t j = gtid(); j < count; j += gstride()) {
        const uint64_t p = phys_index(L, lo + j);
        ore[j] = re[p];
        oim[j] = im[p];
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scatter(T *__restrict__ re, T *__restrict__ im, Layout L, uint64_t lo,
                                                    uint64_t count, const T *__restrict__ ire,
                                                    const T *__restrict__ iim) {
    for (uint64_t j = gtid(); j < count; j += gstride()) {
        const uint64_t p = phys_index(L, lo + j);
        re[p] = ire[j];
        im[p] = iim[j];
    }
}

}  // namespace dev
}  // namespace zeno

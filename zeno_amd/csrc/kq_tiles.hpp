// kq_tiles.hpp — dense ops on 2 to 6 qubits over LDS tiles: a workgroup of
// 512 lanes takes 2^12 amplitudes whose 12 state bits are the op's K bits,
// the low bits of a 128-byte run and the lowest free bits, i.e. G =
// 2^(12-K) whole groups of the op.  Holds the tile walk that the exact VALU
// kernel (k_kq_tiled, below) and the matrix-core kernel (k_kq_mfma,
// kernels.hpp) share, and the host-side description of a tile.  Plain C++
// apart from the HIP built-ins, so its kernel also builds for the host.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "devcommon.hpp"

namespace zeno {
namespace dev {

constexpr int kKqTileBits = 12;
constexpr int kKqThreads = 512;
constexpr uint32_t kRoleOp = 0x100;  // role[i]: kRoleOp | b for op bit b, else group bit index

struct KqTile {
    uint32_t tq[kKqTileBits];    // local bit i <-> state bit tq[i] (ascending)
    uint32_t role[kKqTileBits];  // what local bit i indexes in X
    uint32_t swz[6];             // X column XOR mask of op bit b (0: none)
};

// Streams every tile of 2^12 amplitudes through LDS as X[2D][SP] and back:
// op() reads X and, after a barrier of its own, writes Y over it.  The re
// of element j of group g sits at j*SP + (g ^ f(j)) and its im at (D + j)*SP
// + (g ^ f(j)), f(j) the XOR of swz[b] over the op bits b set in j (f = 0
// keeps the plain layout).  The next tile's
// loads are issued before op() and the previous tile's stores after this
// tile's loads landed, so a wave keeps one tile in flight each way while it
// computes.
template <typename T, int K, int SP, typename Op>
__device__ __forceinline__ void kq_tile_loop(T *__restrict__ re, T *__restrict__ im, const KqTile &mt,
                                             uint64_t tiles, T *sx, Op op) {
    constexpr int TB = kKqTileBits, R = (1 << TB) / kKqThreads, LB = TB - 3;
    constexpr int D = 1 << K;
    static_assert(R == 8, "tile shape");
    const int t = threadIdx.x;
    // lane part and register part of a tile element's state offset, X row
    // offset and X column; the slot is lrow + rrow[r] + (lcol ^ rcol[r])
    uint64_t lo = 0;
    int lrow = 0, lcol = 0;
#pragma unroll
    for (int i = 0; i < LB; ++i) {
        if (!((t >> i) & 1)) continue;
        lo |= 1ull << mt.tq[i];
        const uint32_t ro = mt.role[i];
        if (ro & kRoleOp) lrow += (1 << (ro & 7)) * SP, lcol ^= (int)mt.swz[ro & 7];
        else lcol ^= 1 << ro;
    }
    uint64_t hi[R];
    int rrow[R], rcol[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        hi[r] = 0;
        rrow[r] = rcol[r] = 0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if (!((r >> i) & 1)) continue;
            hi[r] |= 1ull << mt.tq[LB + i];
            const uint32_t ro = mt.role[LB + i];
            if (ro & kRoleOp) rrow[r] += (1 << (ro & 7)) * SP, rcol[r] ^= (int)mt.swz[ro & 7];
            else rcol[r] ^= 1 << ro;
        }
    }
    int rofs[R];
#pragma unroll
    for (int r = 0; r < R; ++r) rofs[r] = lrow + rrow[r] + (lcol ^ rcol[r]);
    auto tile_base = [&](uint64_t tile) {
        uint64_t b = tile;
#pragma unroll
        for (int i = 0; i < TB; ++i) b = insert_zero(b, mt.tq[i]);
        return b;
    };
    T pr[R], pi[R], wr[R], wi[R];
    uint64_t tile = blockIdx.x, base = tile < tiles ? tile_base(tile) : 0, wbase = 0;
    bool pending = false;
    if (tile < tiles) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            pr[r] = re[base | hi[r] | lo];
            pi[r] = im[base | hi[r] | lo];
        }
    }
    while (tile < tiles) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            sx[rofs[r]] = pr[r];
            sx[rofs[r] + D * SP] = pi[r];
        }
        if (pending) {  // previous tile's stores, after this tile's loads landed
#pragma unroll
            for (int r = 0; r < R; ++r) {
                re[wbase | hi[r] | lo] = wr[r];
                im[wbase | hi[r] | lo] = wi[r];
            }
        }
        __syncthreads();
        const uint64_t next = tile + gridDim.x;
        const uint64_t nbase = next < tiles ? tile_base(next) : 0;
        if (next < tiles) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                pr[r] = re[nbase | hi[r] | lo];
                pi[r] = im[nbase | hi[r] | lo];
            }
        }
        op();
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            wr[r] = sx[rofs[r]];
            wi[r] = sx[rofs[r] + D * SP];
        }
        wbase = base;
        pending = true;
        __syncthreads();
        tile = next;
        base = nbase;
    }
    if (pending) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            re[wbase | hi[r] | lo] = wr[r];
            im[wbase | hi[r] | lo] = wi[r];
        }
    }
}

// ------------------------------------------------------ kq through LDS
// kern_kq_scalar (src/state.rs:355-392) on the matrix-core kernel's tiles,
// for K = 2..6: the same products and sums as k_kq, in the same order (per
// output row, ascending columns, from 0), so the result is bit for bit
// k_kq's; what changes is the data path.  k_kq gathers a group's 2^K
// amplitudes into one lane's registers, which leaves loads uncoalesced when
// the op's bits are low and needs 4*2^K registers of amplitudes.  Here the
// tile is read and written in whole 128-byte runs and transposed in LDS.
// Wave w computes the RW = max(D/8, 1) output rows of row band w / WS over
// a slice of G/WS groups (WS = max(8/D, 1) waves share a band when D < 8),
// 8 outputs per lane (lane l holds the slice's groups l + 64i), reading X
// column by column.
// X rows are not padded: the op bits that lie on a wave's lane bits XOR
// column bits that no group bit there uses (swz), so the transposing LDS
// accesses of a wave fall on distinct columns mod 64 whichever bits the op
// has.  The matrix is packed per row band and column ([band][c][row][re,
// im]), so the 2*RW weights of one column are one uniform run read by
// scalar loads.
template <typename T, int K, bool F>
__global__ __launch_bounds__(kKqThreads) void k_kq_tiled(T *__restrict__ re, T *__restrict__ im,
                                                           const T *__restrict__ pk, KqTile mt, uint64_t tiles) {
    constexpr int TB = kKqTileBits;
    constexpr int D = 1 << K, G = 1 << (TB - K), SP = G;
    constexpr int RW = D >= 8 ? D / 8 : 1, WS = D >= 8 ? 1 : 8 / D, GL = G / 64 / WS;
    static_assert(RW * GL == 8 && GL >= 1, "tile shape");
    __shared__ T sx[2 * D * SP];
    const int ln = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int band = wv / WS, g0 = (wv % WS) * (G / WS);  // row band, first group of the slice
    const T *__restrict__ wp = pk + (size_t)band * D * 2 * RW;
    auto fsw = [&](int j) {  // column XOR of X row j (< 64, uniform)
        int f = 0;
#pragma unroll
        for (int b = 0; b < K; ++b)
            if ((j >> b) & 1) f ^= (int)mt.swz[b];
        return f;
    };
    kq_tile_loop<T, K, SP>(re, im, mt, tiles, sx, [&] {
        T ar[RW][GL], ai[RW][GL];
#pragma unroll
        for (int rr = 0; rr < RW; ++rr)
#pragma unroll
            for (int gi = 0; gi < GL; ++gi) ar[rr][gi] = ai[rr][gi] = T(0);
#pragma unroll 4
        for (int c = 0; c < D; ++c) {
            const int col = ln ^ fsw(c);
            T xr[GL], xi[GL];
#pragma unroll
            for (int gi = 0; gi < GL; ++gi) {
                xr[gi] = sx[c * SP + g0 + col + 64 * gi];
                xi[gi] = sx[(D + c) * SP + g0 + col + 64 * gi];
            }
            const T *__restrict__ wc = wp + c * 2 * RW;
#pragma unroll
            for (int rr = 0; rr < RW; ++rr) {
                const T mr = wc[2 * rr], mi = wc[2 * rr + 1];
#pragma unroll
                for (int gi = 0; gi < GL; ++gi) {
                    ar[rr][gi] = msc<F>(mac<F>(ar[rr][gi], mr, xr[gi]), mi, xi[gi]);
                    ai[rr][gi] = mac<F>(mac<F>(ai[rr][gi], mr, xi[gi]), mi, xr[gi]);
                }
            }
        }
        __syncthreads();  // every wave has read X
#pragma unroll
        for (int rr = 0; rr < RW; ++rr) {
            const int row = RW * band + rr, col = ln ^ fsw(row);
#pragma unroll
            for (int gi = 0; gi < GL; ++gi) {
                sx[row * SP + g0 + col + 64 * gi] = ar[rr][gi];
                sx[(D + row) * SP + g0 + col + 64 * gi] = ai[rr][gi];
            }
        }
    });
}

}  // namespace dev

// ------------------------------------------------------------------ host
// Tile of an op on state bits pq (op order) in a state of 2^n amplitudes
// (n >= 12) whose 128-byte run holds 2^run amplitudes: the op's bits, the
// run's bits, then the lowest free bits, ascending; X row bit b <-> op bit
// b, the other tile bits index groups.  No column swizzle.
inline dev::KqTile kq_tile(uint32_t run, const uint32_t *pq, uint32_t k) {
    uint64_t m = 0;
    for (uint32_t b = 0; b < k; ++b) m |= 1ull << pq[b];
    m |= (1ull << run) - 1;
    for (uint32_t b = 0; __builtin_popcountll(m) < dev::kKqTileBits; ++b) m |= 1ull << b;
    dev::KqTile t{};
    uint32_t i = 0, g = 0;
    for (uint32_t b = 0; b < 64; ++b) {
        if (!((m >> b) & 1)) continue;
        t.tq[i] = b;
        t.role[i] = g;
        for (uint32_t j = 0; j < k; ++j)
            if (pq[j] == b) t.role[i] = dev::kRoleOp | j;
        if (!(t.role[i] & dev::kRoleOp)) ++g;
        ++i;
    }
    return t;
}

// Column swizzle of the exact tiled kernel: the op bits on a wave's lane
// bits (tile bits 0..5) take, in order, the column bits 0..5 that no group
// bit there takes, so the 64 lanes of a transposing LDS access see 64
// distinct columns mod 64.
inline void kq_swizzle(dev::KqTile &t) {
    uint32_t used = 0;
    for (int i = 0; i < 6; ++i)
        if (!(t.role[i] & dev::kRoleOp)) used |= 1u << t.role[i];
    int next = 0;
    for (int i = 0; i < 6; ++i) {
        if (!(t.role[i] & dev::kRoleOp)) continue;
        while ((used >> next) & 1) ++next;
        t.swz[t.role[i] & 7] = 1u << next;
        used |= 1u << next;
    }
}

// The exact tiled kernel's matrix, packed per row band and column:
// [band][column][row of the band's max(2^K/8, 1)][re, im] from row-major
// mr, mi.
template <typename T>
std::vector<T> kq_tiled_pack(int K, const T *mr, const T *mi) {
    const int D = 1 << K, RW = D >= 8 ? D / 8 : 1;
    std::vector<T> pk((size_t)2 * D * D);
    for (int w = 0; w < D / RW; ++w)
        for (int c = 0; c < D; ++c)
            for (int rr = 0; rr < RW; ++rr) {
                const size_t o = (((size_t)w * D + c) * RW + rr) * 2;
                pk[o] = mr[(size_t)(RW * w + rr) * D + c];
                pk[o + 1] = mi[(size_t)(RW * w + rr) * D + c];
            }
    return pk;
}

}  // namespace zeno

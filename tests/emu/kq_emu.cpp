// Test infrastructure: runs the exact tiled dense kernel's source
// (k_kq_tiled in zeno_amd/csrc/kq_tiles.hpp) on the CPU, every workgroup of
// the launch in turn, 512 lanes each (tests/emu/lanes.hpp), with the tile
// description and matrix packing the library builds on the host.  Built by
// tests/test_kq_emulation.py with the host stand-ins of
// tests/emu/hip/hip_runtime.h, under AddressSanitizer and
// UndefinedBehaviorSanitizer.
//
//   kq_emu N K QS(csv, op order) GRID ORDER STATE MAT OUT
//
// Built with -DEMU_PREC=32 or 64.  STATE and OUT: re[2^N] then im[2^N] in
// the build's precision.  MAT: 4^K complex doubles (re, im interleaved,
// row-major, row and column bit b <-> QS[b]).  ORDER: lane order in every
// barrier phase (0 forward, 1 backward, >= 2 a seeded shuffle).
#include <cstdio>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "lanes.hpp"
#include "kq_tiles.hpp"

using namespace zeno;

namespace {

template <typename T, int K>
void launch(T *re, T *im, const T *pk, const dev::KqTile &t, uint64_t tiles, unsigned grid, int order) {
    emu::grid_dim.x = grid;
    emu::block_dim.x = dev::kKqThreads;
    for (unsigned b = 0; b < grid; ++b) {
        emu::block_idx.x = b;
        emu::run_group(dev::kKqThreads, order, [&] { dev::k_kq_tiled<T, K, false>(re, im, pk, t, tiles); },
                       128 << 10);
    }
}

template <typename T>
int run(uint32_t n, int k, const std::vector<uint32_t> &qs, unsigned grid, int order, const char *state_path,
        const char *mat_path, const char *out_path) {
    const size_t len = (size_t)1 << n, dim = (size_t)1 << k;
    std::vector<T> st(2 * len);
    std::vector<double> mat(2 * dim * dim);
    FILE *f = std::fopen(state_path, "rb");
    if (!f || std::fread(st.data(), sizeof(T), st.size(), f) != st.size()) return 2;
    std::fclose(f);
    f = std::fopen(mat_path, "rb");
    if (!f || std::fread(mat.data(), sizeof(double), mat.size(), f) != mat.size()) return 2;
    std::fclose(f);
    std::vector<T> mr(dim * dim), mi(dim * dim);
    for (size_t i = 0; i < dim * dim; ++i) mr[i] = (T)mat[2 * i], mi[i] = (T)mat[2 * i + 1];
    const std::vector<T> pk = kq_tiled_pack<T>(k, mr.data(), mi.data());
    dev::KqTile t = kq_tile(sizeof(T) == 8 ? 4 : 5, qs.data(), (uint32_t)k);
    kq_swizzle(t);
    const uint64_t tiles = len >> dev::kKqTileBits;
    T *re = st.data(), *im = st.data() + len;
    switch (k) {
        case 2: launch<T, 2>(re, im, pk.data(), t, tiles, grid, order); break;
        case 3: launch<T, 3>(re, im, pk.data(), t, tiles, grid, order); break;
        case 4: launch<T, 4>(re, im, pk.data(), t, tiles, grid, order); break;
        case 5: launch<T, 5>(re, im, pk.data(), t, tiles, grid, order); break;
        case 6: launch<T, 6>(re, im, pk.data(), t, tiles, grid, order); break;
        default: return 2;
    }
    f = std::fopen(out_path, "wb");
    if (!f) return 2;
    std::fwrite(st.data(), sizeof(T), st.size(), f);
    std::fclose(f);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 9) {
        std::fprintf(stderr, "usage: kq_emu N K QS GRID ORDER STATE MAT OUT\n");
        return 2;
    }
    const uint32_t n = (uint32_t)std::atoi(argv[1]);
    const int k = std::atoi(argv[2]);
    std::vector<uint32_t> qs;
    for (const char *p = argv[3]; *p;) {
        char *end = nullptr;
        qs.push_back((uint32_t)std::strtoul(p, &end, 10));
        if (end == p) return 2;
        p = *end == ',' ? end + 1 : end;
    }
    const unsigned grid = (unsigned)std::atoi(argv[4]);
    const int order = std::atoi(argv[5]);
    if (n < (uint32_t)dev::kKqTileBits || n > 20 || (int)qs.size() != k || grid < 1) return 2;
    for (uint32_t q : qs)
        if (q >= n) return 2;
    using T = std::conditional_t<EMU_PREC == 64, double, float>;
    return run<T>(n, k, qs, grid, order, argv[6], argv[7], argv[8]);
}

// Test infrastructure: the sampler's jump-ahead maps (zeno_amd/csrc/draws.hpp)
// against plain sequential stepping of xoshiro256++.  For several seeded
// states, the state reached by jumping k * kDrawBlock steps with the maps of
// the set bits of k must equal the state after k * kDrawBlock single steps,
// for every block index k below 1500, and the first outputs of each block
// must be the sequential stream's.
#include <cstdio>
#include <cstring>

#include "draws.hpp"

using namespace zeno::draws;

static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static uint64_t next(uint64_t *s) {
    const uint64_t out = rotl(s[0] + s[3], 23) + s[0];
    gf2_step(s);
    return out;
}

int main() {
    const uint64_t seeds[3][4] = {{1, 2, 3, 4},
                                  {0x123456789abcdefull, 0xfedcba9876543210ull, 0x0f0f0f0f12345678ull, 0xdeadbeefcafef00dull},
                                  {~0ull, 0, 0x8000000000000000ull, 1}};
    const int levels = draw_levels(1500 * kDrawBlock);
    std::vector<uint64_t> maps((size_t)levels * 1024);
    copy_jump_maps(levels, maps.data());
    int bad = 0;
    for (const auto &s0 : seeds) {
        uint64_t seq[4];
        std::memcpy(seq, s0, 32);
        for (uint64_t k = 0; k < 1500; ++k) {
            uint64_t v[4];
            std::memcpy(v, s0, 32);
            for (int i = 0; i < levels; ++i)
                if ((k >> i) & 1) {
                    Gf2Map m;
                    std::memcpy(&m, maps.data() + (size_t)i * 1024, sizeof m);
                    uint64_t o[4];
                    gf2_apply(m, v, o);
                    std::memcpy(v, o, 32);
                }
            if (std::memcmp(v, seq, 32) != 0) ++bad;
            if (next(v) != next(seq)) ++bad;  // the block's first output
            for (uint64_t t = 1; t < kDrawBlock; ++t) next(seq);
        }
    }
    std::printf("levels %d, mismatches %d\n", levels, bad);
    return bad != 0;
}

// draws.hpp — jump-ahead for the sampler's xoshiro256++ stream (host side).
//
// xoshiro256++'s state update (s2 ^= s0; s3 ^= s1; s1 ^= s2; s0 ^= s3;
// s2 ^= s1 << 17; s3 = rotl(s3, 45)) is linear over GF(2) on the 256-bit
// state, so k steps are a 256 x 256 bit matrix: the map of each state bit,
// col[j] (4 words).  The device draws (sort.hip, k_draws) start thread k
// kDrawBlock * k steps into the stream, reached with the jump maps
// M^(kDrawBlock * 2^i) of the set bits of k.
#pragma once

#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

namespace zeno {
namespace draws {

constexpr uint64_t kDrawBlock = 256;

struct Gf2Map {
    uint64_t col[256][4];
};

inline void gf2_step(uint64_t *s) {
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = (s[3] << 45) | (s[3] >> 19);
}

inline void gf2_apply(const Gf2Map &a, const uint64_t *v, uint64_t *out) {
    uint64_t r[4] = {0, 0, 0, 0};
    for (int j = 0; j < 256; ++j)
        if ((v[j >> 6] >> (j & 63)) & 1)
            for (int w = 0; w < 4; ++w) r[w] ^= a.col[j][w];
    for (int w = 0; w < 4; ++w) out[w] = r[w];
}

// a after b
inline Gf2Map gf2_compose(const Gf2Map &a, const Gf2Map &b) {
    Gf2Map c;
    for (int j = 0; j < 256; ++j) gf2_apply(a, b.col[j], c.col[j]);
    return c;
}

// Copies the jump maps M^(kDrawBlock * 2^i), i < levels, to dst (1024 words
// each); they are built once per process and extended on demand.
inline void copy_jump_maps(int levels, uint64_t *dst) {
    static std::mutex mu;
    static std::vector<Gf2Map> maps;
    std::lock_guard<std::mutex> lk(mu);
    if (maps.empty()) {
        Gf2Map m;
        for (int j = 0; j < 256; ++j) {
            uint64_t v[4] = {0, 0, 0, 0};
            v[j >> 6] = 1ull << (j & 63);
            gf2_step(v);
            for (int w = 0; w < 4; ++w) m.col[j][w] = v[w];
        }
        for (uint64_t k = 1; k < kDrawBlock; k <<= 1) m = gf2_compose(m, m);
        maps.push_back(m);
    }
    while ((int)maps.size() < levels) maps.push_back(gf2_compose(maps.back(), maps.back()));
    for (int i = 0; i < levels; ++i) std::memcpy(dst + (size_t)i * 1024, &maps[i], sizeof(Gf2Map));
}

// Jump levels for `shots` draws: enough bits for the last block's index.
inline int draw_levels(uint64_t shots) {
    const uint64_t blocks = (shots + kDrawBlock - 1) / kDrawBlock;
    return blocks > 1 ? 64 - __builtin_clzll(blocks - 1) : 0;
}

}  // namespace draws
}  // namespace zeno

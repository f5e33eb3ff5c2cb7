// Test infrastructure: runs the batched trajectory kernel's source
// (zeno_amd/csrc/trajectories.hpp) on the CPU, one workgroup of 256 threads,
// for a compiled program read from a file, and writes every shot's clbit
// word.  Built by tests/test_traj_emulation.py with the host stand-ins of
// tests/emu/hip/hip_runtime.h, under AddressSanitizer and
// UndefinedBehaviorSanitizer, with the kernel's helpers inlined and outlined.
//
//   traj_emu PROG N CREGS(csv) FUSION PRECISION SHOTS SEED NOISE(7 csv) ORDER OUT
//
// Built with -DEMU_PREC=32 or 64; PRECISION must match.
//
// ORDER: lane order in every barrier phase (0 forward, 1 backward, >= 2 a
// seeded shuffle).
//
// PROG holds zr_instr records (72 bytes each); OUT gets SHOTS uint64 words.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "lanes.hpp"
#include "host.hpp"
#include "traj_pack.hpp"
#include "trajectories.hpp"

using namespace zeno;

namespace {

std::vector<double> csv(const char *s) {
    std::vector<double> v;
    for (const char *p = s; *p;) {
        char *end = nullptr;
        v.push_back(std::strtod(p, &end));
        p = *end == ',' ? end + 1 : end;
        if (end == p && *p) break;
    }
    return v;
}

template <typename T>
int run(const Compiled &c, uint64_t shots, uint64_t seed, const std::vector<double> &nzv, int order,
        const char *out_path) {
    const Tables<T> t = pack<T>(c);
    const uint64_t slot = traj_slot_bytes(c.n_qubits, sizeof(T)) / sizeof(T);
    std::vector<T> states(slot);  // one workgroup: one slot (state, block sums)
    std::vector<uint64_t> out(shots ? shots : 1, ~0ull);
    dev::TrajLaunch<T> L{};
    L.states = states.data();
    L.ops = t.ops.data();
    L.fm = t.fm.data();
    L.mats = t.mats.data();
    L.out = out.data();
    L.seed = seed;
    L.shots = shots;
    L.slot = slot;
    L.n = c.n_qubits;
    L.nops = (uint32_t)t.nops;
    L.nfm = (uint32_t)t.nfm;
    L.nz = dev::TrajNoise{nzv[0], nzv[1], nzv[2], nzv[3], nzv[4], nzv[5], nzv[6]};
    L.noisy = 0;
    for (double v : nzv) L.noisy |= v != 0.0;
    emu::block_idx.x = 0;
    emu::grid_dim.x = 1;
    emu::block_dim.x = dev::kBlock;
    if (traj_wide(c)) emu::run_group(dev::kBlock, order, [&] { dev::k_trajectories<T, dev::kTrajWideDense>(&L); });
    else emu::run_group(dev::kBlock, order, [&] { dev::k_trajectories<T, dev::kTrajMaxK>(&L); });
    FILE *f = std::fopen(out_path, "wb");
    if (!f) return 2;
    std::fwrite(out.data(), 8, shots, f);
    std::fclose(f);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 11) {
        std::fprintf(stderr, "usage: traj_emu PROG N CREGS FUSION PRECISION SHOTS SEED NOISE ORDER OUT\n");
        return 2;
    }
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    std::vector<zr_instr> prog;
    zr_instr rec;
    while (std::fread(&rec, sizeof rec, 1, f) == 1) prog.push_back(rec);
    std::fclose(f);
    const uint32_t n = (uint32_t)std::atoi(argv[2]);
    std::vector<uint32_t> cregs;
    for (double v : csv(argv[3])) cregs.push_back((uint32_t)v);
    const int fusion = std::atoi(argv[4]), prec = std::atoi(argv[5]);
    const uint64_t shots = std::strtoull(argv[6], nullptr, 10), seed = std::strtoull(argv[7], nullptr, 10);
    const std::vector<double> nz = csv(argv[8]);
    if (nz.size() != 7) return 2;
    const Compiled c = compile(prog.data(), (int64_t)prog.size(), n, cregs.data(), (uint32_t)cregs.size(), fusion, 10);
    if (!traj_fits(c)) return 3;
    const int order = std::atoi(argv[9]);
    // built for one precision (EMU_PREC), so that each build instantiates
    // half the kernels
    if (prec != EMU_PREC) return 4;
    using T = std::conditional_t<EMU_PREC == 64, double, float>;
    return run<T>(c, shots, seed, nz, order, argv[10]);
}

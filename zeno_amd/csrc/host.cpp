// host.cpp — RNG, gate library, circuit generators and the fusion compiler
// of the zeno HIP backend (host C++).  Arithmetic reproduces the reference
// expression trees so the op stream (fused matrices included) is
// bit-identical to the CPU path; build with -ffp-contract=off.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "host.hpp"

namespace zeno {

static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }
const char *last_error() { return g_error.c_str(); }

// ------------------------------------------------------------------ RNG

uint64_t splitmix64(uint64_t x) {  // exec.rs:227-232
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// rand_xoshiro 0.6.0 seed_from_u64: the 32-byte seed is four consecutive
// SplitMix64 outputs starting from `seed` (the stream's k-th output is the
// finaliser applied to seed + k*golden).
Rng::Rng(uint64_t seed) {
    for (int i = 0; i < 4; ++i) s_[i] = splitmix64(seed + 0x9E3779B97F4A7C15ull * (uint64_t)i);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t Rng::next() {  // xoshiro256++ (Blackman & Vigna)
    const uint64_t out = rotl(s_[0] + s_[3], 23) + s_[0];
    const uint64_t t = s_[1] << 17;
    s_[2] ^= s_[0];
    s_[3] ^= s_[1];
    s_[1] ^= s_[2];
    s_[0] ^= s_[3];
    s_[2] ^= t;
    s_[3] = rotl(s_[3], 45);
    return out;
}

// rand 0.8 Standard f64: top 53 bits scaled by 2^-53.
double Rng::f64() { return (double)(next() >> 11) * 0x1.0p-53; }

// rand 0.8 UniformInt<u64>::sample_single: widening multiply, accept when
// the low word is inside the zone (range << lz) - 1.
uint64_t Rng::range(uint64_t lo, uint64_t hi) {
    const uint64_t r = hi - lo;
    if (r == 0) return next();
    const uint64_t zone = (r << __builtin_clzll(r)) - 1;
    while (true) {
        const unsigned __int128 m = (unsigned __int128)next() * r;
        if ((uint64_t)m <= zone) return lo + (uint64_t)(m >> 64);
    }
}

// ---------------------------------------------------------------- gates

struct GateSig {
    const char *name;
    int arity, n_params;
};
static const GateSig kGates[] = {  // gates.rs:101-115
    {"id", 1, 0},  {"u0", 1, 0},  {"x", 1, 0},    {"y", 1, 0},   {"z", 1, 0},   {"h", 1, 0},
    {"s", 1, 0},   {"sdg", 1, 0}, {"t", 1, 0},    {"tdg", 1, 0}, {"sx", 1, 0},  {"sxdg", 1, 0},
    {"rx", 1, 1},  {"ry", 1, 1},  {"rz", 1, 1},   {"u1", 1, 1},  {"p", 1, 1},   {"u2", 1, 2},
    {"u3", 1, 3},  {"u", 1, 3},   {"cx", 2, 0},   {"cy", 2, 0},  {"cz", 2, 0},  {"ch", 2, 0},
    {"swap", 2, 0}, {"cp", 2, 1}, {"cu1", 2, 1},  {"crx", 2, 1}, {"cry", 2, 1}, {"crz", 2, 1},
    {"rxx", 2, 1}, {"rzz", 2, 1}, {"cu3", 2, 3},  {"ccx", 3, 0}, {"cswap", 3, 0},
};

int gate_lookup(const char *name, int *arity, int *n_params) {
    for (const auto &g : kGates)
        if (!std::strcmp(g.name, name)) {
            *arity = g.arity;
            *n_params = g.n_params;
            return 0;
        }
    return -1;
}

static cplx expi(double t) { return mk(std::cos(t), std::sin(t)); }

static void u3(double th, double ph, double la, cplx *m) {  // gates.rs:43-51
    const double ct = std::cos(th / 2.0), st = std::sin(th / 2.0);
    const cplx el = expi(la), ep = expi(ph), epl = expi(ph + la);
    m[0] = mk(ct, 0.0);
    m[1] = mk(-el.re * st, -el.im * st);
    m[2] = mk(ep.re * st, ep.im * st);
    m[3] = mk(epl.re * ct, epl.im * ct);
}
static void rx(double th, cplx *m) {
    const double ct = std::cos(th / 2.0), st = std::sin(th / 2.0);
    m[0] = mk(ct, 0.0), m[1] = mk(0.0, -st), m[2] = mk(0.0, -st), m[3] = mk(ct, 0.0);
}
static void ry(double th, cplx *m) {
    const double ct = std::cos(th / 2.0), st = std::sin(th / 2.0);
    m[0] = mk(ct, 0.0), m[1] = mk(-st, 0.0), m[2] = mk(st, 0.0), m[3] = mk(ct, 0.0);
}
static void rz(double th, cplx *d) { d[0] = expi(-th / 2.0), d[1] = expi(th / 2.0); }

// Controlled 1-qubit unitary, control = local bit 0 (gates.rs:69-83).
static void controlled(const cplx *u, cplx *m) {
    for (int i = 0; i < 16; ++i) m[i] = mk(0.0, 0.0);
    for (int j = 0; j < 4; ++j) {
        if ((j & 1) == 0) {
            m[j * 4 + j] = mk(1.0, 0.0);
            continue;
        }
        const int t = (j >> 1) & 1;
        for (int tp = 0; tp < 2; ++tp) m[(1 | (tp << 1)) * 4 + j] = u[tp * 2 + t];
    }
}

template <class F>
static void permutation(int dim, F perm, cplx *m) {  // gates.rs:92-98
    for (int i = 0; i < dim * dim; ++i) m[i] = mk(0.0, 0.0);
    for (int j = 0; j < dim; ++j) m[perm(j) * dim + j] = mk(1.0, 0.0);
}

int gate_build(const char *nm, const double *p, int np, cplx *m, bool *is_diag) {
    int arity, want;
    if (gate_lookup(nm, &arity, &want) < 0 || np != want) return -1;
    const std::string n(nm);
    const cplx one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
    const double sq = 0.70710678118654752440;  // FRAC_1_SQRT_2
    *is_diag = false;
    auto diag2 = [&](cplx a, cplx b) { *is_diag = true, m[0] = a, m[1] = b; };
    auto mat2 = [&](cplx a, cplx b, cplx c, cplx d) { m[0] = a, m[1] = b, m[2] = c, m[3] = d; };
    if (n == "id" || n == "u0") diag2(one, one);
    else if (n == "x") mat2(zero, one, one, zero);
    else if (n == "y") mat2(zero, mk(0.0, -1.0), mk(0.0, 1.0), zero);
    else if (n == "z") diag2(one, mk(-1.0, 0.0));
    else if (n == "h") mat2(mk(sq, 0.0), mk(sq, 0.0), mk(sq, 0.0), mk(-sq, 0.0));
    else if (n == "s") diag2(one, mk(0.0, 1.0));
    else if (n == "sdg") diag2(one, mk(0.0, -1.0));
    else if (n == "t") diag2(one, expi(0.78539816339744830962));
    else if (n == "tdg") diag2(one, expi(-0.78539816339744830962));
    else if (n == "sx") mat2(mk(0.5, 0.5), mk(0.5, -0.5), mk(0.5, -0.5), mk(0.5, 0.5));
    else if (n == "sxdg") mat2(mk(0.5, -0.5), mk(0.5, 0.5), mk(0.5, 0.5), mk(0.5, -0.5));
    else if (n == "rx") rx(p[0], m);
    else if (n == "ry") ry(p[0], m);
    else if (n == "rz") *is_diag = true, rz(p[0], m);
    else if (n == "u1" || n == "p") diag2(one, expi(p[0]));
    else if (n == "u2") u3(1.57079632679489661923, p[0], p[1], m);
    else if (n == "u3" || n == "u") u3(p[0], p[1], p[2], m);
    else if (n == "cx") permutation(4, [](int j) { return (j & 1) ? j ^ 2 : j; }, m);
    else if (n == "cy") {
        const cplx y[4] = {zero, mk(0.0, -1.0), mk(0.0, 1.0), zero};
        controlled(y, m);
    } else if (n == "cz") *is_diag = true, m[0] = one, m[1] = one, m[2] = one, m[3] = mk(-1.0, 0.0);
    else if (n == "ch") {
        const cplx h[4] = {mk(sq, 0.0), mk(sq, 0.0), mk(sq, 0.0), mk(-sq, 0.0)};
        controlled(h, m);
    } else if (n == "swap") permutation(4, [](int j) { return ((j & 1) << 1) | ((j >> 1) & 1); }, m);
    else if (n == "cp" || n == "cu1") *is_diag = true, m[0] = one, m[1] = one, m[2] = one, m[3] = expi(p[0]);
    else if (n == "crx" || n == "cry" || n == "cu3") {
        cplx u[4];
        if (n == "crx") rx(p[0], u);
        else if (n == "cry") ry(p[0], u);
        else u3(p[0], p[1], p[2], u);
        controlled(u, m);
    } else if (n == "crz") {  // gates.rs:85-89 cdiag1
        cplx d[2];
        rz(p[0], d);
        *is_diag = true, m[0] = one, m[1] = d[0], m[2] = one, m[3] = d[1];
    } else if (n == "rxx") {
        const double ct = std::cos(p[0] / 2.0), st = std::sin(p[0] / 2.0);
        for (int i = 0; i < 16; ++i) m[i] = zero;
        for (int j = 0; j < 4; ++j) m[j * 4 + j] = mk(ct, 0.0), m[(j ^ 3) * 4 + j] = mk(0.0, -st);
    } else if (n == "rzz") {
        const cplx e0 = expi(-p[0] / 2.0), e1 = expi(p[0] / 2.0);
        *is_diag = true, m[0] = e0, m[1] = e1, m[2] = e1, m[3] = e0;
    } else if (n == "ccx") permutation(8, [](int j) { return (j & 3) == 3 ? j ^ 4 : j; }, m);
    else if (n == "cswap")
        permutation(8, [](int j) { return (j & 1) ? ((j & 1) | ((j & 2) << 1) | ((j & 4) >> 1)) : j; }, m);
    else return -1;
    return arity;
}

// ------------------------------------------------------------ generators

namespace {
constexpr double kTau = 6.28318530717958647692;

struct ProgramWriter {
    zr_instr *out;
    int64_t cap, n = 0;
    void gate(const char *name, std::initializer_list<uint32_t> qs, std::initializer_list<double> ps) {
        if (n < cap) {
            zr_instr &r = out[n];
            std::memset(&r, 0, sizeof r);
            std::strncpy(r.name, name, 7);
            r.nq = (int32_t)qs.size();
            r.np = (int32_t)ps.size();
            int i = 0;
            for (uint32_t q : qs) r.q[i++] = q;
            i = 0;
            for (double p : ps) r.p[i++] = p;
            r.cond_creg = -1;
        }
        ++n;
    }
    void u3_layer(uint32_t nq, Rng &rng) {  // draw order of circuit.rs:158-164
        for (uint32_t q = 0; q < nq; ++q) {
            const double a = rng.f64() * kTau;
            const double b = rng.f64() * kTau;
            const double c = rng.f64() * kTau;
            gate("u3", {q}, {a, b, c});
        }
    }
};
}  // namespace

}  // namespace zeno

using namespace zeno;

// circuit.rs:152-174: U3 on every qubit, then the shifting CX ladder.
extern "C" int64_t zr_gen_random(uint32_t n, uint32_t depth, uint64_t seed, zr_instr *out, int64_t cap) {
    ProgramWriter w{out, cap};
    Rng rng(seed);
    for (uint32_t layer = 0; layer < depth; ++layer) {
        w.u3_layer(n, rng);
        for (uint32_t q = layer % 2; q + 1 < n; q += 2) w.gate("cx", {q, q + 1}, {});
    }
    return w.n;
}

// circuit.rs:138-148 (no final swaps).
extern "C" int64_t zr_gen_qft(uint32_t n, zr_instr *out, int64_t cap) {
    ProgramWriter w{out, cap};
    for (int64_t i = (int64_t)n - 1; i >= 0; --i) {
        w.gate("h", {(uint32_t)i}, {});
        for (int64_t j = i - 1; j >= 0; --j)
            w.gate("cp", {(uint32_t)j, (uint32_t)i}, {3.14159265358979323846 / (double)(1u << (i - j))});
    }
    return w.n;
}

// A seeded U3 layer, then qft(n).
extern "C" int64_t zr_gen_u3_qft(uint32_t n, uint64_t seed, zr_instr *out, int64_t cap) {
    ProgramWriter w{out, cap};
    Rng rng(seed);
    w.u3_layer(n, rng);
    const int64_t m = zr_gen_qft(n, w.n < cap ? out + w.n : nullptr, w.n < cap ? cap - w.n : 0);
    return w.n + m;
}

// Random-matching circuit: per layer a U3 layer, then CX on the pairs
// (p[2i], p[2i+1]) of a Fisher-Yates shuffle p of 0..n drawn from the same
// stream (for i = n-1..1: j = min(floor(u*(i+1)), i)); layers with
// layer % rzz_period == rzz_period-1 end with rzz(q, q+1, u*tau) for every
// neighbour pair.
extern "C" int64_t zr_gen_matching(uint32_t n, uint32_t depth, uint64_t seed, uint32_t rzz_period,
                                   zr_instr *out, int64_t cap) {
    ProgramWriter w{out, cap};
    Rng rng(seed);
    std::vector<uint32_t> perm(n);
    for (uint32_t layer = 0; layer < depth; ++layer) {
        w.u3_layer(n, rng);
        for (uint32_t i = 0; i < n; ++i) perm[i] = i;
        for (int64_t i = (int64_t)n - 1; i >= 1; --i) {
            uint64_t j = (uint64_t)(rng.f64() * (double)(i + 1));
            if (j > (uint64_t)i) j = (uint64_t)i;
            std::swap(perm[i], perm[j]);
        }
        for (uint32_t i = 0; i + 1 < n; i += 2) w.gate("cx", {perm[i], perm[i + 1]}, {});
        if (rzz_period > 0 && layer % rzz_period == rzz_period - 1)
            for (uint32_t q = 0; q + 1 < n; ++q) w.gate("rzz", {q, q + 1}, {rng.f64() * kTau});
    }
    return w.n;
}

extern "C" int zr_gate_build(const char *name, const double *params, int np, double *out, int *is_diag) {
    cplx m[64];
    bool d = false;
    const int k = gate_build(name, params, np, m, &d);
    if (k < 0) return -1;
    *is_diag = d;
    std::memcpy(out, m, sizeof(cplx) * (d ? (1u << k) : (1u << (2 * k))));
    return k;
}

// -------------------------------------------------------------- compiler

namespace zeno {
namespace {

bool self_inverse(const char *nm) {  // gates.rs:30-32
    static const char *L[] = {"x", "y", "z", "h", "cx", "cy", "cz", "ch", "swap", "ccx", "cswap"};
    for (const char *s : L)
        if (!std::strcmp(nm, s)) return true;
    return false;
}

bool is_gate(const zr_instr &r) { return r.kind == 0 && r.cond_creg < 0; }

// compiler.rs:215-239: drop adjacent self-inverse pairs, looking through
// gates on disjoint qubits; non-gates (and if-wrapped ops) fence.
std::vector<zr_instr> cancel(const zr_instr *in, int64_t n) {
    std::vector<zr_instr> out;
    out.reserve((size_t)n);
    for (int64_t x = 0; x < n; ++x) {
        const zr_instr &g = in[x];
        bool dropped = false;
        if (is_gate(g) && g.np == 0 && self_inverse(g.name)) {
            for (size_t i = out.size(); i-- > 0;) {
                const zr_instr &pg = out[i];
                if (!is_gate(pg)) break;
                bool overlap = false;
                for (int a = 0; a < pg.nq; ++a)
                    for (int b = 0; b < g.nq; ++b) overlap |= pg.q[a] == g.q[b];
                if (!overlap) continue;
                bool same = !std::strcmp(pg.name, g.name) && pg.nq == g.nq;
                for (int b = 0; same && b < g.nq; ++b) same = pg.q[b] == g.q[b];
                if (same) {
                    out.erase(out.begin() + (std::ptrdiff_t)i);
                    dropped = true;
                }
                break;
            }
        }
        if (!dropped) out.push_back(g);
    }
    return out;
}

int64_t count_gates(const zr_instr *p, int64_t n) {
    int64_t c = 0;
    for (int64_t i = 0; i < n; ++i) c += p[i].kind == 0;
    return c;
}

// compiler.rs:355-398: permute a matrix/diagonal from argument order to the
// sorted-qubit convention.
void to_sorted(const std::vector<uint32_t> &args, const cplx *m, bool diag, std::vector<uint32_t> &sorted,
               std::vector<cplx> &out) {
    const int k = (int)args.size(), dim = 1 << k;
    sorted = args;
    std::sort(sorted.begin(), sorted.end());
    const size_t len = diag ? (size_t)dim : (size_t)dim * dim;
    out.assign(m, m + len);
    if (sorted == args) return;
    int pos[12] = {0};
    for (int b = 0; b < k; ++b) pos[b] = (int)(std::find(sorted.begin(), sorted.end(), args[b]) - sorted.begin());
    auto arg_index = [&](int js) {
        int j = 0;
        for (int b = 0; b < k; ++b) j |= ((js >> pos[b]) & 1) << b;
        return j;
    };
    if (diag) {
        for (int j = 0; j < dim; ++j) out[j] = m[arg_index(j)];
    } else {
        for (int i = 0; i < dim; ++i) {
            const int ia = arg_index(i);
            for (int j = 0; j < dim; ++j) out[(size_t)i * dim + j] = m[(size_t)ia * dim + arg_index(j)];
        }
    }
}

std::vector<cplx> cx_sorted(uint32_t c, uint32_t t, std::vector<uint32_t> &qs) {
    cplx m[16];
    bool d;
    gate_build("cx", nullptr, 0, m, &d);
    std::vector<cplx> out;
    to_sorted({c, t}, m, false, qs, out);
    return out;
}

void fail(const std::string &msg) { throw std::runtime_error("invalid circuit: " + msg); }

// compiler.rs:242-347
void lower(const zr_instr &r, const Compiled &cc, std::vector<Op> &out) {
    Op op;
    if (r.cond_creg >= 0) {
        if ((uint32_t)r.cond_creg >= cc.creg_sizes.size()) fail("creg #" + std::to_string(r.cond_creg) + " out of range");
        uint32_t off = 0;
        for (int i = 0; i < r.cond_creg; ++i) off += cc.creg_sizes[i];
        op.is_if = true;
        op.creg_off = off;
        op.creg_len = cc.creg_sizes[r.cond_creg];
        op.value = r.cond_value;
    }
    auto check_q = [&](uint32_t q) {
        if (q >= cc.n_qubits)
            fail("qubit " + std::to_string(q) + " out of range (have " + std::to_string(cc.n_qubits) + ")");
    };
    if (r.kind == 1) {
        check_q(r.q[0]);
        if (r.clbit >= cc.n_clbits)
            fail("clbit " + std::to_string(r.clbit) + " out of range (have " + std::to_string(cc.n_clbits) + ")");
        op.kind = kMeasure, op.a = r.q[0], op.b = r.clbit;
        out.push_back(op);
        return;
    }
    if (r.kind == 2) {
        check_q(r.q[0]);
        op.kind = kReset, op.a = r.q[0];
        out.push_back(op);
        return;
    }
    if (r.kind == 3) {
        if (!op.is_if) op.kind = kBarrier, out.push_back(op);
        return;
    }
    char name[9] = {0};
    std::memcpy(name, r.name, 8);
    int arity, np;
    if (gate_lookup(name, &arity, &np) < 0) fail(std::string("unknown gate '") + name + "'");
    if (r.nq != arity)
        fail(std::string("gate '") + name + "' expects " + std::to_string(arity) + " qubit(s), got " +
             std::to_string(r.nq));
    for (int i = 0; i < r.nq; ++i) check_q(r.q[i]);
    for (int i = 0; i < r.nq; ++i)
        for (int j = 0; j < i; ++j)
            if (r.q[i] == r.q[j]) fail(std::string("gate '") + name + "' applied to duplicate qubits");
    if (!std::strcmp(name, "id") || !std::strcmp(name, "u0")) return;
    if (!std::strcmp(name, "cx")) {
        op.kind = kCx, op.a = r.q[0], op.b = r.q[1];
        out.push_back(op);
        return;
    }
    cplx m[64];
    bool diag;
    if (gate_build(name, r.p, r.np, m, &diag) < 0)
        fail(std::string("gate '") + name + "' expects " + std::to_string(np) + " parameter(s), got " +
             std::to_string(r.np));
    op.kind = diag ? kDiagonal : kUnitary;
    to_sorted(std::vector<uint32_t>(r.q, r.q + arity), m, diag, op.qs, op.data);
    out.push_back(op);
}

std::vector<uint32_t> union_sorted(const std::vector<uint32_t> &a, const std::vector<uint32_t> &b) {
    std::vector<uint32_t> u(a);
    u.insert(u.end(), b.begin(), b.end());
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    return u;
}

// compiler.rs:407-415: index over `from` of an index over `to`.
struct SubsetMap {
    int pos[24];
    int nf;
    SubsetMap(const std::vector<uint32_t> &from, const std::vector<uint32_t> &to) : nf((int)from.size()) {
        for (int b = 0; b < nf; ++b) pos[b] = (int)(std::find(to.begin(), to.end(), from[b]) - to.begin());
    }
    int operator()(int i) const {
        int j = 0;
        for (int b = 0; b < nf; ++b) j |= ((i >> pos[b]) & 1) << b;
        return j;
    }
};

// compiler.rs:419-447
std::vector<cplx> embed_unitary(const std::vector<cplx> &m, const std::vector<uint32_t> &from,
                                const std::vector<uint32_t> &to) {
    if (from == to) return m;
    const int dt = 1 << to.size(), df = 1 << from.size();
    SubsetMap sub(from, to);
    int fmask = 0;
    for (int b = 0; b < sub.nf; ++b) fmask |= 1 << sub.pos[b];
    const int rest = (dt - 1) & ~fmask;
    std::vector<cplx> out((size_t)dt * dt, mk(0.0, 0.0));
    for (int i = 0; i < dt; ++i) {
        const int ifr = sub(i), ir = i & rest;
        for (int j = 0; j < dt; ++j) {
            if ((j & rest) != ir) continue;
            out[(size_t)i * dt + j] = m[(size_t)ifr * df + sub(j)];
        }
    }
    return out;
}

// compiler.rs:449-455
std::vector<cplx> embed_diag(const std::vector<cplx> &d, const std::vector<uint32_t> &from,
                             const std::vector<uint32_t> &to) {
    if (from == to) return d;
    const int dt = 1 << to.size();
    SubsetMap sub(from, to);
    std::vector<cplx> out((size_t)dt);
    for (int i = 0; i < dt; ++i) out[i] = d[sub(i)];
    return out;
}

// compiler.rs:457-471: loop order i -> l -> j, zero a[i,l] skipped, f64.
std::vector<cplx> matmul(const std::vector<cplx> &a, const std::vector<cplx> &b, int dim) {
    std::vector<cplx> out((size_t)dim * dim, mk(0.0, 0.0));
    for (int i = 0; i < dim; ++i)
        for (int l = 0; l < dim; ++l) {
            const cplx ail = a[(size_t)i * dim + l];
            if (is_zero(ail)) continue;
            cplx *row = &out[(size_t)i * dim];
            const cplx *brow = &b[(size_t)l * dim];
            for (int j = 0; j < dim; ++j) {
                const cplx p = ail * brow[j];
                row[j].re = row[j].re + p.re;
                row[j].im = row[j].im + p.im;
            }
        }
    return out;
}

// compiler.rs:477-535
std::vector<Op> fuse_diagonals(std::vector<Op> in, size_t diag_max) {
    std::vector<Op> out;
    out.reserve(in.size());
    bool have = false;
    Op cur;
    uint64_t blocked = 0;
    auto flush = [&]() {
        if (have) out.push_back(std::move(cur)), have = false;
    };
    for (Op &op : in) {
        if (!op.is_if && op.kind == kDiagonal) {
            bool clean = true;
            for (uint32_t q : op.qs) clean &= !((blocked >> q) & 1);
            if (have && clean) {
                std::vector<uint32_t> u = union_sorted(cur.qs, op.qs);
                if (u.size() <= diag_max) {
                    std::vector<cplx> a = embed_diag(cur.data, cur.qs, u);
                    const std::vector<cplx> b = embed_diag(op.data, op.qs, u);
                    for (size_t i = 0; i < a.size(); ++i) a[i] = a[i] * b[i];
                    cur.data = std::move(a);
                    cur.qs = std::move(u);
                    continue;
                }
            }
            flush();
            blocked = 0;
            cur = std::move(op);
            have = true;
        } else if (op.gate_like()) {
            if (have) {
                bool hit = false;
                const std::vector<uint32_t> s = op.support();
                for (uint32_t q : s) hit |= std::find(cur.qs.begin(), cur.qs.end(), q) != cur.qs.end();
                if (hit) {
                    flush();
                    blocked = 0;
                } else {
                    for (uint32_t q : s) blocked |= 1ull << q;
                }
            }
            out.push_back(std::move(op));
        } else {
            flush();
            blocked = 0;
            out.push_back(std::move(op));
        }
    }
    flush();
    return out;
}

// compiler.rs:540-547 to_matrix
std::vector<cplx> to_matrix(const Op &op, std::vector<uint32_t> &qs) {
    if (op.kind == kUnitary) {
        qs = op.qs;
        return op.data;
    }
    if (op.kind == kDiagonal) {
        qs = op.qs;
        const int dim = 1 << qs.size();
        std::vector<cplx> m((size_t)dim * dim, mk(0.0, 0.0));
        for (int j = 0; j < dim; ++j) m[(size_t)j * dim + j] = op.data[j];
        return m;
    }
    return cx_sorted(op.a, op.b, qs);
}

// compiler.rs:540-626
std::vector<Op> fuse_general(std::vector<Op> in, size_t kmax) {
    std::vector<Op> out;
    out.reserve(in.size());
    bool have = false;
    int count = 0;
    Op first;
    std::vector<uint32_t> gq;
    std::vector<cplx> gm;
    auto flush = [&]() {
        if (!have) return;
        if (count == 1) {
            out.push_back(std::move(first));
        } else {
            Op f;
            f.kind = kUnitary;
            f.qs = gq;
            f.data = std::move(gm);
            out.push_back(std::move(f));
        }
        have = false;
    };
    for (Op &op : in) {
        if (!(op.gate_like() && op.support().size() <= kmax)) {
            flush();
            out.push_back(std::move(op));
            continue;
        }
        std::vector<uint32_t> qs;
        std::vector<cplx> m = to_matrix(op, qs);
        if (have) {
            std::vector<uint32_t> u = union_sorted(gq, qs);
            if (u.size() <= kmax) {
                const std::vector<cplx> ce = embed_unitary(gm, gq, u);
                const std::vector<cplx> oe = embed_unitary(m, qs, u);
                gm = matmul(oe, ce, 1 << u.size());
                gq = std::move(u);
                ++count;
                continue;
            }
            flush();
        }
        have = true;
        count = 1;
        first = std::move(op);
        gq = std::move(qs);
        gm = std::move(m);
    }
    flush();
    return out;
}

}  // namespace

// compiler.rs:118-200
Compiled compile(const zr_instr *prog, int64_t n, uint32_t n_qubits, const uint32_t *creg_sizes, uint32_t n_cregs,
                 int fusion_max, int diag_max) {
    Compiled cc;
    cc.n_qubits = n_qubits;
    cc.creg_sizes.assign(creg_sizes, creg_sizes + n_cregs);
    for (uint32_t s : cc.creg_sizes) cc.n_clbits += s;
    if (n_qubits == 0) fail("no qubits declared");
    if (n_qubits > 48) fail(std::to_string(n_qubits) + " qubits is beyond any state-vector budget");
    if (cc.n_clbits > 64) fail("more than 64 classical bits are not supported");
    cc.stats.input_gates = count_gates(prog, n);
    const std::vector<zr_instr> kept = cancel(prog, n);
    cc.stats.cancelled = cc.stats.input_gates - count_gates(kept.data(), (int64_t)kept.size());
    std::vector<Op> ops;
    ops.reserve(kept.size());
    for (const zr_instr &r : kept) lower(r, cc, ops);
    fusion_max = std::min(std::max(fusion_max, 0), 6);
    diag_max = std::min(std::max(diag_max, 0), 12);
    if (fusion_max >= 1) {
        ops = fuse_diagonals(std::move(ops), (size_t)diag_max);
        ops = fuse_general(std::move(ops), (size_t)fusion_max);
    }
    ops.erase(std::remove_if(ops.begin(), ops.end(), [](const Op &o) { return o.kind == kBarrier; }), ops.end());
    size_t split = ops.size();
    while (split > 0 && ops[split - 1].kind == kMeasure && !ops[split - 1].is_if) --split;
    bool dyn = false;
    for (size_t i = 0; i < split; ++i) dyn |= ops[i].is_if || ops[i].kind == kMeasure || ops[i].kind == kReset;
    if (!dyn) {
        for (size_t i = split; i < ops.size(); ++i) cc.final_measures.emplace_back(ops[i].a, ops[i].b);
        ops.resize(split);
    }
    cc.dynamic = dyn;
    cc.stats.output_ops = (int64_t)ops.size();
    for (const Op &o : ops)
        if (o.gate_like()) cc.stats.max_fused = std::max<int64_t>(cc.stats.max_fused, (int64_t)o.support().size());
    cc.ops = std::move(ops);
    return cc;
}

}  // namespace zeno

/*
 * zeno_oracle.c — CPU restatement of cleoanka/zeno's state-vector path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Every function cites the reference file:line it restates.  Pins:
 *   rng        — rand_xoshiro's published xoshiro256++ test vector and the
 *                SplitMix64 reference outputs (tests/test_oracle.py);
 *                gen_range (rand 0.8.7 UniformInt<u64>, not in the tree):
 *                restated from its published algorithm and held by
 *                crafted-state vectors on the rejection zone's edges plus a
 *                uniformity check; parity against the crate's own outputs
 *                is unpinned.
 *   gates      — hand-coded standard matrices (tests/reference.rs:157-283
 *                restated independently in numpy), unitarity.
 *   compiler   — compiler.rs:645-726 KATs; RefSim at every fusion level.
 *   kernels    — RefSim / numpy Kronecker unitaries; state.rs:1214-1295
 *                KATs; closed forms (QFT = DFT, GHZ, Bell).
 *   sampler    — brute force inverse-CDF on small states; 6-sigma binomial.
 *   reductions — prob_one / norm_sqr use the fixed order of
 *                zo_fixed_reduce; checked against math.fsum.
 */
#include "zeno_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct {
    double re, im;
} zc;

static zc cmul(zc a, zc b) {
    /* num-complex 0.4 Mul: (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re) */
    zc r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}
static zc cmk(double re, double im) {
    zc r;
    r.re = re;
    r.im = im;
    return r;
}
static zc expi(double t) { return cmk(cos(t), sin(t)); } /* gates.rs:38-40 */
static int czero(zc a) { return a.re == 0.0 && a.im == 0.0; }

static __thread char g_err[512];
const char *zo_error(void) { return g_err; }
static void set_err(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

/* ================================================================== RNG */

/* exec.rs:227-232 */
uint64_t zo_splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

typedef struct {
    uint64_t s[4];
} rng_t;

static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* rand_xoshiro 0.6.0 Xoshiro256PlusPlus::seed_from_u64: four outputs of
 * the SplitMix64 stream started at `seed` fill s[0..4] (little-endian
 * read of the 32-byte seed); used at circuit.rs:154, exec.rs:214, :486. */
static void rng_seed(rng_t *r, uint64_t seed) {
    uint64_t st = seed;
    for (int i = 0; i < 4; i++) {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        r->s[i] = z ^ (z >> 31);
    }
}

/* xoshiro256++ (Blackman & Vigna 2019). */
static uint64_t rng_next(rng_t *r) {
    uint64_t *s = r->s;
    uint64_t res = rotl(s[0] + s[3], 23) + s[0];
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return res;
}

/* rand 0.8 Standard for f64: 53 high bits times 2^-53, in [0,1). */
static double rng_f64(rng_t *r) {
    return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0);
}

/* rand 0.8 UniformInt<u64>::sample_single(lo, hi) = sample_single_inclusive
 * (lo, hi-1): widening multiply with the "leading zeros" rejection zone.
 * Used by noise.rs:158 (see the file header for what pins it). */
static uint64_t rng_range(rng_t *r, uint64_t lo, uint64_t hi) {
    uint64_t range = hi - lo;
    if (range == 0) return rng_next(r);
    int lz = __builtin_clzll(range);
    uint64_t zone = (range << lz) - 1;
    for (;;) {
        uint64_t v = rng_next(r);
        unsigned __int128 m = (unsigned __int128)v * range;
        uint64_t mlo = (uint64_t)m, mhi = (uint64_t)(m >> 64);
        if (mlo <= zone) return lo + mhi;
    }
}

void zo_rng_u64_stream(uint64_t seed, uint64_t n, uint64_t *out) {
    rng_t r;
    rng_seed(&r, seed);
    for (uint64_t i = 0; i < n; i++) out[i] = rng_next(&r);
}
/* Raw xoshiro256++ outputs from an explicit state (for the published
 * test vector of the generator). */
void zo_xoshiro_from_state(const uint64_t *st, uint64_t n, uint64_t *out) {
    rng_t r;
    for (int i = 0; i < 4; i++) r.s[i] = st[i];
    for (uint64_t i = 0; i < n; i++) out[i] = rng_next(&r);
}
void zo_rng_f64_stream(uint64_t seed, uint64_t n, double *out) {
    rng_t r;
    rng_seed(&r, seed);
    for (uint64_t i = 0; i < n; i++) out[i] = rng_f64(&r);
}
/* rng_range draws from an explicit generator state (the boundary tests
 * craft states whose first output lands on the rejection zone's edge). */
void zo_rng_range_from_state(const uint64_t *st, uint64_t lo, uint64_t hi, uint64_t n,
                             uint64_t *out) {
    rng_t r;
    for (int i = 0; i < 4; i++) r.s[i] = st[i];
    for (uint64_t i = 0; i < n; i++) out[i] = rng_range(&r, lo, hi);
}
void zo_rng_range_stream(uint64_t seed, uint64_t lo, uint64_t hi, uint64_t n,
                         uint64_t *out) {
    rng_t r;
    rng_seed(&r, seed);
    for (uint64_t i = 0; i < n; i++) out[i] = rng_range(&r, lo, hi);
}

/* ================================================================ gates */

/* gates.rs:101-115 */
int zo_gate_lookup(const char *nm, int *arity, int *np) {
#define G(s, a, p)                                                             \
    if (!strcmp(nm, s)) {                                                      \
        *arity = a;                                                            \
        *np = p;                                                               \
        return 0;                                                              \
    }
    G("id", 1, 0) G("u0", 1, 0) G("x", 1, 0) G("y", 1, 0) G("z", 1, 0)
    G("h", 1, 0) G("s", 1, 0) G("sdg", 1, 0) G("t", 1, 0) G("tdg", 1, 0)
    G("sx", 1, 0) G("sxdg", 1, 0) G("rx", 1, 1) G("ry", 1, 1) G("rz", 1, 1)
    G("u1", 1, 1) G("p", 1, 1) G("u2", 1, 2) G("u3", 1, 3) G("u", 1, 3)
    G("cx", 2, 0) G("cy", 2, 0) G("cz", 2, 0) G("ch", 2, 0) G("swap", 2, 0)
    G("cp", 2, 1) G("cu1", 2, 1) G("crx", 2, 1) G("cry", 2, 1) G("crz", 2, 1)
    G("rxx", 2, 1) G("rzz", 2, 1) G("cu3", 2, 3) G("ccx", 3, 0)
    G("cswap", 3, 0)
#undef G
    return -1;
}

/* gates.rs:43-51 */
static void g_u3(double th, double ph, double la, zc *m) {
    double ct = cos(th / 2.0), st = sin(th / 2.0);
    zc el = expi(la), ep = expi(ph), epl = expi(ph + la);
    m[0] = cmk(ct, 0.0);
    m[1] = cmk(-el.re * st, -el.im * st);
    m[2] = cmk(ep.re * st, ep.im * st);
    m[3] = cmk(epl.re * ct, epl.im * ct);
}
static void g_rx(double th, zc *m) { /* gates.rs:53-56 */
    double ct = cos(th / 2.0), st = sin(th / 2.0);
    m[0] = cmk(ct, 0.0);
    m[1] = cmk(0.0, -st);
    m[2] = cmk(0.0, -st);
    m[3] = cmk(ct, 0.0);
}
static void g_ry(double th, zc *m) { /* gates.rs:58-61 */
    double ct = cos(th / 2.0), st = sin(th / 2.0);
    m[0] = cmk(ct, 0.0);
    m[1] = cmk(-st, 0.0);
    m[2] = cmk(st, 0.0);
    m[3] = cmk(ct, 0.0);
}
static void g_rz(double th, zc *d) { /* gates.rs:63-65 */
    d[0] = expi(-th / 2.0);
    d[1] = expi(th / 2.0);
}
/* gates.rs:69-83: control = bit 0, target = bit 1 */
static void g_ctrl1(const zc *u, zc *m) {
    for (int i = 0; i < 16; i++) m[i] = cmk(0.0, 0.0);
    for (int j = 0; j < 4; j++) {
        int ctrl = j & 1, t = (j >> 1) & 1;
        if (ctrl == 0) {
            m[j * 4 + j] = cmk(1.0, 0.0);
        } else {
            for (int tp = 0; tp < 2; tp++) {
                int i = 1 | (tp << 1);
                m[i * 4 + j] = u[tp * 2 + t];
            }
        }
    }
}
static int perm_cx(int j) { return (j & 1) ? j ^ 2 : j; }
static int perm_swap(int j) { return ((j & 1) << 1) | ((j >> 1) & 1); }
static int perm_ccx(int j) { return (j & 3) == 3 ? j ^ 4 : j; }
static int perm_cswap(int j) {
    return (j & 1) ? ((j & 1) | ((j & 2) << 1) | ((j & 4) >> 1)) : j;
}
/* gates.rs:92-98 */
static void g_perm(int dim, int (*perm)(int), zc *m) {
    for (int i = 0; i < dim * dim; i++) m[i] = cmk(0.0, 0.0);
    for (int j = 0; j < dim; j++) m[perm(j) * dim + j] = cmk(1.0, 0.0);
}

/* gates.rs:119-188.  Writes zc[4^k] (unitary) or zc[2^k] (diagonal). */
static int gate_build(const char *nm, const double *p, int np, zc *m,
                      int *is_diag) {
    int arity, nparams;
    if (zo_gate_lookup(nm, &arity, &nparams) < 0) return -1;
    if (np != nparams) return -1;
    const zc one = cmk(1.0, 0.0), zero = cmk(0.0, 0.0);
    const double sq = 0.70710678118654752440; /* FRAC_1_SQRT_2 */
    *is_diag = 0;
#define D1(a, b) (*is_diag = 1, m[0] = (a), m[1] = (b))
#define U1(a, b, c, d) (m[0] = (a), m[1] = (b), m[2] = (c), m[3] = (d))
    if (!strcmp(nm, "id") || !strcmp(nm, "u0")) D1(one, one);
    else if (!strcmp(nm, "x")) U1(zero, one, one, zero);
    else if (!strcmp(nm, "y")) U1(zero, cmk(0.0, -1.0), cmk(0.0, 1.0), zero);
    else if (!strcmp(nm, "z")) D1(one, cmk(-1.0, 0.0));
    else if (!strcmp(nm, "h"))
        U1(cmk(sq, 0.0), cmk(sq, 0.0), cmk(sq, 0.0), cmk(-sq, 0.0));
    else if (!strcmp(nm, "s")) D1(one, cmk(0.0, 1.0));
    else if (!strcmp(nm, "sdg")) D1(one, cmk(0.0, -1.0));
    else if (!strcmp(nm, "t")) D1(one, expi(0.78539816339744830962));
    else if (!strcmp(nm, "tdg")) D1(one, expi(-0.78539816339744830962));
    else if (!strcmp(nm, "sx"))
        U1(cmk(0.5, 0.5), cmk(0.5, -0.5), cmk(0.5, -0.5), cmk(0.5, 0.5));
    else if (!strcmp(nm, "sxdg"))
        U1(cmk(0.5, -0.5), cmk(0.5, 0.5), cmk(0.5, 0.5), cmk(0.5, -0.5));
    else if (!strcmp(nm, "rx")) g_rx(p[0], m);
    else if (!strcmp(nm, "ry")) g_ry(p[0], m);
    else if (!strcmp(nm, "rz")) {
        *is_diag = 1;
        g_rz(p[0], m);
    } else if (!strcmp(nm, "u1") || !strcmp(nm, "p")) D1(one, expi(p[0]));
    else if (!strcmp(nm, "u2")) g_u3(1.57079632679489661923, p[0], p[1], m);
    else if (!strcmp(nm, "u3") || !strcmp(nm, "u")) g_u3(p[0], p[1], p[2], m);
    else if (!strcmp(nm, "cx")) g_perm(4, perm_cx, m);
    else if (!strcmp(nm, "cy")) {
        zc y[4] = {zero, cmk(0.0, -1.0), cmk(0.0, 1.0), zero};
        g_ctrl1(y, m);
    } else if (!strcmp(nm, "cz")) {
        *is_diag = 1;
        m[0] = one, m[1] = one, m[2] = one, m[3] = cmk(-1.0, 0.0);
    } else if (!strcmp(nm, "ch")) {
        zc h[4] = {cmk(sq, 0.0), cmk(sq, 0.0), cmk(sq, 0.0), cmk(-sq, 0.0)};
        g_ctrl1(h, m);
    } else if (!strcmp(nm, "swap")) g_perm(4, perm_swap, m);
    else if (!strcmp(nm, "cp") || !strcmp(nm, "cu1")) {
        *is_diag = 1;
        m[0] = one, m[1] = one, m[2] = one, m[3] = expi(p[0]);
    } else if (!strcmp(nm, "crx")) {
        zc u[4];
        g_rx(p[0], u);
        g_ctrl1(u, m);
    } else if (!strcmp(nm, "cry")) {
        zc u[4];
        g_ry(p[0], u);
        g_ctrl1(u, m);
    } else if (!strcmp(nm, "crz")) { /* gates.rs:87-89 cdiag1 */
        zc d[2];
        g_rz(p[0], d);
        *is_diag = 1;
        m[0] = one, m[1] = d[0], m[2] = one, m[3] = d[1];
    } else if (!strcmp(nm, "cu3")) {
        zc u[4];
        g_u3(p[0], p[1], p[2], u);
        g_ctrl1(u, m);
    } else if (!strcmp(nm, "rxx")) {
        double ct = cos(p[0] / 2.0), st = sin(p[0] / 2.0);
        for (int i = 0; i < 16; i++) m[i] = zero;
        for (int j = 0; j < 4; j++) {
            m[j * 4 + j] = cmk(ct, 0.0);
            m[(j ^ 3) * 4 + j] = cmk(0.0, -st);
        }
    } else if (!strcmp(nm, "rzz")) {
        zc e0 = expi(-p[0] / 2.0), e1 = expi(p[0] / 2.0);
        *is_diag = 1;
        m[0] = e0, m[1] = e1, m[2] = e1, m[3] = e0;
    } else if (!strcmp(nm, "ccx")) g_perm(8, perm_ccx, m);
    else if (!strcmp(nm, "cswap")) g_perm(8, perm_cswap, m);
    else return -1;
#undef D1
#undef U1
    return arity;
}

int zo_gate_build(const char *name, const double *params, int np,
                  double *out, int *is_diag) {
    zc m[64];
    int k = gate_build(name, params, np, m, is_diag);
    if (k < 0) return -1;
    int len = *is_diag ? (1 << k) : (1 << (2 * k));
    memcpy(out, m, sizeof(zc) * (size_t)len);
    return k;
}

/* =========================================================== generators */

static void put_gate(zo_instr *o, const char *nm, int nq, const uint32_t *q,
                     int np, const double *p) {
    memset(o, 0, sizeof *o);
    strncpy(o->name, nm, 7);
    o->kind = 0;
    o->nq = nq;
    o->np = np;
    for (int i = 0; i < nq; i++) o->q[i] = q[i];
    for (int i = 0; i < np; i++) o->p[i] = p[i];
    o->cond_creg = -1;
}

#define TAU 6.28318530717958647692

/* circuit.rs:152-174 (brickwork ladder) */
int64_t zo_gen_random(uint32_t n, uint32_t depth, uint64_t seed,
                      zo_instr *out, int64_t cap) {
    rng_t r;
    rng_seed(&r, seed);
    int64_t c = 0;
    for (uint32_t layer = 0; layer < depth; layer++) {
        for (uint32_t q = 0; q < n; q++) {
            double p[3];
            p[0] = rng_f64(&r) * TAU;
            p[1] = rng_f64(&r) * TAU;
            p[2] = rng_f64(&r) * TAU;
            if (c < cap) put_gate(&out[c], "u3", 1, &q, 3, p);
            c++;
        }
        for (uint32_t q = layer % 2; q + 1 < n; q += 2) {
            uint32_t qq[2] = {q, q + 1};
            if (c < cap) put_gate(&out[c], "cx", 2, qq, 0, NULL);
            c++;
        }
    }
    return c;
}

/* circuit.rs:138-148 (no final swaps) */
int64_t zo_gen_qft(uint32_t n, zo_instr *out, int64_t cap) {
    int64_t c = 0;
    for (int64_t i = (int64_t)n - 1; i >= 0; i--) {
        uint32_t qi = (uint32_t)i;
        if (c < cap) put_gate(&out[c], "h", 1, &qi, 0, NULL);
        c++;
        for (int64_t j = i - 1; j >= 0; j--) {
            double ang = 3.14159265358979323846 / (double)(1u << (i - j));
            uint32_t qq[2] = {(uint32_t)j, qi};
            if (c < cap) put_gate(&out[c], "cp", 2, qq, 1, &ang);
            c++;
        }
    }
    return c;
}

/* A seeded U3 layer (draw order of circuit.rs:158-164) followed by qft(n). */
int64_t zo_gen_u3_qft(uint32_t n, uint64_t seed, zo_instr *out, int64_t cap) {
    rng_t r;
    rng_seed(&r, seed);
    int64_t c = 0;
    for (uint32_t q = 0; q < n; q++) {
        double p[3];
        p[0] = rng_f64(&r) * TAU;
        p[1] = rng_f64(&r) * TAU;
        p[2] = rng_f64(&r) * TAU;
        if (c < cap) put_gate(&out[c], "u3", 1, &q, 3, p);
        c++;
    }
    int64_t m = zo_gen_qft(n, c < cap ? out + c : NULL, c < cap ? cap - c : 0);
    return c + m;
}

/* Random-matching generator (not in the reference).  Per layer: U3 on
 * every qubit with the circuit.rs:158-164 draw order; a Fisher-Yates
 * shuffle of 0..n (for i = n-1 .. 1: j = min(floor(u*(i+1)), i), u the
 * next f64 draw) pairs (p[2i], p[2i+1]) with CX; every `rzz_period`-th
 * layer (layer % period == period-1) appends rzz(q, q+1, u*tau) for
 * q = 0..n-2. */
int64_t zo_gen_matching(uint32_t n, uint32_t depth, uint64_t seed,
                        uint32_t rzz_period, zo_instr *out, int64_t cap) {
    rng_t r;
    rng_seed(&r, seed);
    int64_t c = 0;
    uint32_t *perm = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t layer = 0; layer < depth; layer++) {
        for (uint32_t q = 0; q < n; q++) {
            double p[3];
            p[0] = rng_f64(&r) * TAU;
            p[1] = rng_f64(&r) * TAU;
            p[2] = rng_f64(&r) * TAU;
            if (c < cap) put_gate(&out[c], "u3", 1, &q, 3, p);
            c++;
        }
        for (uint32_t i = 0; i < n; i++) perm[i] = i;
        for (int64_t i = (int64_t)n - 1; i >= 1; i--) {
            uint64_t j = (uint64_t)(rng_f64(&r) * (double)(i + 1));
            if (j > (uint64_t)i) j = (uint64_t)i;
            uint32_t t = perm[i];
            perm[i] = perm[j];
            perm[j] = t;
        }
        for (uint32_t i = 0; i + 1 < n; i += 2) {
            uint32_t qq[2] = {perm[i], perm[i + 1]};
            if (c < cap) put_gate(&out[c], "cx", 2, qq, 0, NULL);
            c++;
        }
        if (rzz_period > 0 && layer % rzz_period == rzz_period - 1) {
            for (uint32_t q = 0; q + 1 < n; q++) {
                uint32_t qq[2] = {q, q + 1};
                double th = rng_f64(&r) * TAU;
                if (c < cap) put_gate(&out[c], "rzz", 2, qq, 1, &th);
                c++;
            }
        }
    }
    free(perm);
    return c;
}

/* ============================================================= compiler */

enum { OP_U = 0, OP_D = 1, OP_CX = 2, OP_MEAS = 3, OP_RESET = 4, OP_BARRIER = 5 };

typedef struct {
    int kind;
    int k; /* support size (U/D) */
    uint32_t qs[12];
    zc *data;     /* U: 4^k, D: 2^k */
    uint32_t c, t; /* CX control/target; MEAS/RESET qubit in c, clbit in t */
    int is_if;
    uint32_t creg_off, creg_len;
    uint64_t value;
} op_t;

typedef struct {
    op_t *v;
    int64_t n, cap;
} opvec;

static void ov_push(opvec *o, op_t op) {
    if (o->n == o->cap) {
        o->cap = o->cap ? o->cap * 2 : 64;
        o->v = (op_t *)realloc(o->v, sizeof(op_t) * (size_t)o->cap);
    }
    o->v[o->n++] = op;
}

struct zo_compiled {
    uint32_t n_qubits, n_clbits;
    opvec ops;
    int dynamic;
    uint32_t *fm; /* final measures (qubit, clbit) pairs */
    int64_t nfm;
    int64_t input_gates, cancelled, output_ops, max_fused;
    uint32_t *creg_sizes;
    uint32_t n_cregs;
};

static int is_self_inverse(const char *nm) { /* gates.rs:30-32 */
    static const char *L[] = {"x", "y", "z", "h", "cx", "cy", "cz", "ch",
                              "swap", "ccx", "cswap"};
    for (size_t i = 0; i < sizeof L / sizeof *L; i++)
        if (!strcmp(nm, L[i])) return 1;
    return 0;
}

static int64_t count_gates(const zo_instr *p, int64_t n) { /* :202-211 */
    int64_t c = 0;
    for (int64_t i = 0; i < n; i++)
        if (p[i].kind == 0) c++;
    return c;
}

static int qs_overlap(const zo_instr *a, const zo_instr *b) {
    for (int i = 0; i < a->nq; i++)
        for (int j = 0; j < b->nq; j++)
            if (a->q[i] == b->q[j]) return 1;
    return 0;
}

/* compiler.rs:215-239.  An `if`-wrapped instruction is not Instr::Gate
 * and fences like any non-gate. */
static int64_t cancel(const zo_instr *in, int64_t n, zo_instr *out) {
    int64_t m = 0;
    for (int64_t x = 0; x < n; x++) {
        const zo_instr *g = &in[x];
        int dropped = 0;
        if (g->kind == 0 && g->cond_creg < 0 && g->np == 0 &&
            is_self_inverse(g->name)) {
            for (int64_t i = m - 1; i >= 0; i--) {
                const zo_instr *pg = &out[i];
                if (pg->kind != 0 || pg->cond_creg >= 0) break;
                if (qs_overlap(pg, g)) {
                    int same = !strcmp(pg->name, g->name) && pg->nq == g->nq;
                    for (int b = 0; same && b < g->nq; b++)
                        same = pg->q[b] == g->q[b];
                    if (same) {
                        memmove(&out[i], &out[i + 1],
                                sizeof(zo_instr) * (size_t)(m - i - 1));
                        m--;
                        dropped = 1;
                    }
                    break;
                }
            }
        }
        if (!dropped) out[m++] = *g;
    }
    return m;
}

static void sort_u32(uint32_t *a, int n) {
    for (int i = 1; i < n; i++)
        for (int j = i; j > 0 && a[j - 1] > a[j]; j--) {
            uint32_t t = a[j];
            a[j] = a[j - 1];
            a[j - 1] = t;
        }
}

/* compiler.rs:355-364 arg_index */
static int arg_index(int js, const int *pos, int k) {
    int j = 0;
    for (int b = 0; b < k; b++) j |= ((js >> pos[b]) & 1) << b;
    return j;
}

/* compiler.rs:366-398.  Returns sorted qubits in sq and the permuted data
 * (newly allocated). */
static zc *permute_to_sorted(const uint32_t *q, int k, const zc *m, int diag,
                             uint32_t *sq) {
    int pos[12] = {0};
    for (int i = 0; i < k; i++) sq[i] = q[i];
    sort_u32(sq, k);
    for (int b = 0; b < k; b++)
        for (int i = 0; i < k; i++)
            if (sq[i] == q[b]) pos[b] = i;
    int dim = 1 << k;
    size_t len = diag ? (size_t)dim : (size_t)dim * dim;
    zc *out = (zc *)malloc(sizeof(zc) * len);
    int same = 1;
    for (int i = 0; i < k; i++) same &= sq[i] == q[i];
    if (same) {
        memcpy(out, m, sizeof(zc) * len);
        return out;
    }
    if (diag) {
        for (int j = 0; j < dim; j++) out[j] = m[arg_index(j, pos, k)];
    } else {
        for (int i = 0; i < dim; i++) {
            int ia = arg_index(i, pos, k);
            for (int j = 0; j < dim; j++)
                out[i * dim + j] = m[ia * dim + arg_index(j, pos, k)];
        }
    }
    return out;
}

static int check_q(uint32_t q, uint32_t n) {
    if (q >= n) {
        set_err("invalid circuit: qubit %u out of range (have %u)", q, n);
        return -1;
    }
    return 0;
}

/* compiler.rs:242-347: lower one (possibly if-wrapped) instruction. */
static int lower(const zo_instr *ins, const zo_compiled *cc, opvec *out) {
    op_t op;
    memset(&op, 0, sizeof op);
    if (ins->cond_creg >= 0) {
        if ((uint32_t)ins->cond_creg >= cc->n_cregs) {
            set_err("invalid circuit: creg #%d out of range", ins->cond_creg);
            return -1;
        }
        uint32_t off = 0;
        for (int i = 0; i < ins->cond_creg; i++) off += cc->creg_sizes[i];
        op.is_if = 1;
        op.creg_off = off;
        op.creg_len = cc->creg_sizes[ins->cond_creg];
        op.value = ins->cond_value;
    }
    if (ins->kind == 1) {
        if (check_q(ins->q[0], cc->n_qubits)) return -1;
        if (ins->clbit >= cc->n_clbits) {
            set_err("invalid circuit: clbit %u out of range (have %u)",
                    ins->clbit, cc->n_clbits);
            return -1;
        }
        op.kind = OP_MEAS;
        op.c = ins->q[0];
        op.t = ins->clbit;
        ov_push(out, op);
        return 0;
    }
    if (ins->kind == 2) {
        if (check_q(ins->q[0], cc->n_qubits)) return -1;
        op.kind = OP_RESET;
        op.c = ins->q[0];
        ov_push(out, op);
        return 0;
    }
    if (ins->kind == 3) {
        if (op.is_if) return 0; /* an if-wrapped barrier lowers to nothing */
        op.kind = OP_BARRIER;
        ov_push(out, op);
        return 0;
    }
    int arity, np;
    if (zo_gate_lookup(ins->name, &arity, &np) < 0) {
        set_err("invalid circuit: unknown gate '%s'", ins->name);
        return -1;
    }
    if (ins->nq != arity) {
        set_err("invalid circuit: gate '%s' expects %d qubit(s), got %d",
                ins->name, arity, ins->nq);
        return -1;
    }
    for (int i = 0; i < ins->nq; i++)
        if (check_q(ins->q[i], cc->n_qubits)) return -1;
    for (int i = 0; i < ins->nq; i++)
        for (int j = 0; j < i; j++)
            if (ins->q[i] == ins->q[j]) {
                set_err("invalid circuit: gate '%s' applied to duplicate "
                        "qubits",
                        ins->name);
                return -1;
            }
    if (!strcmp(ins->name, "id") || !strcmp(ins->name, "u0")) return 0;
    if (!strcmp(ins->name, "cx")) {
        op.kind = OP_CX;
        op.c = ins->q[0];
        op.t = ins->q[1];
        ov_push(out, op);
        return 0;
    }
    zc m[64];
    int diag;
    if (gate_build(ins->name, ins->p, ins->np, m, &diag) < 0) {
        set_err("invalid circuit: gate '%s' expects %d parameter(s), got %d",
                ins->name, np, ins->np);
        return -1;
    }
    op.kind = diag ? OP_D : OP_U;
    op.k = arity;
    op.data = permute_to_sorted(ins->q, arity, m, diag, op.qs);
    ov_push(out, op);
    return 0;
}

static int union_sorted(const uint32_t *a, int na, const uint32_t *b, int nb,
                        uint32_t *u) {
    int n = 0;
    for (int i = 0; i < na; i++) u[n++] = a[i];
    for (int i = 0; i < nb; i++) {
        int dup = 0;
        for (int j = 0; j < n; j++) dup |= u[j] == b[i];
        if (!dup) {
            if (n >= 24) return 99;
            u[n++] = b[i];
        }
    }
    sort_u32(u, n);
    return n;
}

/* compiler.rs:407-415 */
static int subset_index(int ito, const uint32_t *from, int nf,
                        const uint32_t *to, int nt) {
    int j = 0;
    for (int b = 0; b < nf; b++) {
        int pos = 0;
        for (int i = 0; i < nt; i++)
            if (to[i] == from[b]) pos = i;
        j |= ((ito >> pos) & 1) << b;
    }
    return j;
}

/* compiler.rs:419-447 */
static zc *embed_unitary(const zc *m, const uint32_t *from, int nf,
                         const uint32_t *to, int nt) {
    int dt = 1 << nt, df = 1 << nf;
    zc *out = (zc *)calloc((size_t)dt * dt, sizeof(zc));
    int same = nf == nt;
    for (int i = 0; same && i < nf; i++) same = from[i] == to[i];
    if (same) {
        memcpy(out, m, sizeof(zc) * (size_t)dt * dt);
        return out;
    }
    int fmask = 0;
    for (int b = 0; b < nf; b++)
        for (int i = 0; i < nt; i++)
            if (to[i] == from[b]) fmask |= 1 << i;
    int rest = (dt - 1) & ~fmask;
    for (int i = 0; i < dt; i++) {
        int ifr = subset_index(i, from, nf, to, nt);
        int ir = i & rest;
        for (int j = 0; j < dt; j++) {
            if ((j & rest) != ir) continue;
            int jfr = subset_index(j, from, nf, to, nt);
            out[i * dt + j] = m[ifr * df + jfr];
        }
    }
    return out;
}

/* compiler.rs:449-455 */
static zc *embed_diag(const zc *d, const uint32_t *from, int nf,
                      const uint32_t *to, int nt) {
    int dt = 1 << nt;
    zc *out = (zc *)malloc(sizeof(zc) * (size_t)dt);
    for (int i = 0; i < dt; i++) out[i] = d[subset_index(i, from, nf, to, nt)];
    return out;
}

/* compiler.rs:457-471: i -> l -> j, skipping zero a[i,l]. */
static zc *matmul(const zc *a, const zc *b, int dim) {
    zc *out = (zc *)calloc((size_t)dim * dim, sizeof(zc));
    for (int i = 0; i < dim; i++)
        for (int l = 0; l < dim; l++) {
            zc ail = a[i * dim + l];
            if (czero(ail)) continue;
            for (int j = 0; j < dim; j++) {
                zc p = cmul(ail, b[l * dim + j]);
                out[i * dim + j].re = out[i * dim + j].re + p.re;
                out[i * dim + j].im = out[i * dim + j].im + p.im;
            }
        }
    return out;
}

static int gate_like(const op_t *o) {
    return !o->is_if && (o->kind == OP_U || o->kind == OP_D || o->kind == OP_CX);
}
/* COp::gate_support (compiler.rs:58-65): Cx gives [control, target]. */
static int support(const op_t *o, uint32_t *s) {
    if (o->kind == OP_CX) {
        s[0] = o->c;
        s[1] = o->t;
        return 2;
    }
    for (int i = 0; i < o->k; i++) s[i] = o->qs[i];
    return o->k;
}

/* compiler.rs:477-535 */
static opvec fuse_diagonals(opvec in, int diag_max) {
    opvec out = {0};
    int have = 0;
    op_t cur;
    memset(&cur, 0, sizeof cur);
    uint64_t blocked = 0; /* qubit set (n <= 48 < 64) */
    for (int64_t x = 0; x < in.n; x++) {
        op_t op = in.v[x];
        if (!op.is_if && op.kind == OP_D) {
            int clean = 1;
            for (int i = 0; i < op.k; i++)
                if (blocked >> op.qs[i] & 1) clean = 0;
            uint32_t u[32];
            int nu = have ? union_sorted(cur.qs, cur.k, op.qs, op.k, u) : 99;
            if (have && clean && nu <= diag_max) {
                zc *a = embed_diag(cur.data, cur.qs, cur.k, u, nu);
                zc *b = embed_diag(op.data, op.qs, op.k, u, nu);
                for (int i = 0; i < (1 << nu); i++) a[i] = cmul(a[i], b[i]);
                free(b);
                free(cur.data);
                free(op.data);
                cur.data = a;
                cur.k = nu;
                memcpy(cur.qs, u, sizeof(uint32_t) * (size_t)nu);
            } else {
                if (have) ov_push(&out, cur);
                blocked = 0;
                cur = op;
                have = 1;
            }
        } else if (gate_like(&op)) {
            if (have) {
                uint32_t s[12];
                int ns = support(&op, s);
                int hit = 0;
                for (int i = 0; i < ns; i++)
                    for (int j = 0; j < cur.k; j++) hit |= s[i] == cur.qs[j];
                if (hit) {
                    ov_push(&out, cur);
                    have = 0;
                    blocked = 0;
                } else {
                    for (int i = 0; i < ns; i++) blocked |= 1ull << s[i];
                }
            }
            ov_push(&out, op);
        } else {
            if (have) ov_push(&out, cur);
            have = 0;
            blocked = 0;
            ov_push(&out, op);
        }
    }
    if (have) ov_push(&out, cur);
    free(in.v);
    return out;
}

/* compiler.rs:540-547 to_matrix */
static zc *to_matrix(const op_t *o, uint32_t *qs, int *k) {
    if (o->kind == OP_U) {
        *k = o->k;
        memcpy(qs, o->qs, sizeof(uint32_t) * (size_t)o->k);
        size_t len = (size_t)1 << (2 * o->k);
        zc *m = (zc *)malloc(sizeof(zc) * len);
        memcpy(m, o->data, sizeof(zc) * len);
        return m;
    }
    if (o->kind == OP_D) {
        *k = o->k;
        memcpy(qs, o->qs, sizeof(uint32_t) * (size_t)o->k);
        int dim = 1 << o->k;
        zc *m = (zc *)calloc((size_t)dim * dim, sizeof(zc));
        for (int j = 0; j < dim; j++) m[j * dim + j] = o->data[j];
        return m;
    }
    zc cx[16];
    int diag;
    gate_build("cx", NULL, 0, cx, &diag);
    uint32_t q[2] = {o->c, o->t};
    *k = 2;
    return permute_to_sorted(q, 2, cx, 0, qs);
}

/* compiler.rs:540-626 */
static opvec fuse_general(opvec in, int kmax) {
    opvec out = {0};
    int have = 0, count = 0;
    op_t first;
    uint32_t gs[32];
    int gk = 0;
    zc *gm = NULL;
    memset(&first, 0, sizeof first);
#define FLUSH()                                                                \
    do {                                                                       \
        if (have) {                                                            \
            if (count == 1) {                                                  \
                ov_push(&out, first);                                          \
                free(gm);                                                      \
            } else {                                                           \
                op_t f;                                                        \
                memset(&f, 0, sizeof f);                                       \
                f.kind = OP_U;                                                 \
                f.k = gk;                                                      \
                memcpy(f.qs, gs, sizeof(uint32_t) * (size_t)gk);               \
                f.data = gm;                                                   \
                ov_push(&out, f);                                              \
                if (first.data) free(first.data);                              \
            }                                                                  \
            have = 0;                                                          \
            gm = NULL;                                                         \
        }                                                                      \
    } while (0)
    for (int64_t x = 0; x < in.n; x++) {
        op_t op = in.v[x];
        uint32_t s[12];
        int fusable = gate_like(&op) && support(&op, s) <= kmax;
        if (!fusable) {
            FLUSH();
            ov_push(&out, op);
            continue;
        }
        uint32_t qs[12];
        int k;
        zc *m = to_matrix(&op, qs, &k);
        if (!have) {
            have = 1;
            count = 1;
            first = op;
            gk = k;
            memcpy(gs, qs, sizeof(uint32_t) * (size_t)k);
            gm = m;
            continue;
        }
        uint32_t u[32];
        int nu = union_sorted(gs, gk, qs, k, u);
        if (nu <= kmax) {
            zc *ce = embed_unitary(gm, gs, gk, u, nu);
            zc *oe = embed_unitary(m, qs, k, u, nu);
            zc *prod = matmul(oe, ce, 1 << nu);
            free(ce);
            free(oe);
            free(m);
            free(gm);
            if (op.data) free(op.data);
            gm = prod;
            gk = nu;
            memcpy(gs, u, sizeof(uint32_t) * (size_t)nu);
            count++;
        } else {
            FLUSH();
            have = 1;
            count = 1;
            first = op;
            gk = k;
            memcpy(gs, qs, sizeof(uint32_t) * (size_t)k);
            gm = m;
        }
    }
    FLUSH();
#undef FLUSH
    free(in.v);
    return out;
}

/* compiler.rs:118-200 */
zo_compiled *zo_compile(const zo_instr *prog, int64_t n_instr,
                        uint32_t n_qubits, const uint32_t *creg_sizes,
                        uint32_t n_cregs, int fusion_max, int diag_max) {
    uint32_t n_clbits = 0;
    for (uint32_t i = 0; i < n_cregs; i++) n_clbits += creg_sizes[i];
    if (n_qubits == 0) {
        set_err("invalid circuit: no qubits declared");
        return NULL;
    }
    if (n_qubits > 48) {
        set_err("invalid circuit: %u qubits is beyond any state-vector budget",
                n_qubits);
        return NULL;
    }
    if (n_clbits > 64) {
        set_err("invalid circuit: more than 64 classical bits are not "
                "supported");
        return NULL;
    }
    zo_compiled *cc = (zo_compiled *)calloc(1, sizeof *cc);
    cc->n_qubits = n_qubits;
    cc->n_clbits = n_clbits;
    cc->n_cregs = n_cregs;
    cc->creg_sizes = (uint32_t *)malloc(sizeof(uint32_t) * (n_cregs + 1));
    if (n_cregs) memcpy(cc->creg_sizes, creg_sizes, sizeof(uint32_t) * n_cregs);
    cc->input_gates = count_gates(prog, n_instr);
    zo_instr *cn = (zo_instr *)malloc(sizeof(zo_instr) * (size_t)(n_instr + 1));
    int64_t m = cancel(prog, n_instr, cn);
    cc->cancelled = cc->input_gates - count_gates(cn, m);
    opvec ops = {0};
    for (int64_t i = 0; i < m; i++) {
        if (lower(&cn[i], cc, &ops) < 0) {
            free(cn);
            cc->ops = ops;
            zo_compiled_free(cc);
            return NULL;
        }
    }
    free(cn);
    if (fusion_max > 6) fusion_max = 6;
    if (diag_max > 12) diag_max = 12;
    if (fusion_max >= 1) {
        ops = fuse_diagonals(ops, diag_max);
        ops = fuse_general(ops, fusion_max);
    }
    /* strip barriers */
    int64_t w = 0;
    for (int64_t i = 0; i < ops.n; i++)
        if (!(ops.v[i].kind == OP_BARRIER)) ops.v[w++] = ops.v[i];
    ops.n = w;
    /* finalize: trailing measures vs dynamic body */
    int64_t split = ops.n;
    while (split > 0 && ops.v[split - 1].kind == OP_MEAS && !ops.v[split - 1].is_if)
        split--;
    int dyn = 0;
    for (int64_t i = 0; i < split; i++) {
        const op_t *o = &ops.v[i];
        if (o->is_if || o->kind == OP_MEAS || o->kind == OP_RESET) dyn = 1;
    }
    cc->fm = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (size_t)(ops.n + 1));
    if (!dyn) {
        for (int64_t i = split; i < ops.n; i++) {
            cc->fm[2 * cc->nfm] = ops.v[i].c;
            cc->fm[2 * cc->nfm + 1] = ops.v[i].t;
            cc->nfm++;
        }
        ops.n = split;
    }
    cc->dynamic = dyn;
    cc->ops = ops;
    cc->output_ops = ops.n;
    int64_t mf = 0;
    for (int64_t i = 0; i < ops.n; i++) {
        const op_t *o = &ops.v[i];
        if (!gate_like(o)) continue;
        int s = o->kind == OP_CX ? 2 : o->k;
        if (s > mf) mf = s;
    }
    cc->max_fused = mf;
    return cc;
}

void zo_compiled_free(zo_compiled *c) {
    if (!c) return;
    for (int64_t i = 0; i < c->ops.n; i++) free(c->ops.v[i].data);
    free(c->ops.v);
    free(c->fm);
    free(c->creg_sizes);
    free(c);
}
int64_t zo_compiled_nops(const zo_compiled *c) { return c->ops.n; }
int zo_compiled_op(const zo_compiled *c, int64_t i, int32_t *kind, int32_t *k,
                   uint32_t *qs, double *data, int32_t *is_if,
                   uint64_t *cond) {
    if (i < 0 || i >= c->ops.n) return -1;
    const op_t *o = &c->ops.v[i];
    *kind = o->kind;
    *is_if = o->is_if;
    cond[0] = o->creg_off;
    cond[1] = o->creg_len;
    cond[2] = o->value;
    if (o->kind == OP_U || o->kind == OP_D) {
        *k = o->k;
        memcpy(qs, o->qs, sizeof(uint32_t) * (size_t)o->k);
        size_t len = o->kind == OP_U ? (size_t)1 << (2 * o->k) : (size_t)1 << o->k;
        if (data) memcpy(data, o->data, sizeof(zc) * len);
    } else {
        *k = 2;
        qs[0] = o->c;
        qs[1] = o->t;
    }
    return 0;
}
void zo_compiled_stats(const zo_compiled *c, int64_t *s) {
    s[0] = c->input_gates;
    s[1] = c->cancelled;
    s[2] = c->output_ops;
    s[3] = c->max_fused;
    s[4] = c->dynamic;
    s[5] = c->nfm;
}
int64_t zo_compiled_final_measures(const zo_compiled *c, uint32_t *pairs) {
    memcpy(pairs, c->fm, sizeof(uint32_t) * 2 * (size_t)c->nfm);
    return c->nfm;
}

/* ============================================================ threading */

static int g_threads = 1;
/* Set on the shot-parallel workers of zo_run_shots: each owns a state, so
 * its sweeps run on the calling thread (rayon nests them the same way). */
static __thread int tl_inner_serial = 0;
void zo_set_threads(int t) { g_threads = t < 1 ? 1 : t; }
int zo_get_threads(void) { return g_threads; }

typedef void (*task_fn)(void *ctx, uint64_t lo, uint64_t hi);
typedef struct {
    task_fn fn;
    void *ctx;
    uint64_t per, total, ntasks;
    uint64_t next; /* guarded by mu */
    pthread_mutex_t mu;
} pool_t;

static void *pool_worker(void *arg) {
    pool_t *p = (pool_t *)arg;
    for (;;) {
        pthread_mutex_lock(&p->mu);
        uint64_t t = p->next++;
        pthread_mutex_unlock(&p->mu);
        if (t >= p->ntasks) return NULL;
        uint64_t lo = t * p->per, hi = (t + 1) * p->per;
        if (hi > p->total) hi = p->total;
        if (lo < hi) p->fn(p->ctx, lo, hi);
    }
}

/* state.rs:90-105 task_ranges, run on g_threads POSIX threads.  The
 * sweeps below are elementwise, so results do not depend on the split. */
static void par_for(uint64_t total, task_fn fn, void *ctx) {
    if (total == 0) return;
    uint64_t th = tl_inner_serial ? 1 : (uint64_t)g_threads;
    uint64_t nt = total <= 8192 ? 1 : total / 8192;
    if (nt > th * 8) nt = th * 8;
    if (nt < 1) nt = 1;
    uint64_t per = (total + nt - 1) / nt;
    if (nt == 1 || th == 1) {
        for (uint64_t t = 0; t < nt; t++) {
            uint64_t lo = t * per, hi = (t + 1) * per;
            if (hi > total) hi = total;
            if (lo < hi) fn(ctx, lo, hi);
        }
        return;
    }
    pool_t p;
    p.fn = fn;
    p.ctx = ctx;
    p.per = per;
    p.total = total;
    p.ntasks = nt;
    p.next = 0;
    pthread_mutex_init(&p.mu, NULL);
    uint64_t nw = th < nt ? th : nt;
    pthread_t *ids = (pthread_t *)malloc(sizeof(pthread_t) * nw);
    for (uint64_t i = 0; i < nw; i++) pthread_create(&ids[i], NULL, pool_worker, &p);
    for (uint64_t i = 0; i < nw; i++) pthread_join(ids[i], NULL);
    free(ids);
    pthread_mutex_destroy(&p.mu);
}

/* ============================================================== kernels */

struct zo_state {
    uint32_t n;
    int prec;
    uint64_t len;
    void *re, *im;
};

/* state.rs:159-169 */
static inline uint64_t insert_zeros(uint64_t g, const uint32_t *qs, int k) {
    uint64_t x = g;
    for (int b = 0; b < k; b++) {
        uint32_t q = qs[b];
        uint64_t low = x & ((1ull << q) - 1);
        x = ((x >> q) << (q + 1)) | low;
    }
    return x;
}

typedef struct {
    void *re, *im;
    uint32_t q;
    int k;
    const uint32_t *qs;
    const uint64_t *offs;
    const void *mre, *mim;
    double d64[8];
    float d32[8];
    uint64_t coff, toff;
    int outcome;
    double scale64;
    float scale32;
} kctx;

/* Kernel bodies instantiated for T = double and T = float. */
#define DEFINE_KERNELS(T, SFX)                                                 \
    /* state.rs:232-260 kern_1q_scalar */                                      \
    static void k1q_##SFX(void *vc, uint64_t h0, uint64_t h1) {                \
        kctx *c = (kctx *)vc;                                                  \
        T *re = (T *)c->re, *im = (T *)c->im;                                  \
        const T *mr = (const T *)c->mre, *mi = (const T *)c->mim;              \
        T m00r = mr[0], m00i = mi[0], m01r = mr[1], m01i = mi[1];              \
        T m10r = mr[2], m10i = mi[2], m11r = mr[3], m11i = mi[3];              \
        uint32_t q = c->q;                                                     \
        uint64_t s = 1ull << q, mask = s - 1;                                  \
        for (uint64_t h = h0; h < h1; h++) {                                   \
            uint64_t a = ((h >> q) << (q + 1)) | (h & mask), b = a + s;        \
            T ar = re[a], ai = im[a], br = re[b], bi = im[b];                  \
            re[a] = m00r * ar - m00i * ai + m01r * br - m01i * bi;             \
            im[a] = m00r * ai + m00i * ar + m01r * bi + m01i * br;             \
            re[b] = m10r * ar - m10i * ai + m11r * br - m11i * bi;             \
            im[b] = m10r * ai + m10i * ar + m11r * bi + m11i * br;             \
        }                                                                      \
    }                                                                          \
    /* state.rs:355-392 kern_kq_scalar */                                      \
    static void kkq_##SFX(void *vc, uint64_t g0, uint64_t g1) {                \
        kctx *c = (kctx *)vc;                                                  \
        T *re = (T *)c->re, *im = (T *)c->im;                                  \
        const T *mr = (const T *)c->mre, *mi = (const T *)c->mim;              \
        int dim = 1 << c->k;                                                   \
        T xr[64], xi[64], yr[64], yi[64];                                      \
        for (uint64_t g = g0; g < g1; g++) {                                   \
            uint64_t base = insert_zeros(g, c->qs, c->k);                      \
            for (int j = 0; j < dim; j++) {                                    \
                uint64_t p = base | c->offs[j];                                \
                xr[j] = re[p];                                                 \
                xi[j] = im[p];                                                 \
            }                                                                  \
            for (int r = 0; r < dim; r++) {                                    \
                T ar = 0, ai = 0;                                              \
                for (int cc = 0; cc < dim; cc++) {                             \
                    T wr = mr[r * dim + cc], wi = mi[r * dim + cc];            \
                    ar = ar + wr * xr[cc] - wi * xi[cc];                       \
                    ai = ai + wr * xi[cc] + wi * xr[cc];                       \
                }                                                              \
                yr[r] = ar;                                                    \
                yi[r] = ai;                                                    \
            }                                                                  \
            for (int j = 0; j < dim; j++) {                                    \
                uint64_t p = base | c->offs[j];                                \
                re[p] = yr[j];                                                 \
                im[p] = yi[j];                                                 \
            }                                                                  \
        }                                                                      \
    }                                                                          \
    /* state.rs:394-417 apply_cx: swap inside the control=1 half */            \
    static void kcx_##SFX(void *vc, uint64_t g0, uint64_t g1) {                \
        kctx *c = (kctx *)vc;                                                  \
        T *re = (T *)c->re, *im = (T *)c->im;                                  \
        for (uint64_t g = g0; g < g1; g++) {                                   \
            uint64_t a = insert_zeros(g, c->qs, 2) | c->coff, b = a | c->toff; \
            T t = re[a];                                                       \
            re[a] = re[b];                                                     \
            re[b] = t;                                                         \
            t = im[a];                                                         \
            im[a] = im[b];                                                     \
            im[b] = t;                                                         \
        }                                                                      \
    }                                                                          \
    /* state.rs:498-527 kern_diag1_scalar */                                   \
    static void kd1_##SFX(void *vc, uint64_t h0, uint64_t h1) {                \
        kctx *c = (kctx *)vc;                                                  \
        T *re = (T *)c->re, *im = (T *)c->im;                                  \
        const T *d = (const T *)(sizeof(T) == 8 ? (const void *)c->d64         \
                                                : (const void *)c->d32);       \
        uint32_t q = c->q;                                                     \
        uint64_t s = 1ull << q, mask = s - 1;                                  \
        for (uint64_t h = h0; h < h1; h++) {                                   \
            uint64_t a = ((h >> q) << (q + 1)) | (h & mask), b = a + s;        \
            T ar = re[a], ai = im[a];                                          \
            re[a] = d[0] * ar - d[1] * ai;                                     \
            im[a] = d[0] * ai + d[1] * ar;                                     \
            T br = re[b], bi = im[b];                                          \
            re[b] = d[2] * br - d[3] * bi;                                     \
            im[b] = d[2] * bi + d[3] * br;                                     \
        }                                                                      \
    }                                                                          \
    /* state.rs:577-596 kern_diagk_scalar */                                   \
    static void kdk_##SFX(void *vc, uint64_t i0, uint64_t i1) {                \
        kctx *c = (kctx *)vc;                                                  \
        T *re = (T *)c->re, *im = (T *)c->im;                                  \
        const T *dr = (const T *)c->mre, *di = (const T *)c->mim;              \
        for (uint64_t i = i0; i < i1; i++) {                                   \
            uint64_t j = 0;                                                    \
            for (int b = 0; b < c->k; b++) j |= ((i >> c->qs[b]) & 1) << b;    \
            T wr = dr[j], wi = di[j], ar = re[i], ai = im[i];                  \
            re[i] = wr * ar - wi * ai;                                         \
            im[i] = wr * ai + wi * ar;                                         \
        }                                                                      \
    }                                                                          \
    /* state.rs:628-661 collapse */                                            \
    static void kcol_##SFX(void *vc, uint64_t h0, uint64_t h1) {               \
        kctx *c = (kctx *)vc;                                                  \
        T *re = (T *)c->re, *im = (T *)c->im;                                  \
        T scale = sizeof(T) == 8 ? (T)c->scale64 : (T)c->scale32;              \
        uint32_t q = c->q;                                                     \
        uint64_t s = 1ull << q, mask = s - 1;                                  \
        for (uint64_t h = h0; h < h1; h++) {                                   \
            uint64_t a = ((h >> q) << (q + 1)) | (h & mask);                   \
            uint64_t keep = c->outcome ? a + s : a;                            \
            uint64_t zero = c->outcome ? a : a + s;                            \
            re[keep] = re[keep] * scale;                                       \
            im[keep] = im[keep] * scale;                                       \
            re[zero] = 0;                                                      \
            im[zero] = 0;                                                      \
        }                                                                      \
    }

DEFINE_KERNELS(double, f64)
DEFINE_KERNELS(float, f32)

zo_state *zo_state_new(uint32_t n, int prec) {
    if (n > 40 || (prec != 32 && prec != 64)) {
        set_err("unsupported state: n=%u prec=%d", n, prec);
        return NULL;
    }
    zo_state *s = (zo_state *)calloc(1, sizeof *s);
    s->n = n;
    s->prec = prec;
    s->len = 1ull << n;
    size_t bytes = (size_t)s->len * (prec == 64 ? 8 : 4);
    s->re = malloc(bytes);
    s->im = malloc(bytes);
    if (!s->re || !s->im) {
        free(s->re);
        free(s->im);
        free(s);
        set_err("out of host memory");
        return NULL;
    }
    zo_reset_all(s);
    return s;
}

void zo_state_free(zo_state *s) {
    if (!s) return;
    free(s->re);
    free(s->im);
    free(s);
}

typedef struct {
    zo_state *s;
} zctx;
static void kzero(void *vc, uint64_t lo, uint64_t hi) {
    zo_state *s = ((zctx *)vc)->s;
    size_t w = s->prec == 64 ? 8 : 4;
    memset((char *)s->re + lo * w, 0, (hi - lo) * w);
    memset((char *)s->im + lo * w, 0, (hi - lo) * w);
}

/* state.rs:109-117 zero_state / :128-132 reset_zero */
void zo_reset_all(zo_state *s) {
    zctx c = {s};
    par_for(s->len, kzero, &c);
    if (s->prec == 64)
        ((double *)s->re)[0] = 1.0;
    else
        ((float *)s->re)[0] = 1.0f;
}

static void ctx_init(kctx *c, zo_state *s) {
    memset(c, 0, sizeof *c);
    c->re = s->re;
    c->im = s->im;
}

/* state.rs:171-188 apply_1q and :275-314 apply_kq: the f64 matrix is
 * converted element-wise to T (T::fr = `as f32`, round to nearest). */
void zo_apply_unitary(zo_state *s, const uint32_t *qs, int k,
                      const double *mat) {
    int dim = 1 << k;
    size_t len = (size_t)dim * dim;
    kctx c;
    ctx_init(&c, s);
    void *mr = malloc(len * 8), *mi = malloc(len * 8);
    for (size_t i = 0; i < len; i++) {
        if (s->prec == 64) {
            ((double *)mr)[i] = mat[2 * i];
            ((double *)mi)[i] = mat[2 * i + 1];
        } else {
            ((float *)mr)[i] = (float)mat[2 * i];
            ((float *)mi)[i] = (float)mat[2 * i + 1];
        }
    }
    c.mre = mr;
    c.mim = mi;
    if (k == 1) {
        c.q = qs[0];
        par_for(s->len >> 1, s->prec == 64 ? k1q_f64 : k1q_f32, &c);
    } else {
        uint64_t offs[64];
        for (int j = 0; j < dim; j++) {
            uint64_t o = 0;
            for (int b = 0; b < k; b++)
                if ((j >> b) & 1) o |= 1ull << qs[b];
            offs[j] = o;
        }
        c.k = k;
        c.qs = qs;
        c.offs = offs;
        par_for(s->len >> k, s->prec == 64 ? kkq_f64 : kkq_f32, &c);
    }
    free(mr);
    free(mi);
}

/* state.rs:419-448 apply_diag */
void zo_apply_diagonal(zo_state *s, const uint32_t *qs, int k,
                       const double *d) {
    kctx c;
    ctx_init(&c, s);
    if (k == 1) {
        for (int i = 0; i < 4; i++) {
            c.d64[i] = d[i];
            c.d32[i] = (float)d[i];
        }
        c.q = qs[0];
        par_for(s->len >> 1, s->prec == 64 ? kd1_f64 : kd1_f32, &c);
        return;
    }
    int dim = 1 << k;
    void *dr = malloc((size_t)dim * 8), *di = malloc((size_t)dim * 8);
    for (int i = 0; i < dim; i++) {
        if (s->prec == 64) {
            ((double *)dr)[i] = d[2 * i];
            ((double *)di)[i] = d[2 * i + 1];
        } else {
            ((float *)dr)[i] = (float)d[2 * i];
            ((float *)di)[i] = (float)d[2 * i + 1];
        }
    }
    c.mre = dr;
    c.mim = di;
    c.k = k;
    c.qs = qs;
    par_for(s->len, s->prec == 64 ? kdk_f64 : kdk_f32, &c);
    free(dr);
    free(di);
}

/* state.rs:394-417 */
void zo_apply_cx(zo_state *s, uint32_t control, uint32_t target) {
    uint32_t qs[2] = {control < target ? control : target,
                      control < target ? target : control};
    kctx c;
    ctx_init(&c, s);
    c.qs = qs;
    c.coff = 1ull << control;
    c.toff = 1ull << target;
    par_for(s->len >> 2, s->prec == 64 ? kcx_f64 : kcx_f32, &c);
}

static inline double amp2(const zo_state *s, uint64_t i) {
    double r, m;
    if (s->prec == 64) {
        r = ((const double *)s->re)[i];
        m = ((const double *)s->im)[i];
    } else {
        r = (double)((const float *)s->re)[i];
        m = (double)((const float *)s->im)[i];
    }
    return r * r + m * m;
}

/* Fixed reduction order shared with the HIP backend.  Elements e[0..N)
 * are cut into blocks of 4096; inside a block, lane t (0..255) adds
 * e[blk*4096 + j*256 + t] for j = 0..15 serially from 0.0, then lanes
 * combine pairwise (stride 128, 64, .., 1: s[t] = s[t] + s[t+stride]).  The
 * block sums are reduced the same way until one value is left.  The
 * reference sums per rayon task and then in rayon's tree order
 * (state.rs:143-156, :598-626), which depends on the schedule. */
typedef double (*elem_fn)(const void *ctx, uint64_t i);
static double reduce_blocks(const void *ctx, elem_fn e, uint64_t N) {
    uint64_t nb = (N + 4095) / 4096;
    double *bs = (double *)malloc(sizeof(double) * (size_t)(nb ? nb : 1));
    for (uint64_t b = 0; b < nb; b++) {
        double s[256];
        for (int t = 0; t < 256; t++) {
            double a = 0.0;
            for (int j = 0; j < 16; j++) {
                uint64_t idx = b * 4096 + (uint64_t)j * 256 + (uint64_t)t;
                if (idx < N) a = a + e(ctx, idx);
            }
            s[t] = a;
        }
        for (int st = 128; st >= 1; st >>= 1)
            for (int t = 0; t < st; t++) s[t] = s[t] + s[t + st];
        bs[b] = s[0];
    }
    while (nb > 1) {
        uint64_t nn = (nb + 4095) / 4096;
        double *ns = (double *)malloc(sizeof(double) * (size_t)nn);
        for (uint64_t b = 0; b < nn; b++) {
            double s[256];
            for (int t = 0; t < 256; t++) {
                double a = 0.0;
                for (int j = 0; j < 16; j++) {
                    uint64_t idx = b * 4096 + (uint64_t)j * 256 + (uint64_t)t;
                    if (idx < nb) a = a + bs[idx];
                }
                s[t] = a;
            }
            for (int st = 128; st >= 1; st >>= 1)
                for (int t = 0; t < st; t++) s[t] = s[t] + s[t + st];
            ns[b] = s[0];
        }
        free(bs);
        bs = ns;
        nb = nn;
    }
    double r = bs[0];
    free(bs);
    return r;
}

typedef struct {
    const zo_state *s;
    uint32_t q;
} pctx;
static double e_prob(const void *vc, uint64_t h) {
    const pctx *c = (const pctx *)vc;
    uint64_t s = 1ull << c->q, mask = s - 1;
    uint64_t a = ((h >> c->q) << (c->q + 1)) | (h & mask);
    return amp2(c->s, a + s);
}
static double e_norm(const void *vc, uint64_t i) {
    return amp2(((const pctx *)vc)->s, i);
}

/* state.rs:598-626 prob_one (fixed order, see reduce_blocks) */
double zo_prob_one(zo_state *s, uint32_t q) {
    pctx c = {s, q};
    return reduce_blocks(&c, e_prob, s->len >> 1);
}

/* state.rs:142-156 norm_sqr (fixed order) */
double zo_norm_sqr(zo_state *s) {
    pctx c = {s, 0};
    return reduce_blocks(&c, e_norm, s->len);
}

/* state.rs:628-661 collapse; exec.rs:201-207 CpuBackend::measure */
int zo_measure(zo_state *s, uint32_t q, double u) {
    double p1 = zo_prob_one(s, q);
    int outcome = u < p1;
    double p = outcome ? p1 : 1.0 - p1;
    kctx c;
    ctx_init(&c, s);
    c.q = q;
    c.outcome = outcome;
    c.scale64 = 1.0 / sqrt(p);
    c.scale32 = (float)c.scale64;
    par_for(s->len >> 1, s->prec == 64 ? kcol_f64 : kcol_f32, &c);
    return outcome;
}

/* state.rs:134-140 to_c64 */
void zo_statevector(const zo_state *s, double *out) {
    for (uint64_t i = 0; i < s->len; i++) {
        if (s->prec == 64) {
            out[2 * i] = ((const double *)s->re)[i];
            out[2 * i + 1] = ((const double *)s->im)[i];
        } else {
            out[2 * i] = (double)((const float *)s->re)[i];
            out[2 * i + 1] = (double)((const float *)s->im)[i];
        }
    }
}
void zo_state_raw(const zo_state *s, void *re, void *im) {
    size_t w = s->prec == 64 ? 8 : 4;
    memcpy(re, s->re, (size_t)s->len * w);
    memcpy(im, s->im, (size_t)s->len * w);
}
void zo_state_set_raw(zo_state *s, const void *re, const void *im) {
    size_t w = s->prec == 64 ? 8 : 4;
    memcpy(s->re, re, (size_t)s->len * w);
    memcpy(s->im, im, (size_t)s->len * w);
}

/* exec.rs:398-405 apply_op, over ops [lo, hi) */
void zo_apply_compiled(zo_state *s, const zo_compiled *c, int64_t lo,
                       int64_t hi) {
    if (hi > c->ops.n) hi = c->ops.n;
    for (int64_t i = lo; i < hi; i++) {
        const op_t *o = &c->ops.v[i];
        if (o->kind == OP_U)
            zo_apply_unitary(s, o->qs, o->k, (const double *)o->data);
        else if (o->kind == OP_D)
            zo_apply_diagonal(s, o->qs, o->k, (const double *)o->data);
        else if (o->kind == OP_CX)
            zo_apply_cx(s, o->c, o->t);
    }
}

/* ============================================================== sampler */

#define CHUNK (1ull << 16) /* sample.rs:13 */

typedef struct {
    const zo_state *s;
    double *sums;
    uint64_t chunk;
} csctx;
static void kchunk(void *vc, uint64_t c0, uint64_t c1) {
    csctx *c = (csctx *)vc;
    for (uint64_t cc = c0; cc < c1; cc++) {
        uint64_t lo = cc * c->chunk, hi = lo + c->chunk;
        if (hi > c->s->len) hi = c->s->len;
        double acc = 0.0;
        for (uint64_t i = lo; i < hi; i++) acc += amp2(c->s, i);
        c->sums[cc] = acc;
    }
}

void zo_chunk_sums(const zo_state *s, double *out) {
    uint64_t chunk = s->len < CHUNK ? s->len : CHUNK;
    uint64_t nch = (s->len + chunk - 1) / chunk;
    csctx c = {s, out, chunk};
    /* each chunk is one task: the serial per-chunk order never changes */
    for (uint64_t i = 0; i < nch; i++) kchunk(&c, i, i + 1);
}

static int cmp_f64(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* sample.rs:29-111 sample_indices, with rng = seed_from_u64(seed)
 * (exec.rs:212-215). */
void zo_sample(const zo_state *s, uint64_t shots, uint64_t seed,
               uint64_t *out) {
    if (shots == 0) return;
    uint64_t n = s->len;
    uint64_t chunk = n < CHUNK ? n : CHUNK;
    uint64_t nch = (n + chunk - 1) / chunk;
    double *sums = (double *)malloc(sizeof(double) * nch);
    csctx cs = {s, sums, chunk};
    par_for(nch, kchunk, &cs);
    double *cum = (double *)malloc(sizeof(double) * (nch + 1));
    cum[0] = 0.0;
    for (uint64_t c = 0; c < nch; c++) cum[c + 1] = cum[c] + sums[c];
    double total = cum[nch];
    rng_t r;
    rng_seed(&r, seed);
    double *us = (double *)malloc(sizeof(double) * shots);
    for (uint64_t i = 0; i < shots; i++) us[i] = rng_f64(&r) * total;
    qsort(us, shots, sizeof(double), cmp_f64);
    uint64_t *starts = (uint64_t *)malloc(sizeof(uint64_t) * (nch + 1));
    for (uint64_t c = 0; c <= nch; c++) { /* partition_point(u < cum[c]) */
        uint64_t lo = 0, hi = shots;
        while (lo < hi) {
            uint64_t mid = lo + (hi - lo) / 2;
            if (us[mid] < cum[c])
                lo = mid + 1;
            else
                hi = mid;
        }
        starts[c] = lo;
    }
    for (uint64_t c = 0; c < nch; c++) {
        uint64_t s0 = starts[c], s1 = starts[c + 1];
        if (s0 >= s1) continue;
        uint64_t lo = c * chunk, hi = lo + chunk;
        if (hi > n) hi = n;
        double acc = cum[c];
        uint64_t ptr = s0;
        for (uint64_t i = lo; i < hi; i++) {
            acc += amp2(s, i);
            while (ptr < s1 && us[ptr] < acc) out[ptr++] = i;
            if (ptr == s1) break;
        }
        if (ptr < s1) {
            uint64_t last = hi - 1;
            for (uint64_t i = hi; i-- > lo;)
                if (amp2(s, i) > 0.0) {
                    last = i;
                    break;
                }
            while (ptr < s1) out[ptr++] = last;
        }
    }
    free(sums);
    free(cum);
    free(us);
    free(starts);
}

/* ============================================================= executor */

typedef struct {
    uint64_t key, n;
} kv;
static int cmp_kv(const void *a, const void *b) {
    uint64_t x = ((const kv *)a)->key, y = ((const kv *)b)->key;
    return (x > y) - (x < y);
}

/* noise.rs:91-100 */
static void apply_pauli(zo_state *s, uint32_t q, uint64_t code) {
    double x[8] = {0, 0, 1, 0, 1, 0, 0, 0};
    double y[8] = {0, 0, 0, -1, 0, 1, 0, 0};
    double z[4] = {1, 0, -1, 0};
    if (code == 1) zo_apply_unitary(s, &q, 1, x);
    if (code == 2) zo_apply_unitary(s, &q, 1, y);
    if (code == 3) zo_apply_diagonal(s, &q, 1, z);
}

/* noise.rs:114-127 */
static void amplitude_damp(zo_state *s, uint32_t q, double gamma, double u) {
    double pj = gamma * zo_prob_one(s, q);
    if (u < pj) {
        zo_measure(s, q, 0.0);
        apply_pauli(s, q, 1);
    } else {
        double sc = 1.0 / sqrt(1.0 - pj);
        double d[4] = {sc, 0.0, sc * sqrt(1.0 - gamma), 0.0};
        zo_apply_diagonal(s, &q, 1, d);
    }
}

/* noise.rs:145-175 */
static void gate_noise(zo_state *s, const zo_noise *m, const uint32_t *qs,
                       int k, rng_t *r) {
    double pd = k == 1 ? m->depolarizing_1q : m->depolarizing_2q;
    if (pd > 0.0 && rng_f64(r) < pd) {
        uint64_t idx = rng_range(r, 1, 1ull << (2 * k));
        for (int b = 0; b < k; b++) apply_pauli(s, qs[b], (idx >> (2 * b)) & 3);
    }
    for (int b = 0; b < k; b++) {
        uint32_t q = qs[b];
        if (m->bit_flip > 0.0 && rng_f64(r) < m->bit_flip) apply_pauli(s, q, 1);
        if (m->phase_flip > 0.0 && rng_f64(r) < m->phase_flip)
            apply_pauli(s, q, 3);
        if (m->amplitude_damping > 0.0) {
            double u = rng_f64(r);
            amplitude_damp(s, q, m->amplitude_damping, u);
        }
    }
}

/* noise.rs:181-192 */
static int readout(const zo_noise *m, int bit, rng_t *r) {
    double p = bit ? m->readout_flip_1to0 : m->readout_flip_0to1;
    if (p > 0.0 && rng_f64(r) < p) return !bit;
    return bit;
}

/* exec.rs:556-619 */
static void exec_shot_op(zo_state *s, const op_t *o, const zo_noise *nz,
                         rng_t *r, uint64_t *clbits) {
    if (o->is_if) {
        uint64_t mask = o->creg_len >= 64 ? ~0ull : ((1ull << o->creg_len) - 1);
        if (((*clbits >> o->creg_off) & mask) != o->value) return;
    }
    if (o->kind == OP_U) {
        zo_apply_unitary(s, o->qs, o->k, (const double *)o->data);
        if (nz) gate_noise(s, nz, o->qs, o->k, r);
    } else if (o->kind == OP_D) {
        zo_apply_diagonal(s, o->qs, o->k, (const double *)o->data);
        if (nz) gate_noise(s, nz, o->qs, o->k, r);
    } else if (o->kind == OP_CX) {
        zo_apply_cx(s, o->c, o->t);
        if (nz) {
            uint32_t qs[2] = {o->c < o->t ? o->c : o->t, o->c < o->t ? o->t : o->c};
            gate_noise(s, nz, qs, 2, r);
        }
    } else if (o->kind == OP_MEAS) {
        int bit = zo_measure(s, o->c, rng_f64(r));
        if (nz) bit = readout(nz, bit, r);
        *clbits = (*clbits & ~(1ull << o->t)) | ((uint64_t)bit << o->t);
    } else if (o->kind == OP_RESET) {
        if (zo_measure(s, o->c, rng_f64(r))) {
            double x[8] = {0, 0, 1, 0, 1, 0, 0, 0};
            uint32_t q = o->c;
            zo_apply_unitary(s, &q, 1, x);
        }
    }
}

static int noise_trivial(const zo_noise *m) {
    return !m || (m->depolarizing_1q == 0.0 && m->depolarizing_2q == 0.0 &&
                  m->bit_flip == 0.0 && m->phase_flip == 0.0 &&
                  m->amplitude_damping == 0.0 && m->readout_flip_0to1 == 0.0 &&
                  m->readout_flip_1to0 == 0.0);
}

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* exec.rs:485-500 run_shot: one trajectory from |0...0> with its own
 * Xoshiro256++ stream seeded with splitmix64(seed ^ shot). */
static uint64_t run_one_shot(zo_state *s, const zo_compiled *c, const zo_noise *nz,
                             uint64_t seed, uint64_t shot) {
    zo_reset_all(s);
    rng_t r;
    rng_seed(&r, zo_splitmix64(seed ^ shot));
    uint64_t cl = 0;
    for (int64_t i = 0; i < c->ops.n; i++) exec_shot_op(s, &c->ops.v[i], nz, &r, &cl);
    for (int64_t f = 0; f < c->nfm; f++) {
        int bit = zo_measure(s, c->fm[2 * f], rng_f64(&r));
        if (nz) bit = readout(nz, bit, &r);
        uint32_t cb = c->fm[2 * f + 1];
        cl = (cl & ~(1ull << cb)) | ((uint64_t)bit << cb);
    }
    return cl;
}

typedef struct {
    const zo_compiled *c;
    const zo_noise *nz;
    int prec;
    uint64_t shots, seed, next; /* next: guarded by mu */
    uint64_t *words;
    int failed;
    pthread_mutex_t mu;
} shot_pool_t;

static void *shot_worker(void *arg) {
    shot_pool_t *p = (shot_pool_t *)arg;
    tl_inner_serial = 1;
    zo_state *s = zo_state_new(p->c->n_qubits, p->prec);
    if (!s) {
        pthread_mutex_lock(&p->mu);
        p->failed = 1;
        pthread_mutex_unlock(&p->mu);
        return NULL;
    }
    for (;;) {
        pthread_mutex_lock(&p->mu);
        uint64_t shot = p->next++;
        pthread_mutex_unlock(&p->mu);
        if (shot >= p->shots) break;
        p->words[shot] = run_one_shot(s, p->c, p->nz, p->seed, shot);
    }
    zo_state_free(s);
    return NULL;
}

/* exec.rs:459-528 run_per_shot with its `parallel` fold: shots are dealt
 * to `threads` workers, each with its own state (one backend per rayon
 * worker); words[shot] gets shot's clbit word.  threads <= 1 runs the
 * shots in order on one state.  Returns 0, or -1 when a state could not be
 * allocated. */
int zo_run_shots(const zo_compiled *c, int prec, uint64_t shots, uint64_t seed,
                 const zo_noise *noise, int threads, uint64_t *words) {
    const zo_noise *nz = noise_trivial(noise) ? NULL : noise;
    if (shots == 0) return 0;
    if (threads <= 1 || shots == 1) {
        zo_state *s = zo_state_new(c->n_qubits, prec);
        if (!s) return -1;
        for (uint64_t shot = 0; shot < shots; shot++) words[shot] = run_one_shot(s, c, nz, seed, shot);
        zo_state_free(s);
        return 0;
    }
    shot_pool_t p;
    p.c = c;
    p.nz = nz;
    p.prec = prec;
    p.shots = shots;
    p.seed = seed;
    p.next = 0;
    p.words = words;
    p.failed = 0;
    pthread_mutex_init(&p.mu, NULL);
    uint64_t nw = (uint64_t)threads < shots ? (uint64_t)threads : shots;
    pthread_t *ids = (pthread_t *)malloc(sizeof(pthread_t) * nw);
    for (uint64_t i = 0; i < nw; i++) pthread_create(&ids[i], NULL, shot_worker, &p);
    for (uint64_t i = 0; i < nw; i++) pthread_join(ids[i], NULL);
    free(ids);
    pthread_mutex_destroy(&p.mu);
    if (p.failed) {
        set_err("state allocation failed");
        return -1;
    }
    return 0;
}

/* exec.rs:347-549: run_sampled for static circuits, run_per_shot for
 * dynamic or noisy ones (shot-parallel over the oracle's thread count, as
 * the reference folds shots over rayon's threads).  Output: unique clbit
 * words with their counts, ascending by word. */
int64_t zo_run(const zo_compiled *c, int prec, uint64_t shots, uint64_t seed,
               const zo_noise *noise, uint64_t *keys, uint64_t *counts,
               int64_t cap, double *sim_seconds) {
    const zo_noise *nz = noise_trivial(noise) ? NULL : noise;
    double t0 = now_s();
    kv *raw = NULL;
    uint64_t nraw = 0;
    if (!nz && !c->dynamic) {
        zo_state *s = zo_state_new(c->n_qubits, prec);
        if (!s) return -1;
        zo_apply_compiled(s, c, 0, c->ops.n);
        if (c->nfm > 0 && shots > 0) {
            uint64_t *smp = (uint64_t *)malloc(sizeof(uint64_t) * shots);
            zo_sample(s, shots, zo_splitmix64(seed), smp);
            raw = (kv *)malloc(sizeof(kv) * shots);
            for (uint64_t i = 0; i < shots; i++) {
                uint64_t basis = smp[i], cl = 0;
                for (int64_t f = 0; f < c->nfm; f++) {
                    uint64_t bit = (basis >> c->fm[2 * f]) & 1;
                    uint32_t cb = c->fm[2 * f + 1];
                    cl = (cl & ~(1ull << cb)) | (bit << cb);
                }
                raw[nraw].key = cl;
                raw[nraw++].n = 1;
            }
            free(smp);
        }
        zo_state_free(s);
    } else if (c->dynamic || c->nfm > 0) {
        uint64_t *words = (uint64_t *)malloc(sizeof(uint64_t) * (shots ? shots : 1));
        if (zo_run_shots(c, prec, shots, seed, nz, g_threads, words) != 0) {
            free(words);
            return -1;
        }
        raw = (kv *)malloc(sizeof(kv) * (shots ? shots : 1));
        for (uint64_t shot = 0; shot < shots; shot++) {
            raw[nraw].key = words[shot];
            raw[nraw++].n = 1;
        }
        free(words);
    }
    if (sim_seconds) *sim_seconds = now_s() - t0;
    qsort(raw, nraw, sizeof(kv), cmp_kv);
    int64_t m = 0;
    for (uint64_t i = 0; i < nraw; i++) {
        if (m > 0 && keys[m - 1] == raw[i].key) {
            counts[m - 1] += raw[i].n;
            continue;
        }
        if (m >= cap) {
            free(raw);
            set_err("key buffer too small");
            return -1;
        }
        keys[m] = raw[i].key;
        counts[m] = raw[i].n;
        m++;
    }
    free(raw);
    return m;
}

/* ============================================================== RefSim */

/* tests/reference.rs:69-116: interleaved amplitudes, argument-order
 * qubits, explicit per-base gather / dense matvec / scatter. */
void zo_refsim_apply(double *amps, uint32_t n, const uint32_t *qs, int k,
                     const double *mat) {
    int dim = 1 << k;
    uint64_t mask = 0, off[64];
    for (int b = 0; b < k; b++) mask += 1ull << qs[b];
    for (int m = 0; m < dim; m++) {
        uint64_t o = 0;
        for (int b = 0; b < k; b++)
            if ((m >> b) & 1) o += 1ull << qs[b];
        off[m] = o;
    }
    const zc *M = (const zc *)mat;
    zc *A = (zc *)amps;
    zc old[64];
    for (uint64_t base = 0; base < (1ull << n); base++) {
        if (base & mask) continue;
        for (int m = 0; m < dim; m++) old[m] = A[base + off[m]];
        for (int l = 0; l < dim; l++) {
            zc acc = cmk(0.0, 0.0);
            for (int m = 0; m < dim; m++) {
                zc p = cmul(M[l * dim + m], old[m]);
                acc.re = acc.re + p.re;
                acc.im = acc.im + p.im;
            }
            A[base + off[l]] = acc;
        }
    }
}

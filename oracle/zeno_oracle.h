/*
 * zeno_oracle.h — CPU restatement of cleoanka/zeno's state-vector hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * backend: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it; nothing in the zeno_amd package may.  It restates, scalar and in
 * plain C, the reference's
 *   - RNG use        (exec.rs:227-232 splitmix64; rand_xoshiro 0.6.0
 *                     Xoshiro256PlusPlus::seed_from_u64; rand 0.8
 *                     Standard f64 and UniformInt<u64>::sample_single)
 *   - gate library   (gates.rs:29-188)
 *   - generators     (circuit.rs:138-174, plus the matching/RZZ generator
 *                     the benchmark configurations ask for)
 *   - compiler       (compiler.rs:118-626)
 *   - scalar kernels (state.rs:90-661)
 *   - sampler        (sample.rs:29-111)
 *   - executor       (exec.rs:164-253, :398-619; noise.rs:91-192)
 *   - RefSim         (tests/reference.rs:69-116), the independent dense
 *                     simulator.
 *
 * Arithmetic follows the reference expression trees exactly; build with
 * -ffp-contract=off and no fast-math.  Reduction order of prob_one and
 * norm_sqr is fixed here (the reference's rayon order is schedule
 * dependent): see zo_fixed_reduce in zeno_oracle.c.
 */
#ifndef ZENO_ORACLE_H
#define ZENO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Program instruction record (one ir::Instr, ir.rs:36-52).  kind:
 * 0 gate, 1 measure, 2 reset, 3 barrier.  cond_creg >= 0 wraps the
 * instruction in `if (creg == cond_value)`. 72 bytes. */
typedef struct {
    char name[8];
    int32_t kind;
    int32_t nq;
    int32_t np;
    uint32_t q[3];
    uint32_t clbit;
    int32_t cond_creg;
    uint64_t cond_value;
    double p[3];
} zo_instr;

typedef struct {
    double depolarizing_1q, depolarizing_2q, bit_flip, phase_flip,
        amplitude_damping, readout_flip_0to1, readout_flip_1to0;
} zo_noise;

/* ---- RNG ---------------------------------------------------------- */
uint64_t zo_splitmix64(uint64_t x);
void zo_rng_u64_stream(uint64_t seed, uint64_t n, uint64_t *out);
void zo_rng_f64_stream(uint64_t seed, uint64_t n, double *out);
void zo_xoshiro_from_state(const uint64_t *st, uint64_t n, uint64_t *out);
void zo_rng_range_stream(uint64_t seed, uint64_t lo, uint64_t hi, uint64_t n,
                         uint64_t *out);
void zo_rng_range_from_state(const uint64_t *st, uint64_t lo, uint64_t hi, uint64_t n,
                             uint64_t *out);

/* ---- gates -------------------------------------------------------- */
/* Writes the matrix (interleaved re,im; 2*4^k doubles for a unitary,
 * 2*2^k for a diagonal).  Returns arity k, or -1 for unknown name / bad
 * param count.  *is_diag = 1 for a diagonal gate. */
int zo_gate_build(const char *name, const double *params, int np,
                  double *out, int *is_diag);
int zo_gate_lookup(const char *name, int *arity, int *n_params);

/* ---- generators --------------------------------------------------- */
/* Each returns the number of instructions written (<= cap). */
int64_t zo_gen_random(uint32_t n, uint32_t depth, uint64_t seed,
                      zo_instr *out, int64_t cap);
int64_t zo_gen_qft(uint32_t n, zo_instr *out, int64_t cap);
int64_t zo_gen_u3_qft(uint32_t n, uint64_t seed, zo_instr *out, int64_t cap);
int64_t zo_gen_matching(uint32_t n, uint32_t depth, uint64_t seed,
                        uint32_t rzz_period, zo_instr *out, int64_t cap);

/* ---- compiler ----------------------------------------------------- */
typedef struct zo_compiled zo_compiled;
/* Returns NULL on error; zo_error() holds the message. */
zo_compiled *zo_compile(const zo_instr *prog, int64_t n_instr,
                        uint32_t n_qubits, const uint32_t *creg_sizes,
                        uint32_t n_cregs, int fusion_max, int diag_max);
void zo_compiled_free(zo_compiled *c);
int64_t zo_compiled_nops(const zo_compiled *c);
/* Op kinds: 0 unitary, 1 diagonal, 2 cx, 3 measure, 4 reset. */
int zo_compiled_op(const zo_compiled *c, int64_t i, int32_t *kind,
                   int32_t *k, uint32_t *qs, double *data, int32_t *is_if,
                   uint64_t *cond /* offset,len,value */);
void zo_compiled_stats(const zo_compiled *c, int64_t *stats /* 6 */);
int64_t zo_compiled_final_measures(const zo_compiled *c, uint32_t *pairs);
const char *zo_error(void);

/* ---- state + kernels ---------------------------------------------- */
typedef struct zo_state zo_state;
zo_state *zo_state_new(uint32_t n, int prec /* 32 | 64 */);
void zo_state_free(zo_state *s);
void zo_set_threads(int t);
int zo_get_threads(void);
void zo_reset_all(zo_state *s);
void zo_apply_unitary(zo_state *s, const uint32_t *qs, int k,
                      const double *mat);
void zo_apply_diagonal(zo_state *s, const uint32_t *qs, int k,
                       const double *d);
void zo_apply_cx(zo_state *s, uint32_t control, uint32_t target);
double zo_prob_one(zo_state *s, uint32_t q);
int zo_measure(zo_state *s, uint32_t q, double u);
double zo_norm_sqr(zo_state *s);
void zo_statevector(const zo_state *s, double *out);
/* Copies the raw state (T = float or double by precision) */
void zo_state_raw(const zo_state *s, void *re, void *im);
void zo_state_set_raw(zo_state *s, const void *re, const void *im);
void zo_apply_compiled(zo_state *s, const zo_compiled *c, int64_t op_lo,
                       int64_t op_hi);
void zo_sample(const zo_state *s, uint64_t shots, uint64_t seed,
               uint64_t *out);
void zo_chunk_sums(const zo_state *s, double *out);

/* ---- executor ----------------------------------------------------- */
/* Runs a compiled program; writes unique clbit words and their counts,
 * sorted by clbit word.  Returns the number of distinct keys, or -1. */
int64_t zo_run(const zo_compiled *c, int prec, uint64_t shots, uint64_t seed,
               const zo_noise *noise, uint64_t *keys, uint64_t *counts,
               int64_t cap, double *sim_seconds);

/* exec.rs:459-528 run_per_shot with its shot-parallel fold: words[shot]
 * gets each shot's clbit word; shots are dealt to `threads` workers with a
 * state each (threads <= 1: in order on one state).  0 or -1. */
int zo_run_shots(const zo_compiled *c, int prec, uint64_t shots, uint64_t seed,
                 const zo_noise *noise, int threads, uint64_t *words);

/* ---- RefSim (tests/reference.rs:69-116) ---------------------------- */
void zo_refsim_apply(double *amps, uint32_t n, const uint32_t *qs, int k,
                     const double *mat);

#ifdef __cplusplus
}
#endif
#endif

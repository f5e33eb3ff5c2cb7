/*
 * zeno_hip.h — C ABI of the MI355X-native zeno state-vector backend.
 *
 * Two layers, both plain C (pointers and sizes, no C++ or torch types):
 *
 *  1. zb_*  — the drop-in boundary.  One entry point per method of the
 *             reference's `pub trait Backend: Send` (src/exec.rs:134-162)
 *             plus its constructor make_backend (src/exec.rs:306-331).  A
 *             Rust FFI crate binds these 1:1.
 *  2. zr_*  — the host side above the boundary (compiler + executor),
 *             mirroring src/compiler.rs:118-200 compile() and
 *             src/exec.rs:334-452 run_compiled/run_sampled.  The Python
 *             package zeno_amd binds these through ctypes.
 *
 * Conventions (src/ir.rs:7-11, src/exec.rs:134-162):
 *  - qubits are little-endian: bit q of a basis index is qubit q;
 *  - `qs` passed to zb_apply_* is sorted ascending and distinct; bit b of a
 *    local matrix index is qs[b];
 *  - matrices are row-major 2^k x 2^k of complex128, interleaved (re, im)
 *    exactly as a Rust `&[Complex64]` is laid out; diagonals have 2^k
 *    entries;
 *  - the backend converts f64 inputs to its precision with round to nearest
 *    (T::fr, src/state.rs:41-50);
 *  - ops are asynchronous on the backend's HIP stream; every read
 *    (prob_one, measure, sample, statevector, norm_sqr, reset_all, finish)
 *    synchronises, as the Metal backend does (src/metal.rs:562-626).
 *
 * Errors: functions returning int give 0 on success and a negative code on
 * failure; functions returning double give NaN on failure.  zb_last_error()
 * holds the message (thread-local).
 */
#ifndef ZENO_HIP_H
#define ZENO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZB_OK 0
#define ZB_ERR_ARG -1         /* invalid argument (qubit range, k, order) */
#define ZB_ERR_UNSUPPORTED -2 /* Error::Unsupported (src/lib.rs:101-102) */
#define ZB_ERR_MEMORY -3      /* Error::Memory (src/lib.rs:82-95)        */
#define ZB_ERR_DEVICE -4      /* HIP runtime failure                     */

typedef struct zb_backend zb_backend;

/* ---------------------------------------------------------------------
 * Layer 1: the Backend trait (src/exec.rs:134-162)
 * ------------------------------------------------------------------- */

/* make_backend(n_qubits, precision, BackendChoice::Hip)
 * (src/exec.rs:306-331) + StateVec::zero_state (src/state.rs:109-117).
 * precision: 32 or 64.  Allocates the split re/im state in HBM and sets
 * |0...0>.  Fails with ZB_ERR_MEMORY when 2^n * 2 * sizeof(T) exceeds the
 * free device memory, ZB_ERR_UNSUPPORTED when no HIP device is present or
 * n > 35. */
int zb_create(uint32_t n_qubits, int precision, zb_backend **out);

/* Drop for Box<dyn Backend>: frees device memory. */
void zb_destroy(zb_backend *h);
/* make_backend over caller-owned device memory: re and im each hold 2^n
 * values of the precision (8 or 4 bytes), stay owned by the caller and
 * must outlive the handle; the state is set to |0...0>.  For callers that
 * sub-allocate states from one HBM pool (the serving loop). */
int zb_create_bound(uint32_t n_qubits, int precision, void *re, void *im,
                    zb_backend **out);
/* Re-point a handle at other caller-owned state memory, possibly with a
 * different qubit count and precision (finishes queued work first, frees
 * a state the handle owned, sets |0...0>); the stream, parameter arena and
 * scratch buffers are kept.  Every zb_tune knob returns to the library
 * default and then takes the zb_tune_default values of the new shape, as
 * at zb_create: nothing set for an earlier shape carries over. */
int zb_rebind(zb_backend *h, uint32_t n_qubits, int precision, void *re,
              void *im);

/* Backend::name (src/exec.rs:135): "hip-f64" or "hip-f32". */
const char *zb_name(const zb_backend *h);

/* Backend::apply_unitary (src/exec.rs:136).  k in 1..6. */
int zb_apply_unitary(zb_backend *h, const uint32_t *qs, uint32_t k,
                     const double *mat);

/* Backend::apply_diagonal (src/exec.rs:137).  k in 1..12. */
int zb_apply_diagonal(zb_backend *h, const uint32_t *qs, uint32_t k,
                      const double *diag);

/* Backend::apply_cx (src/exec.rs:140-147; CpuBackend override at
 * src/exec.rs:197-199 → src/state.rs:394-417): a zero-arithmetic swap over
 * the control=1 half. */
int zb_apply_cx(zb_backend *h, uint32_t control, uint32_t target);

/* Backend::prob_one (src/exec.rs:150; src/state.rs:598-626), f64
 * accumulation in a fixed order (4096-element blocks: 16 serial adds per
 * lane, then a pairwise tree over 256 lanes, recursively). */
double zb_prob_one(zb_backend *h, uint32_t q);

/* Backend::measure (src/exec.rs:152; CpuBackend at :201-207): outcome =
 * u < P(1); collapses with scale T::fr(1/sqrt(p)).  Returns 0/1 or < 0. */
int zb_measure(zb_backend *h, uint32_t q, double u);

/* Backend::reset_all (src/exec.rs:153; src/state.rs:128-132). */
int zb_reset_all(zb_backend *h);

/* Backend::sample (src/exec.rs:154; src/sample.rs:29-111) with
 * rng = Xoshiro256PlusPlus::seed_from_u64(seed).  Writes `shots` basis
 * indices in ascending order to `out` (host memory).  From 32768 shots the
 * draws (the same xoshiro256++ stream) are generated, sorted and routed on
 * the device, and the handle's device scratch grows to hold them (about 24
 * bytes per shot, plus up to 176 KiB of the stream's jump maps). */
int zb_sample(zb_backend *h, uint64_t shots, uint64_t seed, uint64_t *out);

/* zb_sample without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * that must not call hipMalloc or hipFree, such as tight serving loops or
 * harnesses that forbid it).  The same kernels and the same result as
 * zb_sample (indices equal for the same state, shots and seed), with the
 * sampler's device scratch — chunk sums, prefix, draws, routing, sort temp —
 * in a fixed 64 MiB region per GPU that the library holds until the process
 * ends.  That is enough for 2^20 shots at any n <= 35; a request needing
 * more fails with ZB_ERR_ARG (the message names the bytes needed) before any
 * GPU work.  Every call on one GPU shares that region: calls that overlap on
 * one GPU (from several threads or handles) race and must be serialised by
 * the caller. */
int zb_sample_noalloc(zb_backend *h, uint64_t shots, uint64_t seed,
                      uint64_t *out);

/* Backend::statevector (src/exec.rs:155; src/state.rs:134-140): writes
 * 2^n interleaved complex128 values to `out` (host memory). */
int zb_statevector(zb_backend *h, double *out);

/* Backend::norm_sqr (src/exec.rs:156; src/state.rs:142-156), f64, in the
 * fixed order of zb_prob_one. */
double zb_norm_sqr(zb_backend *h);

/* Backend::finish (src/exec.rs:157-161): waits for queued work. */
int zb_finish(zb_backend *h);

/* Thread-local message of the last failure. */
const char *zb_last_error(void);

/* ---------------------------------------------------------------------
 * Layer 1 extensions (no counterpart method in the reference trait).
 * ------------------------------------------------------------------- */

/* Raw split state copy to host (T = float or double by precision), in
 * logical index order. */
int zb_state_raw(zb_backend *h, void *re, void *im);
/* Raw split state upload from host (logical index order). */
int zb_state_set_raw(zb_backend *h, const void *re, const void *im);
/* Device addresses of the split state (for zero-copy consumers); the
 * arrays are in the physical layout of zb_layout. */
int zb_state_device_ptrs(zb_backend *h, void **re, void **im);
/* Number of qubits and precision of a handle. */
uint32_t zb_n_qubits(const zb_backend *h);
int zb_precision(const zb_backend *h);
/* Per-chunk probability masses of the sampler's pass 1
 * (src/sample.rs:38-50; chunk = min(2^n, 65536)), f64, to host. */
int zb_chunk_sums(zb_backend *h, double *out);
/* Execution mode: 0 = exact (one sweep per op, reference expression
 * trees, bit-identical to the CPU path); 1 = tiled (ops whose support fits
 * one on-chip tile are batched into a single HBM pass with the same
 * per-op arithmetic, still bit-identical).  Default 1. */
int zb_set_mode(zb_backend *h, int mode);
/* The HIP stream the backend's kernels run on (hipStream_t). */
void *zb_stream(zb_backend *h);
/* max_i |a_i - b_i| over two states of equal size (any precisions), f64,
 * computed on the device; used to state the f32 error against f64. */
int zb_max_abs_diff(zb_backend *a, zb_backend *b, double *out);
/* Number of visible HIP devices (0 when there is none or the runtime
 * cannot start); never fails. */
int zb_device_count(void);
/* Free and total device memory in bytes: the HBM counterpart of the
 * reference's RAM budget query (src/mem.rs:11-17). */
int zb_mem_info(uint64_t *free_bytes, uint64_t *total_bytes);
/* Runs a plain device copy of n_doubles f64 values `reps` times and
 * returns the average milliseconds per copy: the known-size kernel
 * (8 B read + 8 B written per element) that scales hardware byte
 * counters.  Allocates and frees its own buffers. */
int zb_calibration_copy(uint64_t n_doubles, int reps, double *ms);
/* Streaming-copy probes of n_doubles f64 values (same bytes as the
 * calibration copy), average ms over `reps` after one untimed copy:
 * kind 0 = zb_calibration_copy (8-byte accesses); 1 = 16-byte accesses;
 * 2 = 16-byte accesses, four loads in flight per lane before the stores;
 * 3 = as 2 with non-temporal stores; 4..7 = kinds 0..3 on data hashed
 * from the index instead of a constant fill (values that change from
 * element to element, as a state's do).  What a read + write pass reaches
 * on the device, beside the 8 TB/s of the data sheet. */
int zb_copy_probe(uint64_t n_doubles, int reps, int kind, double *ms);
/* The backend's qubit layout: phys[q] = the bit of the state index that
 * holds circuit qubit q (n entries).  Every zb_* call takes and returns
 * circuit (logical) order; only zb_state_device_ptrs exposes the physical
 * arrays. */
int zb_layout(const zb_backend *h, uint32_t *phys);
/* Tuning knobs of the tiled kernel (flushes first): "prefetch" (0/1:
 * issue the next tile's HBM loads into registers while the current tile
 * computes), "win_grid" (workgroups per window launch, default 16384),
 * "slots" (3 or 4 register qubits per stage, 0 = 3) and "page_bits" (1..12:
 * most window qubits that select a 2 MiB page of the state, default 6),
 * "tile_bits" (12 or 13: tile width of the tiled kernel, 0 = 12),
 * "defer_cx" (0/1, default 1: a CX whose target is outside the open window
 * waits, bit-exactly, for a window that holds it; a CX is a permutation and
 * commutes with every op on other qubits), "relabel" (0/1, default 1: when
 * ops arrive through zr_apply_ops or a run, a window that closes because
 * the next op does not fit also moves onto the low state bits, on
 * write-back, the qubits of its tile that the coming ops need soonest; a
 * relabelling permutes state index bits only),
 * "layout" (-1 by size, 0 identity, 1 spread, >= 2 a seeded shuffle;
 * resets the state to |0...0>) and "chunk_multi" (>= 0, default 4096:
 * from this many 2^16-value chunks the sampler's pass 1 walks 8 chunks per
 * workgroup instead of one; 0 = never; the sums are the same serial
 * chains either way) and "kq_tiled" (0/1/2, default 1: per-op dense ops on
 * 2 to 6 qubits, n >= 12, stage tiles of 2^12 amplitudes through LDS
 * instead of gathering each group into one lane: 0 never, 1 when
 * min(k, 3) or more op bits lie inside a 128-byte run or for 6 qubits in
 * f64, 2
 * always; the same products and sums in the same order either way).  None
 * of them changes results, except
 * "mfma" (0/1, default 0): dense ops on 4 to 6 qubits run as real GEMMs on
 * the matrix cores (v_mfma_*_16x16x4) when n >= 12.  MFMA accumulation
 * rounds differently from the CPU path's sum, so those ops agree with it
 * within tolerance (max |d amp| <= 1e-12 in f64, <= 1e-5 in f32 on unit
 * states) instead of bit for bit; and "fma" (0/1, default 0): every
 * sweep's multiply-accumulate steps, in the CPU path's expression order,
 * become fused multiply-adds (one rounding instead of two), with the same
 * tolerance. */
int zb_tune(zb_backend *h, const char *key, int64_t value);
/* Process-wide default for a tuning knob of zb_tune, applied to every
 * backend of n_qubits at this precision when it is created or re-bound
 * (value INT64_MIN removes it).  The autotuner installs its choices
 * here. */
int zb_tune_default(uint32_t n_qubits, int precision, const char *key,
                    int64_t value);
/* Launch profiling.  zb_profile_begin flushes, synchronises and starts
 * recording a HIP event pair around every state-sweep kernel launch on the
 * backend's stream.  zb_profile_end flushes, synchronises, stops, and
 * writes per launch: the kernel id (ZB_K_*), its duration in ms and its
 * algorithmic HBM bytes (4*2^n*s per sweep, 2*2^n*s for CX and zero init).
 * Returns the number of records written (<= cap) or < 0. */
#define ZB_K_ZERO 0
#define ZB_K_1Q 1
#define ZB_K_CX 2
#define ZB_K_DIAG1 3
#define ZB_K_DIAGK 4
#define ZB_K_WINDOW 5
#define ZB_K_KQ 10 /* + k for the dense k-qubit kernel */
#define ZB_K_KQ_MFMA 20 /* + k for the dense k-qubit kernel on matrix cores */
int zb_profile_begin(zb_backend *h);
int64_t zb_profile_end(zb_backend *h, int32_t *kernel, double *ms,
                       double *bytes, int64_t cap);
/* For the launches the last zb_profile_end returned, in the same order,
 * 4 values each: a window's ops, stages, diagonal ops and CX ops (zero
 * for other kernels).  Returns the number of launches written. */
int64_t zb_profile_detail(const zb_backend *h, int32_t *detail, int64_t cap);

/* ---------------------------------------------------------------------
 * Layer 2: compiler + executor (the host side above the boundary)
 * ------------------------------------------------------------------- */

/* Program instruction (one ir::Instr, src/ir.rs:36-52).  kind: 0 gate,
 * 1 measure, 2 reset, 3 barrier.  cond_creg >= 0 wraps the instruction in
 * `if (creg == cond_value)`.  72 bytes. */
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
} zr_instr;

/* noise::NoiseModel (src/noise.rs:37-46). */
typedef struct {
    double depolarizing_1q, depolarizing_2q, bit_flip, phase_flip,
        amplitude_damping, readout_flip_0to1, readout_flip_1to0;
} zr_noise;

typedef struct zr_compiled zr_compiled;

/* compiler::compile (src/compiler.rs:118-200): cancel, lower, diagonal
 * fusion (diag_max <= 12), dense fusion (fusion_max <= 6).  Returns
 * NULL on error (zb_last_error()). */
zr_compiled *zr_compile(const zr_instr *prog, int64_t n_instr,
                        uint32_t n_qubits, const uint32_t *creg_sizes,
                        uint32_t n_cregs, int fusion_max, int diag_max);
void zr_compiled_free(zr_compiled *c);
int64_t zr_compiled_nops(const zr_compiled *c);
/* Op kinds: 0 unitary, 1 diagonal, 2 cx, 3 measure, 4 reset.  data gets
 * the matrix or diagonal (interleaved complex128); cond = offset, len,
 * value of an `if` wrapper. */
int zr_compiled_op(const zr_compiled *c, int64_t i, int32_t *kind,
                   int32_t *k, uint32_t *qs, double *data, int32_t *is_if,
                   uint64_t *cond);
/* compiler::Stats (src/compiler.rs:68-78) + dynamic + #final measures. */
void zr_compiled_stats(const zr_compiled *c, int64_t *stats /* 6 */);
int64_t zr_compiled_final_measures(const zr_compiled *c, uint32_t *pairs);

/* Apply ops [lo, hi) of a compiled program to a backend
 * (src/exec.rs:398-405 apply_op in a loop). */
int zr_apply_ops(zb_backend *h, const zr_compiled *c, int64_t lo, int64_t hi);

/* mem::MemPlan (src/mem.rs:95-101) together with the fields of
 * Error::Memory (src/lib.rs:82-95), so a binding can build either. */
typedef struct {
    uint32_t n_qubits;
    int32_t precision;     /* 32 or 64 */
    uint64_t needed_bytes; /* mem::state_bytes: 2^n * (8 or 16), saturated */
    uint64_t budget_bytes;
    uint32_t max_qubits;   /* mem::max_qubits(budget, precision) */
} zr_mem_plan;

/* mem::budget_bytes (src/mem.rs:103-109) with HBM in place of RAM:
 * mem_limit > 0 wins; otherwise fraction * the device's total HBM
 * (hipMemGetInfo), where a numeric ZENO_MEM_BYTES replaces that total as it
 * replaces physical RAM in src/mem.rs:11-17.  Returns 0 when there is no
 * override and no device. */
uint64_t zr_budget_bytes(uint64_t mem_limit, double fraction);

/* mem::plan (src/mem.rs:114-130): fills *plan (may be NULL); ZB_OK when
 * the state fits, else ZB_ERR_MEMORY with Error::Memory's message
 * ("state vector for N qubits at fP needs X but the budget is Y (max M
 * qubits at this precision; ...)") in zb_last_error(). */
int zr_plan(uint32_t n_qubits, int precision, uint64_t budget,
            zr_mem_plan *plan);

/* resolve_precision (src/exec.rs:261-304) on the HBM budget.  requested:
 * 0 = auto (f64 if it fits, else f32 with the reference's auto-precision
 * notice written to `notice`, NUL-terminated, at most notice_cap bytes),
 * 32 or 64 = that precision if it fits.  *precision gets the choice and
 * *plan its mem::plan; ZB_ERR_MEMORY (with *plan filled) when even the
 * chosen precision does not fit. */
int zr_resolve_precision(uint32_t n_qubits, int requested, uint64_t budget,
                         int *precision, zr_mem_plan *plan, char *notice,
                         size_t notice_cap);

/* exec::run_compiled (src/exec.rs:334-549) on the HIP backend: static
 * circuits run once and are sampled (run_sampled); dynamic or noisy ones
 * run per shot (run_per_shot).  Writes the unique clbit words with their
 * counts, ascending by word; returns how many, or < 0.  sim_seconds covers
 * state allocation, every op, finish and sampling (src/exec.rs:372-381).
 * mode: 0 and 1 as zb_set_mode; per-shot circuits in mode 1 run as batched
 * trajectories (zr_run_trajectories) where that is the faster way, else
 * the host's per-shot loop; mode 2 is mode 1 with the per-shot loop always
 * (the words are the same either way). */
int64_t zr_run(const zr_compiled *c, int precision, uint64_t shots,
               uint64_t seed, const zr_noise *noise, int mode,
               uint64_t *keys, uint64_t *counts, int64_t cap,
               double *sim_seconds);

/* Batched noisy / dynamic trajectories (src/exec.rs:459-619,
 * src/noise.rs:91-192): every shot runs in one persistent kernel launch
 * captured in a hipGraph, each workgroup executing whole trajectories on
 * its own state.  Shot s uses Xoshiro256++ seeded with
 * splitmix64(seed ^ s) and the frozen draw order, so out[s] (the shot's
 * clbit word) equals the CPU per-shot path's.  Needs n <= 26, dense ops on
 * <= 6 qubits and diagonals on <= 12 (zr_trajectories_supported);
 * kernel_seconds gets the launch-to-completion time.  The call allocates
 * one workspace for up to two concurrent trajectories per CU, fewer when
 * free device memory is short (ZB_ERR_MEMORY when not even one fits). */
int zr_trajectories_supported(const zr_compiled *c);
int zr_run_trajectories(const zr_compiled *c, int precision, uint64_t shots,
                        uint64_t seed, const zr_noise *noise, uint64_t *out,
                        double *kernel_seconds);

/* Many small static circuits as uneven instances of one persistent kernel
 * launch (no counterpart in the reference, which runs one circuit per
 * call; each instance reproduces run_sampled, src/exec.rs:407-452).  Each
 * instance's state stays in on-chip memory for its whole run: static
 * circuits of n <= 13 qubits in f64 or <= 14 in f32, dense ops on <= 3
 * qubits in f64 or <= 4 in f32, <= 2^17 shots (zr_batch_supported).  The launch is captured in a
 * hipGraph.  words gets, for instance i at offset shots[0] + ... +
 * shots[i-1], its shots' clbit words (instances without final measures
 * write none: their counts are empty); their counts equal zr_run's for the
 * same circuit, precision, shots and seed.  seconds (may be NULL) gets the
 * wall time of the batch. */
int zr_batch_supported(const zr_compiled *c, int precision, uint64_t shots);
int zr_run_batch(const zr_compiled *const *circuits, const int32_t *precisions,
                 const uint64_t *shots, const uint64_t *seeds, int64_t count,
                 uint64_t *words, double *seconds);

/* ---------------------------------------------------------------------
 * Layer 3: serving loop (no counterpart in the reference, which runs one
 * circuit per invocation with mem::plan admission, src/mem.rs:79-130)
 * ------------------------------------------------------------------- */

/* Continuous batching on one GPU: one HBM pool of pool_bytes (0 = 80% of
 * free device memory) is allocated at start; `slots` worker slots each own
 * a backend with its own HIP stream.  A request is admitted, first fit in
 * arrival order, when a slot is idle and the pool has room for its state
 * (2 * 2^n * 4 or 8 bytes); its state lives in the pool until it
 * completes.  A noisy or dynamic request whose shots run as batched
 * trajectories is charged its trajectory workspace instead when that is
 * larger (up to two states per CU, at most 2 GiB of them, plus op tables),
 * and runs in it: no request allocates device memory.  mode: see
 * zb_set_mode. */
typedef struct zs_server zs_server;
int zs_create(uint64_t pool_bytes, int slots, int mode, zs_server **out);
/* Queue a run of a compiled circuit (the server copies it; the caller may
 * free it at once).  Returns the request id (> 0), or < 0 when the state
 * can never fit in the pool (ZB_ERR_MEMORY) or the arguments are bad. */
int64_t zs_submit(zs_server *s, const zr_compiled *c, int precision,
                  uint64_t shots, uint64_t seed, const zr_noise *noise);
/* Block until request `id` completes; writes its clbit-word counts as
 * zr_run does and times[3] = submit, start, end (seconds since the server
 * started).  Each id can be waited for once. */
int64_t zs_wait(zs_server *s, int64_t id, uint64_t *keys, uint64_t *counts,
                int64_t cap, double *times);
/* out[4] = queued requests, running requests, largest free pool block,
 * pool size (bytes). */
void zs_status(zs_server *s, int64_t *out);
/* Small static requests (zr_batch_supported) do not take a slot: a batcher
 * runs all that have arrived, up to max_batch (default 1024), as instances
 * of one persistent launch while the slots run the others; whatever arrives
 * meanwhile forms the next batch.  max_batch 0 sends every request to the
 * slots (one at a time with slots = 1).  Exact mode (mode 0) never batches. */
int zs_set_batch(zs_server *s, int max_batch);
/* out[4] = batches launched, requests served in them, graph replays of the
 * batch launch, max_batch. */
void zs_counters(zs_server *s, int64_t *out);
/* Fails queued requests, waits for running ones, frees the pool.  No
 * zs_wait may still be blocked on this server. */
void zs_destroy(zs_server *s);

/* Seeded circuit generators (src/circuit.rs:138-174 and the benchmark
 * configurations' matching/RZZ variant).  Return the instruction count;
 * write at most `cap`. */
int64_t zr_gen_random(uint32_t n, uint32_t depth, uint64_t seed,
                      zr_instr *out, int64_t cap);
int64_t zr_gen_qft(uint32_t n, zr_instr *out, int64_t cap);
int64_t zr_gen_u3_qft(uint32_t n, uint64_t seed, zr_instr *out, int64_t cap);
int64_t zr_gen_matching(uint32_t n, uint32_t depth, uint64_t seed,
                        uint32_t rzz_period, zr_instr *out, int64_t cap);

/* OpenQASM 2.0 front end (src/qasm.rs:70-76 parse_str; subset and rules
 * of docs/QASM.md): writes up to `cap` instructions, the qubit count and
 * up to cregs_cap classical register sizes; returns the instruction count
 * (call again with a larger cap if it exceeds cap) or < 0 with
 * "line L:C: message" in zb_last_error(). */
int64_t zr_parse_qasm(const char *src, zr_instr *out, int64_t cap,
                      uint32_t *n_qubits, uint32_t *cregs, uint32_t *n_cregs,
                      uint32_t cregs_cap);

/* gates::build (src/gates.rs:119-188): returns the arity or -1. */
int zr_gate_build(const char *name, const double *params, int np,
                  double *out, int *is_diag);

#ifdef __cplusplus
}
#endif
#endif

// host.hpp — host-side types of the zeno HIP backend: complex numbers with
// the reference's exact arithmetic, the seeded RNG, the compiled-op stream
// and the compiler/executor entry points (C ABI in include/zeno_hip.h).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "zeno_hip.h"

namespace zeno {

// Complex number with num-complex 0.4 arithmetic (no FMA, no Annex-G
// NaN recovery): Mul is (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re).
struct cplx {
    double re, im;
};
inline cplx mk(double re, double im) { return cplx{re, im}; }
inline cplx operator*(cplx a, cplx b) {
    return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
inline bool is_zero(cplx a) { return a.re == 0.0 && a.im == 0.0; }

// splitmix64 (src/exec.rs:227-232).
uint64_t splitmix64(uint64_t x);

// Xoshiro256PlusPlus seeded as rand_xoshiro 0.6.0 seed_from_u64 does, with
// rand 0.8's Standard f64 and UniformInt<u64>::sample_single.
class Rng {
  public:
    explicit Rng(uint64_t seed);
    uint64_t next();
    double f64();
    uint64_t range(uint64_t lo, uint64_t hi);

  private:
    uint64_t s_[4];
};

enum OpKind : int { kUnitary = 0, kDiagonal = 1, kCx = 2, kMeasure = 3, kReset = 4, kBarrier = 5 };

// One compiled op (compiler.rs:24-55).  Unitary/Diagonal carry sorted qubits
// and a row-major matrix or a diagonal; Cx carries control/target; Measure
// carries qubit (a) and clbit (b); Reset carries qubit (a).  `cond` wraps
// the op in `if (clbits[off..off+len] == value)`.
struct Op {
    OpKind kind = kUnitary;
    std::vector<uint32_t> qs;
    std::vector<cplx> data;
    uint32_t a = 0, b = 0;
    bool is_if = false;
    uint32_t creg_off = 0, creg_len = 0;
    uint64_t value = 0;
    bool gate_like() const { return !is_if && (kind == kUnitary || kind == kDiagonal || kind == kCx); }
    // COp::gate_support: Cx gives [control, target] (unsorted).
    std::vector<uint32_t> support() const {
        if (kind == kCx) return {a, b};
        return qs;
    }
};

struct Stats {
    int64_t input_gates = 0, cancelled = 0, output_ops = 0, max_fused = 0;
};

struct Compiled {
    uint32_t n_qubits = 0, n_clbits = 0;
    std::vector<uint32_t> creg_sizes;
    std::vector<Op> ops;
    bool dynamic = false;
    std::vector<std::pair<uint32_t, uint32_t>> final_measures;
    Stats stats;
};

// gates::build (gates.rs:119-188).  Returns arity or -1.
int gate_build(const char *name, const double *params, int np, cplx *out, bool *is_diag);
int gate_lookup(const char *name, int *arity, int *n_params);

// compiler::compile (compiler.rs:118-200).  Throws std::runtime_error with
// the reference's message on invalid input.
Compiled compile(const zr_instr *prog, int64_t n, uint32_t n_qubits, const uint32_t *creg_sizes,
                 uint32_t n_cregs, int fusion_max, int diag_max);

}  // namespace zeno

// Opaque handle of the C ABI (include/zeno_hip.h).
struct zr_compiled {
    zeno::Compiled c;
};

namespace zeno {

// exec::run_compiled (src/exec.rs:407-549) on an existing backend whose
// qubit count matches: resets it to |0...0> (unless the caller says it
// already is), runs the circuit (sampled,
// batched trajectories, or per shot) and writes the clbit-word counts,
// ascending by word.  Returns ZB_OK or a negative ZB_ERR_*.
int run_compiled_on(zb_backend *h, const zr_compiled *c, uint64_t shots, uint64_t seed, const zr_noise *noise,
                    int mode, bool state_is_zero, std::vector<std::pair<uint64_t, uint64_t>> &out);

// Window scheduling hint: need[i] is the mask of circuit qubits that the
// i-th coming op needs inside a tile (a dense op's qubits, a CX's target),
// valid until cleared with (nullptr, 0); pos is the op about to be applied.
// With it, windows relabel qubits on write-back (results do not change).
void backend_lookahead(zb_backend *h, const uint64_t *need, size_t count);
void backend_lookahead_pos(zb_backend *h, size_t pos);
// Flushes the run's queued work, the last window homing circuit qubits
// 0..low-1 onto the low state bits; call while the hint is still set.
// Returns a ZB_* code.
int backend_end_run(zb_backend *h);

// Thread-local last-error plumbing for the C ABI.
void set_error(const std::string &msg);
const char *last_error();

}  // namespace zeno

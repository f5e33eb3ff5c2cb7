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
    const uint64_t *state() const { return s_; }  // the next output's state (s0..s3)

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

struct TrajWorkspace;

// exec::run_compiled (src/exec.rs:407-549) on an existing backend whose
// qubit count matches: resets it to |0...0> (unless the caller says it
// already is), runs the circuit (sampled, batched trajectories, or per
// shot) and writes the clbit-word counts, ascending by word.  Batched
// trajectories run in `traj` when given (no device allocation), else in a
// workspace allocated for the call; when that does not fit, the shots run
// one by one on h.  Returns ZB_OK or a negative ZB_ERR_*.
int run_compiled_on(zb_backend *h, const zr_compiled *c, uint64_t shots, uint64_t seed, const zr_noise *noise,
                    int mode, bool state_is_zero, std::vector<std::pair<uint64_t, uint64_t>> &out,
                    TrajWorkspace *traj = nullptr);

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

// Batched trajectories (trajectories.hip) in memory the caller provides:
// `mem` holds the states of the concurrent trajectories plus the op tables
// and per-shot words (traj_workspace_bytes); `stream` is the launch stream.
// With a fixed `launch_block` (device memory of at least 256 bytes that the
// holder keeps), the captured launch in `graph` is replayed by every later
// run at the same grid and precision; the holder frees it with
// traj_graph_free.  No device memory is allocated.
struct TrajGraph;
struct TrajWorkspace {
    void *mem = nullptr;
    uint64_t bytes = 0;
    void *stream = nullptr;        // hipStream_t
    void *launch_block = nullptr;  // optional, see above
    TrajGraph *graph = nullptr;
};
constexpr uint64_t kTrajLaunchBlockBytes = 256;
// Free device memory a standalone batched run leaves to everything else.
constexpr uint64_t kTrajMargin = 256ull << 20;
// Concurrent trajectories worth running: min(shots, 2 per CU), at least 1.
uint64_t traj_max_grid(uint64_t shots);
// Workspace bytes for `grid` concurrent trajectories of c.
uint64_t traj_workspace_bytes(const Compiled &c, int precision, uint64_t shots, uint64_t grid);
// The grid a workspace of `bytes` holds (0: not even one state).
uint64_t traj_grid_for(const Compiled &c, int precision, uint64_t shots, uint64_t bytes);
int run_trajectories_in(const Compiled &c, int precision, uint64_t shots, uint64_t seed, const zr_noise *noise,
                        uint64_t *out, double *secs, TrajWorkspace &w);
void traj_graph_free(TrajGraph *g);
uint64_t traj_graph_replays(const TrajGraph *g);
bool trajectories_supported(const Compiled &c);
// Whether the batched launch is the faster way to run c's shots with a
// workspace of `bytes`: always up to 20 qubits (a trajectory's state stays
// in the L2 / last-level cache); above, where every sweep streams HBM, only
// with at least one concurrent trajectory per CU, since fewer stream less
// than the chip's bandwidth that the serial loop's full-grid sweeps use.
bool traj_preferred(const Compiled &c, int precision, uint64_t shots, uint64_t bytes);

// Small static circuits as instances of one persistent launch (batch.hip).
// A runner owns a stream, a device workspace with its pinned staging, a
// fixed launch block and the graphs it captured (one per precision, state
// size and grid, replayed whenever a batch of that shape comes).
constexpr uint64_t kBatchMaxShots = 1ull << 17;
struct BatchRunner;
struct BatchItem {
    const Compiled *c = nullptr;
    int precision = 64;
    uint64_t shots = 0, seed = 0;   // the run's seed, as zr_run takes it
    std::vector<uint64_t> indices;  // out: sampled basis indices, in draw order
};
// Static, n <= 13 (f64) / 14 (f32), dense ops on <= 3 (f64) / 4 (f32)
// qubits, <= 2^17 shots.
bool batch_supported(const Compiled &c, int precision, uint64_t shots);
// Workspace bytes one instance takes in a batch.
uint64_t batch_item_bytes(const Compiled &c, int precision, uint64_t shots);
BatchRunner *batch_runner_new(uint64_t workspace_bytes);
void batch_runner_free(BatchRunner *r);
uint64_t batch_runner_capacity(const BatchRunner *r);
uint64_t batch_runner_replays(const BatchRunner *r);
// Runs the items (one launch per precision present); ZB_ERR_MEMORY when
// they do not fit the runner's workspace together.
int batch_run(BatchRunner *r, std::vector<BatchItem> &items);
// Sampled indices -> clbit words through the final measures.
void batch_words(const Compiled &c, const std::vector<uint64_t> &idx, uint64_t *words);

// Thread-local last-error plumbing for the C ABI.
void set_error(const std::string &msg);
const char *last_error();

}  // namespace zeno

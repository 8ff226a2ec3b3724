// Device statevector engine: simulator options, plan lowering, execution on
// the GPU, terminal sampling, trajectories and result queries.
//
// Follows src/monolithic/statevector/{shots.rs,simulator.rs} and
// src/engine.rs of the reference for semantics; the kernels are in
// kernels.hip.
#pragma once

#include <cstdint>
#include <exception>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "circuit.hpp"
#include "planner.hpp"

namespace dq {

struct MemErr : std::exception {
  std::string msg;
  explicit MemErr(std::string m) : msg(std::move(m)) {}
  const char* what() const noexcept override { return msg.c_str(); }
};
struct RuntimeErr : std::exception {
  std::string msg;
  explicit RuntimeErr(std::string m) : msg(std::move(m)) {}
  const char* what() const noexcept override { return msg.c_str(); }
};
struct DeviceErr : std::exception {
  std::string msg;
  explicit DeviceErr(std::string m) : msg(std::move(m)) {}
  const char* what() const noexcept override { return msg.c_str(); }
};

enum Fusion { FUSE_NONE = 0, FUSE_GREEDY = 1, FUSE_LEARNED = 2 };

// LDS tile sweeps (complex64 with fusion): tile sizes the kernel is built for.
constexpr int kTileMin = 10, kTileMax = 12, kTileDefault = 12;
// Out-of-place complex64 sweeps also take 13-qubit tiles (the whole 64 KiB
// LDS; in-place sweeps of such a plan use 12).
constexpr int kXTileMax = 13;

struct Options {
  bool has_seed = false;
  uint64_t seed = 0;
  bool profile = false;
  uint64_t budget = 1024ull << 20;  // bytes
  bool has_max_parallel_shots = false;
  uint64_t max_parallel_shots = 0;
  bool sample_terminal = true;
  int prec = 16;  // bytes per amplitude: 16 = complex128, 8 = complex64
  int fusion = FUSE_NONE;
  int max_fused = 4;
  int tile = kTileDefault;  // LDS tile qubits for complex64 fused sweeps; 0 = off
};

struct SimProfile {
  uint64_t oq_calls = 0, nq_calls = 0, mq_calls = 0;
  double oq_time = 0, nq_time = 0, mq_time = 0, total_time = 0;
};

// Streams are reused per device (creating and destroying one costs about
// 0.1 ms, a visible share of a small circuit's call). The pool is never torn
// down: the HIP runtime may be gone by static destruction.
hipStream_t acquire_stream(int* dev);
void release_stream(hipStream_t s, int dev);

class DeviceMem {
 public:
  DeviceMem() = default;
  explicit DeviceMem(size_t bytes) { alloc(bytes); }
  DeviceMem(const DeviceMem&) = delete;
  DeviceMem& operator=(const DeviceMem&) = delete;
  DeviceMem(DeviceMem&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  DeviceMem& operator=(DeviceMem&& o) noexcept;
  ~DeviceMem();
  void alloc(size_t bytes);
  void reset();
  void* get() const { return p_; }
  size_t size() const { return n_; }
  template <typename T>
  T* as() const {
    return (T*)p_;
  }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

struct Result {
  int prec = 16;
  uint32_t n = 0;
  uint64_t nc = 0;
  DeviceMem state;
  std::vector<uint8_t> cbits;
  bool has_profile = false;
  SimProfile prof;
};

using Counts = std::vector<std::pair<std::string, uint64_t>>;

// Per-launch timing record (engine-recorded kernel timings).
struct OpTiming {
  int kind;      // dq_op_timing.kind codes (include/dqsim_sv.h)
  int k;         // qubits touched
  int low;       // lowest qubit touched
  int gates;     // source gates folded into the launch
  double ms;     // device time
  double bytes;  // algorithmic HBM bytes moved
  float feat[mlp::kFeatures];  // planner features of the launch
};

struct EvolveStats {
  std::vector<OpTiming> ops;
  double total_ms = 0;      // event time over the whole evolution
  uint64_t gates = 0;       // unitary gate instructions in the input
  uint64_t launches = 0;
};

std::unique_ptr<Result> simulate(const Options& o, const char* json, size_t len);
// simulate() of a unitary circuit into a caller's device state (state_len =
// 2^n amplitudes), plan constants in the static scratch array of
// noalloc_scratch(); allocates no device memory and returns after `stream`
// has finished.
void evolve_into(const Options& o, const char* json, size_t len, void* state, uint64_t state_len,
                 hipStream_t stream);
// Device address of this GPU's allocation-free scratch array (resolved once
// per device, with that device current) and its size.
void* noalloc_scratch(uint64_t* bytes);
Counts simulate_shots(const Options& o, const char* json, size_t len, uint64_t shots, bool profile);

void result_probabilities(const Result& r, const std::vector<uint64_t>& qubits, double* out);
Counts result_counts(const Result& r, uint64_t shots, const std::vector<uint64_t>& qubits, uint64_t seed);
double result_fidelity(const Result& r, const void* other_c128, uint64_t len);
void result_statevector(const Result& r, void* out_c128);
void result_amplitude(const Result& r, uint64_t index, double* out);
void result_gather(const Result& r, const uint64_t* idx, uint64_t count, double* out);

// Evolve the unitary prefix of a circuit from |0...0> `repeats` times and
// report per-launch device timings of the last repeat. The final state can
// be kept (returned in *keep) for accuracy checks.
EvolveStats evolve_timed(const Options& o, const char* json, size_t len, int repeats, int record,
                         std::unique_ptr<Result>* keep);

// Max |a-b| and sqrt(sum |a-b|^2) between two results of equal size.
void result_compare(const Result& a, const Result& b, double* max_abs, double* l2);

// Compile-only plan description (no device use), as JSON.
std::string plan_description(const char* json, size_t len, int prec, int fusion, int max_fused, int tile,
                             bool permute);
// Sweep mode of complex64 fused plans whose budget holds a second state
// (0 in place, 1 out of place with persistent workgroups, 2 out of place with
// one workgroup per tile); returns the previous mode.
int set_sweep_mode(int mode);
// Arithmetic of 3-qubit blocks in complex64 tile sweeps (0 FP32 matrix cores,
// 1 split precision); returns the previous mode, -1 for an invalid one.
int set_block_arith(int mode);
// Wave-owned block runs in 256-thread out-of-place sweeps on (1, default) or
// a barrier after every block (0); returns the previous mode, -1 for an
// invalid one.
int set_wave_runs(int mode);
// Zero-start of plans that open with out-of-place sweeps on or off
// (process-wide); returns the previous setting.
bool set_zero_start(bool on);
// The FP16 split of split-precision block matrices: hi = RNE(x), lo =
// RNE(x - hi), per element (|x| < 65520).
void split_f16_host(const double* x, uint64_t n, uint16_t* hi, uint16_t* lo);

int device_count();

// ---- batched small circuits (stream pool + single-launch kernel, batch.hip) ----
struct BatchOptions {
  int workers = 0;          // host threads, one HIP stream each; 0 = tuned (default 4)
  int max_fused = 0;        // fusion width 1..4; 0 = tuned (default 3)
  bool autotune = true;     // time (workers, width) candidates on a sample
  int tile = kTileDefault;  // LDS tile qubits of each circuit's plan (0 = off)
  int prec = 8;             // bytes per amplitude: 8 complex64, 16 complex128
  bool dump_states = false;
  bool persistent = false;  // the persistent kernel: removed, raises
  // Single launch: complex64 circuits of up to this many qubits (<= 12) that
  // measure every qubit run one workgroup each, all in one hipGraph launch
  // concurrent with the pool; 0 = every circuit on the pool.
  int lds_qubits = 12;
};
struct BatchOutcome {
  std::vector<uint64_t> nq, nc;             // per circuit
  std::vector<uint64_t> offsets;            // sample offset per circuit
  std::vector<uint32_t> idx;                // outcome per shot over the measured qubits (highest = MSB)
  std::vector<std::vector<std::pair<uint64_t, uint64_t>>> cmap;  // (cbit, bit of the outcome)
  std::vector<double> latency_us;           // per circuit: plan, upload, evolution, sampling
  std::vector<uint64_t> state_off;          // kept final states (amplitude offsets)
  std::vector<double> states;               // interleaved re, im
  double total_ms = 0;                      // wall time of the run (tuning excluded)
  int workers = 0, fused = 0, tile = 0;
  uint32_t streams = 0;
  uint64_t lds_count = 0;                   // circuits run by the single launch
  double lds_ms = 0;                        // its device time (graph launch to completion)
  std::vector<std::pair<std::string, double>> tuning;  // candidate -> ms on the sample
};
std::unique_ptr<BatchOutcome> batch_run(const std::vector<std::pair<const char*, size_t>>& circuits,
                                        const std::vector<uint64_t>& shots, const std::vector<uint64_t>& seeds,
                                        const BatchOptions& opt);
Counts batch_counts(const BatchOutcome& b, uint64_t i);
// Measured FP32 matrix-core throughput (TFLOP/s) of a register-only
// v_mfma_f32_16x16x4f32 loop over every CU; also the time per launch.
double mfma_fp32_peak_tflops(int reps, double* ms_per_launch);
// Relabelled tile sweeps on or off (process-wide); returns the previous setting.
bool set_sweep_relabel(bool on);
// Out-of-place sweeps: qubits 0..3 in every tile (true) or only in each
// run's first and last tile (false, the default); returns the previous value.
bool set_sweep_pinned(bool on);
// Host-only text dump of the program the persistent batched kernel would
// upload (circuits, launch order, fused ops, FP32 matrices; circuits above
// `lmax` qubits use an HBM slot of the printed size).
std::string batch_program_text(const std::vector<std::pair<const char*, size_t>>& circuits,
                               const std::vector<uint64_t>& shots, const std::vector<uint64_t>& seeds, int max_fused,
                               int lmax);
// Device and kernel limits relevant to the batched launcher (no launch);
// fills info[0..5] and info[7..10] in the order of dq_batch_device_info.
void batch_device_info(int block, int* info);

// Learned fusion planner weights (2 hidden layers of 64, FP32). Null clears.
// Host forward pass of the planner MLP (packed weights, one feature vector).
float mlp_eval_host(const float* w, const float* x);
void set_planner_weights(const float* w, size_t count);
// Train the planner MLP on the GPU (Adam, mean squared error); w is in/out.
void mlp_train(float* w, const float* X, const float* T, uint64_t n, int epochs, int batch, float lr, uint64_t seed,
               float* loss_curve);
// One forward + backward pass on the GPU: gradients and loss of a batch.
void mlp_grad(const float* w, const float* X, const float* T, int nb, float* g, float* loss);
// GPU forward pass.
void mlp_predict(const float* w, const float* X, uint64_t n, float* y);
// Host FP32 forward pass (the planner's own inference path).
float mlp_eval_host(const float* w, const float* x);

}  // namespace dq

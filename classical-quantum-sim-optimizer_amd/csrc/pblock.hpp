// P-block backend: the state as a product of independent dense blocks
// (reference src/distributed/pblock/). Block amplitudes are complex128 on
// the GPU; the host keeps the reference's bookkeeping (which block and local
// position holds each wire, which wires are in a known basis state,
// statistics and the memory budget).
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "engine.hpp"

namespace dq {
namespace pb {

// PBlockSimulator options (simulator.rs:45-105).
struct Options {
  uint64_t memory = 1024ull << 20;  // bytes
  bool has_max_block = false;
  uint64_t max_block_qubits = 0;
  uint64_t max_parallel_shots = 1;
  bool sample_terminal = true;
  bool split_separable = false;
  uint64_t max_split_qubits = 12;
  double separation_tolerance = 0.0;
  bool profile = false;
  bool has_seed = false;
  uint64_t seed = 0;
};
// The constructor's checks (ValueErr with the reference's messages).
void check_options(const Options& o);

// Counters of the block pool (model.rs:20-32).
struct Stats {
  uint64_t merge_calls = 0;
  double merge_time = 0;
  uint64_t merge_output_amplitudes = 0;
  uint64_t split_calls = 0;
  uint64_t separation_checks = 0;
  double separation_time = 0;
  double separation_residual = 0;
  uint64_t skipped_controlled_gates = 0;
  uint64_t peak_block_qubits = 0;
  uint64_t peak_live_amplitudes = 0;
  uint64_t peak_working_bytes = 0;
  void add(const Stats& o);
};

class Pool;
struct Result {
  std::shared_ptr<void> stream;                   // HIP stream of the pool (outlives it)
  std::unique_ptr<Pool> pool;
  std::vector<uint64_t> physical;                 // wire i = physical label physical[i]
  std::vector<std::pair<uint64_t, int>> cbits;     // (cbit, value), ascending cbit
  double elapsed = 0;
  Result();
  ~Result();
};

std::unique_ptr<Result> simulate(const Options& o, const char* json, size_t len);
Counts simulate_shots(const Options& o, const char* json, size_t len, uint64_t shots, bool profile);

// How the calling thread's last simulate_shots ran its trajectory shots:
// completed in lockstep batches, run alone (evicted from a batch, or every
// shot when batches are off), and the largest batch's accounted bytes
// (slots x the slot's bytes at its peak, plus a retained prefix).
struct ShotsStats {
  uint64_t lockstep_shots = 0, lone_shots = 0, peak_batch_bytes = 0, batch_size = 0;
};
ShotsStats last_shots_stats();

// Result queries (result.rs). Query qubits are physical labels; a null list
// means every qubit, highest label first.
std::vector<std::vector<uint64_t>> block_qubits(const Result& r);
std::string diagnostics_json(const Result& r, bool with_total_time);
bool profiled(const Result& r);
uint64_t statevector_len(const Result& r);  // 2^n; MemErr beyond the budget
void statevector(const Result& r, void* out_c128);
void amplitude(const Result& r, const std::string& bits, double* out2);
std::vector<std::pair<uint64_t, double>> probabilities(const Result& r, const std::vector<uint64_t>* qubits);
Counts counts(const Result& r, uint64_t shots, const std::vector<uint64_t>* qubits, uint64_t seed);
double expectation_value(const Result& r, const std::string& pauli);

}  // namespace pb
}  // namespace dq

// Device structures and launcher of the single-launch batched-circuit kernel.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>

namespace dq {
namespace bk {

// Largest circuit held whole in one workgroup's LDS: 2^12 complex64
// amplitudes (32 KiB) plus the sampler's 256 FP64 chunk sums stay within the
// default 64 KiB LDS request.
constexpr int kBatchLdsMax = 12;

struct BatchCircuit {
  uint32_t n, nops;
  uint64_t op_off;  // first op in the op array
  uint64_t seed;
  uint32_t shots, pad;
  uint64_t out_off;  // first sample slot in the output array
};

struct BatchOp {
  uint32_t k;        // 1..5 qubits (the single-launch kernel takes 1..4)
  uint32_t mat_off;  // float offset of the 4^k complex FP32 matrix
  uint32_t q[5];     // qubits, q[0] = MSB of the local index
  uint32_t s[5];     // the same qubits sorted ascending
};

// One launch: `count` circuits, circuit list[i] on workgroup i; LDS sized
// for states of lds_n qubits (every circuit of the launch has at most that).
struct LdsLaunch {
  const BatchCircuit* circ;
  const uint32_t* list;
  uint32_t count;
  int lds_n;
  const BatchOp* ops;
  const float* mats;
  uint32_t* out;                // sampled outcomes, at each circuit's out_off
  uint64_t* times;              // per circuit: device wall clock at start and finish
  float2* dump;                 // final states (null: not kept)
  const uint64_t* dump_off;     // amplitude offset of each circuit's kept state
};

// Dynamic LDS bytes of a launch for states of n qubits.
size_t lds_bytes(int n);
// hipFuncGetAttributes of the kernel (private_bytes: per-lane private
// segment); returns the hipError_t.
int func_info(int* max_dynamic_lds, int* static_lds, int* max_threads, int* private_bytes);
// hipOccupancyMaxActiveBlocksPerMultiprocessor at `lds` dynamic LDS bytes.
int occupancy(size_t lds, int* blocks_per_cu);
// Enqueue the launch on `st` (graph-capturable: one kernel node, no memset).
void launch_lds(const LdsLaunch& a, hipStream_t st);

}  // namespace bk
}  // namespace dq

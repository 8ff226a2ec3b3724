// Device structures and launch-limit probes of the (disabled) persistent batched-circuit kernel.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

namespace dq {
namespace bk {

struct BatchCircuit {
  uint32_t n, nops;
  uint64_t op_off;  // first op in the op array
  uint64_t seed;
  uint32_t shots, pad;
  uint64_t out_off;  // first sample slot in the output array
};

struct BatchOp {
  uint32_t k;        // 1..5 qubits
  uint32_t mat_off;  // float offset of the 4^k complex FP32 matrix
  uint32_t q[5];     // qubits, q[0] = MSB of the local index
  uint32_t s[5];     // the same qubits sorted ascending
};

size_t lds_bytes(int lmax, int block);
// hipFuncSetAttribute(MaxDynamicSharedMemorySize); returns the hipError_t.
int set_lds_limit(int block, size_t bytes);
// hipFuncGetAttributes of the kernel instance for `block` (private_bytes: the
// per-lane private segment, its call stack included); returns hipError_t.
int func_info(int block, int* max_dynamic_lds, int* static_lds, int* max_threads, int* private_bytes);
// hipOccupancyMaxActiveBlocksPerMultiprocessor at `lds` dynamic LDS bytes.
int occupancy(int block, size_t lds, int* blocks_per_cu);

}  // namespace bk
}  // namespace dq

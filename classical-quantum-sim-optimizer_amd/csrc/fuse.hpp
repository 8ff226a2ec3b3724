// Gate fusion: groups consecutive compiled unitary operations into dense
// blocks of at most `max_k` qubits, so that one HBM sweep of the state
// applies many gates.
//
// The reference fuses only runs of 1-qubit gates
// (src/monolithic/statevector/execution.rs:368-416); block fusion is new and
// is pinned by amplitude equivalence with the unfused plan.
#pragma once

#include <functional>
#include <vector>

#include "circuit.hpp"

namespace dq {

struct FusedBlock {
  std::vector<uint32_t> qubits;  // sorted descending: qubits[0] = MSB of the local index
  std::vector<int> ops;          // indices into the op list, execution order
  std::vector<cplx> mat;         // 4^k row-major, FP64
  bool diagonal = false;
};

// Qubits an operation touches (targets and controls).
std::vector<uint32_t> op_qubits(const Op& op);

// Apply one unitary operation to a small 2^k vector whose bit
// `local[q]` holds qubit q (reference-kernel arithmetic).
void apply_small(const Op& op, std::vector<cplx>& v, const std::vector<int>& local);

// Optional scorer used by the learned planner: returns the predicted
// benefit of merging (block a into block b); merges with a score <= 0 are
// skipped. A null scorer accepts every merge that fits in max_k.
using MergeScorer = std::function<double(const FusedBlock& a, const FusedBlock& b, const std::vector<Op>& ops)>;

// Greedy fuser over ops[begin, end) (all unitary).
std::vector<FusedBlock> greedy_fuse(const std::vector<Op>& ops, size_t begin, size_t end, int max_k,
                                    const MergeScorer& scorer = nullptr);
// The same over the ops listed in `order` (a valid execution order of
// unitary ops; blocks keep the listed ops' relative order per qubit).
std::vector<FusedBlock> greedy_fuse_list(const std::vector<Op>& ops, const std::vector<int>& order, int max_k,
                                         const MergeScorer& scorer = nullptr);

// Fill block.mat and block.diagonal from its ops.
void build_block_matrix(FusedBlock& block, const std::vector<Op>& ops);

}  // namespace dq

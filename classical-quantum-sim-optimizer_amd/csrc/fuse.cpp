// Greedy block fuser. Two passes:
//  1. Sweep the ops in order. An op joins the block that last touched any of
//     its qubits (the latest such block) when the union still has at most
//     max_k qubits. No later block touches those qubits, so the op may run at
//     that block's position. Otherwise it opens a new block.
//  2. Merge pass: block i joins block j, the first later block touching any
//     of i's qubits, when the union fits. Blocks strictly between them touch
//     none of i's qubits, so running i's ops first is legal.
// The block order stays a topological order of the original op sequence, so
// each block's matrix is the ordered product of its ops.
#include "fuse.hpp"

#include <algorithm>
#include <set>

namespace dq {

std::vector<uint32_t> op_qubits(const Op& op) {
  std::vector<uint32_t> q;
  for (auto& c : op.controls) q.push_back(c.first);
  switch (op.kind) {
    case OpKind::One:
    case OpKind::X: q.push_back(op.target); break;
    case OpKind::Swap:
      q.push_back(op.a);
      q.push_back(op.b);
      break;
    case OpKind::Diagonal:
    case OpKind::Dense:
      for (uint32_t x : op.qubits) q.push_back(x);
      break;
    default: break;
  }
  return q;
}

namespace {
bool controls_hold(const Op& op, size_t i, const std::vector<int>& local) {
  for (auto& c : op.controls)
    if ((((i >> local[c.first]) & 1) != 0) != c.second) return false;
  return true;
}
}  // namespace

void apply_small(const Op& op, std::vector<cplx>& v, const std::vector<int>& local) {
  const size_t len = v.size();
  switch (op.kind) {
    case OpKind::One: {
      size_t tb = (size_t)1 << local[op.target];
      for (size_t i = 0; i < len; ++i) {
        if ((i & tb) || !controls_hold(op, i, local)) continue;
        cplx a = v[i], b = v[i | tb];
        v[i] = cadd(cmul(op.mat[0], a), cmul(op.mat[1], b));
        v[i | tb] = cadd(cmul(op.mat[2], a), cmul(op.mat[3], b));
      }
      break;
    }
    case OpKind::X: {
      size_t tb = (size_t)1 << local[op.target];
      for (size_t i = 0; i < len; ++i)
        if (!(i & tb) && controls_hold(op, i, local)) std::swap(v[i], v[i | tb]);
      break;
    }
    case OpKind::Swap: {
      size_t ab = (size_t)1 << local[op.a], bb = (size_t)1 << local[op.b];
      for (size_t i = 0; i < len; ++i)
        if ((i & ab) && !(i & bb) && controls_hold(op, i, local)) std::swap(v[i], v[(i & ~ab) | bb]);
      break;
    }
    case OpKind::Diagonal: {
      for (size_t i = 0; i < len; ++i) {
        size_t j = 0;
        for (uint32_t q : op.qubits) j = (j << 1) | ((i >> local[q]) & 1);
        v[i] = cmul(v[i], op.mat[j]);
      }
      break;
    }
    case OpKind::Dense: {
      const size_t k = op.qubits.size(), dim = (size_t)1 << k;
      size_t tmask = 0;
      for (uint32_t q : op.qubits) tmask |= (size_t)1 << local[q];
      std::vector<cplx> vals(dim), res(dim);
      for (size_t base = 0; base < len; ++base) {
        if (base & tmask) continue;
        std::vector<size_t> idx(dim);
        for (size_t j = 0; j < dim; ++j) {
          size_t off = 0;
          for (size_t bit = 0; bit < k; ++bit)
            if ((j >> (k - 1 - bit)) & 1) off |= (size_t)1 << local[op.qubits[bit]];
          idx[j] = base | off;
          vals[j] = v[idx[j]];
        }
        for (size_t r = 0; r < dim; ++r) {
          cplx acc(0, 0);
          for (size_t c = 0; c < dim; ++c) acc = cadd(acc, cmul(op.mat[r * dim + c], vals[c]));
          res[r] = acc;
        }
        for (size_t j = 0; j < dim; ++j) v[idx[j]] = res[j];
      }
      break;
    }
    default: break;
  }
}

std::vector<FusedBlock> greedy_fuse(const std::vector<Op>& ops, size_t begin, size_t end, int max_k,
                                    const MergeScorer& scorer) {
  std::vector<int> order;
  for (size_t i = begin; i < end; ++i) order.push_back((int)i);
  return greedy_fuse_list(ops, order, max_k, scorer);
}

std::vector<FusedBlock> greedy_fuse_list(const std::vector<Op>& ops, const std::vector<int>& order, int max_k,
                                         const MergeScorer& scorer) {
  std::vector<FusedBlock> blocks;
  std::vector<std::set<uint32_t>> qsets;
  std::vector<int> last(64, -1);
  for (int i : order) {
    std::vector<uint32_t> q = op_qubits(ops[i]);
    int t = -1;
    for (uint32_t x : q) t = std::max(t, last[x]);
    bool joined = false;
    if (t >= 0) {
      std::set<uint32_t> u = qsets[t];
      u.insert(q.begin(), q.end());
      if ((int)u.size() <= max_k) {
        bool ok = true;
        if (scorer) {
          FusedBlock single;
          single.qubits.assign(q.begin(), q.end());
          single.ops = {(int)i};
          FusedBlock tb;
          tb.qubits.assign(qsets[t].begin(), qsets[t].end());
          tb.ops = blocks[t].ops;
          ok = scorer(tb, single, ops) > 0.0;
        }
        if (ok) {
          qsets[t] = u;
          blocks[t].ops.push_back((int)i);
          joined = true;
        }
      }
    }
    if (!joined) {
      t = (int)blocks.size();
      blocks.emplace_back();
      blocks.back().ops.push_back((int)i);
      qsets.emplace_back(q.begin(), q.end());
    }
    for (uint32_t x : q) last[x] = t;
  }

  // Merge pass, repeated until stable.
  bool changed = true;
  std::vector<char> alive(blocks.size(), 1);
  while (changed) {
    changed = false;
    for (size_t i = 0; i < blocks.size(); ++i) {
      if (!alive[i]) continue;
      size_t j = blocks.size();
      for (size_t c = i + 1; c < blocks.size() && j == blocks.size(); ++c) {
        if (!alive[c]) continue;
        for (uint32_t x : qsets[i])
          if (qsets[c].count(x)) {
            j = c;
            break;
          }
      }
      if (j == blocks.size()) continue;
      std::set<uint32_t> u = qsets[i];
      u.insert(qsets[j].begin(), qsets[j].end());
      if ((int)u.size() > max_k) continue;
      if (scorer) {
        FusedBlock a, b;
        a.qubits.assign(qsets[i].begin(), qsets[i].end());
        a.ops = blocks[i].ops;
        b.qubits.assign(qsets[j].begin(), qsets[j].end());
        b.ops = blocks[j].ops;
        if (scorer(a, b, ops) <= 0.0) continue;
      }
      std::vector<int> merged = blocks[i].ops;
      merged.insert(merged.end(), blocks[j].ops.begin(), blocks[j].ops.end());
      blocks[j].ops = std::move(merged);
      qsets[j] = u;
      alive[i] = 0;
      changed = true;
    }
  }
  std::vector<FusedBlock> out;
  for (size_t i = 0; i < blocks.size(); ++i) {
    if (!alive[i]) continue;
    FusedBlock b = std::move(blocks[i]);
    b.qubits.assign(qsets[i].rbegin(), qsets[i].rend());  // descending
    out.push_back(std::move(b));
  }
  return out;
}

void build_block_matrix(FusedBlock& block, const std::vector<Op>& ops) {
  const int k = (int)block.qubits.size();
  const size_t dim = (size_t)1 << k;
  std::vector<int> local(64, -1);
  for (int i = 0; i < k; ++i) local[block.qubits[i]] = k - 1 - i;
  block.mat.assign(dim * dim, cplx(0, 0));
  for (size_t c = 0; c < dim; ++c) {
    std::vector<cplx> col(dim, cplx(0, 0));
    col[c] = cplx(1, 0);
    for (int oi : block.ops) apply_small(ops[oi], col, local);
    for (size_t r = 0; r < dim; ++r) block.mat[r * dim + c] = col[r];
  }
  block.diagonal = true;
  for (size_t r = 0; r < dim && block.diagonal; ++r)
    for (size_t c = 0; c < dim; ++c)
      if (r != c && (block.mat[r * dim + c].re != 0.0 || block.mat[r * dim + c].im != 0.0)) {
        block.diagonal = false;
        break;
      }
}

}  // namespace dq

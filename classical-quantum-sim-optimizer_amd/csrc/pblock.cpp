// P-block backend (reference src/distributed/pblock/): a product of dense
// complex128 blocks in GPU memory. Gates, measurements and marginals run the
// statevector kernels on one block's buffer (k::one, k::xgate, k::swap,
// k::diag, k::dense, the FP64 reductions, k::meas_decide, k::collapse);
// merges, splits, the dense export, Pauli expectations and product
// sampling have kernels of their own (k::pb_*). The host keeps what the
// reference keeps: wire -> (block, local bit), wires in a known basis state,
// the free-slot stack, statistics and the byte budget, with the reference's
// rules for when blocks merge, split or skip a controlled gate.
//
//   Options / budget / errors ...... simulator.rs:45-115, model.rs:60-131, result.rs:11-20
//   compile ........................ execution.rs:57-120
//   Pool (merge, gates, measure) ... model.rs:84-556
//   amplitude / sampling tables .... model.rs:557-617
//   simulate / simulate_shots ...... simulator.rs:106-322
//   queries ........................ result.rs:21-293
#include "pblock.hpp"

#include <hip/hip_runtime_api.h>
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <optional>
#include <random>
#include <sstream>
#include <thread>
#include <unordered_map>

#include "circuit.hpp"
#include "kernels.hpp"

namespace dq {
namespace pb {
namespace {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw DeviceErr(std::string(what) + ": " + hipGetErrorString(e));
}
#define PBCHK(x) hip_check((x), #x)

// Pool errors: budget and size messages are MemoryError, others RuntimeError
// (result.rs:11-20).
[[noreturn]] void fail(const std::string& m) {
  if (m.find("memory") != std::string::npos || m.find("allocation") != std::string::npos ||
      m.find("max_block_qubits") != std::string::npos)
    throw MemErr(m);
  throw RuntimeErr(m);
}

uint64_t nbytes(uint64_t n, uint64_t size) {
  if (size && n > (uint64_t)INT64_MAX / size) fail("P-block allocation size overflow");
  return n * size;
}
uint64_t dimension(uint64_t n) {
  if (n >= 64) fail("P-block dense allocation dimension exceeds addressable memory");
  const uint64_t d = 1ull << n;
  nbytes(d, 16);
  return d;
}
uint64_t add_checked(uint64_t a, uint64_t b) {
  if (a > UINT64_MAX - b) fail("P-block allocation size overflow");
  return a + b;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

uint64_t random_seed() {
  std::random_device rd;
  return ((uint64_t)rd() << 32) ^ (uint64_t)rd();
}

struct Stream {
  hipStream_t s = nullptr;
  int dev = 0;
  Stream() { s = acquire_stream(&dev); }
  ~Stream() {
    if (s) release_stream(s, dev);
  }
  Stream(const Stream&) = delete;
  Stream& operator=(const Stream&) = delete;
};

k::Expand expand(std::vector<uint32_t> pos, uint64_t fixed) {
  std::sort(pos.begin(), pos.end());
  k::Expand e;
  e.nbits = (uint32_t)pos.size();
  for (size_t i = 0; i < pos.size(); ++i) e.pos[i] = pos[i];
  e.fixed = fixed;
  return e;
}

}  // namespace

void Stats::add(const Stats& o) {
  merge_calls += o.merge_calls;
  merge_time += o.merge_time;
  merge_output_amplitudes =
      o.merge_output_amplitudes > UINT64_MAX - merge_output_amplitudes ? UINT64_MAX
                                                                        : merge_output_amplitudes + o.merge_output_amplitudes;
  split_calls += o.split_calls;
  separation_checks += o.separation_checks;
  separation_time += o.separation_time;
  separation_residual += o.separation_residual;
  skipped_controlled_gates += o.skipped_controlled_gates;
  peak_block_qubits = std::max(peak_block_qubits, o.peak_block_qubits);
  peak_live_amplitudes = std::max(peak_live_amplitudes, o.peak_live_amplitudes);
  peak_working_bytes = std::max(peak_working_bytes, o.peak_working_bytes);
}

void check_options(const Options& o) {
  if (o.memory == 0 || o.memory > (uint64_t)INT64_MAX)
    throw ValueErr("max_memory_mb must be positive and representable");
  if (o.has_max_block && (o.max_block_qubits == 0 || o.max_block_qubits >= 64))
    throw ValueErr(
        "max_block_qubits must be positive and smaller than the machine word width, or None");
  if (o.max_parallel_shots == 0) throw ValueErr("max_parallel_shots must be positive");
  if (o.max_split_qubits == 0 || o.max_split_qubits >= 64)
    throw ValueErr("max_split_qubits must be positive and smaller than the machine word width");
  if (!std::isfinite(o.separation_tolerance) || o.separation_tolerance < 0.0 || o.separation_tolerance >= 1.0)
    throw ValueErr("separation_tolerance must be finite and in [0,1)");
}

// ---------------------------------------------------------------------------
// Compile (execution.rs:57-120): single-qubit gates accumulate per wire;
// any other instruction first emits every pending product (ascending wire,
// identities dropped). Barriers and remote_barrier are no-ops.
// ---------------------------------------------------------------------------
namespace {

enum class PK { Gate, Measure, Reset, Conditional, Noop };
struct POp {
  PK kind = PK::Noop;
  Op gate;  // PK::Gate: a compiled unitary on wire indices
  uint64_t q = 0, c = 0;
  Condition cond;
  std::shared_ptr<POp> inner;
  bool unitary() const { return kind == PK::Gate || kind == PK::Noop; }
  bool terminal() const { return kind == PK::Measure || kind == PK::Noop; }
};

bool is_remote_barrier(const Instruction& in) {
  if (in.kind != Kind::Gate) return false;
  std::string n = in.name;
  for (auto& ch : n) ch = (char)std::tolower((unsigned char)ch);
  return n == "remote_barrier";
}

POp compile_one_op(const Instruction& in) {
  POp p;
  switch (in.kind) {
    case Kind::Measure:
      p.kind = PK::Measure;
      p.q = in.qubits[0];
      p.c = in.cbit;
      return p;
    case Kind::Reset:
      p.kind = PK::Reset;
      p.q = in.qubits[0];
      return p;
    case Kind::Conditional:
      p.kind = PK::Conditional;
      p.cond = in.cond;
      p.inner = std::make_shared<POp>(compile_one_op(*in.op));
      return p;
    case Kind::Barrier: return p;
    default: break;
  }
  if (is_remote_barrier(in)) return p;
  Op op;
  if (!compile_instruction(in, op)) return p;
  p.kind = PK::Gate;
  p.gate = std::move(op);
  return p;
}

std::vector<POp> compile_ops(const std::vector<Instruction>& insts) {
  const Mat2 ident = {cplx(1, 0), cplx(0, 0), cplx(0, 0), cplx(1, 0)};
  std::map<uint32_t, Mat2> pending;
  std::vector<POp> out;
  auto flush_all = [&]() {
    for (auto& kv : pending) {
      bool is_ident = true;
      for (int i = 0; i < 4; ++i) is_ident = is_ident && ceq(kv.second[i], ident[i]);
      if (is_ident) continue;
      POp p;
      p.kind = PK::Gate;
      p.gate.kind = OpKind::One;
      p.gate.target = kv.first;
      p.gate.mat.assign(kv.second.begin(), kv.second.end());
      out.push_back(std::move(p));
    }
    pending.clear();
  };
  for (const Instruction& in : insts) {
    uint32_t q;
    Mat2 m;
    if (one_qubit_matrix(in, q, m)) {
      auto it = pending.find(q);
      const Mat2 old = it == pending.end() ? ident : it->second;
      Mat2 prod;
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
          cplx acc(0, 0);
          for (int kk = 0; kk < 2; ++kk) acc = cadd(acc, cmul(m[i * 2 + kk], old[kk * 2 + j]));
          prod[i * 2 + j] = acc;
        }
      pending[q] = prod;
      continue;
    }
    flush_all();
    POp p = compile_one_op(in);
    if (p.kind != PK::Noop) out.push_back(std::move(p));
  }
  flush_all();
  return out;
}

// ---------------------------------------------------------------------------
// Input (input.rs:30-53): the quantum registers define the wires (sorted
// ranges, flattened); instructions are renumbered onto wires.
// ---------------------------------------------------------------------------
struct Input {
  std::vector<uint64_t> physical;
  std::vector<Instruction> instructions;
  uint64_t num_cbits = 0;
};

void remap(Instruction& in, const std::unordered_map<uint64_t, uint64_t>& wire) {
  for (auto& q : in.qubits) q = wire.at(q);
  if (in.op) {
    auto copy = std::make_shared<Instruction>(*in.op);
    remap(*copy, wire);
    in.op = copy;
  }
}

Input prepare(const char* json, size_t len, uint64_t budget) {
  Circuit c = parse_circuit(json, len);
  validate_circuit_rules(c);
  const auto ranges = quantum_ranges(c);
  uint64_t n = 0;
  for (auto& r : ranges) {
    if (n > UINT64_MAX - (r.second - r.first)) throw ValueErr("Qubit count overflow");
    n += r.second - r.first;
  }
  if (n > UINT64_MAX / 256 || n * 256 > budget) throw MemErr("P-block wire metadata exceeds max_memory_mb");
  Input in;
  std::unordered_map<uint64_t, uint64_t> wire;
  for (auto& r : ranges)
    for (uint64_t q = r.first; q < r.second; ++q) {
      wire[q] = in.physical.size();
      in.physical.push_back(q);
    }
  in.num_cbits = c.num_cbits();
  in.instructions = std::move(c.instructions);
  for (auto& x : in.instructions) remap(x, wire);
  return in;
}

std::string format_cbits(const std::map<uint64_t, int>& cb, uint64_t nc) {
  std::string s(nc, '0');
  for (auto& kv : cb)
    if (kv.first < nc && kv.second) s[nc - 1 - kv.first] = '1';
  return s;
}

}  // namespace

// ---------------------------------------------------------------------------
// Block pool (model.rs:50-617)
// ---------------------------------------------------------------------------
struct Block {
  std::shared_ptr<DeviceMem> mem;  // owner (the initial blocks share one slab)
  double* amps = nullptr;          // 2 * size doubles on the device
  uint64_t size = 0;
  std::vector<uint64_t> qubits;    // local bit i holds wire qubits[i]
  uint64_t nq() const { return qubits.size(); }
};

class Pool {
 public:
  std::vector<std::unique_ptr<Block>> blocks;  // null: free slot
  std::vector<std::pair<uint64_t, uint64_t>> positions;  // wire -> (block, local bit)
  std::vector<int8_t> known;                   // -1 unknown, else the basis bit
  std::vector<uint64_t> free;
  Options opt;
  Stats stats;
  uint64_t live = 0;
  hipStream_t st = nullptr;
  DeviceMem rng, ms, cb, err, red_a, red_b;
  std::vector<std::shared_ptr<DeviceMem>> mats;  // dense matrices of the compiled ops
  std::map<const Op*, const void*> mat_of;
  // Buffers of replaced blocks that queued kernels may still read; released
  // at the next stream synchronisation.
  mutable std::vector<std::shared_ptr<DeviceMem>> graveyard;

  void sync() const {
    PBCHK(hipStreamSynchronize(st));
    graveyard.clear();
  }

  Pool(uint64_t n, const Options& o, hipStream_t s) : opt(o), st(s) {
    if (nbytes(n, 288) > opt.memory) fail("P-block initial state exceeds max_memory_mb");
    positions.resize(n);
    known.assign(n, 0);
    rng.alloc(sizeof(k::RngState));
    ms.alloc(sizeof(k::MeasState));
    cb.alloc(8);
    err.alloc(sizeof(int));
    PBCHK(hipMemsetAsync(err.get(), 0, sizeof(int), st));
    PBCHK(hipMemsetAsync(ms.get(), 0, sizeof(k::MeasState), st));
    reset_zero();
  }
  uint64_t len() const { return positions.size(); }
  uint64_t bytes() const { return live * 16 + len() * 256; }
  uint64_t check(uint64_t extra) const {
    const uint64_t total = add_checked(bytes(), extra);
    if (total > opt.memory) fail("P-block working allocation exceeds max_memory_mb");
    return total;
  }
  void reserve(uint64_t extra) { stats.peak_working_bytes = std::max(stats.peak_working_bytes, check(extra)); }
  void record() {
    stats.peak_live_amplitudes = std::max(stats.peak_live_amplitudes, live);
    stats.peak_working_bytes = std::max(stats.peak_working_bytes, bytes());
  }
  void seed_rng(uint64_t seed, uint64_t stream) { k::rng_init(rng.as<k::RngState>(), 1, seed, stream, st); }

  static std::unique_ptr<Block> new_block(uint64_t size, std::vector<uint64_t> qubits) {
    auto b = std::make_unique<Block>();
    b->mem = std::make_shared<DeviceMem>(16 * size);
    b->amps = b->mem->as<double>();
    b->size = size;
    b->qubits = std::move(qubits);
    return b;
  }
  // A single-wire block holding (v4[0] + i v4[1], v4[2] + i v4[3]), written
  // by a kernel (no host copy to wait for).
  std::unique_ptr<Block> basis_block(uint64_t q, const double* v4) {
    auto b = new_block(2, {q});
    k::pb_set2(b->amps, v4[0], v4[1], v4[2], v4[3], st);
    return b;
  }

  // Every wire its own block |0> (model.rs:184-199), in one device slab.
  void reset_zero() {
    const uint64_t n = len();
    blocks.clear();
    free.clear();
    auto slab = std::make_shared<DeviceMem>(32 * std::max<uint64_t>(n, 1));
    std::vector<double> init(4 * std::max<uint64_t>(n, 1), 0.0);
    for (uint64_t q = 0; q < n; ++q) init[4 * q] = 1.0;
    PBCHK(hipMemcpyAsync(slab->get(), init.data(), 32 * std::max<uint64_t>(n, 1), hipMemcpyHostToDevice, st));
    sync();
    for (uint64_t q = 0; q < n; ++q) {
      auto b = std::make_unique<Block>();
      b->mem = slab;
      b->amps = slab->as<double>() + 4 * q;
      b->size = 2;
      b->qubits = {q};
      blocks.push_back(std::move(b));
      positions[q] = {q, 0};
      known[q] = 0;
    }
    live = 2 * n;
    stats = Stats();
    stats.peak_block_qubits = n > 0 ? 1 : 0;
    record();
  }

  // A copy with its own device buffers (trajectories restart from it).
  std::unique_ptr<Pool> clone() const {
    auto p = std::unique_ptr<Pool>(new Pool(*this, 0));
    return p;
  }
  // Restore this pool to `src` (model.rs:149-183).
  void restore(const Pool& src) {
    bool same = blocks.size() == src.blocks.size();
    for (size_t i = 0; same && i < blocks.size(); ++i) {
      const bool a = (bool)blocks[i], b = (bool)src.blocks[i];
      same = a == b && (!a || blocks[i]->qubits == src.blocks[i]->qubits);
    }
    if (same) {
      for (size_t i = 0; i < blocks.size(); ++i)
        if (blocks[i])
          PBCHK(hipMemcpyAsync(blocks[i]->amps, src.blocks[i]->amps, 16 * blocks[i]->size, hipMemcpyDeviceToDevice,
                               st));
    } else {
      copy_blocks(src);
    }
    positions = src.positions;
    known = src.known;
    free = src.free;
    live = src.live;
    stats = Stats();
    for (auto& b : src.blocks)
      if (b) stats.peak_block_qubits = std::max(stats.peak_block_qubits, b->nq());
    record();
  }

  void read2(const Block& b, cplx& a0, cplx& a1) const {
    double v[4];
    PBCHK(hipMemcpyAsync(v, b.amps, 32, hipMemcpyDeviceToHost, st));
    sync();
    a0 = cplx(v[0], v[1]);
    a1 = cplx(v[2], v[3]);
  }

  // model.rs:199-244
  uint64_t merge(uint64_t a, uint64_t b) {
    if (a == b) return a;
    const uint64_t n = blocks[a]->nq() + blocks[b]->nq();
    if (opt.has_max_block && n > opt.max_block_qubits)
      fail("P-block merge exceeds max_block_qubits; no entanglement was discarded");
    const uint64_t dim = dimension(n);
    reserve(nbytes(dim, 16));
    const double t0 = opt.profile ? now_s() : 0.0;
    const Block& left = *blocks[a];
    const Block& right = *blocks[b];
    std::vector<uint64_t> qubits = left.qubits;
    qubits.insert(qubits.end(), right.qubits.begin(), right.qubits.end());
    auto out = new_block(dim, qubits);
    k::pb_kron(out->amps, left.amps, left.size, right.amps, right.size, st);
    live = live - left.size - right.size + dim;
    for (size_t i = 0; i < qubits.size(); ++i) positions[qubits[i]] = {a, i};
    graveyard.push_back(blocks[a]->mem);
    graveyard.push_back(blocks[b]->mem);
    blocks[a] = std::move(out);
    blocks[b].reset();
    free.push_back(b);
    stats.peak_block_qubits = std::max(stats.peak_block_qubits, n);
    stats.merge_calls += 1;
    stats.merge_output_amplitudes =
        stats.merge_output_amplitudes > UINT64_MAX - dim ? UINT64_MAX : stats.merge_output_amplitudes + dim;
    if (opt.profile) {
      sync();
      stats.merge_time += now_s() - t0;
    }
    record();
    return a;
  }

  // model.rs:245-255: merge into the smallest block holding any of qs.
  uint64_t ensure(const std::vector<uint64_t>& qs) {
    std::vector<uint64_t> bl;
    for (uint64_t q : qs) bl.push_back(positions[q].first);
    std::sort(bl.begin(), bl.end());
    bl.erase(std::unique(bl.begin(), bl.end()), bl.end());
    std::stable_sort(bl.begin(), bl.end(), [&](uint64_t x, uint64_t y) { return blocks[x]->nq() < blocks[y]->nq(); });
    const uint64_t first = bl[0];
    for (size_t i = 1; i < bl.size(); ++i) merge(first, bl[i]);
    return first;
  }

  // model.rs:256-271
  void refresh_known(uint64_t q) {
    const Block& b = *blocks[positions[q].first];
    if (b.nq() != 1) {
      known[q] = -1;
      return;
    }
    cplx a0, a1;
    read2(b, a0, a1);
    const cplx zero(0, 0);
    known[q] = ceq(a1, zero) ? 0 : ceq(a0, zero) ? 1 : -1;
  }

  // model.rs:272-288: known controls that do not hold skip the gate.
  std::optional<std::vector<std::pair<uint64_t, bool>>> controls(const std::vector<std::pair<uint64_t, bool>>& cs) {
    std::vector<std::pair<uint64_t, bool>> active;
    for (auto& c : cs) {
      const int8_t v = known[c.first];
      if (v >= 0) {
        if ((v != 0) != c.second) {
          stats.skipped_controlled_gates += 1;
          return std::nullopt;
        }
      } else {
        active.push_back(c);
      }
    }
    return active;
  }

  k::Expand local_expand(const std::vector<std::pair<uint64_t, bool>>& cs, std::vector<uint32_t> extra) {
    uint64_t fixed = 0;
    std::vector<uint32_t> pos = std::move(extra);
    for (auto& c : cs) {
      const uint32_t l = (uint32_t)positions[c.first].second;
      pos.push_back(l);
      if (c.second) fixed |= 1ull << l;
    }
    return expand(pos, fixed);
  }

  // model.rs:289-330
  void one(uint64_t q, const std::vector<cplx>& m, const std::vector<std::pair<uint64_t, bool>>& cs) {
    auto active = controls(cs);
    if (!active) return;
    if (!active->empty()) {
      const Block& b = *blocks[positions[q].first];
      if (b.nq() == 1) {
        cplx a, z;
        read2(b, a, z);
        if (ceq(cadd(cmul(m[0], a), cmul(m[1], z)), a) && ceq(cadd(cmul(m[2], a), cmul(m[3], z)), z)) {
          stats.skipped_controlled_gates += 1;
          return;
        }
      }
    }
    std::vector<uint64_t> qs;
    for (auto& c : *active) qs.push_back(c.first);
    qs.push_back(q);
    const uint64_t idx = ensure(qs);
    Block& b = *blocks[idx];
    const uint32_t t = (uint32_t)positions[q].second;
    const k::Expand e = local_expand(*active, {t});
    // An exact X matrix moves amplitudes (engine::apply_x), as in the reference.
    const Mat2 xm = gates::X();
    bool is_x = true;
    for (int i = 0; i < 4; ++i) is_x = is_x && ceq(m[i], xm[i]);
    if (is_x) {
      k::xgate(16, b.amps, b.size, 1, (int)b.nq(), e, t, k::Flag{}, st);
    } else {
      double u[8];
      for (int i = 0; i < 4; ++i) {
        u[2 * i] = m[i].re;
        u[2 * i + 1] = m[i].im;
      }
      k::one(16, b.amps, b.size, 1, (int)b.nq(), e, t, u, k::Flag{}, st);
    }
    for (uint64_t w : qs) refresh_known(w);
    maybe_split(qs);
  }

  // model.rs:331-339: a swap without active controls only relabels wires.
  void relabel_swap(uint64_t a, uint64_t b) {
    const auto pa = positions[a], pbp = positions[b];
    blocks[pa.first]->qubits[pa.second] = b;
    blocks[pbp.first]->qubits[pbp.second] = a;
    std::swap(positions[a], positions[b]);
    std::swap(known[a], known[b]);
  }

  // model.rs:340-443
  void gate(const Op& op) {
    std::vector<std::pair<uint64_t, bool>> cs;
    for (auto& c : op.controls) cs.push_back({c.first, c.second});
    switch (op.kind) {
      case OpKind::One: one(op.target, op.mat, cs); break;
      case OpKind::X: {
        const Mat2 x = gates::X();
        one(op.target, std::vector<cplx>(x.begin(), x.end()), cs);
        break;
      }
      case OpKind::Swap: {
        auto active = controls(cs);
        if (!active) return;
        if (active->empty()) {
          relabel_swap(op.a, op.b);
          return;
        }
        std::vector<uint64_t> qs;
        for (auto& c : *active) qs.push_back(c.first);
        qs.push_back(op.a);
        qs.push_back(op.b);
        const uint64_t idx = ensure(qs);
        Block& b = *blocks[idx];
        const uint32_t la = (uint32_t)positions[op.a].second, lb = (uint32_t)positions[op.b].second;
        k::swap(16, b.amps, b.size, 1, (int)b.nq(), local_expand(*active, {la, lb}), la, lb, k::Flag{}, st);
        for (uint64_t w : qs) refresh_known(w);
        maybe_split(qs);
        break;
      }
      case OpKind::Diagonal: {
        const std::vector<uint32_t>& qs = op.qubits;
        std::vector<uint64_t> active;
        for (uint32_t q : qs)
          if (known[q] < 0) active.push_back(q);
        std::vector<cplx> reduced;
        for (uint64_t basis = 0; basis < (1ull << active.size()); ++basis) {
          uint64_t index = 0, cursor = 0;
          for (uint32_t q : qs) {
            uint64_t bit;
            if (known[q] >= 0) bit = (uint64_t)known[q];
            else bit = (basis >> (active.size() - 1 - cursor++)) & 1;
            index = (index << 1) | bit;
          }
          reduced.push_back(op.mat[index]);
        }
        bool uniform = true;
        for (auto& v : reduced) uniform = uniform && ceq(v, reduced[0]);
        if (active.empty() || uniform) {
          Block& b = *blocks[positions[qs[0]].first];
          k::pb_scale(b.amps, b.size, reduced[0].re, reduced[0].im, st);
        } else {
          const uint64_t idx = ensure(active);
          Block& b = *blocks[idx];
          std::vector<uint32_t> local;
          for (uint64_t q : active) local.push_back((uint32_t)positions[q].second);
          const int kq = (int)local.size();
          std::vector<uint64_t> pats;
          std::vector<double> vals;
          for (int j = 0; j < (1 << kq); ++j) {
            if (reduced[j].re == 1.0 && reduced[j].im == 0.0) continue;
            uint64_t pat = 0;
            for (int i = 0; i < kq; ++i)
              if ((j >> (kq - 1 - i)) & 1) pat |= 1ull << local[i];
            pats.push_back(pat);
            vals.push_back(reduced[j].re);
            vals.push_back(reduced[j].im);
          }
          k::diag(16, b.amps, b.size, 1, (int)b.nq(), expand(local, 0), (int)pats.size(), pats.data(), vals.data(),
                  k::Flag{}, st);
          maybe_split(active);
        }
        break;
      }
      case OpKind::Dense: {
        std::vector<uint64_t> qs(op.qubits.begin(), op.qubits.end());
        const uint64_t idx = ensure(qs);
        Block& b = *blocks[idx];
        std::vector<uint32_t> local;
        for (uint64_t q : qs) local.push_back((uint32_t)positions[q].second);
        const int kq = (int)local.size();
        uint64_t offsets[32] = {};
        for (int j = 0; j < (1 << kq); ++j) {
          uint64_t off = 0;
          for (int i = 0; i < kq; ++i)
            if ((j >> (kq - 1 - i)) & 1) off |= 1ull << local[i];
          offsets[j] = off;
        }
        k::dense(16, b.amps, b.size, 1, (int)b.nq(), kq, expand(local, 0), dense_matrix(op), offsets, k::Flag{}, st);
        for (uint64_t w : qs) refresh_known(w);
        maybe_split(qs);
        break;
      }
      case OpKind::Noop: break;
      default: fail("Internal error: nonunitary operation in a P-block gate");
    }
  }

  const void* dense_matrix(const Op& op) {
    auto it = mat_of.find(&op);
    if (it != mat_of.end()) return it->second;
    auto m = std::make_shared<DeviceMem>(16 * op.mat.size());
    std::vector<double> h;
    for (auto& v : op.mat) {
      h.push_back(v.re);
      h.push_back(v.im);
    }
    PBCHK(hipMemcpyAsync(m->get(), h.data(), 8 * h.size(), hipMemcpyHostToDevice, st));
    sync();
    mats.push_back(m);
    mat_of[&op] = m->get();
    return m->get();
  }

  // model.rs:444-470: wire q leaves its block as the factor `single`.
  void install_split(uint64_t q, const double* single4, std::unique_ptr<Block> remaining) {
    const auto pos = positions[q];
    const uint64_t idx = pos.first;
    std::unique_ptr<Block> old = std::move(blocks[idx]);
    graveyard.push_back(old->mem);
    if (free.empty()) fail("Internal error: a multi-qubit block leaves no free slot");
    const uint64_t slot = free.back();
    free.pop_back();
    std::vector<uint64_t> qs = old->qubits;
    qs.erase(qs.begin() + (long)pos.second);
    live = live - old->size + remaining->size + 2;
    for (size_t i = 0; i < qs.size(); ++i) positions[qs[i]] = {idx, i};
    remaining->qubits = qs;
    blocks[idx] = std::move(remaining);
    blocks[slot] = basis_block(q, single4);
    positions[q] = {slot, 0};
    stats.split_calls += 1;
    // refresh_known(q) on the factor just written, from its host values.
    const cplx a0(single4[0], single4[1]), a1(single4[2], single4[3]), zero(0, 0);
    known[q] = ceq(a1, zero) ? 0 : ceq(a0, zero) ? 1 : -1;
    for (uint64_t w : qs) refresh_known(w);
    record();
  }

  // FP64 |a|^2 split by local bit `branch` (-1: none): device double2
  // (p0, p1) after the reduction chain of the statevector kernels.
  const void* reduce_prob(const Block& b, int branch) {
    const int n = (int)b.nq();
    const int seg1 = n < 11 ? n : 11;
    const uint64_t c1 = 1ull << (n - seg1);
    if (red_a.size() < 16 * c1) red_a.alloc(16 * c1);
    if (red_b.size() < 16 * c1) red_b.alloc(16 * c1);
    k::ProbSpec ps;
    ps.branch = branch;
    k::seg_reduce_prob(16, b.amps, b.size, 1, n, ps, seg1, red_a.get(), c1, st);
    uint64_t count = c1;
    void* src = red_a.get();
    void* dst = red_b.get();
    while (count > 1) {
      int lg = 0;
      while ((1ull << lg) < count) ++lg;
      const int s = lg < 11 ? lg : 11;
      k::seg_reduce_f64(src, c1, count, 1, s, dst, c1, st);
      count >>= s;
      std::swap(src, dst);
    }
    return src;
  }

  // model.rs:475-496 with engine::measure_qubit's rule (the statevector
  // measurement kernels; Philox draw from this pool's stream).
  int measure(uint64_t q) {
    const auto pos = positions[q];
    Block& b = *blocks[pos.first];
    const uint64_t size = b.size;
    if (size > 2) reserve(nbytes(size / 2 + 2, 16));
    const void* tot = reduce_prob(b, (int)pos.second);
    k::meas_decide(tot, 0, 1, ms.as<k::MeasState>(), rng.as<k::RngState>(), cb.as<uint8_t>(), 1, -1, k::Flag{},
                   err.as<int>(), st);
    k::collapse(16, b.amps, b.size, 1, (int)b.nq(), (uint32_t)pos.second, ms.as<k::MeasState>(), st);
    k::MeasState m;
    int e = 0;
    PBCHK(hipMemcpyAsync(&m, ms.get(), sizeof m, hipMemcpyDeviceToHost, st));
    PBCHK(hipMemcpyAsync(&e, err.get(), sizeof e, hipMemcpyDeviceToHost, st));
    sync();
    if (e & 1) throw RuntimeErr("Cannot measure a state with nonfinite or zero norm");
    const int value = m.outcome;
    if (size > 2) {
      auto rem = new_block(size / 2, {});
      k::pb_extract(rem->amps, b.amps, size / 2, (uint32_t)pos.second, (uint32_t)value, st);
      double single[4] = {0, 0, 0, 0};
      single[2 * value] = 1.0;
      install_split(q, single, std::move(rem));
    }
    known[q] = (int8_t)value;
    return value;
  }

  // model.rs:497-556
  void maybe_split(const std::vector<uint64_t>& qs) {
    if (!opt.split_separable) return;
    for (uint64_t q : qs) {
      const auto pos = positions[q];
      const Block& b = *blocks[pos.first];
      if (b.nq() < 2 || b.nq() > opt.max_split_qubits) continue;
      const uint64_t size = b.size;
      const uint64_t extra = nbytes(size / 2 + 2, 16);
      try {
        check(extra);
      } catch (const MemErr&) {
        continue;
      }
      reserve(extra);
      const double t0 = opt.profile ? now_s() : 0.0;
      stats.separation_checks += 1;
      auto rem = new_block(size / 2, {});
      DeviceMem res(16 * 10);
      k::pb_separate(b.amps, size / 2, (uint32_t)pos.second, rem->amps, res.as<double>(), st);
      double r[10];
      PBCHK(hipMemcpyAsync(r, res.get(), sizeof r, hipMemcpyDeviceToHost, st));
      sync();
      if (r[9] != 0.0) fail("P-block separation encountered an invalid norm");
      const double relative = r[4] / r[5];
      if (relative <= opt.separation_tolerance) {
        k::pb_scale_real(rem->amps, rem->size, std::sqrt(r[5] / r[6]), st);
        install_split(q, r, std::move(rem));
        stats.separation_residual += relative;
      }
      if (opt.profile) stats.separation_time += now_s() - t0;
    }
  }

  // Per-block CDF tables for sampling (model.rs:578-605); returns device
  // tables (CDF arrays kept alive in `keep`).
  std::vector<k::PbTable> sampling_tables(std::vector<std::shared_ptr<DeviceMem>>& keep) {
    check(add_checked(nbytes(live, 8), nbytes(len(), 64)));
    std::vector<k::PbTable> tabs;
    std::vector<const Block*> live_blocks;
    for (auto& b : blocks)
      if (b) live_blocks.push_back(b.get());
    DeviceMem totals(8 * std::max<size_t>(live_blocks.size(), 1));
    DeviceMem errs(4);
    PBCHK(hipMemsetAsync(errs.get(), 0, 4, st));
    // Blocks of at most 1024 amplitudes (one CDF group): one slab, one launch.
    uint64_t small_len = 0;
    for (const Block* b : live_blocks)
      if (b->size <= 1024) small_len += b->size;
    std::shared_ptr<DeviceMem> slab;
    if (small_len > 0) {
      slab = std::make_shared<DeviceMem>(8 * small_len);
      keep.push_back(slab);
    }
    std::vector<k::PbSmallTable> small;
    tabs.resize(live_blocks.size());
    uint64_t off = 0;
    for (size_t t = 0; t < live_blocks.size(); ++t) {
      const Block& b = *live_blocks[t];
      if (b.size <= 1024) {
        double* cdf = slab->as<double>() + off;
        off += b.size;
        small.push_back({b.amps, cdf, b.size, t});
        tabs[t] = {cdf, b.size, (uint32_t)b.nq(), 0};
        continue;
      }
      // |a|^2 then its CDF in the same buffer (the block scan reads each
      // 2048-entry block into LDS before it writes the block back), so the
      // table stays within the 8 bytes per amplitude the budget counts.
      auto cdf = std::make_shared<DeviceMem>(8 * b.size);
      auto gt = std::make_shared<DeviceMem>(8 * k::cdf_tmp_entries(b.size));
      k::pb_norms(cdf->as<double>(), b.amps, b.size, st);
      k::cdf_build(cdf->as<double>(), 1, b.size, cdf->as<double>(), gt->as<double>(), totals.as<double>() + t,
                   errs.as<int>(), st);
      keep.push_back(cdf);
      keep.push_back(gt);
      tabs[t] = {cdf->get(), b.size, (uint32_t)b.nq(), 0};
    }
    if (!small.empty()) {
      auto desc = std::make_shared<DeviceMem>(sizeof(k::PbSmallTable) * small.size());
      PBCHK(hipMemcpyAsync(desc->get(), small.data(), sizeof(k::PbSmallTable) * small.size(), hipMemcpyHostToDevice,
                           st));
      k::pb_small_tables(desc->as<k::PbSmallTable>(), (uint32_t)small.size(), totals.as<double>(), st);
      keep.push_back(desc);
    }
    std::vector<double> tot(live_blocks.size());
    if (!tot.empty())
      PBCHK(hipMemcpyAsync(tot.data(), totals.get(), 8 * tot.size(), hipMemcpyDeviceToHost, st));
    sync();
    for (double v : tot)
      if (!(v > 0.0) || !std::isfinite(v)) fail("P-block sampling encountered an invalid norm");
    return tabs;
  }

  // Outcome index per (shot, live block) in block order.
  std::vector<uint32_t> sample(const std::vector<k::PbTable>& tabs, uint64_t shots, uint64_t seed,
                               uint64_t stream0) {
    std::vector<uint32_t> out(shots * tabs.size());
    if (out.empty()) return out;
    DeviceMem dt(sizeof(k::PbTable) * tabs.size()), dout(4 * out.size());
    PBCHK(hipMemcpyAsync(dt.get(), tabs.data(), sizeof(k::PbTable) * tabs.size(), hipMemcpyHostToDevice, st));
    k::pb_sample(dout.as<uint32_t>(), shots, dt.as<k::PbTable>(), (uint32_t)tabs.size(), seed, stream0, st);
    PBCHK(hipMemcpyAsync(out.data(), dout.get(), 4 * out.size(), hipMemcpyDeviceToHost, st));
    sync();
    return out;
  }

  std::vector<const Block*> live_blocks() const {
    std::vector<const Block*> v;
    for (auto& b : blocks)
      if (b) v.push_back(b.get());
    return v;
  }

 private:
  Pool(const Pool& o, int) : positions(o.positions), known(o.known), free(o.free), opt(o.opt), stats(o.stats),
                             live(o.live), st(o.st), mats(o.mats), mat_of(o.mat_of) {
    rng.alloc(sizeof(k::RngState));
    ms.alloc(sizeof(k::MeasState));
    cb.alloc(8);
    err.alloc(sizeof(int));
    PBCHK(hipMemsetAsync(err.get(), 0, sizeof(int), st));
    PBCHK(hipMemsetAsync(ms.get(), 0, sizeof(k::MeasState), st));
    copy_blocks(o);
  }
  void copy_blocks(const Pool& src) {
    blocks.clear();
    for (auto& b : src.blocks) {
      if (!b) {
        blocks.emplace_back();
        continue;
      }
      auto c = new_block(b->size, b->qubits);
      PBCHK(hipMemcpyAsync(c->amps, b->amps, 16 * b->size, hipMemcpyDeviceToDevice, st));
      blocks.push_back(std::move(c));
    }
  }
};

// ---------------------------------------------------------------------------
// Lockstep trajectories. `batch` shots share one block structure: block i
// keeps batch slots of blocks[i]->size amplitudes (slot b at amps + 2 * size
// * b), so each gate, merge and measurement is one launch for every shot. The
// host decisions a lone pool takes from its state (known controls, trivial
// controlled gates, reduced diagonals, merges, extractions) are taken once
// for the batch; a shot whose own decision differs from the batch's leaves it
// (`alive` false) and is run again afterwards on a pool of its own from the
// start. Shot s draws from Philox stream s in both places, so every count is
// the one a lone trajectory gives. Slots that left keep evolving unitarily
// (or are flagged off) and are never read.
// A lockstep batch that would exceed the call's memory budget as a whole.
struct BatchTooLarge : std::exception {
  const char* what() const noexcept override { return "lockstep batch exceeds the memory budget"; }
};

class BatchPool {
 public:
  using Mask = std::vector<char>;  // per slot: the current op applies
  uint32_t B;
  std::vector<std::unique_ptr<Block>> blocks;
  std::vector<std::pair<uint64_t, uint64_t>> positions;
  std::vector<uint64_t> free;
  uint64_t live = 0;
  std::vector<int8_t> known;  // slot-major, B x wires
  Mask alive;
  std::vector<std::map<uint64_t, int>> cbits;
  Options opt;
  uint64_t total_budget;   // bytes for all B slots together
  uint64_t peak_total = 0;  // largest B x slot bytes accounted
  hipStream_t st;
  DeviceMem rng, ms, err, red_a, red_b, flags;
  std::vector<std::shared_ptr<DeviceMem>> mats;
  std::map<const Op*, const void*> mat_of;
  std::vector<std::shared_ptr<DeviceMem>> graveyard;
  std::vector<std::shared_ptr<void>> host_keep;  // sources of queued uploads
  void keep_host(std::shared_ptr<void> p) { host_keep.push_back(std::move(p)); }

  // Every slot starts as `base`; slot b is shot streams[b] (its Philox
  // stream). Each slot keeps the per-shot budget o.memory (the reference's
  // per-worker limit, whose errors the lone runs reproduce); the batch as a
  // whole keeps within `total` (BatchTooLarge otherwise: the caller retries
  // with fewer slots).
  BatchPool(const Pool& base, const std::vector<uint64_t>& streams, const Options& o, uint64_t total, uint64_t seed)
      : B((uint32_t)streams.size()), positions(base.positions), free(base.free), live(base.live),
        alive(streams.size(), 1), cbits(streams.size()), opt(o), total_budget(total), st(base.st), mats(base.mats),
        mat_of(base.mat_of) {
    account(0);
    const uint64_t n = positions.size();
    for (auto& b : base.blocks) {
      if (!b) {
        blocks.emplace_back();
        continue;
      }
      auto c = new_block(b->size, b->qubits);
      k::broadcast(16, c->amps, b->size, B, b->amps, (int)b->nq(), st);
      blocks.push_back(std::move(c));
    }
    known.resize((size_t)B * n);
    for (uint32_t s = 0; s < B; ++s) std::copy(base.known.begin(), base.known.end(), known.begin() + (size_t)s * n);
    rng.alloc(sizeof(k::RngState) * B);
    ms.alloc(sizeof(k::MeasState) * B);
    err.alloc(sizeof(int));
    flags.alloc(sizeof(int) * B);
    PBCHK(hipMemsetAsync(err.get(), 0, sizeof(int), st));
    PBCHK(hipMemsetAsync(ms.get(), 0, sizeof(k::MeasState) * B, st));
    auto init = std::make_shared<std::vector<k::RngState>>(B);
    for (uint32_t b = 0; b < B; ++b) (*init)[b] = k::RngState{seed, streams[b], 0};
    PBCHK(hipMemcpyAsync(rng.get(), init->data(), sizeof(k::RngState) * B, hipMemcpyHostToDevice, st));
    keep_host(init);
  }

  void sync() {
    PBCHK(hipStreamSynchronize(st));
    graveyard.clear();
    host_keep.clear();
  }
  uint64_t wires() const { return positions.size(); }
  int8_t& kn(uint32_t s, uint64_t q) { return known[(size_t)s * wires() + q]; }
  uint64_t bytes() const { return live * 16 + wires() * 256; }
  // The whole batch's bytes with `extra` more per slot; BatchTooLarge beyond
  // total_budget.
  void account(uint64_t extra) {
    const uint64_t slot = add_checked(bytes(), extra);
    if (slot > UINT64_MAX / B || slot * B > total_budget) throw BatchTooLarge();
    peak_total = std::max(peak_total, slot * B);
  }
  void reserve(uint64_t extra) {
    if (add_checked(bytes(), extra) > opt.memory) fail("P-block working allocation exceeds max_memory_mb");
    account(extra);
  }
  std::unique_ptr<Block> new_block(uint64_t size, std::vector<uint64_t> qubits) const {
    auto b = std::make_unique<Block>();
    b->mem = std::make_shared<DeviceMem>(16 * size * B);
    b->amps = b->mem->as<double>();
    b->size = size;
    b->qubits = std::move(qubits);
    return b;
  }

  bool any(const Mask& m) const { return std::find(m.begin(), m.end(), 1) != m.end(); }
  bool covers(const Mask& m) const {
    for (uint32_t s = 0; s < B; ++s)
      if (alive[s] && !m[s]) return false;
    return true;
  }
  // Slots of m whose decision differs from the first one's leave the batch.
  template <class D>
  void agree(Mask& m, const std::vector<D>& d) {
    int r = -1;
    for (uint32_t s = 0; s < B; ++s) {
      if (!m[s]) continue;
      if (r < 0) r = (int)s;
      else if (!(d[s] == d[r])) m[s] = alive[s] = 0;
    }
  }
  int first(const Mask& m) const {
    for (uint32_t s = 0; s < B; ++s)
      if (m[s]) return (int)s;
    return -1;
  }
  // A structural change applies to every slot or to none: the slots on the
  // other side of the first live slot leave. Returns whether it applies.
  bool unanimous(Mask& m) {
    const int r = first(alive);
    const bool apply = r >= 0 && m[r];
    for (uint32_t s = 0; s < B; ++s) {
      if (alive[s] && (bool)m[s] != apply) alive[s] = 0;
      m[s] = apply && alive[s];
    }
    return apply;
  }
  k::Flag flag(const Mask& m) {
    if (covers(m)) return k::Flag{};
    auto h = std::make_shared<std::vector<int>>(m.begin(), m.end());
    PBCHK(hipMemcpyAsync(flags.get(), h->data(), sizeof(int) * B, hipMemcpyHostToDevice, st));
    host_keep.push_back(h);
    return k::Flag{flags.as<int>(), 1};
  }
  // (a0, a1) of every slot of a one-wire block.
  std::vector<cplx> read_single(const Block& b) {
    std::vector<double> v(4 * (size_t)B);
    PBCHK(hipMemcpyAsync(v.data(), b.amps, 32 * (size_t)B, hipMemcpyDeviceToHost, st));
    sync();
    std::vector<cplx> out(2 * (size_t)B);
    for (size_t i = 0; i < out.size(); ++i) out[i] = cplx(v[2 * i], v[2 * i + 1]);
    return out;
  }
  // Pool::refresh_known for the slots of m.
  void refresh(const std::vector<uint64_t>& ws, const Mask& m) {
    for (uint64_t w : ws) {
      const Block& b = *blocks[positions[w].first];
      if (b.nq() != 1) {
        for (uint32_t s = 0; s < B; ++s)
          if (m[s]) kn(s, w) = -1;
        continue;
      }
      const std::vector<cplx> v = read_single(b);
      const cplx zero(0, 0);
      for (uint32_t s = 0; s < B; ++s)
        if (m[s]) kn(s, w) = ceq(v[2 * s + 1], zero) ? 0 : ceq(v[2 * s], zero) ? 1 : -1;
    }
  }

  uint64_t merge(uint64_t a, uint64_t b) {
    if (a == b) return a;
    const uint64_t n = blocks[a]->nq() + blocks[b]->nq();
    if (opt.has_max_block && n > opt.max_block_qubits)
      fail("P-block merge exceeds max_block_qubits; no entanglement was discarded");
    const uint64_t dim = dimension(n);
    reserve(nbytes(dim, 16));
    const Block& left = *blocks[a];
    const Block& right = *blocks[b];
    std::vector<uint64_t> qubits = left.qubits;
    qubits.insert(qubits.end(), right.qubits.begin(), right.qubits.end());
    auto out = new_block(dim, qubits);
    k::pb_kron_batch(out->amps, left.amps, left.size, right.amps, right.size, B, st);
    live = live - left.size - right.size + dim;
    for (size_t i = 0; i < qubits.size(); ++i) positions[qubits[i]] = {a, i};
    graveyard.push_back(blocks[a]->mem);
    graveyard.push_back(blocks[b]->mem);
    blocks[a] = std::move(out);
    blocks[b].reset();
    free.push_back(b);
    return a;
  }
  bool needs_merge(const std::vector<uint64_t>& qs) const {
    for (uint64_t q : qs)
      if (positions[q].first != positions[qs[0]].first) return true;
    return false;
  }
  // Pool::ensure; the merges are structural, so m must be unanimous first.
  uint64_t ensure(const std::vector<uint64_t>& qs) {
    std::vector<uint64_t> bl;
    for (uint64_t q : qs) bl.push_back(positions[q].first);
    std::sort(bl.begin(), bl.end());
    bl.erase(std::unique(bl.begin(), bl.end()), bl.end());
    std::stable_sort(bl.begin(), bl.end(), [&](uint64_t x, uint64_t y) { return blocks[x]->nq() < blocks[y]->nq(); });
    for (size_t i = 1; i < bl.size(); ++i) merge(bl[0], bl[i]);
    return bl[0];
  }
  // Per slot: -1 when a known control fails (the gate is skipped), else the
  // set of controls still unknown, as a bit mask over cs.
  std::vector<int64_t> control_codes(const std::vector<std::pair<uint64_t, bool>>& cs, const Mask& m) {
    std::vector<int64_t> code(B, 0);
    for (uint32_t s = 0; s < B; ++s) {
      if (!m[s]) continue;
      for (size_t i = 0; i < cs.size(); ++i) {
        const int8_t v = kn(s, cs[i].first);
        if (v < 0) code[s] |= 1ll << i;
        else if ((v != 0) != cs[i].second) {
          code[s] = -1;
          break;
        }
      }
    }
    return code;
  }
  // Controls, trivial-gate skip and merges shared by one() and the
  // controlled swap; returns the active controls or nullopt (nothing runs).
  std::optional<std::vector<std::pair<uint64_t, bool>>> controls(const std::vector<std::pair<uint64_t, bool>>& cs,
                                                                 Mask& m) {
    if (!any(m)) return std::nullopt;
    const std::vector<int64_t> code = control_codes(cs, m);
    agree(m, code);
    const int r = first(m);
    if (r < 0 || code[r] < 0) return std::nullopt;
    std::vector<std::pair<uint64_t, bool>> active;
    for (size_t i = 0; i < cs.size(); ++i)
      if ((code[r] >> i) & 1) active.push_back(cs[i]);
    return active;
  }
  k::Expand local_expand(const std::vector<std::pair<uint64_t, bool>>& cs, std::vector<uint32_t> extra) const {
    uint64_t fixed = 0;
    std::vector<uint32_t> pos = std::move(extra);
    for (auto& c : cs) {
      const uint32_t l = (uint32_t)positions[c.first].second;
      pos.push_back(l);
      if (c.second) fixed |= 1ull << l;
    }
    return expand(pos, fixed);
  }
  // Merge the blocks of qs when needed (only if every slot applies).
  std::optional<uint64_t> gather(const std::vector<uint64_t>& qs, Mask& m) {
    if (needs_merge(qs) && !unanimous(m)) return std::nullopt;
    if (!any(m)) return std::nullopt;
    return ensure(qs);
  }

  // Pool::one
  void one(uint64_t q, const std::vector<cplx>& mat, const std::vector<std::pair<uint64_t, bool>>& cs, Mask& m) {
    auto active = controls(cs, m);
    if (!active) return;
    if (!active->empty() && blocks[positions[q].first]->nq() == 1) {
      const std::vector<cplx> v = read_single(*blocks[positions[q].first]);
      std::vector<char> trivial(B, 0);
      for (uint32_t s = 0; s < B; ++s) {
        if (!m[s]) continue;
        const cplx a = v[2 * s], z = v[2 * s + 1];
        trivial[s] = ceq(cadd(cmul(mat[0], a), cmul(mat[1], z)), a) && ceq(cadd(cmul(mat[2], a), cmul(mat[3], z)), z);
      }
      agree(m, trivial);
      const int r = first(m);
      if (r < 0 || trivial[r]) return;
    }
    std::vector<uint64_t> qs;
    for (auto& c : *active) qs.push_back(c.first);
    qs.push_back(q);
    const auto idx = gather(qs, m);
    if (!idx) return;
    Block& b = *blocks[*idx];
    const uint32_t t = (uint32_t)positions[q].second;
    const k::Expand e = local_expand(*active, {t});
    const Mat2 xm = gates::X();
    bool is_x = true;
    for (int i = 0; i < 4; ++i) is_x = is_x && ceq(mat[i], xm[i]);
    if (is_x) {
      k::xgate(16, b.amps, b.size, B, (int)b.nq(), e, t, flag(m), st);
    } else {
      double u[8];
      for (int i = 0; i < 4; ++i) {
        u[2 * i] = mat[i].re;
        u[2 * i + 1] = mat[i].im;
      }
      k::one(16, b.amps, b.size, B, (int)b.nq(), e, t, u, flag(m), st);
    }
    refresh(qs, m);
  }

  const void* dense_matrix(const Op& op) {
    auto it = mat_of.find(&op);
    if (it != mat_of.end()) return it->second;
    auto dm = std::make_shared<DeviceMem>(16 * op.mat.size());
    std::vector<double> h;
    for (auto& v : op.mat) {
      h.push_back(v.re);
      h.push_back(v.im);
    }
    PBCHK(hipMemcpyAsync(dm->get(), h.data(), 8 * h.size(), hipMemcpyHostToDevice, st));
    sync();
    mats.push_back(dm);
    mat_of[&op] = dm->get();
    return dm->get();
  }

  // Pool::gate
  void gate(const Op& op, Mask& m) {
    std::vector<std::pair<uint64_t, bool>> cs;
    for (auto& c : op.controls) cs.push_back({c.first, c.second});
    switch (op.kind) {
      case OpKind::One: one(op.target, op.mat, cs, m); break;
      case OpKind::X: {
        const Mat2 x = gates::X();
        one(op.target, std::vector<cplx>(x.begin(), x.end()), cs, m);
        break;
      }
      case OpKind::Swap: {
        auto active = controls(cs, m);
        if (!active) return;
        if (active->empty()) {
          if (!unanimous(m)) return;
          const auto pa = positions[op.a], pbp = positions[op.b];
          blocks[pa.first]->qubits[pa.second] = op.b;
          blocks[pbp.first]->qubits[pbp.second] = op.a;
          std::swap(positions[op.a], positions[op.b]);
          for (uint32_t s = 0; s < B; ++s) std::swap(kn(s, op.a), kn(s, op.b));
          return;
        }
        std::vector<uint64_t> qs;
        for (auto& c : *active) qs.push_back(c.first);
        qs.push_back(op.a);
        qs.push_back(op.b);
        const auto idx = gather(qs, m);
        if (!idx) return;
        Block& b = *blocks[*idx];
        const uint32_t la = (uint32_t)positions[op.a].second, lb = (uint32_t)positions[op.b].second;
        k::swap(16, b.amps, b.size, B, (int)b.nq(), local_expand(*active, {la, lb}), la, lb, flag(m), st);
        refresh(qs, m);
        break;
      }
      case OpKind::Diagonal: {
        if (!any(m)) return;
        const std::vector<uint32_t>& qs = op.qubits;
        // Per slot: the unknown wires (mask over qs) and the reduced values.
        std::vector<std::vector<uint64_t>> key(B);
        for (uint32_t s = 0; s < B; ++s) {
          if (!m[s]) continue;
          uint64_t act = 0, na = 0;
          for (size_t i = 0; i < qs.size(); ++i)
            if (kn(s, qs[i]) < 0) act |= 1ull << i, ++na;
          key[s].push_back(act);
          for (uint64_t basis = 0; basis < (1ull << na); ++basis) {
            uint64_t index = 0, cursor = 0;
            for (uint32_t q : qs) {
              uint64_t bit;
              if (kn(s, q) >= 0) bit = (uint64_t)kn(s, q);
              else bit = (basis >> (na - 1 - cursor++)) & 1;
              index = (index << 1) | bit;
            }
            uint64_t re, im;
            std::memcpy(&re, &op.mat[index].re, 8);
            std::memcpy(&im, &op.mat[index].im, 8);
            key[s].push_back(re);
            key[s].push_back(im);
          }
        }
        agree(m, key);
        const int r = first(m);
        if (r < 0) return;
        std::vector<uint64_t> active;
        for (size_t i = 0; i < qs.size(); ++i)
          if ((key[r][0] >> i) & 1) active.push_back(qs[i]);
        std::vector<cplx> reduced;
        for (size_t j = 1; j < key[r].size(); j += 2) {
          double re, im;
          std::memcpy(&re, &key[r][j], 8);
          std::memcpy(&im, &key[r][j + 1], 8);
          reduced.push_back(cplx(re, im));
        }
        bool uniform = true;
        for (auto& v : reduced) uniform = uniform && ceq(v, reduced[0]);
        if (active.empty() || uniform) {
          Block& b = *blocks[positions[qs[0]].first];
          if (covers(m)) {
            k::pb_scale(b.amps, b.size * B, reduced[0].re, reduced[0].im, st);
          } else {
            for (uint32_t s = 0; s < B; ++s)
              if (m[s]) k::pb_scale(b.amps + 2 * b.size * s, b.size, reduced[0].re, reduced[0].im, st);
          }
        } else {
          const auto idx = gather(active, m);
          if (!idx) return;
          Block& b = *blocks[*idx];
          std::vector<uint32_t> local;
          for (uint64_t q : active) local.push_back((uint32_t)positions[q].second);
          const int kq = (int)local.size();
          std::vector<uint64_t> pats;
          std::vector<double> vals;
          for (int j = 0; j < (1 << kq); ++j) {
            if (reduced[j].re == 1.0 && reduced[j].im == 0.0) continue;
            uint64_t pat = 0;
            for (int i = 0; i < kq; ++i)
              if ((j >> (kq - 1 - i)) & 1) pat |= 1ull << local[i];
            pats.push_back(pat);
            vals.push_back(reduced[j].re);
            vals.push_back(reduced[j].im);
          }
          k::diag(16, b.amps, b.size, B, (int)b.nq(), expand(local, 0), (int)pats.size(), pats.data(), vals.data(),
                  flag(m), st);
        }
        break;
      }
      case OpKind::Dense: {
        std::vector<uint64_t> qs(op.qubits.begin(), op.qubits.end());
        const auto idx = gather(qs, m);
        if (!idx) return;
        Block& b = *blocks[*idx];
        std::vector<uint32_t> local;
        for (uint64_t q : qs) local.push_back((uint32_t)positions[q].second);
        const int kq = (int)local.size();
        uint64_t offsets[32] = {};
        for (int j = 0; j < (1 << kq); ++j) {
          uint64_t off = 0;
          for (int i = 0; i < kq; ++i)
            if ((j >> (kq - 1 - i)) & 1) off |= 1ull << local[i];
          offsets[j] = off;
        }
        k::dense(16, b.amps, b.size, B, (int)b.nq(), kq, expand(local, 0), dense_matrix(op), offsets, flag(m), st);
        refresh(qs, m);
        break;
      }
      case OpKind::Noop: break;
      default: fail("Internal error: nonunitary operation in a P-block gate");
    }
  }

  // Pool::measure for the slots of m; returns the outcome per slot. A
  // measurement that splits its block is structural (unanimous).
  std::vector<int> measure(uint64_t q, Mask& m) {
    std::vector<int> out(B, 0);
    const auto pos = positions[q];
    Block& b = *blocks[pos.first];
    const uint64_t size = b.size;
    if (size > 2 && !unanimous(m)) return out;
    if (!any(m)) return out;
    if (size > 2) reserve(nbytes(size / 2 + 2, 16));
    const int n = (int)b.nq();
    const int seg1 = n < 11 ? n : 11;
    const uint64_t c1 = 1ull << (n - seg1);
    if (red_a.size() < 16 * c1 * B) red_a.alloc(16 * c1 * B);
    if (red_b.size() < 16 * c1 * B) red_b.alloc(16 * c1 * B);
    k::ProbSpec ps;
    ps.branch = (int)pos.second;
    k::seg_reduce_prob(16, b.amps, b.size, B, n, ps, seg1, red_a.get(), c1, st);
    uint64_t count = c1;
    void* src = red_a.get();
    void* dst = red_b.get();
    while (count > 1) {
      int lg = 0;
      while ((1ull << lg) < count) ++lg;
      const int sg = lg < 11 ? lg : 11;
      k::seg_reduce_f64(src, c1, count, B, sg, dst, c1, st);
      count >>= sg;
      std::swap(src, dst);
    }
    k::meas_decide(src, c1, B, ms.as<k::MeasState>(), rng.as<k::RngState>(), nullptr, 0, -1, flag(m),
                   err.as<int>(), st);
    k::collapse(16, b.amps, b.size, B, n, (uint32_t)pos.second, ms.as<k::MeasState>(), st);
    std::vector<k::MeasState> h(B);
    int e = 0;
    PBCHK(hipMemcpyAsync(h.data(), ms.get(), sizeof(k::MeasState) * B, hipMemcpyDeviceToHost, st));
    PBCHK(hipMemcpyAsync(&e, err.get(), sizeof e, hipMemcpyDeviceToHost, st));
    sync();
    if (e & 1) {
      // Only a slot's own run can tell whether the norm was its own: every
      // slot is run again alone, where the error is raised.
      std::fill(alive.begin(), alive.end(), 0);
      std::fill(m.begin(), m.end(), 0);
      return out;
    }
    for (uint32_t s = 0; s < B; ++s) out[s] = h[s].outcome;
    if (size > 2) {
      auto rem = new_block(size / 2, {});
      k::pb_extract_ms(rem->amps, b.amps, size / 2, (uint32_t)pos.second, ms.as<k::MeasState>(), B, st);
      const uint64_t idx = pos.first;
      std::unique_ptr<Block> old = std::move(blocks[idx]);
      graveyard.push_back(old->mem);
      if (free.empty()) fail("Internal error: a multi-qubit block leaves no free slot");
      const uint64_t slot = free.back();
      free.pop_back();
      std::vector<uint64_t> qs = old->qubits;
      qs.erase(qs.begin() + (long)pos.second);
      live = live - old->size + rem->size + 2;
      for (size_t i = 0; i < qs.size(); ++i) positions[qs[i]] = {idx, i};
      rem->qubits = qs;
      blocks[idx] = std::move(rem);
      auto basis = new_block(2, {q});
      k::pb_basis_ms(basis->amps, ms.as<k::MeasState>(), B, st);
      blocks[slot] = std::move(basis);
      positions[q] = {slot, 0};
      refresh(qs, m);
    }
    for (uint32_t s = 0; s < B; ++s)
      if (m[s]) kn(s, q) = (int8_t)out[s];
    return out;
  }
};

Result::Result() = default;
Result::~Result() = default;

namespace {

// One compiled op on a pool (execution.rs:62-84 `run`); cbits is the
// classical register map (unset bits read 0).
void run_op(const POp& op, Pool& pool, std::map<uint64_t, int>& cbits) {
  switch (op.kind) {
    case PK::Conditional: {
      uint64_t actual = 0;
      for (uint64_t bit = 0; bit < op.cond.creg_size; ++bit) {
        auto it = cbits.find(op.cond.creg_base + bit);
        actual |= (uint64_t)(it == cbits.end() ? 0 : it->second) << bit;
      }
      if (actual == op.cond.creg_value) run_op(*op.inner, pool, cbits);
      break;
    }
    case PK::Measure: cbits[op.c] = pool.measure(op.q); break;
    case PK::Reset:
      if (pool.measure(op.q) == 1) {
        const Mat2 x = gates::X();
        pool.one(op.q, std::vector<cplx>(x.begin(), x.end()), {});
      }
      break;
    case PK::Gate: pool.gate(op.gate); break;
    case PK::Noop: break;
  }
}

// run_op for the live slots of a batch (m: the slots the op reaches).
void run_batch(const POp& op, BatchPool& p, BatchPool::Mask& m) {
  switch (op.kind) {
    case PK::Conditional: {
      BatchPool::Mask inner(p.B, 0);
      for (uint32_t s = 0; s < p.B; ++s) {
        if (!m[s]) continue;
        uint64_t actual = 0;
        for (uint64_t bit = 0; bit < op.cond.creg_size; ++bit) {
          auto it = p.cbits[s].find(op.cond.creg_base + bit);
          actual |= (uint64_t)(it == p.cbits[s].end() ? 0 : it->second) << bit;
        }
        inner[s] = actual == op.cond.creg_value;
      }
      run_batch(*op.inner, p, inner);
      break;
    }
    case PK::Measure: {
      const std::vector<int> out = p.measure(op.q, m);
      for (uint32_t s = 0; s < p.B; ++s)
        if (m[s]) p.cbits[s][op.c] = out[s];
      break;
    }
    case PK::Reset: {
      const std::vector<int> out = p.measure(op.q, m);
      BatchPool::Mask x(p.B, 0);
      for (uint32_t s = 0; s < p.B; ++s) x[s] = m[s] && out[s] == 1;
      const Mat2 xm = gates::X();
      p.one(op.q, std::vector<cplx>(xm.begin(), xm.end()), {}, x);
      break;
    }
    case PK::Gate: p.gate(op.gate, m); break;
    case PK::Noop: break;
  }
}

// simulator.rs:171-182
Options working_options(const Options& o, uint64_t nc, bool profile) {
  const uint64_t classical = nbytes(nc, 32);
  if (classical >= o.memory) fail("P-block classical workspace exceeds max_memory_mb");
  Options w = o;
  w.memory -= classical;
  w.profile = profile;
  return w;
}

std::string jnum(double x) {
  char buf[40];
  std::snprintf(buf, sizeof buf, "%.17g", x);
  return buf;
}

void stats_json(std::ostringstream& js, const Stats& s) {
  js << "\"merge_calls\": " << s.merge_calls << ", \"merge_time\": " << jnum(s.merge_time)
     << ", \"merge_output_amplitudes\": " << s.merge_output_amplitudes << ", \"split_calls\": " << s.split_calls
     << ", \"separation_checks\": " << s.separation_checks << ", \"separation_time\": " << jnum(s.separation_time)
     << ", \"separation_residual\": " << jnum(s.separation_residual)
     << ", \"skipped_controlled_gates\": " << s.skipped_controlled_gates
     << ", \"peak_block_qubits\": " << s.peak_block_qubits << ", \"peak_live_amplitudes\": " << s.peak_live_amplitudes
     << ", \"peak_working_bytes\": " << s.peak_working_bytes;
}

void write_profile(double pre, double fusion, double shots_total, double total, uint64_t nshots,
                   const char* strategy, uint64_t parallel, uint64_t prefix, uint64_t budget, const Stats& s) {
  std::ostringstream js;
  js << "{\"preprocessing_time\": " << jnum(pre) << ", \"fusion_time\": " << jnum(fusion)
     << ", \"shots_total_time\": " << jnum(shots_total) << ", \"total_time\": " << jnum(total)
     << ", \"num_shots\": " << nshots << ", \"execution_strategy\": \"" << strategy
     << "\", \"parallel_shots\": " << parallel << ", \"deterministic_prefix_ops\": " << prefix
     << ", \"memory_budget_bytes\": " << budget << ", ";
  stats_json(js, s);
  js << "}";
  if (mkdir("dqsim_profiles", 0777) != 0 && errno != EEXIST)
    throw RuntimeErr(std::string("Failed to write P-block profile: ") + std::strerror(errno));
  const auto ns =
      std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::system_clock::now().time_since_epoch())
          .count();
  const std::string path = "dqsim_profiles/pblock_shots_profile_" + std::to_string(ns) + ".json";
  std::ofstream f(path);
  if (!f) throw RuntimeErr("Failed to write P-block profile: cannot open " + path);
  f << js.str();
  if (!f) throw RuntimeErr("Failed to write P-block profile: write failed");
}

}  // namespace

// simulator.rs:106-118
std::unique_ptr<Result> simulate(const Options& o, const char* json, size_t len) {
  const double t0 = now_s();
  Input in = prepare(json, len, o.memory);
  const std::vector<POp> ops = compile_ops(in.instructions);
  const Options w = working_options(o, in.num_cbits, o.profile);
  const uint64_t seed = o.has_seed ? o.seed : random_seed();
  auto res = std::make_unique<Result>();
  auto stream = std::make_shared<Stream>();
  res->stream = stream;  // queries run on it; released after the pool
  res->pool = std::make_unique<Pool>(in.physical.size(), w, stream->s);
  Pool& pool = *res->pool;
  pool.seed_rng(seed, 0);
  std::map<uint64_t, int> cbits;
  for (const POp& op : ops) run_op(op, pool, cbits);
  PBCHK(hipStreamSynchronize(pool.st));
  res->physical = in.physical;
  res->cbits.assign(cbits.begin(), cbits.end());
  res->elapsed = now_s() - t0;
  return res;
}

// simulator.rs:119-322
thread_local ShotsStats g_shots_stats;
ShotsStats last_shots_stats() { return g_shots_stats; }

Counts simulate_shots(const Options& o, const char* json, size_t len, uint64_t shots, bool profile) {
  g_shots_stats = ShotsStats();
  const double start = now_s();
  Input in = prepare(json, len, o.memory);
  const double pre = now_s() - start;
  const double fs = now_s();
  const std::vector<POp> ops = compile_ops(in.instructions);
  const double fusion = now_s() - fs;
  const Options options = working_options(o, in.num_cbits, profile);
  const uint64_t seed = o.has_seed ? o.seed : random_seed();
  const double exec_start = now_s();
  const uint64_t n = in.physical.size(), nc = in.num_cbits;
  std::map<std::string, uint64_t> counts;
  Stats stats;
  const char* strategy = "empty";
  uint64_t workers = 0, prefix_used = 0;
  auto finish = [&]() {
    if (profile)
      write_profile(pre, fusion, now_s() - exec_start, now_s() - start, shots, strategy, workers, prefix_used,
                    o.memory, stats);
    return Counts(counts.begin(), counts.end());
  };
  if (shots == 0) return finish();
  Stream stream;  // declared before every pool, so it is destroyed after them
  size_t prefix_len = 0;
  while (prefix_len < ops.size() && ops[prefix_len].unitary()) ++prefix_len;
  auto prefix = std::make_unique<Pool>(n, options, stream.s);
  prefix->seed_rng(seed, 0);
  std::map<uint64_t, int> cbits;
  for (size_t i = 0; i < prefix_len; ++i) run_op(ops[i], *prefix, cbits);
  bool all_terminal = true;
  for (size_t i = prefix_len; i < ops.size(); ++i) all_terminal = all_terminal && ops[i].terminal();
  if (o.sample_terminal && all_terminal) {
    if (prefix_len == ops.size()) {
      counts[std::string(nc, '0')] = shots;
      strategy = "terminal_sampling";
      workers = 1;
      prefix_used = prefix_len;
      stats = prefix->stats;
      return finish();
    }
    try {
      std::vector<std::shared_ptr<DeviceMem>> keep;
      const std::vector<k::PbTable> tabs = prefix->sampling_tables(keep);
      uint64_t table_bytes = 0;
      for (auto& t : tabs) table_bytes += t.len * 8;
      prefix->stats.peak_working_bytes =
          std::max(prefix->stats.peak_working_bytes, prefix->bytes() + table_bytes + n * 64);
      const std::vector<const Block*> lb = prefix->live_blocks();
      std::vector<uint32_t> idx = prefix->sample(tabs, shots, seed, 0);
      std::vector<int> bits(n, 0);
      for (uint64_t s = 0; s < shots; ++s) {
        for (size_t t = 0; t < lb.size(); ++t)
          for (size_t i = 0; i < lb[t]->qubits.size(); ++i)
            bits[lb[t]->qubits[i]] = (int)((idx[s * lb.size() + t] >> i) & 1);
        std::map<uint64_t, int> cb;
        for (size_t i = prefix_len; i < ops.size(); ++i)
          if (ops[i].kind == PK::Measure) cb[ops[i].c] = bits[ops[i].q];
        counts[format_cbits(cb, nc)] += 1;
      }
      strategy = "terminal_sampling";
      workers = 1;
      prefix_used = prefix_len;
      stats = prefix->stats;
      return finish();
    } catch (const MemErr&) {
      // Tables over budget: fall back to trajectories (simulator.rs:231-232).
    }
  }
  stats.add(prefix->stats);
  const bool retain = prefix_len > 0 && prefix->bytes() * 2 < options.memory &&
                      prefix->bytes() + prefix->stats.peak_working_bytes <= options.memory;
  const uint64_t reserved = retain ? prefix->bytes() : 0;
  const uint64_t classical = nbytes(nc, 32);
  const uint64_t per_worker =
      std::max<uint64_t>(std::max(prefix->stats.peak_working_bytes, prefix->bytes()) + classical, 4096);
  const uint64_t threads = std::max<unsigned>(1u, std::thread::hardware_concurrency());
  workers = std::min({o.max_parallel_shots, shots, threads,
                      std::max<uint64_t>((options.memory - reserved) / per_worker, 1)});
  Options wopt = options;
  wopt.memory = (wopt.memory - reserved - (workers - 1) * classical) / workers;
  prefix->stats = Stats();
  std::unique_ptr<Pool> kept;
  if (retain) kept = std::move(prefix);
  prefix.reset();
  // Without profiling or separation, shots run in lockstep batches (one
  // launch sequence per batch); the shots a batch evicts, or every shot
  // otherwise, run one after another on a restarted pool. Shot s always
  // draws from Philox stream s, so counts depend neither on the batching nor
  // on the worker count.
  const uint64_t start_op = retain ? prefix_len : 0;
  std::vector<uint64_t> todo;
  if (!profile && !wopt.split_separable && shots > 1) {
    std::unique_ptr<Pool> zero;
    if (kept) {
      if (kept->bytes() > wopt.memory)
        fail("P-block prefix exceeds worker memory budget; reduce max_parallel_shots");
    } else {
      zero = std::make_unique<Pool>(n, wopt, stream.s);
    }
    const Pool& base = kept ? *kept : *zero;
    // The slots of a batch share the call's budget (less a retained
    // prefix): a batch that would outgrow it is retried with half the slots.
    // Shots a batch evicts (their structural decisions differ from the
    // batch's) run again from the start in the next round's batches, where
    // the shots that took the same path stay together; after kRounds rounds,
    // or after an error, the rest run alone.
    constexpr int kRounds = 64;
    const uint64_t total = options.memory - reserved;
    uint64_t bmax = 1024;
    std::vector<uint64_t> pending(shots);
    for (uint64_t s = 0; s < shots; ++s) pending[s] = s;
    bool failed = false;
    for (int round = 0; round < kRounds && !pending.empty() && !failed; ++round) {
      std::vector<uint64_t> evicted;
      for (size_t p0 = 0; p0 < pending.size();) {
        const size_t bsz = std::min<size_t>(bmax, pending.size() - p0);
        const std::vector<uint64_t> streams(pending.begin() + p0, pending.begin() + p0 + bsz);
        std::vector<std::map<uint64_t, int>> done;
        BatchPool::Mask alive;
        try {
          BatchPool bp(base, streams, wopt, total, seed);
          for (size_t i = start_op; i < ops.size() && bp.any(bp.alive); ++i) {
            BatchPool::Mask m = bp.alive;
            run_batch(ops[i], bp, m);
          }
          bp.sync();
          done = std::move(bp.cbits);
          alive = std::move(bp.alive);
          g_shots_stats.peak_batch_bytes = std::max(g_shots_stats.peak_batch_bytes, bp.peak_total + reserved);
          g_shots_stats.batch_size = std::max<uint64_t>(g_shots_stats.batch_size, bsz);
        } catch (const BatchTooLarge&) {
          if (bsz > 1) {
            bmax = std::max<uint64_t>(1, bsz / 2);
            continue;
          }
          failed = true;
        } catch (const std::exception&) {
          failed = true;
        }
        if (failed) {
          // The lone runs raise the error of the first failing shot, as the
          // per-shot path does: this and every later shot run alone.
          for (size_t q = p0; q < pending.size(); ++q) todo.push_back(pending[q]);
          for (uint64_t s : evicted) todo.push_back(s);
          evicted.clear();
          break;
        }
        for (size_t b = 0; b < bsz; ++b) {
          if (alive[b]) {
            counts[format_cbits(done[b], nc)] += 1;
            ++g_shots_stats.lockstep_shots;
          } else {
            evicted.push_back(streams[b]);
          }
        }
        p0 += bsz;
      }
      pending = std::move(evicted);
    }
    if (!failed)
      for (uint64_t s : pending) todo.push_back(s);
    std::sort(todo.begin(), todo.end());
  } else {
    for (uint64_t s = 0; s < shots; ++s) todo.push_back(s);
  }
  std::unique_ptr<Pool> pool;
  g_shots_stats.lone_shots += todo.size();
  for (uint64_t s : todo) {
    if (!pool) {
      if (kept) {
        if (kept->bytes() > wopt.memory)
          fail("P-block prefix exceeds worker memory budget; reduce max_parallel_shots");
        pool = kept->clone();
        pool->opt = wopt;
      } else {
        pool = std::make_unique<Pool>(n, wopt, stream.s);
      }
    } else if (kept) {
      pool->restore(*kept);
    } else {
      pool->reset_zero();
    }
    pool->seed_rng(seed, s);
    std::map<uint64_t, int> cb;
    for (size_t i = start_op; i < ops.size(); ++i) run_op(ops[i], *pool, cb);
    counts[format_cbits(cb, nc)] += 1;
    stats.add(pool->stats);
  }
  strategy = retain ? "prefix_trajectories" : "trajectories";
  prefix_used = retain ? prefix_len : 0;
  return finish();
}

// ---------------------------------------------------------------------------
// Queries (result.rs)
// ---------------------------------------------------------------------------
namespace {

std::vector<uint64_t> query_wires(const Result& r, const std::vector<uint64_t>* qubits) {
  std::vector<uint64_t> qs;
  if (!qubits) qs.assign(r.physical.rbegin(), r.physical.rend());
  else qs = *qubits;
  std::vector<uint64_t> local;
  for (uint64_t q : qs) {
    auto it = std::lower_bound(r.physical.begin(), r.physical.end(), q);
    if (it == r.physical.end() || *it != q)
      throw ValueErr("Physical qubit " + std::to_string(q) + " is not in the result");
    const uint64_t i = (uint64_t)(it - r.physical.begin());
    if (std::find(local.begin(), local.end(), i) != local.end()) throw ValueErr("Query qubits must be distinct");
    local.push_back(i);
  }
  return local;
}

// Per-block FP64 sums over the block (norm_squared).
double block_norm(Pool& p, const Block& b) {
  const void* tot = p.reduce_prob(b, -1);
  double v[2];
  PBCHK(hipMemcpyAsync(v, tot, 16, hipMemcpyDeviceToHost, p.st));
  PBCHK(hipStreamSynchronize(p.st));
  return v[0];
}

}  // namespace

std::vector<std::vector<uint64_t>> block_qubits(const Result& r) {
  std::vector<std::vector<uint64_t>> out;
  for (const Block* b : r.pool->live_blocks()) {
    std::vector<uint64_t> qs;
    for (uint64_t q : b->qubits) qs.push_back(r.physical[q]);
    out.push_back(qs);
  }
  return out;
}

bool profiled(const Result& r) { return r.pool->opt.profile; }

std::string diagnostics_json(const Result& r, bool with_total_time) {
  Pool& p = *r.pool;
  std::ostringstream js;
  js << "{";
  stats_json(js, p.stats);
  js << ", \"block_sizes\": [";
  uint64_t live = 0;
  double norm = 1.0;
  bool first = true;
  for (const Block* b : p.live_blocks()) {
    js << (first ? "" : ", ") << b->nq();
    first = false;
    live += b->size;
    norm = norm * block_norm(p, *b);
  }
  js << "], \"live_amplitudes\": " << live << ", \"norm_squared\": " << jnum(norm);
  if (with_total_time) js << ", \"total_time\": " << jnum(r.elapsed);
  js << "}";
  return js.str();
}

uint64_t statevector_len(const Result& r) {
  const uint64_t n = r.physical.size();
  const uint64_t size = dimension(n);
  r.pool->check(add_checked(nbytes(size, 16), nbytes(n, 8)));
  return size;
}

// Dense export (result.rs:118-140): amplitude(b) = product over blocks of
// the block amplitude at b's bits, local bit i of a block = wire qubits[i].
void statevector(const Result& r, void* out) {
  Pool& p = *r.pool;
  const uint64_t size = statevector_len(r);
  std::vector<k::PbTable> tabs;
  std::vector<uint32_t> shifts;
  for (const Block* b : p.live_blocks()) {
    tabs.push_back({b->amps, b->size, (uint32_t)b->nq(), (uint32_t)shifts.size()});
    for (uint64_t q : b->qubits) shifts.push_back((uint32_t)q);
  }
  DeviceMem dt(sizeof(k::PbTable) * std::max<size_t>(tabs.size(), 1)),
      ds(4 * std::max<size_t>(shifts.size(), 1)), dout(16 * size);
  if (!tabs.empty())
    PBCHK(hipMemcpyAsync(dt.get(), tabs.data(), sizeof(k::PbTable) * tabs.size(), hipMemcpyHostToDevice, p.st));
  if (!shifts.empty()) PBCHK(hipMemcpyAsync(ds.get(), shifts.data(), 4 * shifts.size(), hipMemcpyHostToDevice, p.st));
  k::pb_product_c128(dout.get(), size, dt.as<k::PbTable>(), (uint32_t)tabs.size(), ds.as<uint32_t>(), p.st);
  PBCHK(hipMemcpyAsync(out, dout.get(), 16 * size, hipMemcpyDeviceToHost, p.st));
  PBCHK(hipStreamSynchronize(p.st));
}

// result.rs:141-160 and model.rs:557-577
void amplitude(const Result& r, const std::string& bits, double* out2) {
  Pool& p = *r.pool;
  const uint64_t n = r.physical.size();
  if (bits.size() != n || !std::all_of(bits.begin(), bits.end(), [](char c) { return c == '0' || c == '1'; }))
    throw ValueErr("Expected one binary digit per physical qubit, highest physical index first");
  p.check(nbytes(n, 8));
  std::vector<k::PbTable> tabs;
  for (const Block* b : p.live_blocks()) {
    uint64_t index = 0;
    for (size_t i = 0; i < b->qubits.size(); ++i)
      if (bits[n - 1 - b->qubits[i]] == '1') index |= 1ull << i;
    tabs.push_back({b->amps + 2 * index, 1, 0, 0});
  }
  DeviceMem dt(sizeof(k::PbTable) * std::max<size_t>(tabs.size(), 1)), ds(4), dout(16);
  if (!tabs.empty())
    PBCHK(hipMemcpyAsync(dt.get(), tabs.data(), sizeof(k::PbTable) * tabs.size(), hipMemcpyHostToDevice, p.st));
  k::pb_product_c128(dout.get(), 1, dt.as<k::PbTable>(), (uint32_t)tabs.size(), ds.as<uint32_t>(), p.st);
  PBCHK(hipMemcpyAsync(out2, dout.get(), 16, hipMemcpyDeviceToHost, p.st));
  PBCHK(hipStreamSynchronize(p.st));
}

// result.rs:161-209: per-block marginals (the statevector reduction
// kernels over a virtual index map) combined as a product on the device.
std::vector<std::pair<uint64_t, double>> probabilities(const Result& r, const std::vector<uint64_t>* qubits) {
  Pool& p = *r.pool;
  const std::vector<uint64_t> qs = query_wires(r, qubits);
  const uint64_t size = dimension(qs.size());
  const std::vector<const Block*> lb = p.live_blocks();
  uint64_t table_size = 0;
  std::vector<std::vector<std::pair<uint64_t, uint64_t>>> sel(lb.size());  // (position in qs, local bit)
  for (size_t t = 0; t < lb.size(); ++t) {
    for (size_t i = 0; i < qs.size(); ++i) {
      auto it = std::find(lb[t]->qubits.begin(), lb[t]->qubits.end(), qs[i]);
      if (it != lb[t]->qubits.end()) sel[t].push_back({i, (uint64_t)(it - lb[t]->qubits.begin())});
    }
    table_size = add_checked(table_size, dimension(sel[t].size()));
  }
  p.check(add_checked(nbytes(size, 96), nbytes(table_size, 8)));
  std::vector<k::PbTable> tabs;
  std::vector<uint32_t> shifts;
  std::vector<std::shared_ptr<DeviceMem>> keep;
  for (size_t t = 0; t < lb.size(); ++t) {
    const Block& b = *lb[t];
    const int nb = (int)b.nq(), kk = (int)sel[t].size();
    // Virtual order: unselected local bits low, then the selected ones with
    // the first selected highest (the marginal index has it as MSB).
    std::vector<uint8_t> used(nb, 0), map;
    for (auto& s : sel[t]) used[s.second] = 1;
    for (int q = 0; q < nb; ++q)
      if (!used[q]) map.push_back((uint8_t)q);
    for (int j = kk - 1; j >= 0; --j) map.push_back((uint8_t)sel[t][j].second);
    const int seg = nb - kk;
    const int seg1 = seg < 11 ? seg : 11;
    const uint64_t count1 = 1ull << (nb - seg1);
    auto a = std::make_shared<DeviceMem>(16 * count1);
    auto bb = std::make_shared<DeviceMem>(16 * count1);
    k::ProbSpec ps;
    ps.branch = -1;
    ps.map = map.data();
    k::seg_reduce_prob(16, b.amps, b.size, 1, nb, ps, seg1, a->get(), 0, p.st);
    int remaining = seg - seg1;
    uint64_t count = count1;
    while (remaining > 0) {
      const int s = remaining < 11 ? remaining : 11;
      const uint64_t next = std::max<uint64_t>(count >> s, 1);
      k::seg_reduce_f64(a->get(), 0, count, 1, s, bb->get(), 0, p.st);
      count = next;
      remaining -= s;
      std::swap(a, bb);
    }
    // x channel of the double2 bins into a dense table.
    auto tab = std::make_shared<DeviceMem>(8 * count);
    PBCHK(hipMemcpy2DAsync(tab->get(), 8, a->get(), 16, 8, count, hipMemcpyDeviceToDevice, p.st));
    keep.push_back(a);
    keep.push_back(bb);
    keep.push_back(tab);
    tabs.push_back({tab->get(), count, (uint32_t)kk, (uint32_t)shifts.size()});
    // Table index bit j (LSB 0) is selected entry kk-1-j, whose query
    // position pos sits at bit (|qs| - 1 - pos) of the output index.
    for (int j = 0; j < kk; ++j) shifts.push_back((uint32_t)(qs.size() - 1 - sel[t][kk - 1 - j].first));
  }
  DeviceMem dt(sizeof(k::PbTable) * std::max<size_t>(tabs.size(), 1)),
      ds(4 * std::max<size_t>(shifts.size(), 1)), dout(8 * size);
  if (!tabs.empty())
    PBCHK(hipMemcpyAsync(dt.get(), tabs.data(), sizeof(k::PbTable) * tabs.size(), hipMemcpyHostToDevice, p.st));
  if (!shifts.empty()) PBCHK(hipMemcpyAsync(ds.get(), shifts.data(), 4 * shifts.size(), hipMemcpyHostToDevice, p.st));
  k::pb_product_f64(dout.as<double>(), size, dt.as<k::PbTable>(), (uint32_t)tabs.size(), ds.as<uint32_t>(), p.st);
  std::vector<double> probs(size);
  PBCHK(hipMemcpyAsync(probs.data(), dout.get(), 8 * size, hipMemcpyDeviceToHost, p.st));
  PBCHK(hipStreamSynchronize(p.st));
  std::vector<std::pair<uint64_t, double>> out;
  for (uint64_t i = 0; i < size; ++i)
    if (probs[i] > 0.0) out.push_back({i, probs[i]});
  return out;
}

// result.rs:210-240: shot s draws Philox stream s, one draw per block.
Counts counts(const Result& r, uint64_t shots, const std::vector<uint64_t>* qubits, uint64_t seed) {
  Pool& p = *r.pool;
  const std::vector<uint64_t> qs = query_wires(r, qubits);
  std::map<std::string, uint64_t> c;
  if (shots == 0) return {};
  std::vector<std::shared_ptr<DeviceMem>> keep;
  const std::vector<k::PbTable> tabs = p.sampling_tables(keep);
  const std::vector<const Block*> lb = p.live_blocks();
  const std::vector<uint32_t> idx = p.sample(tabs, shots, seed, 0);
  std::vector<int> bits(r.physical.size(), 0);
  for (uint64_t s = 0; s < shots; ++s) {
    for (size_t t = 0; t < lb.size(); ++t)
      for (size_t i = 0; i < lb[t]->qubits.size(); ++i)
        bits[lb[t]->qubits[i]] = (int)((idx[s * lb.size() + t] >> i) & 1);
    std::string key;
    for (uint64_t q : qs) key.push_back(bits[q] ? '1' : '0');
    c[key] += 1;
  }
  return Counts(c.begin(), c.end());
}

// result.rs:241-293
double expectation_value(const Result& r, const std::string& pauli) {
  Pool& p = *r.pool;
  const uint64_t n = r.physical.size();
  if (pauli.size() != n ||
      !std::all_of(pauli.begin(), pauli.end(), [](char c) { return c == 'I' || c == 'X' || c == 'Y' || c == 'Z'; }))
    throw ValueErr("Expected one I, X, Y or Z per physical qubit");
  p.check(nbytes(n, 1));
  cplx acc(1, 0);
  for (const Block* b : p.live_blocks()) {
    uint64_t xm = 0, ym = 0, zm = 0;
    for (size_t i = 0; i < b->qubits.size(); ++i) {
      const char op = pauli[n - 1 - b->qubits[i]];
      if (op == 'X') xm |= 1ull << i;
      else if (op == 'Y') ym |= 1ull << i;
      else if (op == 'Z') zm |= 1ull << i;
    }
    DeviceMem part(16 * k::pb_pauli_blocks(b->size)), out(16);
    k::pb_pauli(b->amps, b->size, xm, ym, zm, (uint32_t)b->nq(), part.get(), out.get(), p.st);
    double v[2];
    PBCHK(hipMemcpyAsync(v, out.get(), 16, hipMemcpyDeviceToHost, p.st));
    PBCHK(hipStreamSynchronize(p.st));
    acc = cmul(acc, cplx(v[0], v[1]));
  }
  return acc.re;
}

}  // namespace pb
}  // namespace dq

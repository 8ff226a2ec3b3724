// Device statevector engine (host side).
//
// Reference semantics followed here:
//   state_layout / zero_state / check_budget ... src/monolithic/statevector/shots.rs:22-48
//   run_shots (terminal + trajectories) ........ shots.rs:58-260
//   Operation::execute, conditionals ........... execution.rs:276-366
//   simulate / SimulationResult queries ........ simulator.rs:85-202,319-371
//   write_shots_profile ........................ src/profiling.rs:8-18
//
// All device work of one call runs on one HIP stream; measurement outcomes,
// conditions and RNG draws stay on the device, so a trajectory never waits
// for the host between operations.
#include "engine.hpp"

#include <hip/hip_runtime_api.h>
#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <exception>
#include <fstream>
#include <functional>
#include <map>
#include <mutex>
#include <random>
#include <sstream>
#include <thread>
#include <unordered_map>

#include "batch.hpp"
#include "fuse.hpp"
#include "kernels.hpp"
#include "planner.hpp"
#include "philox.hpp"

namespace dq {

// ---------------------------------------------------------------------------
// HIP helpers
// ---------------------------------------------------------------------------
namespace {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw DeviceErr(std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPCHK(x) hip_check((x), #x)

struct Stream {
  hipStream_t s = nullptr;
  int dev = 0;
  Stream() { s = acquire_stream(&dev); }
  ~Stream() {
    if (s) release_stream(s, dev);
  }
  Stream(const Stream&) = delete;
  Stream& operator=(const Stream&) = delete;
  void sync() { HIPCHK(hipStreamSynchronize(s)); }
};

struct Event {
  hipEvent_t e = nullptr;
  Event() { HIPCHK(hipEventCreate(&e)); }
  ~Event() {
    if (e) (void)hipEventDestroy(e);
  }
  Event(const Event&) = delete;
  Event& operator=(const Event&) = delete;
  Event(Event&& o) noexcept : e(o.e) { o.e = nullptr; }
};

float elapsed_ms(const Event& a, const Event& b) {
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, a.e, b.e));
  return ms;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

DeviceMem& DeviceMem::operator=(DeviceMem&& o) noexcept {
  if (this != &o) {
    reset();
    p_ = o.p_;
    n_ = o.n_;
    o.p_ = nullptr;
    o.n_ = 0;
  }
  return *this;
}
// Device allocation cache: freed blocks of at least 1 MiB are kept (at most
// kCacheBlocks of them) and handed out again for a request of the same size,
// so repeated simulations of one size do not pay for hipMalloc/hipFree of
// multi-GiB states. The cache is emptied and the allocation retried when
// hipMalloc reports out-of-memory.
namespace {
constexpr size_t kCacheMin = 1 << 20;
constexpr size_t kCacheBlocks = 4;
std::mutex g_cache_mu;
std::vector<std::pair<size_t, void*>> g_cache;

void* cache_take(size_t bytes) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  for (size_t i = 0; i < g_cache.size(); ++i)
    if (g_cache[i].first == bytes) {
      void* p = g_cache[i].second;
      g_cache.erase(g_cache.begin() + (long)i);
      return p;
    }
  return nullptr;
}
bool cache_put(size_t bytes, void* p) {
  if (bytes < kCacheMin) return false;
  // hipFree waits for the device; keep that guarantee for cached blocks so a
  // block is never handed out while queued work may still touch it.
  if (hipDeviceSynchronize() != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  std::lock_guard<std::mutex> lk(g_cache_mu);
  if (g_cache.size() >= kCacheBlocks) {
    (void)hipFree(g_cache.front().second);
    g_cache.erase(g_cache.begin());
  }
  g_cache.push_back({bytes, p});
  return true;
}
void cache_release() {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  for (auto& b : g_cache) (void)hipFree(b.second);
  g_cache.clear();
}
}  // namespace

// Stream-ordered arena of one batch worker thread. While a thread has one
// installed, its DeviceMem blocks come from and go back to the arena's free
// lists (power-of-two size classes) with no hipFree and no device-wide
// synchronisation: everything the worker runs is ordered on its one stream,
// so a block handed out again is only touched after the earlier work on it.
// The blocks are freed when the arena is destroyed, after the stream drains.
struct DeviceArena {
  std::vector<std::pair<size_t, void*>> free_blocks;
  std::vector<void*> owned;
  ~DeviceArena() {
    for (void* p : owned) (void)hipFree(p);
  }
  bool owns(const void* p) const { return std::find(owned.begin(), owned.end(), p) != owned.end(); }
};
namespace {
thread_local DeviceArena* t_arena = nullptr;
size_t arena_class(size_t bytes) {
  size_t c = 256;
  while (c < bytes) c <<= 1;
  return c;
}
}  // namespace

namespace {
struct StreamPool {
  std::mutex mu;
  std::map<int, std::vector<hipStream_t>> idle;  // device -> streams
};
StreamPool& stream_pool() {
  static StreamPool* p = new StreamPool;
  return *p;
}
constexpr size_t kIdleStreams = 32;
}  // namespace

hipStream_t acquire_stream(int* dev) {
  HIPCHK(hipGetDevice(dev));
  {
    StreamPool& p = stream_pool();
    std::lock_guard<std::mutex> lk(p.mu);
    auto& v = p.idle[*dev];
    if (!v.empty()) {
      hipStream_t s = v.back();
      v.pop_back();
      return s;
    }
  }
  hipStream_t s = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  return s;
}

void release_stream(hipStream_t s, int dev) {
  {
    StreamPool& p = stream_pool();
    std::lock_guard<std::mutex> lk(p.mu);
    auto& v = p.idle[dev];
    if (v.size() < kIdleStreams) {
      v.push_back(s);
      return;
    }
  }
  (void)hipStreamDestroy(s);
}

DeviceMem::~DeviceMem() { reset(); }
void DeviceMem::reset() {
  if (p_ && t_arena && t_arena->owns(p_)) t_arena->free_blocks.push_back({n_, p_});
  else if (p_ && !cache_put(n_, p_)) (void)hipFree(p_);
  p_ = nullptr;
  n_ = 0;
}
void DeviceMem::alloc(size_t bytes) {
  reset();
  if (bytes == 0) bytes = 16;
  if (t_arena) {
    const size_t c = arena_class(bytes);
    auto& fb = t_arena->free_blocks;
    for (size_t i = 0; i < fb.size(); ++i)
      if (fb[i].first == c) {
        p_ = fb[i].second;
        n_ = c;
        fb.erase(fb.begin() + (long)i);
        return;
      }
    hipError_t e = hipMalloc(&p_, c);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
      (void)hipGetLastError();
      p_ = nullptr;
      throw MemErr("Unable to allocate statevector");
    }
    hip_check(e, "hipMalloc");
    t_arena->owned.push_back(p_);
    n_ = c;
    return;
  }
  if (void* c = cache_take(bytes)) {
    p_ = c;
    n_ = bytes;
    return;
  }
  hipError_t e = hipMalloc(&p_, bytes);
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
    (void)hipGetLastError();
    cache_release();
    e = hipMalloc(&p_, bytes);
  }
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
    (void)hipGetLastError();
    p_ = nullptr;
    throw MemErr("Unable to allocate statevector");
  }
  hip_check(e, "hipMalloc");
  n_ = bytes;
}

int device_count() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

// Relabelled tile sweeps (emit_sweeps, emit_exact_sweeps).
std::atomic<bool> g_sweep_relabel{false};
bool set_sweep_relabel(bool on) { return g_sweep_relabel.exchange(on); }
std::atomic<bool> g_sweep_pinned{false};
bool set_sweep_pinned(bool on) { return g_sweep_pinned.exchange(on); }

// Complex64 fused sweeps when the budget holds a second state: 0 in place
// (k_tile), out of place (k_xsweep) with 1 persistent 256-thread, 2 one-shot
// 256-thread, 3 one-shot 512-thread, 4 persistent 512-thread workgroups.
std::atomic<int> g_sweep_mode{2};
int set_sweep_mode(int mode) { return mode < 0 || mode > 4 ? -1 : g_sweep_mode.exchange(mode); }

// Arithmetic of 3-qubit blocks in complex64 tile sweeps: 0 FP32 matrix cores
// (16x16x4 f32), 1 split precision (each FP32 operand as two FP16 parts on
// 16x16x32 f16, all four partial products accumulated in FP32; the default:
// a quarter of the matrix-core cycles at about 1.1-1.5x the FP32 path's
// amplitude error).
std::atomic<int> g_block_arith{1};
int set_block_arith(int mode) { return mode < 0 || mode > 1 ? -1 : g_block_arith.exchange(mode); }
// Blocks of a tile grouped into wave-owned runs (1, the default) or a
// barrier after every block (0); read by the 256-thread out-of-place sweeps.
std::atomic<int> g_wave_runs{1};
int set_wave_runs(int mode) { return mode < 0 || mode > 1 ? -1 : g_wave_runs.exchange(mode); }

// Plans that open with out-of-place sweeps start from a synthesised |0...0>
// and skip the reads of all-zero tiles (Exec::start_zero); off: initialise
// the state and read it.
std::atomic<bool> g_zero_start{true};
bool set_zero_start(bool on) { return g_zero_start.exchange(on); }

double mfma_fp32_peak_tflops(int reps, double* ms_per_launch) {
  int dev = 0, cus = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const uint32_t grid = (uint32_t)cus * 8;  // 8 waves per SIMD
  const int iters = 20000;
  DeviceMem out;
  out.alloc((size_t)k::mfma_peak_out_floats(grid) * sizeof(float));
  Stream st;
  k::mfma_peak(out.as<float>(), grid, iters, st.s);  // warm-up
  Event a, b;
  HIPCHK(hipEventRecord(a.e, st.s));
  for (int r = 0; r < std::max(1, reps); ++r) k::mfma_peak(out.as<float>(), grid, iters, st.s);
  HIPCHK(hipEventRecord(b.e, st.s));
  st.sync();
  HIPCHK(hipGetLastError());
  const double ms = elapsed_ms(a, b) / std::max(1, reps);
  if (ms_per_launch) *ms_per_launch = ms;
  return (double)k::mfma_peak_flops_per_launch(grid, iters) / (ms * 1e-3) / 1e12;
}

// ---------------------------------------------------------------------------
// Budget (shots.rs:22-48)
// ---------------------------------------------------------------------------
namespace {

struct Layout {
  uint64_t len, bytes;
};

Layout state_layout(uint64_t n, int prec) {
  if (n >= 64) throw MemErr("State dimension overflow");
  uint64_t len = 1ull << n;
  if (len > (uint64_t)INT64_MAX / (uint64_t)prec) throw MemErr("State byte size overflow");
  return {len, len * (uint64_t)prec};
}

void check_budget(uint64_t required, uint64_t budget) {
  if (required > budget)
    throw MemErr("Simulation needs at least " + std::to_string(required) +
                 " working bytes, exceeding max_memory_mb budget (" + std::to_string(budget) + " bytes)");
}

uint64_t random_seed() {
  std::random_device rd;
  return ((uint64_t)rd() << 32) ^ (uint64_t)rd();
}

// ---------------------------------------------------------------------------
// Device plan
// ---------------------------------------------------------------------------
enum DKind { D_ONE, D_X, D_SWAP, D_DIAG, D_DENSE, D_FUSED, D_MEAS, D_RESET, D_NOOP, D_TILE, D_XSWEEP };

struct DevOp {
  DKind kind = D_NOOP;
  int category = -1;  // profile bucket (execution.rs:297-347)
  k::Expand e;
  uint32_t target = 0, a = 0, b = 0, qubit = 0;
  int64_t cbit = -1;
  double u[8] = {};
  std::vector<uint64_t> pat;
  std::vector<double> dval;
  int kq = 0;
  uint64_t offsets[32] = {};
  size_t mat_off = 0;
  bool mfma = false;
  size_t emb_off = 0;
  std::vector<k::Cond> conds;
  size_t cond_off = 0;
  int gates = 1;
  int low = 0;
  int nctl = 0;
  float feat[mlp::kFeatures] = {};
  std::vector<uint32_t> touched;  // qubits the launch touches (plan description)
  std::vector<int> src;           // compiled ops the launch applies, in order
  // D_TILE: tile qubits and the blocks applied in LDS (TileBlock array in the blob).
  k::TileArgs targs{};
  size_t tile_off = 0;
  std::vector<std::vector<uint32_t>> tile_blocks;  // qubits of each block (plan description)
  std::vector<int> tile_mfma;  // block kind of each tile block (0: no matrix core; 4, 5: split precision)
  std::vector<uint32_t> xtile_q;  // out-of-place sweeps: the qubit of each tile-index bit
  bool exact = false;  // complex128 sweep of unfused ops (k_tile128)
  int chains = 0;      // exact sweeps: register chains / single-op passes
  int fixed_chains = 0;  // exact sweeps: chains run by compiled-signature code
  std::vector<uint32_t> tile_lq;  // qubit of each local bit at the load (plan description)
  // D_XSWEEP: out-of-place sweep from the current state buffer to the other
  // one (blob at tile_off, as for D_TILE); `persist` picks the launch shape.
  k::XSweepArgs xargs{};
  bool persist = true;
};

// Planner features of a block of compiled ops (also used for single ops):
//  0 k/5             1 lowest qubit/34     2 highest qubit/34   3 qubits < 4, /4
//  4 ops folded/32   5 diagonal-op share   6 Clifford-op share  7 all ops diagonal
//  8 n/34            9 k >= 4 (MFMA)      10 span/34           11 touched share of amplitudes
// 12-15 stay zero here; a tile sweep sets 12 = 1, 13 = L/12, 14 = blocks on
// the 16x16x4 path / 16, 15 = 4-qubit blocks / 16, plus 4 (ops/32) and 8 (n/34).
bool clifford_entry(const cplx& e) {
  const double r = std::fabs(e.re), i = std::fabs(e.im), t = 1e-12, h = 0.70710678118654752440;
  auto is = [&](double a, double b) { return std::fabs(r - a) < t && std::fabs(i - b) < t; };
  return is(0, 0) || is(1, 0) || is(0, 1) || is(h, 0) || is(0, h) || is(0.5, 0.5);
}
void block_features(const std::vector<uint32_t>& qubits, const std::vector<int>& opidx, const std::vector<Op>& ops,
                    int n, float* f) {
  const int k = (int)qubits.size();
  uint32_t lo = 63, hi = 0;
  int low4 = 0;
  for (uint32_t q : qubits) {
    lo = std::min(lo, q);
    hi = std::max(hi, q);
    low4 += q < 4;
  }
  int diag = 0, cliff = 0;
  for (int oi : opidx) {
    const Op& o = ops[oi];
    diag += o.kind == OpKind::Diagonal;
    bool c = o.kind == OpKind::X || o.kind == OpKind::Swap;
    if (!c && !o.mat.empty()) {
      c = true;
      for (const cplx& e : o.mat) c = c && clifford_entry(e);
    }
    cliff += c;
  }
  const float no = (float)std::max<size_t>(1, opidx.size());
  double touched = 1.0;
  if (opidx.size() == 1) {
    const Op& o = ops[opidx[0]];
    if (o.kind == OpKind::One || o.kind == OpKind::X) touched = std::ldexp(1.0, -(int)o.controls.size());
    else if (o.kind == OpKind::Swap) touched = std::ldexp(1.0, -1 - (int)o.controls.size());
    else if (o.kind == OpKind::Diagonal) {
      int ne = 0;
      for (const cplx& e : o.mat) ne += !(e.re == 1.0 && e.im == 0.0);
      touched = (double)ne / (double)o.mat.size();
    }
  }
  f[0] = k / 5.f;
  f[1] = lo / 34.f;
  f[2] = hi / 34.f;
  f[3] = low4 / 4.f;
  f[4] = opidx.size() / 32.f;
  f[5] = diag / no;
  f[6] = cliff / no;
  f[7] = diag == (int)opidx.size() ? 1.f : 0.f;
  f[8] = n / 34.f;
  f[9] = k >= 4 ? 1.f : 0.f;
  f[10] = (hi - lo) / 34.f;
  f[11] = (float)touched;
  for (int i = 12; i < mlp::kFeatures; ++i) f[i] = 0.f;  // tile-sweep features
}

struct Plan {
  std::vector<DevOp> ops;
  std::vector<uint8_t> blob;  // matrices and condition chains
  DeviceMem dev;
  bool uses_alt = false;  // out-of-place sweeps: the executor needs a second state buffer
  // Learned choices of out-of-place sweep runs: 100 * (qubits 0..3 in every
  // tile) + block width, one per unitary run that the model planned.
  std::vector<int> learned_choices;
  size_t push(const void* p, size_t bytes) {
    size_t off = (blob.size() + 255) & ~(size_t)255;
    blob.resize(off + bytes);
    if (bytes) std::memcpy(blob.data() + off, p, bytes);
    return off;
  }
  void upload(hipStream_t st) {
    if (blob.empty()) return;
    dev.alloc(blob.size());
    HIPCHK(hipMemcpyAsync(dev.get(), blob.data(), blob.size(), hipMemcpyHostToDevice, st));
  }
  const uint8_t* base() const { return dev.as<const uint8_t>(); }
};

k::Expand make_expand(std::vector<uint32_t> pos, uint64_t fixed) {
  std::sort(pos.begin(), pos.end());
  k::Expand e;
  e.nbits = (uint32_t)pos.size();
  for (size_t i = 0; i < pos.size(); ++i) e.pos[i] = pos[i];
  e.fixed = fixed;
  return e;
}

void set_dense_offsets(DevOp& d, const std::vector<uint32_t>& qubits) {
  const int k = (int)qubits.size();
  d.kq = k;
  for (int j = 0; j < (1 << k); ++j) {
    uint64_t off = 0;
    for (int i = 0; i < k; ++i)
      if ((j >> (k - 1 - i)) & 1) off |= 1ull << qubits[i];
    d.offsets[j] = off;
  }
  d.e = make_expand(qubits, 0);
  d.low = (int)*std::min_element(qubits.begin(), qubits.end());
}

void push_matrix(Plan& plan, DevOp& d, const std::vector<cplx>& mat, int prec) {
  if (prec == 16) {
    std::vector<double> v;
    for (auto& c : mat) {
      v.push_back(c.re);
      v.push_back(c.im);
    }
    d.mat_off = plan.push(v.data(), v.size() * sizeof(double));
  } else {
    std::vector<float> v;
    for (auto& c : mat) {
      v.push_back((float)c.re);
      v.push_back((float)c.im);
    }
    d.mat_off = plan.push(v.data(), v.size() * sizeof(float));
  }
}

void set_diag(DevOp& d, const std::vector<uint32_t>& qubits, const std::vector<cplx>& diag) {
  const int k = (int)qubits.size();
  d.kind = D_DIAG;
  d.e = make_expand(qubits, 0);
  d.kq = k;
  d.low = (int)*std::min_element(qubits.begin(), qubits.end());
  for (int j = 0; j < (1 << k); ++j) {
    if (diag[j].re == 1.0 && diag[j].im == 0.0) continue;
    uint64_t pat = 0;
    for (int i = 0; i < k; ++i)
      if ((j >> (k - 1 - i)) & 1) pat |= 1ull << qubits[i];
    d.pat.push_back(pat);
    d.dval.push_back(diag[j].re);
    d.dval.push_back(diag[j].im);
  }
}

// Lower one compiled op (execution.rs:276-347) to a device op.
DevOp lower(const Op& op, Plan& plan, int prec) {
  DevOp d;
  if (op.kind == OpKind::Conditional) {
    const Op* cur = &op;
    std::vector<k::Cond> conds;
    while (cur->kind == OpKind::Conditional) {
      conds.push_back({cur->cond.creg_base, cur->cond.creg_size, cur->cond.creg_value});
      cur = cur->inner.get();
    }
    d = lower(*cur, plan, prec);
    d.conds = conds;
    d.cond_off = plan.push(conds.data(), conds.size() * sizeof(k::Cond));
    return d;
  }
  std::vector<uint32_t> pos;
  uint64_t fixed = 0;
  for (auto& c : op.controls) {
    pos.push_back(c.first);
    if (c.second) fixed |= 1ull << c.first;
  }
  d.nctl = (int)op.controls.size();
  switch (op.kind) {
    case OpKind::One:
      d.kind = D_ONE;
      d.category = op.controls.empty() ? 0 : 1;
      d.target = op.target;
      pos.push_back(op.target);
      d.e = make_expand(pos, fixed);
      for (int i = 0; i < 4; ++i) {
        d.u[2 * i] = op.mat[i].re;
        d.u[2 * i + 1] = op.mat[i].im;
      }
      d.kq = 1 + d.nctl;
      d.low = (int)*std::min_element(pos.begin(), pos.end());
      break;
    case OpKind::X:
      d.kind = D_X;
      d.category = op.controls.empty() ? 0 : 1;
      d.target = op.target;
      pos.push_back(op.target);
      d.e = make_expand(pos, fixed);
      d.kq = 1 + d.nctl;
      d.low = (int)*std::min_element(pos.begin(), pos.end());
      break;
    case OpKind::Swap:
      d.kind = D_SWAP;
      d.category = 1;
      d.a = op.a;
      d.b = op.b;
      pos.push_back(op.a);
      pos.push_back(op.b);
      d.e = make_expand(pos, fixed);
      d.kq = 2 + d.nctl;
      d.low = (int)*std::min_element(pos.begin(), pos.end());
      break;
    case OpKind::Diagonal:
      set_diag(d, op.qubits, op.mat);
      d.category = op.qubits.size() > 1 ? 1 : 0;
      break;
    case OpKind::Dense:
      d.kind = D_DENSE;
      d.category = 1;
      set_dense_offsets(d, op.qubits);
      push_matrix(plan, d, op.mat, prec);
      break;
    case OpKind::Measure:
      d.kind = D_MEAS;
      d.category = 2;
      d.qubit = op.qubit;
      d.cbit = (int64_t)op.cbit;
      d.kq = 1;
      d.low = (int)op.qubit;
      break;
    case OpKind::Reset:
      d.kind = D_RESET;
      d.category = 2;
      d.qubit = op.qubit;
      d.kq = 1;
      d.low = (int)op.qubit;
      break;
    default:
      d.kind = D_NOOP;
      d.category = -1;
      break;
  }
  return d;
}

// Real embedding of a complex block for the MFMA kernel: [[Ur,-Ui],[Ui,Ur]].
std::vector<float> real_embedding(const std::vector<cplx>& m, int k) {
  const int D = 1 << k, R = 2 * D;
  std::vector<float> e((size_t)R * R);
  for (int r = 0; r < D; ++r)
    for (int c = 0; c < D; ++c) {
      const cplx v = m[(size_t)r * D + c];
      e[(size_t)r * R + c] = (float)v.re;
      e[(size_t)r * R + c + D] = (float)-v.im;
      e[(size_t)(r + D) * R + c] = (float)v.im;
      e[(size_t)(r + D) * R + c + D] = (float)v.re;
    }
  return e;
}

std::vector<float> g_planner;  // learned planner weights (empty = greedy)

double planner_score(const FusedBlock& a, const FusedBlock& b, const std::vector<Op>& ops, int n);
// Predicted device time of a launch (ms): the model is trained on
// log(ms per 2^24 amplitudes).
double predicted_ms(const DevOp& d, int n) {
  if (d.kind == D_NOOP) return 0.0;
  return std::exp((double)mlp_eval_host(g_planner.data(), d.feat)) * std::ldexp(1.0, n - 24);
}

// One fused block as its own launch.
void emit_block(Plan& plan, FusedBlock& blk, const std::vector<Op>& ops, int prec, int n) {
  float feat[mlp::kFeatures];
  block_features(blk.qubits, blk.ops, ops, n, feat);
  if (blk.ops.size() == 1) {
    plan.ops.push_back(lower(ops[blk.ops[0]], plan, prec));
    std::copy(feat, feat + mlp::kFeatures, plan.ops.back().feat);
    plan.ops.back().touched = blk.qubits;
    plan.ops.back().src = blk.ops;
    return;
  }
  if (blk.mat.empty()) build_block_matrix(blk, ops);
  DevOp d;
  d.category = 1;
  d.gates = (int)blk.ops.size();
  if (blk.diagonal) {
    std::vector<cplx> dg;
    const size_t dim = (size_t)1 << blk.qubits.size();
    for (size_t r = 0; r < dim; ++r) dg.push_back(blk.mat[r * dim + r]);
    set_diag(d, blk.qubits, dg);
  } else {
    d.kind = prec == 8 ? D_FUSED : D_DENSE;
    set_dense_offsets(d, blk.qubits);
    push_matrix(plan, d, blk.mat, prec);
    if (prec == 8 && d.kq >= 4) {
      std::vector<float> emb = real_embedding(blk.mat, d.kq);
      d.emb_off = plan.push(emb.data(), emb.size() * sizeof(float));
      d.mfma = true;
    }
  }
  std::copy(feat, feat + mlp::kFeatures, d.feat);
  d.touched = blk.qubits;
  d.src = blk.ops;
  plan.ops.push_back(std::move(d));
}

uint64_t qubit_mask(const std::vector<uint32_t>& q) {
  uint64_t m = 0;
  for (uint32_t x : q) m |= 1ull << x;
  return m;
}

// Tile-sweep matrix form of a block: kind 3 (k = 2; 1-qubit blocks are
// padded to 2 qubits) the 8x8 real embedding as A fragments of the 4x4x1
// MFMA, kind 1 (k = 3) permuted 16x16 real embedding for the 16x16x4 MFMA,
// kind 2 (k = 4) the real embedding [[Ur, -Ui], [Ui, Ur]] for the 32x32x2 MFMA.
// (kind 0, a complex FP32 matrix, serves no tile block.)
int tile_kind(int k, bool split = false) {
  return k <= 2 ? 3 : k == 3 ? (split ? 4 : 1) : (k == 4 && split) ? 5 : 2;
}

// IEEE binary16 encoding of x rounded to nearest, ties to even (|x| < 65520).
// The rounding is one FP64 add and subtract of 1.5 * 2^52 on x scaled so
// that FP16's last place is 1 (scaling by a power of two is exact, and the
// add rounds to an integer, to nearest even, since |x * s| < 2^12); the
// encoding is then read off the bits of the rounded double.
uint16_t half_bits(double x) {
  uint64_t u;
  std::memcpy(&u, &x, 8);
  const uint16_t sign = (uint16_t)((u >> 48) & 0x8000);
  const int be = (int)((u >> 52) & 0x7FF);
  if (be == 0) return sign;  // zero or a double subnormal: rounds to zero
  const int e = be - 1023;   // |x| in [2^e, 2^(e+1))
  if (e < -25) return sign;  // below half the smallest FP16 subnormal
  if (e > 15) throw RuntimeErr("internal: FP16 overflow in a split-precision matrix");
  const int ex = e < -14 ? -14 : e;  // FP16 exponent; subnormals share -14
  double s;                          // 2^(10 - ex)
  const uint64_t su = (uint64_t)(1023 + 10 - ex) << 52;
  std::memcpy(&s, &su, 8);
  const double magic = 6755399441055744.0;  // 1.5 * 2^52
  const double a = x < 0 ? -x : x;
  const double q = (a * s + magic) - magic;  // RNE(|x| / 2^(ex - 10)), an integer < 2^12
  const uint32_t qi = (uint32_t)q;
  if (qi < 1024) return (uint16_t)(sign | (uint16_t)qi);  // subnormal (ex = -14), or zero
  if (qi >= 2048) {                                        // rounding carried into the next binade
    if (ex + 1 > 15) throw RuntimeErr("internal: FP16 overflow in a split-precision matrix");
    return (uint16_t)(sign | (uint16_t)((ex + 16) << 10));
  }
  return (uint16_t)(sign | (uint16_t)((ex + 15) << 10) | (uint16_t)(qi - 1024));
}
double half_value(uint16_t h) {
  const int ex = (h >> 10) & 31, m = h & 1023;
  double v;
  if (ex == 0) {
    v = (double)m * 0x1p-24;
  } else {
    const uint64_t u = ((uint64_t)(ex - 15 + 1023) << 52) | ((uint64_t)m << 42);
    std::memcpy(&v, &u, 8);
  }
  return (h & 0x8000) ? -v : v;
}
// FP16 split of an FP64 value: hi = RNE(x), lo = RNE(x - hi).
void half_split(double x, uint16_t& hi, uint16_t& lo) {
  hi = half_bits(x);
  lo = half_bits(x - half_value(hi));
}
size_t tile_matrix_bytes(int k) {
  const size_t d = (size_t)1 << std::max(k, 2);
  return 4 * (2 * d) * (2 * d);
}
std::vector<float> tile_matrix(const std::vector<cplx>& m, int k, bool split = false) {
  const int kind = tile_kind(k, split);
  if (kind == 4) {
    // Split precision: per lane l (row l % 16, u = l / 16), dword s of
    // [s][lane] packs two FP16 values (low half first): s = 0, 1 the high
    // parts Eh of the real- and imaginary-part columns of input 2u + s, s =
    // 2, 3 the low parts El = E - Eh of the same, with E the embedding of the
    // FP32 path (row 2c+p: output c, part p) taken from the FP64 matrix.
    std::vector<float> v(256, 0.f);
    for (int s = 0; s < 2; ++s)
      for (int l = 0; l < 64; ++l) {
        const int row = l % 16, c = row >> 1, p = row & 1, cin = 2 * (l / 16) + s;
        const cplx x = m[(size_t)c * 8 + cin];
        const double ev[2] = {p == 0 ? x.re : x.im, p == 0 ? -x.im : x.re};  // input part 0 (real), 1 (imaginary)
        uint32_t wh = 0, wl = 0;
        for (int q = 0; q < 2; ++q) {
          uint16_t hi, lo;
          half_split(ev[q], hi, lo);
          wh |= (uint32_t)hi << (16 * q);
          wl |= (uint32_t)lo << (16 * q);
        }
        std::memcpy(&v[(size_t)s * 64 + l], &wh, 4);
        std::memcpy(&v[(size_t)(s + 2) * 64 + l], &wl, 4);
      }
    return v;
  }
  if (kind == 5) {
    // Split precision, k = 4: the embedding E = [[Ur, -Ui], [Ui, Ur]] of the
    // FP32 path (rows: real parts of outputs 0..15, then imaginary parts;
    // columns: real parts of inputs, then imaginary parts) taken from the
    // FP64 matrix. Per (t, lane l), row l % 32, h = l / 32, inputs a = 4t +
    // h and b = 4t + 2 + h: dwords [4 (64 t + l) + w] pack (real-part column,
    // imaginary-part column) of Eh for a (w = 0) and b (w = 1), of El = E -
    // Eh for a (w = 2) and b (w = 3), low half first.
    std::vector<float> v(1024, 0.f);
    for (int t = 0; t < 4; ++t)
      for (int l = 0; l < 64; ++l) {
        const int row = l % 32, c = row % 16, p = row / 16, h = l / 32;
        for (int w = 0; w < 2; ++w) {
          const int cin = 4 * t + 2 * w + h;
          const cplx x = m[(size_t)c * 16 + cin];
          const double ev[2] = {p == 0 ? x.re : x.im, p == 0 ? -x.im : x.re};
          uint32_t wh = 0, wl = 0;
          for (int q = 0; q < 2; ++q) {
            uint16_t hi, lo;
            half_split(ev[q], hi, lo);
            wh |= (uint32_t)hi << (16 * q);
            wl |= (uint32_t)lo << (16 * q);
          }
          std::memcpy(&v[(size_t)4 * (64 * t + l) + w], &wh, 4);
          std::memcpy(&v[(size_t)4 * (64 * t + l) + w + 2], &wl, 4);
        }
      }
    return v;
  }
  if (kind == 3) {
    // E[2c+p][2c'+s]: output c part p, input c' part s (0 real, 1 imaginary).
    // Stored as A fragments: [i][h][kk] = E[4 h + i][kk].
    std::vector<float> f(64, 0.f);
    for (int row = 0; row < 8; ++row)
      for (int col = 0; col < 8; ++col) {
        const int c = row >> 1, p = row & 1, cin = col >> 1, q = col & 1;
        const cplx x = m[(size_t)c * 4 + cin];
        const double val = p == 0 ? (q == 0 ? x.re : -x.im) : (q == 0 ? x.im : x.re);
        f[(size_t)(row & 3) * 16 + (row >> 2) * 8 + col] = (float)val;
      }
    return f;
  }
  if (kind == 2) {
    // Fragment-major: [t][q][lane] = E[32 t + lane % 32][2 q + lane / 32].
    const int R = 2 << k, T = R / 32, STEPS = R / 2;
    const std::vector<float> e = real_embedding(m, k);
    std::vector<float> f((size_t)R * R);
    for (int t = 0; t < T; ++t)
      for (int q = 0; q < STEPS; ++q)
        for (int l = 0; l < 64; ++l)
          f[((size_t)t * STEPS + q) * 64 + l] = e[(size_t)(32 * t + l % 32) * R + 2 * q + l / 32];
    return f;
  }
  std::vector<float> v;
  if (kind == 0) {
    for (auto& c : m) {
      v.push_back((float)c.re);
      v.push_back((float)c.im);
    }
    return v;
  }
  // Embedding row 2c+p: output c, part p (0 real, 1 imaginary). Column
  // 4s+u: input 2u + s/2, part s%2. Entry = d(out part)/d(in part) of
  // out = U in. Stored fragment-major: [s][lane] = E[lane % 16][4 s + lane / 16].
  std::vector<float> e(256, 0.f);
  for (int row = 0; row < 16; ++row)
    for (int col = 0; col < 16; ++col) {
      const int c = row >> 1, p = row & 1, s = col >> 2, u = col & 3;
      const int cin = 2 * u + (s >> 1), q = s & 1;
      const cplx x = m[(size_t)c * 8 + cin];
      const double val = p == 0 ? (q == 0 ? x.re : -x.im) : (q == 0 ? x.im : x.re);
      e[(size_t)row * 16 + col] = (float)val;
    }
  v.assign(256, 0.f);
  for (int s = 0; s < 4; ++s)
    for (int l = 0; l < 64; ++l) v[(size_t)s * 64 + l] = e[(size_t)(l % 16) * 16 + 4 * s + l / 16];
  return v;
}

// ---------------------------------------------------------------------------
// Relabelled tile sweeps
// ---------------------------------------------------------------------------
// During a run of tile sweeps a bit position of the state may hold another
// qubit than its own: a sweep may store its tile with the positions of some
// of its qubits exchanged (in place: the same addresses it read), and the
// next sweep's tile then has to contain whatever qubits sit at positions
// 0..F-1 (the contiguous runs its loads move). Each sweep can bring up to two
// of its qubits into those low positions, chosen as the ones the next blocks
// need first, so the low positions stop pinning qubits 0..F-1 into every
// tile. A run ends with the identity layout (restore passes if needed).
// Off by default: at 32 qubits it cuts the sweeps by 9 % (67 -> 61 complex64,
// 150 -> 138 exact complex128), but the sweeps it leaves carry more blocks
// on kernels whose arithmetic and HBM time are nearly equal, and the lane
// swaps add VALU work, so the whole run measured about 5 % slower.
// Invariant kept by every relabelling: low position j holds qubit j or a
// qubit >= F (a qubit < F only ever returns to its own position).

struct BitLayout {
  uint32_t pos[64];  // bit position holding qubit q
  uint32_t at[64];   // qubit held at bit position p
  int n;
  explicit BitLayout(int n_) : n(n_) {
    for (uint32_t q = 0; q < 64; ++q) pos[q] = at[q] = q;
  }
  bool identity() const {
    for (int q = 0; q < n; ++q)
      if (pos[q] != (uint32_t)q) return false;
    return true;
  }
  uint64_t low_mask(int F) const {
    uint64_t m = 0;
    for (int p = 0; p < F; ++p) m |= 1ull << at[p];
    return m;
  }
};

// A sweep's tile as the kernel sees it (k::TileArgs): load-order local bit
// i holds qubit lq[i], loaded from position spos[i]; after the lane swaps
// (as the tile goes into LDS) local bit i holds qubit cq[i], which the
// blocks address and which is stored at position opos[i].
struct TileMap {
  int L = 0, F = 0;
  uint32_t lq[12] = {}, cq[12] = {}, spos[12] = {}, opos[12] = {}, srt[12] = {};
  uint32_t nswap = 0, sbit[2] = {0, 0};
  bool relabels() const {
    if (nswap) return true;
    for (int i = 0; i < L; ++i)
      if (opos[i] != spos[i]) return true;
    return false;
  }
  TileMap plain() const {  // same tile, stored where it was read
    TileMap t = *this;
    std::copy(spos, spos + 12, t.opos);
    std::copy(lq, lq + 12, t.cq);
    t.nswap = 0;
    return t;
  }
  void to_args(k::TileArgs& a) const {
    a.L = (uint32_t)L;
    std::copy(spos, spos + 12, a.spos);
    std::copy(opos, opos + 12, a.opos);
    std::copy(srt, srt + 12, a.srt);
    a.nswap = nswap;
    a.sbit[0] = sbit[0];
    a.sbit[1] = sbit[1];
  }
};

// Tile of the qubits in `tmask` (exactly L, including those at positions
// 0..F-1) in layout `lay`, stored with `swaps` = (qubit x, low position j):
// x (at a position >= F) sits at local bit 8 + s and trades places with lane
// bit j. Then every other content of the tile's high positions goes to its
// own position when that is in the tile, else keeps its position when free,
// else takes a free one. Updates `lay` to the layout after the store.
TileMap sweep_map(BitLayout& lay, uint64_t tmask, int L, int F,
                  const std::vector<std::pair<uint32_t, uint32_t>>& swaps) {
  if (__builtin_popcountll(tmask) != L || (tmask & lay.low_mask(F)) != lay.low_mask(F) || swaps.size() > 2 ||
      (int)swaps.size() > L - 8)
    throw RuntimeErr("internal: bad tile for a sweep");
  TileMap tm;
  tm.L = L;
  tm.F = F;
  std::vector<uint32_t> rest;
  for (int p = F; p < lay.n; ++p) {
    const uint32_t q = lay.at[p];
    bool sw = false;
    for (auto& x : swaps) sw |= x.first == q;
    if (((tmask >> q) & 1) && !sw) rest.push_back(q);
  }
  for (auto& x : swaps)
    if (lay.pos[x.first] < (uint32_t)F || !((tmask >> x.first) & 1) || x.second >= (uint32_t)F)
      throw RuntimeErr("internal: bad lane swap for a sweep");
  if (rest.size() + swaps.size() != (size_t)(L - F)) throw RuntimeErr("internal: tile size mismatch");
  for (int b = 0; b < F; ++b) tm.lq[b] = lay.at[b];
  for (size_t s = 0; s < swaps.size(); ++s) tm.lq[8 + s] = swaps[s].first;
  size_t r = 0;
  for (int b = F; b < L; ++b)
    if (b < 8 || b >= 8 + (int)swaps.size()) tm.lq[b] = rest[r++];
  for (int b = 0; b < L; ++b) tm.spos[b] = lay.pos[tm.lq[b]];
  std::copy(tm.spos, tm.spos + L, tm.srt);
  std::sort(tm.srt, tm.srt + L);
  // Contents after the lane swaps.
  uint32_t content[12];
  std::copy(tm.lq, tm.lq + L, content);
  tm.nswap = (uint32_t)swaps.size();
  for (size_t s = 0; s < swaps.size(); ++s) {
    const uint32_t j = swaps[s].second;
    tm.sbit[s] = j;
    std::swap(content[j], content[8 + s]);
  }
  std::vector<char> high(64, 0), used(64, 0);
  for (int b = F; b < L; ++b) high[tm.spos[b]] = 1;
  std::vector<int> dest(L, -1);
  for (int b = 0; b < F; ++b) dest[b] = b;
  for (int b = F; b < L; ++b)
    if (high[content[b]] && !used[content[b]]) {
      dest[b] = (int)content[b];
      used[content[b]] = 1;
    }
  for (int b = F; b < L; ++b) {
    const uint32_t cur = lay.pos[content[b]];
    if (dest[b] < 0 && high[cur] && !used[cur]) {
      dest[b] = (int)cur;
      used[cur] = 1;
    }
  }
  for (int b = F; b < L; ++b)
    if (dest[b] < 0)
      for (int c = F; c < L; ++c)
        if (!used[tm.spos[c]]) {
          dest[b] = (int)tm.spos[c];
          used[tm.spos[c]] = 1;
          break;
        }
  for (int b = 0; b < L; ++b) {
    tm.cq[b] = content[b];
    tm.opos[b] = (uint32_t)dest[b];
    lay.pos[content[b]] = tm.opos[b];
    lay.at[tm.opos[b]] = content[b];
  }
  return tm;
}

// Low-position choice for the next sweep: the tile's qubits at positions >= F
// whose first use (`first[q]`, smaller is sooner) comes before that of a low
// occupant replace the latest-used occupants, at most two per sweep. A qubit
// < F may only take its own position. At the end of a run (`last`) qubits
// 0..F-1 in the tile return to their positions instead.
std::vector<std::pair<uint32_t, uint32_t>> choose_swaps(const BitLayout& lay, uint64_t tmask, int L, int F,
                                                         const std::vector<uint32_t>& first, bool last) {
  std::vector<std::pair<uint32_t, uint32_t>> sw;
  const size_t cap = (size_t)std::min(2, L - 8);
  if (last) {
    for (int j = 0; j < F && sw.size() < cap; ++j)
      if (lay.at[j] != (uint32_t)j && ((tmask >> j) & 1)) sw.push_back({(uint32_t)j, (uint32_t)j});
    return sw;
  }
  std::vector<uint32_t> cand;
  for (int p = F; p < lay.n; ++p) {
    const uint32_t q = lay.at[p];
    if (((tmask >> q) & 1) && first[q] != UINT32_MAX) cand.push_back(q);
  }
  std::stable_sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) { return first[a] < first[b]; });
  std::vector<char> slot_used(F, 0);
  for (uint32_t x : cand) {
    if (sw.size() >= cap) break;
    int best = -1;
    for (int j = 0; j < F; ++j) {
      if (slot_used[j] || first[lay.at[j]] <= first[x]) continue;
      if (x < (uint32_t)F && (uint32_t)j != x) continue;
      if (best < 0 || first[lay.at[j]] > first[lay.at[best]]) best = j;
    }
    if (best < 0) continue;
    slot_used[best] = 1;
    sw.push_back({x, (uint32_t)best});
  }
  return sw;
}

// Passes of tile sweeps with no blocks that bring the state back to the
// identity layout. Each pass fixes up to two low positions (the qubit of
// position j trades places with the occupant, which goes to its own position)
// and sends every other qubit whose current and own positions both fit in
// the tile to its own position; no qubit at its own position moves, so every
// pass places at least one more qubit and the loop ends.
void restore_layout(BitLayout& lay, int L, int F, const std::function<void(const TileMap&)>& emit) {
  for (int pass = 0; !lay.identity(); ++pass) {
    if (pass >= 64) throw RuntimeErr("internal: tile layout restore did not converge");
    std::vector<char> inP(64, 0);
    int room = L - F;
    std::vector<std::pair<uint32_t, uint32_t>> sw;
    auto fits = [&](uint32_t a, uint32_t b) {
      const int need = (a >= (uint32_t)F && !inP[a]) + (b >= (uint32_t)F && !inP[b] && b != a);
      return need <= room;
    };
    auto add = [&](uint32_t p) {
      if (p >= (uint32_t)F && !inP[p]) {
        inP[p] = 1;
        --room;
      }
    };
    for (int j = 0; j < F && (int)sw.size() < std::min(2, L - 8); ++j) {
      if (lay.at[j] == (uint32_t)j) continue;
      const uint32_t from = lay.pos[j], occ = lay.at[j];
      if (from < (uint32_t)F || occ < (uint32_t)F) throw RuntimeErr("internal: low-position invariant broken");
      if (!fits(from, occ)) continue;
      add(from);
      add(occ);
      sw.push_back({(uint32_t)j, (uint32_t)j});
    }
    for (int q = F; q < lay.n; ++q) {
      const uint32_t cur = lay.pos[q];
      if (cur == (uint32_t)q || cur < (uint32_t)F) continue;
      if (fits(cur, (uint32_t)q)) {
        add(cur);
        add((uint32_t)q);
      }
    }
    for (int p = F; p < lay.n && room > 0; ++p) add((uint32_t)p);
    uint64_t tmask = lay.low_mask(F);
    for (int p = F; p < lay.n; ++p)
      if (inP[p]) tmask |= 1ull << lay.at[p];
    emit(sweep_map(lay, tmask, L, F, sw));
  }
}

// Blob of a tile sweep (TileBlock descriptors, then the block matrices) for
// the tile whose LDS local bit b holds qubit cq[b]; fills the launch's block
// bookkeeping (gates, source ops, features) and returns the blob's offset in
// the plan, its 16-byte units and the widest block.
struct TileBlob {
  size_t off = 0;
  uint32_t blob16 = 0, nblocks = 0, kmax = 0;
  bool split = false;  // some block runs in split precision (the tile is held scaled)
};
// Runs of a tile's blocks for the wave-owned group split of the 256-thread
// out-of-place sweeps: consecutive blocks of at most 3 qubits share two wave
// bits (pad[2] = w0 | w1 << 8, w0 < w1; pad[0] = the same two bits with each
// block's bits below them removed) that no block of the run touches and that
// lie above every block's lane bits (the 4 lowest bits outside a 3-qubit
// block, the 6 lowest outside a 2-qubit one: the lanes then address LDS as
// with the per-block split, bank for bank), the two highest such bits; only
// the last block of a run has a barrier after it (pad[1] bit 0). A 4-qubit
// block is a run of its own. With `runs` false every block is a run of one.
// Other kernels ignore the fields.
void wave_runs(std::vector<k::TileBlock>& tb, int L, bool runs) {
  auto lane_top = [](const k::TileBlock& t) {  // highest lane bit of the block
    uint32_t m = 0;
    for (uint32_t q = 0; q < t.k; ++q) m |= 1u << t.loc[q];
    int left = t.k == 3 ? 4 : 6, bit = 0;
    for (;; ++bit)
      if (!((m >> bit) & 1u) && --left == 0) return bit;
  };
  auto pick = [L](uint32_t used, int above, uint32_t* w) {  // two highest free bits above `above`
    int c = 0;
    for (int bit = L - 1; bit > above && c < 2; --bit)
      if (!((used >> bit) & 1u)) w[c++] = (uint32_t)bit;
    return c == 2;
  };
  size_t i = 0;
  while (i < tb.size()) {
    if (tb[i].k > 3) {
      tb[i].pad[0] = tb[i].pad[2] = 0;
      tb[i].pad[1] = 1;
      ++i;
      continue;
    }
    uint32_t used = 0, w[2] = {0, 0};
    int above = -1;
    size_t j = i;
    while (j < tb.size() && tb[j].k <= 3) {
      uint32_t u = used, cand[2];
      for (uint32_t q = 0; q < tb[j].k; ++q) u |= 1u << tb[j].loc[q];
      const int a = std::max(above, lane_top(tb[j]));
      if (!pick(u, a, cand)) break;
      used = u;
      above = a;
      w[0] = cand[0];
      w[1] = cand[1];
      ++j;
      if (!runs) break;
    }
    if (j == i) throw std::logic_error("wave_runs: no two free tile bits above a block's lane bits");
    const uint32_t w0 = w[1], w1 = w[0];
    for (size_t m = i; m < j; ++m) {
      k::TileBlock& t = tb[m];
      uint32_t c0 = w0, c1 = w1;
      for (uint32_t q = 0; q < t.k; ++q) {
        c0 -= t.loc[q] < w0;
        c1 -= t.loc[q] < w1;
      }
      t.pad[0] = c0 | c1 << 8;
      t.pad[1] = m + 1 == j ? 1u : 0u;
      t.pad[2] = w0 | w1 << 8;
    }
    i = j;
  }
}

TileBlob tile_blob(Plan& plan, DevOp& d, std::vector<FusedBlock*>& blocks, const uint32_t* cq, int L,
                   const std::vector<Op>& ops, int n, bool split) {
  TileBlob out;
  std::vector<int> local(64, -1);
  for (int b = 0; b < L; ++b) local[cq[b]] = b;
  std::vector<k::TileBlock> tb;
  std::vector<float> mats;
  const size_t head = blocks.size() * sizeof(k::TileBlock);
  for (FusedBlock* orig : blocks) {
    // 1-qubit blocks become 2-qubit blocks (U kron I on the highest free
    // tile qubit) for the 4x4x1 MFMA path.
    FusedBlock padded;
    FusedBlock* b = orig;
    if (orig->qubits.size() < 2) {
      padded.ops = orig->ops;
      std::vector<uint32_t> q = orig->qubits;
      for (int i = L - 1; i >= 0 && q.size() < 2; --i)
        if (std::find(q.begin(), q.end(), cq[i]) == q.end()) q.push_back(cq[i]);
      std::sort(q.rbegin(), q.rend());
      padded.qubits = q;
      b = &padded;
    }
    if (b->mat.empty()) build_block_matrix(*b, ops);
    k::TileBlock t{};
    t.k = (uint32_t)b->qubits.size();
    t.kind = (uint32_t)tile_kind((int)t.k, split);
    t.mat_off = (uint32_t)(head + mats.size() * sizeof(float));
    std::vector<float> m = tile_matrix(b->mat, (int)t.k, split);
    out.split |= t.kind == 4 || t.kind == 5;
    mats.insert(mats.end(), m.begin(), m.end());
    while (mats.size() % 4) mats.push_back(0.f);  // 16-byte alignment
    for (uint32_t i = 0; i < t.k; ++i) {
      t.loc[i] = (uint32_t)local[b->qubits[i]];
      t.srt[i] = t.loc[i];
    }
    std::sort(t.srt, t.srt + t.k);
    tb.push_back(t);
    out.kmax = std::max(out.kmax, t.k);
    d.gates += (int)b->ops.size();
    d.src.insert(d.src.end(), b->ops.begin(), b->ops.end());
    d.tile_blocks.push_back(orig->qubits);
    d.tile_mfma.push_back((int)t.kind);
  }
  wave_runs(tb, L, g_wave_runs.load() == 1);
  std::vector<uint8_t> blob(head + mats.size() * sizeof(float));
  std::memcpy(blob.data(), tb.data(), head);
  std::memcpy(blob.data() + head, mats.data(), mats.size() * sizeof(float));
  out.blob16 = (uint32_t)(blob.size() / 16);
  out.nblocks = (uint32_t)blocks.size();
  out.off = plan.push(blob.data(), blob.size());
  // Matrix-core work in units of a 3-qubit block (a 2-qubit block on the
  // 4x4x1 path costs half of one, a 4-qubit block counts apart).
  float n3 = 0.f;
  int n4 = 0;
  for (const k::TileBlock& t : tb) {
    if (t.kind == 2) ++n4;
    else n3 += t.kind == 3 ? 0.5f : 1.f;
  }
  d.feat[4] = d.gates / 32.f;
  d.feat[8] = n / 34.f;
  d.feat[12] = 1.f;
  d.feat[13] = L / 12.f;
  d.feat[14] = n3 / 16.f;
  d.feat[15] = n4 / 16.f;
  return out;
}

// Several fused blocks applied to 2^L-amplitude tiles held in LDS: one HBM
// read and write of the state for all of them. The blob (descriptors, then
// matrices) is what every workgroup copies into LDS.
void emit_tile(Plan& plan, std::vector<FusedBlock*>& blocks, const TileMap& tm, const std::vector<Op>& ops,
               int n) {
  DevOp d;
  d.kind = D_TILE;
  d.category = 1;
  d.gates = 0;
  const int L = tm.L;
  d.touched.assign(tm.lq, tm.lq + L);
  std::sort(d.touched.begin(), d.touched.end());
  d.tile_lq.assign(tm.lq, tm.lq + L);
  tm.to_args(d.targs);
  const TileBlob tbl = tile_blob(plan, d, blocks, tm.cq, L, ops, n, g_block_arith.load() == 1);  // LDS order
  d.targs.split = tbl.split ? 1 : 0;
  d.targs.nblocks = tbl.nblocks;
  d.targs.kmax = tbl.kmax;
  d.targs.blob16 = tbl.blob16;
  d.tile_off = tbl.off;
  d.kq = L;
  d.low = 0;
  plan.ops.push_back(std::move(d));
}

// Tile sweeps of a unitary run ops[begin, end):
//  1. fuse to blocks of <= 2 qubits (wider compiled ops stay single blocks);
//  2. group the blocks into sweeps: a sweep's qubits always include those at
//     bit positions 0..3 (a tile row is then a 128-byte run); it repeatedly
//     takes, among the blocks whose earlier untaken blocks share no qubit
//     with them (so the taken order is a valid reordering), the one adding
//     the fewest new qubits, while the union stays within L qubits. Each
//     sweep tries this from the low qubits alone and seeded with each of the
//     first 8 available blocks, and keeps the trial that takes the most
//     blocks. The tile is padded to L qubits with the qubits the next blocks
//     use first, and (relabelling on) the sweep's store brings up to two of
//     them into positions 0..3;
//  3. fuse each sweep's ops again into blocks of <= kt qubits (fewer passes
//     over the tile in LDS) and emit one tile launch per sweep, split where the
//     blob would exceed its LDS budget. A sweep of one block that leaves the
//     layout alone is an ordinary launch; blocks wider than 4 qubits always
//     are, after the identity layout is restored.
void emit_sweeps(Plan& plan, const std::vector<Op>& ops, size_t begin, size_t end, int prec, int n, int L, int kt,
                 const MergeScorer& scorer) {
  constexpr size_t kWindow = 512;
  constexpr int F = 4;
  // LDS per workgroup = 8 * 2^L + blob; the blob cap keeps it within 64 KiB.
  constexpr size_t kTileBlobMax = 32768, kTileBlocksMax = 64;
  const bool relabel = g_sweep_relabel;
  BitLayout lay(n);
  std::vector<FusedBlock> pre = greedy_fuse(ops, begin, end, 2);
  std::vector<uint64_t> mask(pre.size());
  for (size_t b = 0; b < pre.size(); ++b) mask[b] = qubit_mask(pre[b].qubits);
  auto tileable = [&](size_t b) { return !pre[b].qubits.empty() && pre[b].qubits.size() <= 4; };
  std::vector<int> rem(pre.size());
  for (size_t b = 0; b < pre.size(); ++b) rem[b] = (int)b;
  std::vector<char> taken(pre.size(), 0);
  std::vector<int> mark(pre.size(), -1);  // trial number that took the block
  int trial_no = 0;
  // One min-growth trial from `seed` (-1: the low qubits only); returns the
  // blocks it would take, in a valid execution order.
  auto trial = [&](size_t head, int seed, uint64_t& smask) {
    const int gen = ++trial_no;
    std::vector<int> take;
    smask = lay.low_mask(F);
    if (seed >= 0) {
      mark[seed] = gen;
      take.push_back(seed);
      smask |= mask[seed];
    }
    for (;;) {
      uint64_t blocked = 0;
      int best = -1, best_grow = 99;
      size_t seen = 0;
      for (size_t r = head; r < rem.size() && seen < kWindow; ++r) {
        const int b = rem[r];
        if (taken[b] || mark[b] == gen) continue;
        ++seen;
        if (!(mask[b] & blocked) && tileable((size_t)b) && __builtin_popcountll(smask | mask[b]) <= L) {
          const int grow = __builtin_popcountll(mask[b] & ~smask);
          if (grow < best_grow) {
            best = b;
            best_grow = grow;
            if (grow == 0) break;
          }
        }
        blocked |= mask[b];
      }
      if (best < 0 || take.size() >= 4 * kTileBlocksMax) break;
      mark[best] = gen;
      take.push_back(best);
      smask |= mask[best];
    }
    return take;
  };
  auto restore = [&]() {
    restore_layout(lay, L, F, [&](const TileMap& tm) {
      std::vector<FusedBlock*> none;
      emit_tile(plan, none, tm, ops, n);
    });
  };
  size_t head = 0;
  while (head < rem.size()) {
    // rem[head..] holds the untaken blocks in program order.
    const int h = rem[head];
    if (!tileable((size_t)h)) {
      restore();
      emit_block(plan, pre[h], ops, prec, n);
      taken[h] = 1;
      ++head;
      continue;
    }
    // Multi-start: grow from the low qubits alone and from each of the first
    // kStarts available blocks; keep the trial that takes the most blocks.
    constexpr int kStarts = 8;
    uint64_t smask = 0;
    std::vector<int> sweep = trial(head, -1, smask);
    {
      uint64_t blocked = 0;
      int starts = 0;
      for (size_t r = head; r < rem.size() && starts < kStarts; ++r) {
        const int b = rem[r];
        if (taken[b]) continue;
        if (!(mask[b] & blocked) && tileable((size_t)b) &&
            __builtin_popcountll(lay.low_mask(F) | mask[b]) <= L) {
          ++starts;
          uint64_t m2 = 0;
          std::vector<int> t = trial(head, b, m2);
          if (t.size() > sweep.size()) {
            sweep = std::move(t);
            smask = m2;
          }
        }
        blocked |= mask[b];
      }
    }
    for (int b : sweep) taken[b] = 1;
    // The head block is always available and fits (its <= 4 qubits with the
    // 4 low ones), so the sweep is never empty.
    // First use of each qubit among the blocks still to come (window).
    std::vector<uint32_t> first(64, UINT32_MAX);
    bool more = false;
    {
      size_t seen = 0;
      for (size_t r = head; r < rem.size() && seen < kWindow; ++r) {
        const int b = rem[r];
        if (taken[b]) continue;
        more = true;
        for (uint32_t q : pre[b].qubits)
          if (first[q] == UINT32_MAX) first[q] = (uint32_t)seen;
        ++seen;
      }
    }
    if (relabel) {
      std::vector<uint32_t> cand;
      for (int p = 0; p < n; ++p) cand.push_back(lay.at[p]);
      std::stable_sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) { return first[a] < first[b]; });
      for (uint32_t q : cand) {
        if (__builtin_popcountll(smask) >= L) break;
        smask |= 1ull << q;
      }
    } else {
      for (int p = 0; __builtin_popcountll(smask) < L; ++p) smask |= 1ull << lay.at[p];
    }
    const bool was_identity = lay.identity();
    const TileMap tm =
        sweep_map(lay, smask, L, F, relabel ? choose_swaps(lay, smask, L, F, first, !more)
                                            : std::vector<std::pair<uint32_t, uint32_t>>{});
    std::vector<int> order;
    for (int b : sweep) order.insert(order.end(), pre[b].ops.begin(), pre[b].ops.end());
    std::vector<FusedBlock> blocks = greedy_fuse_list(ops, order, std::max(1, std::min(kt, 4)), scorer);
    size_t i = 0;
    while (i < blocks.size()) {
      size_t blob = 0, j = i;
      while (j < blocks.size() && j - i < kTileBlocksMax) {
        const size_t bb = sizeof(k::TileBlock) + tile_matrix_bytes((int)blocks[j].qubits.size());
        if (blob + bb > kTileBlobMax) break;
        blob += bb;
        ++j;
      }
      // Only the sweep's last launch relabels.
      const TileMap& m = j == blocks.size() ? tm : tm.plain();
      if (j == i + 1 && was_identity && !m.relabels()) {
        emit_block(plan, blocks[i], ops, prec, n);
      } else {
        std::vector<FusedBlock*> bp;
        for (size_t b = i; b < j; ++b) bp.push_back(&blocks[b]);
        emit_tile(plan, bp, m, ops, n);
      }
      i = j;
    }
    while (head < rem.size() && taken[rem[head]]) ++head;
  }
  restore();
}

// One out-of-place sweep launch: tile of the qubits local[0..L) (local bit j
// holds local[j]), read from the layout `in_pos` (qubit -> bit position;
// contiguous when local bit j sits at position j) and written to the layout
// `out_pos`, with `blocks` applied in LDS.
// Cost-model features of an out-of-place sweep launch (the learned planner's
// inputs; they replace tile_blob's in-place sweep features):
//   0 1 (out-of-place sweep)       1 3-qubit blocks with a tile bit < 4, /16
//   2 2-qubit blocks with a tile bit < 6, /16   3 2-qubit blocks, /16
//   4 gates/32                     5 gather (first sweep of a run)
//   6 destination positions 0..3 stored from tile bits >= 4, /4
//   7 lowest destination positions (from 0) stored from tile bits, /12
//   8 n/34                         9 512-thread workgroups
//  10 persistent workgroups       11 blob bytes/32768
//  12 1 (tile sweep)              13 L/12
//  14 3-qubit blocks, /16         15 4-qubit blocks, /16
// The matrix-core block routines spread a wave's lanes over the lowest tile
// bits outside the block (16 lanes for 3 qubits, 64 for 2), so blocks on low
// bits read LDS in strided slots; so does a store whose lowest destination
// positions come from high tile bits.
void xsweep_features(DevOp& d, const std::vector<FusedBlock*>& blocks, const uint32_t* local, int L, int n) {
  const k::XSweepArgs& a = d.xargs;
  std::vector<int> pos(64, -1);
  for (int j = 0; j < L; ++j) pos[local[j]] = j;
  int n3 = 0, n3lo = 0, n2 = 0, n2lo = 0, n4 = 0;
  for (const FusedBlock* b : blocks) {
    int lo = 99;
    for (uint32_t q : b->qubits) lo = std::min(lo, pos[q]);
    const size_t k = b->qubits.size();
    if (k <= 2) {
      ++n2;
      n2lo += lo < 6;
    } else if (k == 3) {
      ++n3;
      n3lo += lo < 4;
    } else {
      ++n4;
    }
  }
  int strided = 0, run = 0;
  for (int j = 0; j < L; ++j) {
    if (a.wpos[j] < 4 && a.wlocal[j] >= 4) ++strided;
    if (a.wpos[j] == (uint32_t)run) ++run;
  }
  float* f = d.feat;
  std::fill(f, f + mlp::kFeatures, 0.f);
  f[0] = 1.f;
  f[1] = n3lo / 16.f;
  f[2] = n2lo / 16.f;
  f[3] = n2 / 16.f;
  f[4] = d.gates / 32.f;
  f[5] = (float)a.gather;
  f[6] = strided / 4.f;
  f[7] = run / 12.f;
  f[8] = n / 34.f;
  f[9] = a.threads == 512 ? 1.f : 0.f;
  f[10] = d.persist ? 1.f : 0.f;
  f[11] = 16.f * a.blob16 / 32768.f;
  f[12] = 1.f;
  f[13] = L / 12.f;
  f[14] = n3 / 16.f;
  f[15] = n4 / 16.f;
}

// The layout of an out-of-place sweep in d.xargs: tile local bit j holds
// qubit local[j], read from position in_pos[local[j]] and stored at
// out_pos[local[j]]; the tile-index bits are the source positions outside
// the tile (d.xtile_q: their qubits), stored at their out_pos.
void xsweep_layout(DevOp& d, const uint32_t* local, int L, const uint32_t* in_pos, const uint32_t* out_pos, int n,
                   bool persist, int threads) {
  d.kind = D_XSWEEP;
  d.category = 1;
  d.gates = 0;
  d.persist = persist;
  k::XSweepArgs& a = d.xargs;
  a.L = (uint32_t)L;
  a.ntb = (uint32_t)(n - L);
  a.threads = (uint32_t)threads;
  std::vector<int> in_at(64, -1);
  for (int q = 0; q < n; ++q) in_at[in_pos[q]] = q;
  std::vector<char> is_local(64, 0), src_local(64, 0);
  bool contiguous = true;
  for (int j = 0; j < L; ++j) {
    is_local[local[j]] = 1;
    src_local[in_pos[local[j]]] = 1;
    a.rpos[j] = in_pos[local[j]];
    contiguous = contiguous && in_pos[local[j]] == (uint32_t)j;
  }
  a.gather = contiguous ? 0u : 1u;
  std::copy(a.rpos, a.rpos + L, a.rsrt);
  std::sort(a.rsrt, a.rsrt + L);
  // Store order: the tile's destination positions ascending.
  std::vector<std::pair<uint32_t, uint32_t>> dst;  // (destination position, local bit)
  for (int j = 0; j < L; ++j) dst.push_back({out_pos[local[j]], (uint32_t)j});
  std::sort(dst.begin(), dst.end());
  for (int j = 0; j < L; ++j) {
    a.wpos[j] = dst[j].first;
    a.wlocal[j] = dst[j].second;
  }
  // Tile-index bit k: the k-th lowest source position outside the tile.
  uint32_t k = 0;
  for (int p = 0; p < n; ++p)
    if (!src_local[p]) {
      d.xtile_q.push_back((uint32_t)in_at[p]);
      a.wtile[k++] = out_pos[in_at[p]];
    }
  // Both maps must be bijections onto positions 0..n-1 (the kernel's reads
  // and writes then stay inside the state and cover it once).
  uint64_t rseen = 0, wseen = 0;
  for (int j = 0; j < L; ++j) {
    rseen |= 1ull << a.rpos[j];
    wseen |= 1ull << a.wpos[j];
  }
  for (uint32_t t = 0; t < k; ++t) wseen |= 1ull << a.wtile[t];
  const uint64_t all = n == 64 ? ~0ull : (1ull << n) - 1;
  if ((int)k != n - L || __builtin_popcountll(rseen) != L || (rseen & ~all) || wseen != all)
    throw RuntimeErr("internal: bad out-of-place sweep layout");
  d.touched.assign(local, local + L);
  std::sort(d.touched.begin(), d.touched.end());
  d.tile_lq.assign(local, local + L);
}

void emit_xlaunch(Plan& plan, std::vector<FusedBlock*>& blocks, const uint32_t* local, int L, const uint32_t* in_pos,
                  const uint32_t* out_pos, int n, const std::vector<Op>& ops, bool persist, int threads,
                  bool dry = false) {
  DevOp d;
  xsweep_layout(d, local, L, in_pos, out_pos, n, persist, threads);
  k::XSweepArgs& a = d.xargs;
  if (dry) {
    // Cost-model features only: no matrices, nothing in the plan's blob.
    size_t bytes = 0;
    for (const FusedBlock* b : blocks) {
      d.gates += (int)b->ops.size();
      bytes += sizeof(k::TileBlock) + tile_matrix_bytes((int)b->qubits.size());
      a.kmax = std::max<uint32_t>(a.kmax, (uint32_t)std::max<size_t>(2, b->qubits.size()));
    }
    a.nblocks = (uint32_t)blocks.size();
    a.blob16 = (uint32_t)(bytes / 16);
  } else {
    // Split precision runs one workgroup per tile (not in persistent launches).
    const TileBlob tbl = tile_blob(plan, d, blocks, local, L, ops, n, g_block_arith.load() == 1 && !persist);
    a.split = tbl.split ? 1 : 0;
    a.nblocks = tbl.nblocks;
    a.kmax = tbl.kmax;
    a.blob16 = tbl.blob16;
    d.tile_off = tbl.off;
  }
  d.kq = L;
  d.low = 0;
  xsweep_features(d, blocks, local, L, n);
  plan.uses_alt = true;
  plan.ops.push_back(std::move(d));
}

// Out-of-place (permuting) tile sweeps of a unitary run ops[begin, end),
// complex64 with fusion. Sweeps read their tiles from one state buffer and
// write them to the other, so the bit layout (qubit -> position) may change
// at every sweep: each sweep writes its tile so that the next sweep's qubits
// sit at positions 0..L-1, whose tiles are then single contiguous runs of
// 2^L amplitudes. Positions 0..3 between two sweeps hold 4 qubits that both
// tiles share (the first 4 kept qubits in the earlier sweep's tile order):
// a store then moves runs of at least 16 amplitudes (128 bytes), longer by
// the further kept qubits (those go right above them). A run starts and ends
// in the identity layout, so its first and last tiles hold qubits 0..3: the
// first sweep gathers its tile (runs of 16 amplitudes), the last stores to
// the qubits' own positions. With `pinned`, qubits 0..3 are in every tile
// (as in emit_sweeps), which keeps every store's LDS reads in 128-byte runs
// but leaves L - 4 tile qubits for gates; otherwise a tile only has to share
// 4 qubits with its neighbours, and the shared qubits are taken from the
// earlier tile's lowest 4 positions where possible (then its store reads LDS
// in the same 128-byte runs). Blocks are grouped into sweeps greedily (as in
// emit_sweeps); the tile is padded with the qubits the next blocks use first,
// which keeps consecutive sweeps overlapping.
struct XGroup {
  std::vector<int> blocks;  // pre-fused blocks in execution order
  uint64_t q = 0;           // tile qubits (exactly L)
  int standalone = -1;      // a block run as its own launch in the identity layout
};
// Sweeps of a unitary run: the run's ops pre-fused into blocks of <= 2
// qubits, and those blocks grouped into tiles (independent of the block
// width the sweeps fuse to afterwards).
struct XPartition {
  std::vector<FusedBlock> pre;
  std::vector<XGroup> groups;
};
constexpr size_t kXTileBlocksMax = 64;

XPartition group_xsweeps(std::vector<FusedBlock> pre_blocks, int n, int L, bool pinned) {
  constexpr size_t kWindow = 512;
  constexpr size_t kTileBlocksMax = kXTileBlocksMax;
  constexpr int kShared = 4;
  XPartition part;
  const uint64_t anchor = (1ull << std::min(4, n)) - 1;
  std::vector<FusedBlock>& pre = part.pre;
  pre = std::move(pre_blocks);
  std::vector<uint64_t> mask(pre.size());
  for (size_t b = 0; b < pre.size(); ++b) mask[b] = qubit_mask(pre[b].qubits);
  auto tileable = [&](size_t b) { return !pre[b].qubits.empty() && pre[b].qubits.size() <= 4; };
  std::vector<int> rem(pre.size());
  for (size_t b = 0; b < pre.size(); ++b) rem[b] = (int)b;
  std::vector<char> taken(pre.size(), 0);
  std::vector<int> mark(pre.size(), -1);
  int trial_no = 0;
  // A tile that must hold `m` fits when m, plus the qubits still needed to
  // share kShared with the previous tile `prev` (0: the run's first tile),
  // is at most L qubits.
  auto fits = [&](uint64_t m, uint64_t prev) {
    const int need = prev ? std::max(0, kShared - __builtin_popcountll(m & prev)) : 0;
    return __builtin_popcountll(m) + need <= L;
  };
  auto trial = [&](size_t head, int seed, uint64_t base, uint64_t prev, uint64_t& smask) {
    const int gen = ++trial_no;
    std::vector<int> take;
    smask = base;
    if (seed >= 0) {
      mark[seed] = gen;
      take.push_back(seed);
      smask |= mask[seed];
    }
    for (;;) {
      uint64_t blocked = 0;
      int best = -1, best_grow = 99;
      size_t seen = 0;
      for (size_t r = head; r < rem.size() && seen < kWindow; ++r) {
        const int b = rem[r];
        if (taken[b] || mark[b] == gen) continue;
        ++seen;
        if (!(mask[b] & blocked) && tileable((size_t)b) && fits(smask | mask[b], prev)) {
          const int grow = __builtin_popcountll(mask[b] & ~smask);
          if (grow < best_grow) {
            best = b;
            best_grow = grow;
            if (grow == 0) break;
          }
        }
        blocked |= mask[b];
      }
      if (best < 0 || take.size() >= 4 * kTileBlocksMax) break;
      mark[best] = gen;
      take.push_back(best);
      smask |= mask[best];
    }
    return take;
  };
  // The best of a trial from the head and trials seeded by each of the first
  // kStarts blocks that could open the sweep.
  auto select = [&](size_t head, uint64_t base, uint64_t prev, uint64_t& smask) {
    constexpr int kStarts = 8;
    std::vector<int> sweep = trial(head, -1, base, prev, smask);
    uint64_t blocked = 0;
    int starts = 0;
    for (size_t r = head; r < rem.size() && starts < kStarts; ++r) {
      const int b = rem[r];
      if (taken[b]) continue;
      if (!(mask[b] & blocked) && tileable((size_t)b) && fits(base | mask[b], prev)) {
        ++starts;
        uint64_t m2 = 0;
        std::vector<int> t = trial(head, b, base, prev, m2);
        if (t.size() > sweep.size()) {
          sweep = std::move(t);
          smask = m2;
        }
      }
      blocked |= mask[b];
    }
    return sweep;
  };
  // Whether taking `sweep` leaves the run without a next tileable block.
  auto ends_run = [&](size_t head, const std::vector<int>& sweep) {
    const int gen = ++trial_no;
    for (int b : sweep) mark[b] = gen;
    for (size_t r = head; r < rem.size(); ++r) {
      const int b = rem[r];
      if (taken[b] || mark[b] == gen) continue;
      return !tileable((size_t)b);
    }
    return true;
  };
  using Group = XGroup;
  std::vector<Group>& groups = part.groups;
  // The previous tile of the current run (0 at a run's start) and its local
  // order (tile position -> qubit), as the emission below lays them out.
  uint64_t prev = 0;
  std::vector<uint32_t> prev_order;
  size_t head = 0;
  while (head < rem.size()) {
    const int h = rem[head];
    if (!tileable((size_t)h)) {
      Group g;
      g.standalone = h;
      groups.push_back(g);
      taken[h] = 1;
      ++head;
      prev = 0;
      prev_order.clear();
      continue;
    }
    uint64_t base = pinned || !prev ? anchor : 0;
    uint64_t smask = 0;
    std::vector<int> sweep = select(head, base, prev, smask);
    if ((smask & anchor) != anchor && ends_run(head, sweep)) {
      base |= anchor;  // the run's last tile stores to the identity layout
      sweep = select(head, base, prev, smask);
    }
    for (int b : sweep) taken[b] = 1;
    std::vector<uint32_t> first(64, UINT32_MAX);
    {
      size_t seen = 0;
      for (size_t r = head; r < rem.size() && seen < kWindow; ++r) {
        const int b = rem[r];
        if (taken[b]) continue;
        for (uint32_t q : pre[b].qubits)
          if (first[q] == UINT32_MAX) first[q] = (uint32_t)seen;
        ++seen;
      }
    }
    std::vector<uint32_t> cand;
    for (int q = 0; q < n; ++q) cand.push_back((uint32_t)q);
    std::stable_sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) { return first[a] < first[b]; });
    if (prev) {
      // Shared qubits: the previous tile's lowest positions first, then its
      // other qubits in order of next use.
      for (size_t j = 0; j < prev_order.size() && j < (size_t)kShared; ++j)
        if (__builtin_popcountll(smask & prev) < kShared) smask |= 1ull << prev_order[j];
      for (uint32_t q : cand)
        if (__builtin_popcountll(smask & prev) < kShared && ((prev >> q) & 1)) smask |= 1ull << q;
    }
    // Pad the tile with the qubits the next blocks use first.
    for (uint32_t q : cand) {
      if (__builtin_popcountll(smask) >= L) break;
      smask |= 1ull << q;
    }
    if (__builtin_popcountll(smask) != L || (prev && __builtin_popcountll(smask & prev) < kShared))
      throw RuntimeErr("internal: out-of-place sweep tile");
    std::vector<uint32_t> order;
    for (uint32_t q : prev_order)
      if ((smask >> q) & 1) order.push_back(q);
    for (int q = 0; q < n; ++q)
      if (((smask >> q) & 1) && !((prev >> q) & 1)) order.push_back((uint32_t)q);
    prev = smask;
    prev_order = std::move(order);
    Group g;
    g.blocks = std::move(sweep);
    g.q = smask;
    groups.push_back(std::move(g));
    while (head < rem.size() && taken[rem[head]]) ++head;
  }
  return part;
}

// Launches of a partition's sweeps with blocks of <= kt qubits. `dry` makes
// the cost-model features only (no matrices, nothing in the plan's blob) and
// leaves out the standalone blocks, which every candidate shares. The
// layouts are laid out first (each sweep's store layout is the next one's
// read layout); the sweeps' blocks and launches are then independent, and a
// large plan builds them on host threads, each sweep into a plan of its own
// whose blob is appended in sweep order afterwards.
void emit_xpartition(Plan& plan, XPartition& part, const std::vector<Op>& ops, int prec, int n, int L, int kt,
                     const MergeScorer& scorer, bool persist, int threads, bool dry) {
  constexpr size_t kTileBlobMax = 32768, kTileBlocksMax = kXTileBlocksMax;
  std::vector<FusedBlock>& pre = part.pre;
  const std::vector<XGroup>& groups = part.groups;
  struct Job {
    size_t s;
    std::vector<uint32_t> order, cur_pos, next_pos;
    Plan out;
  };
  std::vector<uint32_t> ident(64);
  for (uint32_t q = 0; q < 64; ++q) ident[q] = q;
  size_t nsweeps = 0;
  for (const XGroup& g : groups) nsweeps += g.standalone < 0;
  std::vector<Job> jobs(nsweeps);
  size_t gi = 0, ji = 0;
  while (gi < groups.size()) {
    if (groups[gi].standalone >= 0) {
      ++gi;
      continue;
    }
    size_t gj = gi;
    while (gj < groups.size() && groups[gj].standalone < 0) ++gj;
    // Sweep gi reads the identity layout: local order = its qubits ascending
    // (qubits 0..3 first).
    std::vector<uint32_t> cur_pos = ident, order;
    for (int q = 0; q < n; ++q)
      if ((groups[gi].q >> q) & 1) order.push_back((uint32_t)q);
    for (size_t s = gi; s < gj; ++s) {
      const XGroup& g = groups[s];
      std::vector<uint32_t> next_pos = ident, next_order;
      if (s + 1 < gj) {
        const uint64_t nq = groups[s + 1].q;
        for (uint32_t q : order)
          if ((nq >> q) & 1) next_order.push_back(q);  // kept qubits, in this sweep's order
        for (int q = 0; q < n; ++q)
          if (((nq >> q) & 1) && !((g.q >> q) & 1)) next_order.push_back((uint32_t)q);
        if ((int)next_order.size() != L) throw RuntimeErr("internal: out-of-place sweep tile size");
        uint32_t p = 0;
        for (uint32_t q : next_order) next_pos[q] = p++;
        for (int q = 0; q < n; ++q)
          if (!((nq >> q) & 1)) next_pos[q] = p++;
      }
      Job& jb = jobs[ji++];
      jb.s = s;
      jb.order = order;
      jb.cur_pos = cur_pos;
      jb.next_pos = next_pos;
      cur_pos = std::move(next_pos);
      order = std::move(next_order);
    }
    gi = gj;
  }
  // One sweep: its ops fused again into blocks of <= kt qubits, split into
  // launches where the blob would exceed its LDS budget; only the last
  // launch changes the layout.
  auto build = [&](Job& jb) {
    std::vector<int> opsorder;
    for (int b : groups[jb.s].blocks) opsorder.insert(opsorder.end(), pre[b].ops.begin(), pre[b].ops.end());
    std::vector<FusedBlock> blocks = greedy_fuse_list(ops, opsorder, std::max(1, std::min(kt, 4)), scorer);
    size_t i = 0;
    do {
      size_t blob = 0, j = i;
      while (j < blocks.size() && j - i < kTileBlocksMax) {
        const size_t bb = sizeof(k::TileBlock) + tile_matrix_bytes((int)blocks[j].qubits.size());
        if (blob + bb > kTileBlobMax) break;
        blob += bb;
        ++j;
      }
      std::vector<FusedBlock*> bp;
      for (size_t b = i; b < j; ++b) bp.push_back(&blocks[b]);
      const bool last = j >= blocks.size();
      emit_xlaunch(jb.out, bp, jb.order.data(), L, jb.cur_pos.data(), last ? jb.next_pos.data() : jb.cur_pos.data(),
                   n, ops, persist, threads, dry);
      i = j;
    } while (i < blocks.size());
  };
  const unsigned hw = std::thread::hardware_concurrency();
  const size_t nth = !dry && n >= 24 && jobs.size() >= 8 ? std::min<size_t>({hw ? (size_t)hw : 1, 8, jobs.size()}) : 1;
  if (nth <= 1) {
    for (Job& jb : jobs) build(jb);
  } else {
    std::atomic<size_t> next{0};
    std::vector<std::exception_ptr> errs(nth);
    std::vector<std::thread> th;
    for (size_t t = 0; t < nth; ++t)
      th.emplace_back([&, t] {
        try {
          for (size_t x; (x = next++) < jobs.size();) build(jobs[x]);
        } catch (...) {
          errs[t] = std::current_exception();
        }
      });
    for (std::thread& t : th) t.join();
    for (const std::exception_ptr& e : errs)
      if (e) std::rethrow_exception(e);
  }
  // Into the plan in group order: standalone blocks emitted here, each
  // sweep's launches appended with their blob offsets moved by where the
  // sweep's blob lands (both 256-byte aligned).
  ji = 0;
  for (const XGroup& g : groups) {
    if (g.standalone >= 0) {
      if (!dry) emit_block(plan, pre[g.standalone], ops, prec, n);
      continue;
    }
    Job& jb = jobs[ji++];
    const size_t base = jb.out.blob.empty() ? 0 : plan.push(jb.out.blob.data(), jb.out.blob.size());
    for (DevOp& d : jb.out.ops) {
      d.tile_off += base;
      plan.ops.push_back(std::move(d));
    }
    plan.uses_alt = plan.uses_alt || jb.out.uses_alt;
  }
}

void emit_xsweeps(Plan& plan, const std::vector<Op>& ops, size_t begin, size_t end, int prec, int n, int L, int kt,
                  const MergeScorer& scorer, bool persist, int threads, bool pinned) {
  XPartition part = group_xsweeps(greedy_fuse(ops, begin, end, 2), n, L, pinned);
  emit_xpartition(plan, part, ops, prec, n, L, kt, scorer, persist, threads, false);
}

// Exact complex128 sweeps: maximal runs of consecutive compiled ops (program
// order, nothing reordered) whose qubits fit in L tile qubits together with
// qubits 0..2; each op keeps its per-op kernel's arithmetic, so results are
// bit-identical to one launch per op. Dense ops wider than 3 qubits and
// diagonals with more than 32 entries run as their own launches.
bool exact_tileable(const Op& op) {
  switch (op.kind) {
    case OpKind::One:
    case OpKind::X:
    case OpKind::Swap:
    case OpKind::Noop: return true;
    case OpKind::Diagonal: {
      int ne = 0;
      for (const cplx& e : op.mat) ne += !(e.re == 1.0 && e.im == 0.0);
      return ne <= 32;
    }
    case OpKind::Dense: return op.qubits.size() <= 3;
    default: return false;
  }
}

size_t exact_op_bytes(const Op& op) {
  size_t data = sizeof(k::TileChain);  // at most one chain per op
  if (op.kind == OpKind::One) data = 64;
  else if (op.kind == OpKind::Diagonal) data = 128 + 16 * op.mat.size();
  else if (op.kind == OpKind::Dense) data = 16 * op.mat.size();
  return sizeof(k::TileOp) + ((data + 15) & ~(size_t)15);
}

// The chains and op records of the exact ops [i, j) on a tile whose local
// bit of qubit q is local[q], as k_tile128 and k_xsweep128 read them: adds
// the ops to d's gate count, sources and per-op qubit lists and returns the
// blob (*nchains chains, *nfixed of them on compiled signatures).
std::vector<uint8_t> exact_tile_blob(Plan& plan, DevOp& d, const std::vector<Op>& ops, size_t i, size_t j,
                                     const std::vector<int>& local, uint32_t* nchains, int* nfixed) {
  auto lmask = [&](uint64_t g) {
    uint32_t m = 0;
    for (uint32_t q = 0; q < 64; ++q)
      if ((g >> q) & 1) m |= 1u << local[q];
    return m;
  };
  std::vector<k::TileOp> tops;
  std::vector<uint8_t> data;
  std::vector<std::vector<uint8_t>> datas;
  for (size_t o = i; o < j; ++o) {
    if (ops[o].kind == OpKind::Noop) {
      d.src.push_back((int)o);
      continue;
    }
    DevOp x = lower(ops[o], plan, 16);
    k::TileOp t{};
    t.nbits = x.e.nbits;
    for (uint32_t b = 0; b < x.e.nbits; ++b) {
      t.pos[b] = (uint32_t)local[x.e.pos[b]];
      t.qmask |= 1u << t.pos[b];
    }
    std::sort(t.pos, t.pos + t.nbits);  // local bits need not follow qubit order
    t.fixed = lmask(x.e.fixed);
    std::vector<uint8_t> dd;
    auto put = [&](const void* p, size_t bytes) {
      const uint8_t* c = (const uint8_t*)p;
      dd.insert(dd.end(), c, c + bytes);
    };
    switch (x.kind) {
      case D_ONE:
        t.kind = 0;
        t.bbit = 1u << local[x.target];
        put(x.u, sizeof(x.u));
        break;
      case D_X:
        t.kind = 1;
        t.abit = 0;
        t.bbit = 1u << local[x.target];
        break;
      case D_SWAP:
        t.kind = 1;
        t.abit = 1u << local[x.a];
        t.bbit = 1u << local[x.b];
        break;
      case D_DIAG: {
        t.kind = 2;
        t.k = (uint32_t)x.pat.size();
        uint32_t pats[32] = {};
        for (size_t e = 0; e < x.pat.size(); ++e) pats[e] = lmask(x.pat[e]);
        put(pats, sizeof(pats));
        put(x.dval.data(), x.dval.size() * sizeof(double));
        break;
      }
      case D_DENSE: {
        t.kind = 3;
        t.k = (uint32_t)x.kq;
        for (int b = 0; b < x.kq; ++b) t.q[b] = (uint32_t)local[ops[o].qubits[b]];
        put(plan.blob.data() + x.mat_off, sizeof(double) * 2 * ((size_t)1 << (2 * x.kq)));
        break;
      }
      default: throw RuntimeErr("internal: op kind not allowed in an exact sweep");
    }
    while (dd.size() % 16) dd.push_back(0);
    tops.push_back(t);
    datas.push_back(std::move(dd));
    d.gates += 1;
    d.src.push_back((int)o);
    d.tile_blocks.push_back(op_qubits(ops[o]));
    d.tile_mfma.push_back(0);
  }
  // Register chains: consecutive pair updates, X permutations and diagonals
  // whose qubits together number at most 2 (a 3-qubit chain's registers made
  // the kernel spill to scratch memory); other ops run alone through LDS.
  std::vector<k::TileChain> chains;
  uint32_t umask = 0;
  auto chainable = [](const k::TileOp& t) {
    return (t.kind == 0 || (t.kind == 1 && t.abit == 0) || t.kind == 2) && __builtin_popcount(t.qmask) <= 2;
  };
  for (size_t t = 0; t < tops.size(); ++t) {
    const bool ok = chainable(tops[t]);
    if (ok && !chains.empty() && chains.back().nq > 0 && __builtin_popcount(umask | tops[t].qmask) <= 2) {
      chains.back().count += 1;
      umask |= tops[t].qmask;
    } else {
      chains.push_back({ok ? 1u : 0u, (uint32_t)t, 1u, 0u});
      umask = ok ? tops[t].qmask : 0;
    }
    if (chains.back().nq > 0) {
      uint32_t nq = 0, cb = 0;
      for (uint32_t b = 0; b < 32; ++b)
        if ((umask >> b) & 1) cb |= b << (8 * nq++);
      chains.back().nq = nq;
      chains.back().cbits = cb;
    }
  }
  // A diagonal inside a chain gets its multiplier per chain register
  // (register j holds chain-bit pattern j): bit j of q[0] says whether
  // amplitude j is multiplied, and the (re, im) pairs follow the op's data at
  // byte q[1]. The kernel then multiplies without matching patterns.
  for (const k::TileChain& ch : chains) {
    if (ch.nq == 0) continue;
    uint32_t c[3] = {ch.cbits & 0xFF, (ch.cbits >> 8) & 0xFF, (ch.cbits >> 16) & 0xFF};
    auto cidx = [&](uint32_t m) {
      uint32_t r = 0;
      for (uint32_t b = 0; b < ch.nq; ++b)
        if ((m >> c[b]) & 1) r |= 1u << b;
      return r;
    };
    for (uint32_t o = ch.first; o < ch.first + ch.count; ++o) {
      k::TileOp& t = tops[o];
      if (t.kind != 2) {
        // Pair update or swap: bit p of q[0] marks the register pairs it
        // applies to, pairs (0,1) (2,3) (0,2) (1,3) of the chain registers.
        static const uint32_t pa[4] = {0, 2, 0, 1}, pb[4] = {1, 3, 2, 3};
        const uint32_t tj = cidx(t.bbit), cm = cidx(t.qmask & ~t.bbit), cv = cidx(t.fixed);
        uint32_t pm = 0;
        for (uint32_t pp = 0; pp < 4; ++pp)
          if (pb[pp] < (1u << ch.nq) && pb[pp] == (pa[pp] | tj) && (pa[pp] & tj) == 0 && (pa[pp] & cm) == cv)
            pm |= 1u << pp;
        t.q[0] = pm;
        continue;
      }
      const uint32_t* pats = (const uint32_t*)datas[o].data();
      const double* vals = (const double*)(datas[o].data() + 32 * sizeof(uint32_t));
      const uint32_t qm = cidx(t.qmask);
      double dl[8] = {};
      uint32_t mask = 0;
      for (uint32_t jj = 0; jj < (1u << ch.nq); ++jj)
        for (uint32_t e = 0; e < t.k; ++e)
          if (cidx(pats[e]) == (jj & qm)) {
            mask |= 1u << jj;
            dl[2 * jj] = vals[2 * e];
            dl[2 * jj + 1] = vals[2 * e + 1];
          }
      t.q[0] = mask;
      t.q[1] = (uint32_t)datas[o].size();
      const uint8_t* b = (const uint8_t*)dl;
      datas[o].insert(datas[o].end(), b, b + sizeof(dl));
    }
  }
  // Chains whose op sequence is one of the kernel's compiled signatures
  // (two qubits, k::kChainSigs: updates on either register bit, a controlled
  // swap of registers 1,3 or 2,3, a multiplier on register 3 such as a
  // controlled phase; or one qubit and a single update) run code with every
  // register index fixed at compile time; the signature number goes in bits
  // 8.. of nq.
  int fixed_chains = 0;
  for (k::TileChain& ch : chains) {
    if (ch.nq == 1 && ch.count == 1 && tops[ch.first].kind == 0 && tops[ch.first].q[0] == 0x1) {
      ch.nq = 1u | (1u << 8);
      ++fixed_chains;
      continue;
    }
    if (ch.nq != 2 || ch.count > 3) continue;
    int code[3] = {0, 0, 0};
    bool ok = true;
    for (uint32_t o = 0; o < ch.count && ok; ++o) {
      const k::TileOp& t = tops[ch.first + o];
      if (t.kind == 0 && t.q[0] == 0x3) code[o] = k::CH_U1;
      else if (t.kind == 0 && t.q[0] == 0xC) code[o] = k::CH_U2;
      else if (t.kind == 1 && t.q[0] == 0x8) code[o] = k::CH_S13;
      else if (t.kind == 1 && t.q[0] == 0x2) code[o] = k::CH_S23;
      else if (t.kind == 2 && t.q[0] == 0x8) code[o] = k::CH_D3;
      else ok = false;
    }
    if (!ok) continue;
    for (uint32_t sgn = 1; sgn < k::kChainSigCount; ++sgn)
      if (k::kChainSigs[sgn][0] == code[0] && k::kChainSigs[sgn][1] == code[1] && k::kChainSigs[sgn][2] == code[2]) {
        ch.nq = 2u | (sgn << 8);
        ++fixed_chains;
        break;
      }
  }
  const size_t chead = chains.size() * sizeof(k::TileChain);
  const size_t head = chead + tops.size() * sizeof(k::TileOp);
  size_t off = head;
  for (size_t t = 0; t < tops.size(); ++t) {
    tops[t].data_off = (uint32_t)off;
    off += datas[t].size();
  }
  std::vector<uint8_t> blob(off);
  std::memcpy(blob.data(), chains.data(), chead);
  std::memcpy(blob.data() + chead, tops.data(), tops.size() * sizeof(k::TileOp));
  off = head;
  for (auto& dd : datas) {
    std::memcpy(blob.data() + off, dd.data(), dd.size());
    off += dd.size();
  }
  *nchains = (uint32_t)chains.size();
  *nfixed = fixed_chains;
  return blob;
}

void emit_tile128(Plan& plan, const std::vector<Op>& ops, size_t i, size_t j, const TileMap& tm) {
  DevOp d;
  d.kind = D_TILE;
  d.exact = true;
  d.category = 1;
  d.gates = 0;
  const int L = tm.L;
  std::vector<int> local(64, -1);
  for (int b = 0; b < L; ++b) local[tm.cq[b]] = b;  // LDS order (after the lane swaps)
  d.touched.assign(tm.lq, tm.lq + L);
  std::sort(d.touched.begin(), d.touched.end());
  d.tile_lq.assign(tm.lq, tm.lq + L);
  tm.to_args(d.targs);
  uint32_t nchains = 0;
  int nfixed = 0;
  const std::vector<uint8_t> blob = exact_tile_blob(plan, d, ops, i, j, local, &nchains, &nfixed);
  d.targs.L = (uint32_t)L;
  d.targs.nblocks = nchains;
  d.chains = (int)nchains;
  d.fixed_chains = nfixed;
  d.targs.blob16 = (uint32_t)(blob.size() / 16);
  d.tile_off = plan.push(blob.data(), blob.size());
  d.kq = L;
  d.low = 0;
  d.feat[4] = d.gates / 32.f;
  d.feat[12] = 1.f;
  d.feat[13] = L / 12.f;
  d.feat[14] = d.tile_blocks.size() / 16.f;
  plan.ops.push_back(std::move(d));
}

void emit_exact_sweeps(Plan& plan, const std::vector<Op>& ops, size_t begin, size_t end, int n, int L) {
  // LDS per workgroup = 16 * 2^L + blob; the blob cap keeps it within 64 KiB.
  constexpr size_t kBlobMax = 32768, kOpsMax = 128, kWindow = 512;
  // The qubits at positions 0..2 are always tile qubits: 8 lanes move one
  // contiguous 128 B run. Fixing 0..3 (256 B runs) needed 171 sweeps instead
  // of 150 at 32 qubits and measured 3.6 % slower.
  constexpr int F = 3;
  const bool relabel = g_sweep_relabel;
  BitLayout lay(n);
  auto plain = [&](size_t o) {
    plan.ops.push_back(lower(ops[o], plan, 16));
    plan.ops.back().touched = op_qubits(ops[o]);
    block_features(plan.ops.back().touched, {(int)o}, ops, n, plan.ops.back().feat);
    plan.ops.back().src = {(int)o};
  };
  auto restore = [&]() {
    restore_layout(lay, L, F, [&](const TileMap& tm) { emit_tile128(plan, ops, 0, 0, tm); });
  };
  size_t i = begin;
  while (i < end) {
    if (!exact_tileable(ops[i])) {
      restore();
      plain(i++);
      continue;
    }
    uint64_t smask = lay.low_mask(F) | qubit_mask(op_qubits(ops[i]));
    size_t blob = exact_op_bytes(ops[i]);
    size_t j = i + 1;
    while (j < end && j - i < kOpsMax && exact_tileable(ops[j])) {
      const uint64_t m = qubit_mask(op_qubits(ops[j]));
      const size_t bb = exact_op_bytes(ops[j]);
      if (__builtin_popcountll(smask | m) > L || blob + bb > kBlobMax) break;
      smask |= m;
      blob += bb;
      ++j;
    }
    // First use of each qubit among the ops that follow (window).
    std::vector<uint32_t> first(64, UINT32_MAX);
    for (size_t o = j; o < end && o < j + kWindow; ++o)
      for (uint32_t q : op_qubits(ops[o]))
        if (first[q] == UINT32_MAX) first[q] = (uint32_t)(o - j);
    const bool last = j == end || !exact_tileable(ops[j]);
    if (relabel) {
      std::vector<uint32_t> cand;
      for (int p = 0; p < n; ++p) cand.push_back(lay.at[p]);
      std::stable_sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) { return first[a] < first[b]; });
      for (uint32_t q : cand) {
        if (__builtin_popcountll(smask) >= L) break;
        smask |= 1ull << q;
      }
    } else {
      for (int p = 0; __builtin_popcountll(smask) < L; ++p) smask |= 1ull << lay.at[p];
    }
    const bool was_identity = lay.identity();
    const TileMap tm =
        sweep_map(lay, smask, L, F, relabel ? choose_swaps(lay, smask, L, F, first, last)
                                            : std::vector<std::pair<uint32_t, uint32_t>>{});
    if (j - i == 1 && was_identity && !tm.relabels()) plain(i);
    else emit_tile128(plan, ops, i, j, tm);
    i = j;
  }
  restore();
}

// One exact out-of-place sweep: ops [i, j) in program order on the tile
// whose local bit b holds qubit local[b], read in layout in_pos and stored in
// layout out_pos.
void emit_exact_xlaunch(Plan& plan, const std::vector<Op>& ops, size_t i, size_t j, const uint32_t* local, int L,
                        const uint32_t* in_pos, const uint32_t* out_pos, int n) {
  DevOp d;
  xsweep_layout(d, local, L, in_pos, out_pos, n, false, 256);
  d.exact = true;
  std::vector<int> lpos(64, -1);
  for (int b = 0; b < L; ++b) lpos[local[b]] = b;
  uint32_t nchains = 0;
  int nfixed = 0;
  const std::vector<uint8_t> blob = exact_tile_blob(plan, d, ops, i, j, lpos, &nchains, &nfixed);
  d.xargs.nblocks = nchains;
  d.xargs.blob16 = (uint32_t)(blob.size() / 16);
  d.chains = (int)nchains;
  d.fixed_chains = nfixed;
  d.tile_off = plan.push(blob.data(), blob.size());
  d.kq = L;
  d.low = 0;
  d.feat[4] = d.gates / 32.f;
  d.feat[12] = 1.f;
  d.feat[13] = L / 12.f;
  d.feat[14] = d.tile_blocks.size() / 16.f;
  plan.uses_alt = true;
  plan.ops.push_back(std::move(d));
}

// Exact complex128 sweeps out of place (a second state buffer): runs of
// consecutive tileable compiled ops, in program order as in
// emit_exact_sweeps (each op keeps its per-op kernel's arithmetic), on tiles
// of any L qubits. Each sweep reads its tile as one contiguous run (a run's
// first sweep gathers it from the identity layout) and stores it in the
// layout the next sweep reads; a run's last sweep stores the identity
// layout. Stores move runs of at least 8 amplitudes (128 bytes): consecutive
// tiles share at least 3 qubits, and a run's first and last tiles hold
// qubits 0..2. Ops on more than 5 qubits run as their own launches (with the
// anchor and the shared qubits they would not fit a tile).
void emit_exact_xsweeps(Plan& plan, const std::vector<Op>& ops, size_t begin, size_t end, int n, int L) {
  constexpr size_t kBlobMax = 32768, kOpsMax = 128, kWindow = 512;
  constexpr int kShared = 3;
  const uint64_t anchor = 7;
  auto tileable = [&](size_t o) { return exact_tileable(ops[o]) && op_qubits(ops[o]).size() <= 5; };
  auto plain = [&](size_t o) {
    plan.ops.push_back(lower(ops[o], plan, 16));
    plan.ops.back().touched = op_qubits(ops[o]);
    block_features(plan.ops.back().touched, {(int)o}, ops, n, plan.ops.back().feat);
    plan.ops.back().src = {(int)o};
  };
  std::vector<uint32_t> ident(64);
  for (uint32_t q = 0; q < 64; ++q) ident[q] = q;
  size_t i = begin;
  while (i < end) {
    if (!tileable(i)) {
      plain(i++);
      continue;
    }
    size_t e = i;
    while (e < end && tileable(e)) ++e;
    // Sweeps of the run [i, e): op ranges and tiles.
    struct Sweep {
      size_t a, b;
      uint64_t q;
    };
    std::vector<Sweep> sw;
    uint64_t prev = 0;
    std::vector<uint32_t> prev_order;
    for (size_t a = i; a < e;) {
      const uint64_t base = prev ? 0 : anchor;
      auto fits = [&](uint64_t m, bool last) {
        m |= base | (last ? anchor : 0);
        const int need = prev ? std::max(0, kShared - __builtin_popcountll(m & prev)) : 0;
        return __builtin_popcountll(m) + need <= L;
      };
      uint64_t m = qubit_mask(op_qubits(ops[a]));
      size_t blob = exact_op_bytes(ops[a]), b = a + 1;
      if (!fits(m, b == e)) throw RuntimeErr("internal: exact out-of-place sweep tile");
      while (b < e && b - a < kOpsMax) {
        const uint64_t mb = qubit_mask(op_qubits(ops[b]));
        const size_t bb = exact_op_bytes(ops[b]);
        if (blob + bb > kBlobMax || !fits(m | mb, b + 1 == e)) break;
        m |= mb;
        blob += bb;
        ++b;
      }
      m |= base | (b == e ? anchor : 0);
      // Shared qubits: the previous tile's lowest positions first, then its
      // other qubits in order of next use; then the qubits the next ops use
      // first.
      std::vector<uint32_t> first(64, UINT32_MAX);
      for (size_t o = b; o < e && o < b + kWindow; ++o)
        for (uint32_t q : op_qubits(ops[o]))
          if (first[q] == UINT32_MAX) first[q] = (uint32_t)(o - b);
      std::vector<uint32_t> cand;
      for (int q = 0; q < n; ++q) cand.push_back((uint32_t)q);
      std::stable_sort(cand.begin(), cand.end(), [&](uint32_t x, uint32_t y) { return first[x] < first[y]; });
      if (prev) {
        for (size_t t = 0; t < prev_order.size() && t < (size_t)kShared; ++t)
          if (__builtin_popcountll(m & prev) < kShared) m |= 1ull << prev_order[t];
        for (uint32_t q : cand)
          if (__builtin_popcountll(m & prev) < kShared && ((prev >> q) & 1)) m |= 1ull << q;
      }
      for (uint32_t q : cand) {
        if (__builtin_popcountll(m) >= L) break;
        m |= 1ull << q;
      }
      if (__builtin_popcountll(m) != L || (prev && __builtin_popcountll(m & prev) < kShared))
        throw RuntimeErr("internal: exact out-of-place sweep tile");
      std::vector<uint32_t> order;
      for (uint32_t q : prev_order)
        if ((m >> q) & 1) order.push_back(q);
      for (int q = 0; q < n; ++q)
        if (((m >> q) & 1) && !((prev >> q) & 1)) order.push_back((uint32_t)q);
      sw.push_back({a, b, m});
      prev = m;
      prev_order = std::move(order);
      a = b;
    }
    // Layouts (as emit_xpartition lays them out) and launches.
    std::vector<uint32_t> cur_pos = ident, order;
    for (int q = 0; q < n; ++q)
      if ((sw[0].q >> q) & 1) order.push_back((uint32_t)q);
    for (size_t s = 0; s < sw.size(); ++s) {
      std::vector<uint32_t> next_pos = ident, next_order;
      if (s + 1 < sw.size()) {
        const uint64_t nq = sw[s + 1].q;
        for (uint32_t q : order)
          if ((nq >> q) & 1) next_order.push_back(q);
        for (int q = 0; q < n; ++q)
          if (((nq >> q) & 1) && !((sw[s].q >> q) & 1)) next_order.push_back((uint32_t)q);
        if ((int)next_order.size() != L) throw RuntimeErr("internal: exact out-of-place sweep tile size");
        uint32_t p = 0;
        for (uint32_t q : next_order) next_pos[q] = p++;
        for (int q = 0; q < n; ++q)
          if (!((nq >> q) & 1)) next_pos[q] = p++;
      }
      emit_exact_xlaunch(plan, ops, sw[s].a, sw[s].b, order.data(), L, cur_pos.data(), next_pos.data(), n);
      cur_pos = std::move(next_pos);
      order = std::move(next_order);
    }
    i = e;
  }
}

// `permute`: complex64 fused sweeps may run out of place (emit_xsweeps) when
// the process-wide sweep mode allows it; the caller then gives the executor
// room for a second state buffer.
Plan build_plan(const std::vector<Op>& ops, size_t begin, size_t end, int prec, int fusion, int max_fused, int n,
                int tile, bool permute = false) {
  Plan plan;
  const int mode = g_sweep_mode.load();
  permute = permute && mode != 0;
  size_t i = begin;
  const int lc = std::min(tile, 11);  // complex128 tiles: 16 * 2^L bytes of LDS
  while (i < end) {
    if (prec == 16 && fusion == FUSE_NONE && tile >= kTileMin && n >= lc + 4 && ops[i].is_unitary()) {
      size_t j = i;
      while (j < end && ops[j].is_unitary()) ++j;
      if (permute) emit_exact_xsweeps(plan, ops, i, j, n, lc);
      else emit_exact_sweeps(plan, ops, i, j, n, lc);
      i = j;
      continue;
    }
    if (!ops[i].is_unitary() || fusion == FUSE_NONE) {
      plan.ops.push_back(lower(ops[i], plan, prec));
      if (ops[i].is_unitary()) {
        plan.ops.back().touched = op_qubits(ops[i]);
        block_features(plan.ops.back().touched, {(int)i}, ops, n, plan.ops.back().feat);
      }
      plan.ops.back().src = {(int)i};
      ++i;
      continue;
    }
    size_t j = i;
    while (j < end && ops[j].is_unitary()) ++j;
    MergeScorer scorer = nullptr;
    if (fusion == FUSE_LEARNED && !g_planner.empty())
      scorer = [n](const FusedBlock& a, const FusedBlock& b, const std::vector<Op>& o) {
        return planner_score(a, b, o, n);
      };
    if (prec == 8 && tile >= kTileMin && tile <= kXTileMax && n >= tile + 4) {
      if (permute) {
        // 13-qubit tiles: 512 threads, one workgroup per tile.
        const bool persist = tile <= kTileMax && (mode == 1 || mode == 4);
        const int threads = tile > kTileMax || (mode >= 3 && tile >= 11) ? 512 : 256;
        const int def_k = std::max(1, std::min(max_fused, 4));
        const bool def_pin = g_sweep_pinned;
        int best_k = def_k;
        bool best_pin = def_pin;
        if (fusion == FUSE_LEARNED && !g_planner.empty()) {
          // Choose (tile rule, block width) by the model's predicted total
          // time: the candidate predicted fastest (the default, the
          // process-wide tile rule with blocks of up to max_fused qubits, on
          // a tie).
          // Host threads: the two tile rules' groupings, then one dry
          // emission per (rule, width) candidate.
          const std::vector<FusedBlock> pre = greedy_fuse(ops, i, j, 2);
          XPartition parts[2];
          double pred[2][5] = {};
          std::exception_ptr err[2][5];
          auto run_all = [&](std::vector<std::pair<int, int>> jobs, const std::function<void(int, int)>& fn) {
            std::vector<std::thread> th;
            for (size_t t = 1; t < jobs.size(); ++t)
              th.emplace_back([&, t] {
                try {
                  fn(jobs[t].first, jobs[t].second);
                } catch (...) {
                  err[jobs[t].first][jobs[t].second] = std::current_exception();
                }
              });
            try {
              fn(jobs[0].first, jobs[0].second);
            } catch (...) {
              err[jobs[0].first][jobs[0].second] = std::current_exception();
            }
            for (std::thread& t : th) t.join();
            for (auto& row : err)
              for (const std::exception_ptr& e : row)
                if (e) std::rethrow_exception(e);
          };
          run_all({{0, 0}, {1, 0}}, [&](int pc, int) { parts[pc] = group_xsweeps(pre, n, tile, pc == 1); });
          std::vector<std::pair<int, int>> cands;
          for (int pc = 0; pc < 2; ++pc)
            for (int kt = std::min(2, def_k); kt <= def_k; ++kt) cands.push_back({pc, kt});
          // Dry emissions (features only; candidates share the partition),
          // then the chosen one emitted in full.
          run_all(cands, [&](int pc, int kt) {
            Plan cand;
            emit_xpartition(cand, parts[pc], ops, prec, n, tile, kt, nullptr, persist, threads, true);
            for (const DevOp& d : cand.ops) pred[pc][kt] += predicted_ms(d, n);
          });
          double best = pred[def_pin ? 1 : 0][def_k];
          for (int pc = 0; pc < 2; ++pc)
            for (int kt = std::min(2, def_k); kt <= def_k; ++kt)
              if (pred[pc][kt] < best) {
                best = pred[pc][kt];
                best_k = kt;
                best_pin = pc == 1;
              }
          plan.learned_choices.push_back((best_pin ? 100 : 0) + best_k);
          emit_xpartition(plan, parts[best_pin ? 1 : 0], ops, prec, n, tile, best_k, nullptr, persist, threads,
                          false);
        } else {
          emit_xsweeps(plan, ops, i, j, prec, n, tile, best_k, nullptr, persist, threads, best_pin);
        }
      } else if (fusion == FUSE_LEARNED && !g_planner.empty()) {
        // Choose (block width, tile size) by the model's predicted total
        // time; leave the default (blocks <= 3 qubits, the largest tile) only
        // for a predicted gain of at least 3 %.
        const int def_k = std::min(max_fused, 3), def_l = std::min(tile, kTileMax);
        int best_k = def_k, best_l = def_l;
        double best = 1e300, def_t = 1e300;
        for (int kt = 2; kt <= std::max(2, std::min(max_fused, 4)); ++kt)
          for (int l = kTileMin; l <= def_l; ++l) {
            if (n < l + 4) continue;
            Plan cand;
            const int k = std::min(kt, max_fused);
            emit_sweeps(cand, ops, i, j, prec, n, l, k, nullptr);
            double t = 0;
            for (const DevOp& d : cand.ops) t += predicted_ms(d, n);
            if (k == def_k && l == def_l) def_t = t;
            if (t < best) {
              best = t;
              best_k = k;
              best_l = l;
            }
          }
        if (best > 0.97 * def_t) {
          best_k = def_k;
          best_l = def_l;
        }
        emit_sweeps(plan, ops, i, j, prec, n, best_l, best_k, nullptr);
      } else {
        emit_sweeps(plan, ops, i, j, prec, n, std::min(tile, kTileMax), max_fused, scorer);
      }
    } else {
      std::vector<FusedBlock> blocks = greedy_fuse(ops, i, j, std::max(1, std::min(max_fused, 5)), scorer);
      for (FusedBlock& blk : blocks) emit_block(plan, blk, ops, prec, n);
    }
    i = j;
  }
  return plan;
}

// ---------------------------------------------------------------------------
// Executor
// ---------------------------------------------------------------------------
// Launches of a unitary plan op on `nb` states of `stride` amplitudes at `s`
// (plan constants at `cbase`); false for measure, reset and no-op, which need
// the executor's buffers or do nothing.
bool run_unitary(const DevOp& d, int prec, void* s, uint64_t stride, uint32_t nb, int n, const uint8_t* cbase,
                 k::Flag f, bool mfma_enabled, hipStream_t st) {
  switch (d.kind) {
    case D_ONE: k::one(prec, s, stride, nb, n, d.e, d.target, d.u, f, st); return true;
    case D_X: k::xgate(prec, s, stride, nb, n, d.e, d.target, f, st); return true;
    case D_SWAP: k::swap(prec, s, stride, nb, n, d.e, d.a, d.b, f, st); return true;
    case D_DIAG:
      k::diag(prec, s, stride, nb, n, d.e, (int)d.pat.size(), d.pat.data(), d.dval.data(), f, st);
      return true;
    case D_DENSE: k::dense(prec, s, stride, nb, n, d.kq, d.e, cbase + d.mat_off, d.offsets, f, st); return true;
    case D_FUSED:
      if (d.mfma && mfma_enabled &&
          k::fused_block_mfma(s, stride, nb, n, d.kq, d.e, (const float*)(cbase + d.emb_off), d.offsets, st))
        return true;
      k::fused_block(s, stride, nb, n, d.kq, d.e, (const float*)(cbase + d.mat_off), d.offsets, st);
      return true;
    case D_TILE:
      if (d.exact) k::tile_sweep_c128(s, stride, nb, n, d.targs, cbase + d.tile_off, st);
      else k::tile_sweep(s, stride, nb, n, d.targs, cbase + d.tile_off, st);
      return true;
    default: return false;
  }
}

// Zero start of a plan that runs from |0...0> with a second state buffer:
// whether its first launch is a one-shot out-of-place sweep that synthesises
// its source, and (masks) the zero-tile masks of its leading sweeps in run
// order. While some qubit has not yet been in any tile it is 0 in every
// amplitude, so a sweep's tiles whose index sets such a qubit hold zeros
// (bit b of a mask: tile-index bit b is such a qubit).
bool zero_start_masks(const Plan& plan, std::vector<uint64_t>* masks) {
  masks->clear();
  const DevOp* first = nullptr;
  for (const DevOp& d : plan.ops)
    if (d.kind != D_NOOP) {
      first = &d;
      break;
    }
  if (!g_zero_start.load() || !first || first->kind != D_XSWEEP || first->persist || !first->conds.empty())
    return false;
  uint64_t support = 0;  // qubits that may be non-zero in some amplitude
  for (const DevOp& d : plan.ops) {
    if (d.kind == D_NOOP) continue;
    if (d.kind != D_XSWEEP || d.persist || !d.conds.empty()) break;
    uint64_t mask = 0;
    for (size_t b = 0; b < d.xtile_q.size(); ++b)
      if (!((support >> d.xtile_q[b]) & 1)) mask |= 1ull << b;
    if (mask == 0) break;
    masks->push_back(mask);
    for (uint32_t q : d.tile_lq) support |= 1ull << q;
  }
  return true;
}

// Fractions of out-of-place sweep number i (run order) under a zero start:
// of the source read from HBM (0 for the synthesised first source), and of
// the tiles whose blocks run (the others store zeros).
double sweep_read_fraction(bool started, const std::vector<uint64_t>& masks, size_t i) {
  if (i == 0 && started) return 0.0;
  return i < masks.size() ? std::ldexp(1.0, -__builtin_popcountll(masks[i])) : 1.0;
}
double sweep_work_fraction(const std::vector<uint64_t>& masks, size_t i) {
  return i < masks.size() ? std::ldexp(1.0, -__builtin_popcountll(masks[i])) : 1.0;
}

struct Exec {
  int prec, n;
  uint64_t nc, ncs, stride, c1;
  uint32_t batch;
  hipStream_t st;
  DeviceMem owned, cbits, rng, ms, flags, err, red_a, red_b;
  DeviceMem alt_mem;      // second state buffer of out-of-place sweeps
  void* sp = nullptr;     // state of slot 0; slots follow at `stride` amplitudes
  void* alt = nullptr;    // the other buffer (out-of-place sweeps write here, then the two trade roles)
  const uint8_t* cbase = nullptr;
  bool mfma_enabled = true;
  bool zero_source = false;         // the next out-of-place sweep reads |0...0> (nothing was written)
  bool zero_started = false;        // the last start_zero left the state unwritten
  std::vector<uint64_t> zero_masks;  // zero-tile masks of the leading out-of-place sweeps (run order)
  size_t next_sweep = 0;


  // Every slot to |0...0> before a plan runs from its start. When the plan
  // opens with one-shot out-of-place sweeps, nothing is written here: the
  // first sweep synthesises its source (tile 0 holds amplitude 1 at local
  // index 0, every other tile stores zeros), and while some qubit has not
  // yet been in any tile it is 0 in every amplitude, so the next sweeps'
  // tiles whose index sets it store zeros without reading (zero_mask). The
  // result is the same bit for bit: every block maps a zero tile to zeros.
  void start_zero(const Plan& plan) {
    next_sweep = 0;
    zero_source = zero_start_masks(plan, &zero_masks) && alt != nullptr;
    if (!zero_source) zero_masks.clear();
    zero_started = zero_source;
    if (!zero_source) k::init_zero(prec, sp, stride, batch, n, st);
  }


  // Plan constants, and the second state buffer when the plan sweeps out of
  // place (only for states the executor owns).
  void prepare(const Plan& plan) {
    cbase = plan.base();
    if (plan.uses_alt && alt == nullptr) {
      if (!owned.get()) throw RuntimeErr("internal: out-of-place sweeps on a borrowed state");
      alt_mem.alloc(stride * (uint64_t)prec * batch);
      alt = alt_mem.get();
    }
  }
  // After the plan: `owned` holds the current state again.
  void settle() {
    if (alt != nullptr && sp == alt_mem.get()) {
      std::swap(owned, alt_mem);
      sp = owned.get();
      alt = alt_mem.get();
    }
  }

  // With `external` set, the executor works on a caller-owned state.
  Exec(int prec_, int n_, uint64_t nc_, uint32_t batch_, hipStream_t st_, void* external = nullptr)
      : prec(prec_), n(n_), nc(nc_), batch(batch_), st(st_) {
    ncs = nc ? nc : 1;
    stride = 1ull << n;
    const int seg1 = n < 11 ? n : 11;
    c1 = 1ull << (n - seg1);
    if (external) {
      sp = external;
    } else {
      owned.alloc(stride * (uint64_t)prec * batch);
      sp = owned.get();
    }
    cbits.alloc(ncs * batch);
    rng.alloc(sizeof(k::RngState) * batch);
    ms.alloc(sizeof(k::MeasState) * batch);
    flags.alloc(sizeof(int) * batch);
    err.alloc(sizeof(int));
    red_a.alloc(16 * c1 * batch);
    red_b.alloc(16 * c1 * batch);
    HIPCHK(hipMemsetAsync(cbits.get(), 0, ncs * batch, st));
    HIPCHK(hipMemsetAsync(err.get(), 0, sizeof(int), st));
    HIPCHK(hipMemsetAsync(ms.get(), 0, sizeof(k::MeasState) * batch, st));
  }

  void check_err() {
    HIPCHK(hipGetLastError());
    int e = 0;
    HIPCHK(hipMemcpyAsync(&e, err.get(), sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (e & 1) throw RuntimeErr("Cannot measure a state with nonfinite or zero norm");
    if (e & 2) throw RuntimeErr("Invalid state normalization when sampling");
  }

  // Total |a|^2 split by `branch` for each slot; returns the device pointer
  // to the double2 totals (stride c1).
  const void* reduce_prob(int branch, uint32_t nb) {
    k::ProbSpec ps;
    ps.branch = branch;
    const int seg1 = n < 11 ? n : 11;
    k::seg_reduce_prob(prec, sp, stride, nb, n, ps, seg1, red_a.get(), c1, st);
    uint64_t count = 1ull << (n - seg1);
    void* src = red_a.get();
    void* dst = red_b.get();
    while (count > 1) {
      int lg = 0;
      while ((1ull << lg) < count) ++lg;
      int s = lg < 11 ? lg : 11;
      k::seg_reduce_f64(src, c1, count, nb, s, dst, c1, st);
      count >>= s;
      std::swap(src, dst);
    }
    return src;
  }

  void measure(uint32_t q, int64_t cbit, k::Flag f, uint32_t nb) {
    const void* tot = reduce_prob((int)q, nb);
    k::meas_decide(tot, c1, nb, ms.as<k::MeasState>(), rng.as<k::RngState>(), cbits.as<uint8_t>(), ncs, cbit, f,
                   err.as<int>(), st);
    k::collapse(prec, sp, stride, nb, n, q, ms.as<k::MeasState>(), st);
  }

  void run(const DevOp& d, uint32_t nb) {
    k::Flag f;
    if (!d.conds.empty()) {
      k::cond_eval(cbits.as<uint8_t>(), ncs, nb, (const k::Cond*)(cbase + d.cond_off), (int)d.conds.size(),
                   flags.as<int>(), st);
      f.ptr = flags.as<int>();
      f.stride = 1;
    }
    void* s = sp;
    if (run_unitary(d, prec, s, stride, nb, n, cbase, f, mfma_enabled, st)) return;
    switch (d.kind) {
      case D_XSWEEP: {
        if (!d.conds.empty() || alt == nullptr) throw RuntimeErr("internal: out-of-place sweep without its buffers");
        k::XSweepArgs a = d.xargs;
        a.from_zero = zero_source ? 1u : 0u;
        a.zero_mask = next_sweep < zero_masks.size() ? zero_masks[next_sweep] : 0;
        zero_source = false;
        ++next_sweep;
        if (d.exact) k::xsweep128(sp, alt, stride, nb, n, a, cbase + d.tile_off, st);
        else k::xsweep(sp, alt, stride, nb, n, a, cbase + d.tile_off, d.persist, st);
        std::swap(sp, alt);
        break;
      }
      case D_MEAS: measure(d.qubit, d.cbit, f, nb); break;
      case D_RESET: {
        measure(d.qubit, -1, f, nb);
        k::Expand e = make_expand({d.qubit}, 0);
        k::Flag xf;
        xf.ptr = (const int*)((const char*)ms.get() + offsetof(k::MeasState, xflag));
        xf.stride = sizeof(k::MeasState) / sizeof(int);
        k::xgate(prec, s, stride, nb, n, e, d.qubit, xf, st);
        break;
      }
      default: break;
    }
  }
};

double op_bytes(const DevOp& d, int n, int prec) {
  const double len = std::ldexp(1.0, n);
  switch (d.kind) {
    case D_ONE:
    case D_X: return 2.0 * prec * len / std::ldexp(1.0, d.nctl);
    case D_SWAP: return 2.0 * prec * len / std::ldexp(1.0, 1 + d.nctl);
    case D_DIAG: return 2.0 * prec * len / std::ldexp(1.0, d.kq) * (double)d.pat.size();
    case D_DENSE:
    case D_FUSED:
    case D_TILE:
    case D_XSWEEP: return 2.0 * prec * len;
    case D_MEAS:
    case D_RESET: return 3.0 * prec * len;
    default: return 0.0;
  }
}

int op_kind_code(const DevOp& d) {
  switch (d.kind) {
    case D_XSWEEP: return d.exact ? 12 : 11;
    case D_ONE: return 0;
    case D_X: return 1;
    case D_SWAP: return 2;
    case D_DIAG: return 3;
    case D_DENSE: return 4;
    case D_FUSED: return d.mfma ? 8 : 5;
    case D_TILE: return d.exact ? 10 : 9;
    case D_MEAS: return 6;
    case D_RESET: return 7;
    default: return -1;
  }
}

struct Parsed {
  Circuit c;
  std::vector<Op> ops;
  uint64_t n = 0, nc = 0;
};

Parsed read_and_validate(const char* json, size_t len) {
  Parsed p;
  p.c = parse_circuit(json, len);
  validate_circuit(p.c);
  p.n = p.c.num_qubits();
  p.nc = p.c.num_cbits();
  return p;
}

size_t prefix_end_of(const std::vector<Op>& ops) {
  size_t i = 0;
  while (i < ops.size() && ops[i].is_unitary()) ++i;
  return i;
}

uint64_t count_gates(const std::vector<Instruction>& insts) {
  uint64_t g = 0;
  for (auto& in : insts) {
    switch (in.kind) {
      case Kind::Measure:
      case Kind::Reset:
      case Kind::Barrier:
      case Kind::Classical:
      case Kind::Conditional: break;
      case Kind::Gate:
        if (in.name != "remote_barrier") ++g;
        break;
      default: ++g;
    }
  }
  return g;
}

// ---------------------------------------------------------------------------
// Marginals, CDF and sampling
// ---------------------------------------------------------------------------

// Virtual-index map for the marginal over `qubits` (qubits[0] = MSB of the
// output): low virtual bits walk the other qubits in ascending order.
std::vector<uint8_t> marginal_map(int n, const std::vector<uint64_t>& qubits) {
  std::vector<char> used(n, 0);
  for (auto q : qubits) used[q] = 1;
  std::vector<uint8_t> map;
  for (int q = 0; q < n; ++q)
    if (!used[q]) map.push_back((uint8_t)q);
  for (size_t i = 0; i < qubits.size(); ++i) map.push_back((uint8_t)qubits[qubits.size() - 1 - i]);
  return map;
}

// Reduce the state in virtual order over segments of 2^seg elements.
// Returns a device buffer with 2^(n-seg) double2 values (channel x).
DeviceMem reduce_virtual(Exec& ex, const std::vector<uint8_t>& map, int seg) {
  const int n = ex.n;
  const int seg1 = seg < 11 ? seg : 11;
  const uint64_t count1 = 1ull << (n - seg1);
  DeviceMem a(16 * count1), b;
  k::ProbSpec ps;
  ps.branch = -1;
  ps.map = map.data();
  k::seg_reduce_prob(ex.prec, ex.sp, ex.stride, 1, n, ps, seg1, a.get(), 0, ex.st);
  int remaining = seg - seg1;
  uint64_t count = count1;
  while (remaining > 0) {
    int s = remaining < 11 ? remaining : 11;
    uint64_t next = count >> s;
    if (next == 0) next = 1;
    b.alloc(16 * next);
    k::seg_reduce_f64(a.get(), 0, count, 1, s, b.get(), 0, ex.st);
    count = next;
    remaining -= s;
    std::swap(a, b);
  }
  return a;
}

constexpr int kMatMax = 24;  // largest marginal width sampled from a materialised CDF

uint64_t cdf_bytes(uint64_t m) {
  if (m <= (uint64_t)kMatMax) return (1ull << m) * 8;
  return (1ull << (m - 11)) * 8;
}

// Sample `shots` indices over the marginal of `qubits` (qubits[0] = MSB).
std::vector<uint64_t> sample_indices(Exec& ex, const std::vector<uint64_t>& qubits, uint64_t shots, uint64_t seed) {
  const int n = ex.n;
  const int m = (int)qubits.size();
  std::vector<uint8_t> map = marginal_map(n, qubits);
  DeviceMem idx(8 * std::max<uint64_t>(shots, 1));
  DeviceMem tot(8);
  if (m <= kMatMax) {
    DeviceMem bins = reduce_virtual(ex, map, n - m);
    const uint64_t count = 1ull << m;
    DeviceMem C(8 * count), gt(8 * k::cdf_tmp_entries(count));
    k::cdf_build(bins.as<double>(), 2, count, C.as<double>(), gt.as<double>(), tot.as<double>(), ex.err.as<int>(),
                 ex.st);
    k::sample_cdf(C.as<double>(), count, tot.as<double>(), seed, 0, shots, idx.as<uint64_t>(), ex.st);
    ex.check_err();
  } else {
    if (n - m > 11) throw ValueErr("marginal sampling over more than 2^11 amplitudes per outcome is not supported");
    DeviceMem blk = reduce_virtual(ex, map, n - m + 11);
    const uint64_t nblocks = 1ull << (m - 11);
    DeviceMem C(8 * nblocks), gt(8 * k::cdf_tmp_entries(nblocks));
    k::cdf_build(blk.as<double>(), 2, nblocks, C.as<double>(), gt.as<double>(), tot.as<double>(), ex.err.as<int>(),
                 ex.st);
    k::sample_blocked(ex.prec, ex.sp, n, map.data(), n - m, C.as<double>(), nblocks, tot.as<double>(), seed,
                      0, shots, idx.as<uint64_t>(), ex.st);
    ex.check_err();
  }
  std::vector<uint64_t> out(shots);
  if (shots) HIPCHK(hipMemcpyAsync(out.data(), idx.get(), 8 * shots, hipMemcpyDeviceToHost, ex.st));
  HIPCHK(hipStreamSynchronize(ex.st));
  return out;
}

std::string format_bits(const uint8_t* cb, uint64_t nc) {
  std::string s(nc, '0');
  for (uint64_t i = 0; i < nc; ++i) s[nc - 1 - i] = (char)('0' + cb[i]);
  return s;
}

std::string json_num(double x) {
  if (!std::isfinite(x)) return "null";
  char buf[40];
  std::snprintf(buf, sizeof buf, "%.17g", x);
  return buf;
}

void write_shots_profile(double pre, double fusion, double shots_total, double total, uint64_t nshots,
                         const std::vector<double>& shot_times, const char* strategy, uint64_t parallel,
                         uint64_t working, uint64_t prefix_ops) {
  std::ostringstream js;
  js << "{\n  \"preprocessing_time\": " << json_num(pre) << ",\n  \"fusion_time\": " << json_num(fusion)
     << ",\n  \"shots_total_time\": " << json_num(shots_total) << ",\n  \"total_time\": " << json_num(total)
     << ",\n  \"num_shots\": " << nshots << ",\n  \"shot_times\": [";
  for (size_t i = 0; i < shot_times.size(); ++i) js << (i ? ",\n    " : "\n    ") << json_num(shot_times[i]);
  js << (shot_times.empty() ? "]" : "\n  ]");
  js << ",\n  \"execution_strategy\": \"" << strategy << "\",\n  \"parallel_shots\": " << parallel
     << ",\n  \"working_bytes\": " << working << ",\n  \"deterministic_prefix_ops\": " << prefix_ops << "\n}";
  if (mkdir("dqsim_profiles", 0777) != 0 && errno != EEXIST)
    throw RuntimeErr(std::string("Failed to write shots profile: ") + std::strerror(errno));
  auto ns = std::chrono::duration_cast<std::chrono::nanoseconds>(
                std::chrono::system_clock::now().time_since_epoch())
                .count();
  std::string path = "dqsim_profiles/statevector_shots_profile_" + std::to_string(ns) + ".json";
  std::ofstream f(path);
  if (!f) throw RuntimeErr("Failed to write shots profile: cannot open " + path);
  f << js.str();
  if (!f) throw RuntimeErr("Failed to write shots profile: write failed");
}

}  // namespace

// ---------------------------------------------------------------------------
// simulate() (simulator.rs:319-371)
// ---------------------------------------------------------------------------
std::unique_ptr<Result> simulate(const Options& o, const char* json, size_t len) {
  const double t0 = now_s();
  Parsed p = read_and_validate(json, len);
  Layout lay = state_layout(p.n, o.prec);
  if (lay.bytes > UINT64_MAX - p.nc) throw MemErr("Working memory size overflow");
  check_budget(lay.bytes + p.nc, o.budget);
  p.ops = compile(p.c.instructions);
  const uint64_t seed = o.has_seed ? o.seed : random_seed();
  Stream st;
  // A profiled run times each compiled op by its reference category, so it
  // runs one launch per op.
  // Out-of-place sweeps need room for a second state within the budget.
  const bool permute = !o.profile && lay.bytes <= o.budget - (lay.bytes + p.nc);
  Plan plan = build_plan(p.ops, 0, p.ops.size(), o.prec, o.fusion, o.max_fused, (int)p.n, o.profile ? 0 : o.tile,
                         permute);
  plan.upload(st.s);
  Exec ex(o.prec, (int)p.n, p.nc, 1, st.s);
  ex.prepare(plan);
  ex.start_zero(plan);
  k::rng_init(ex.rng.as<k::RngState>(), 1, seed, 0, st.s);
  auto res = std::make_unique<Result>();
  res->prec = o.prec;
  res->n = (uint32_t)p.n;
  res->nc = p.nc;
  if (o.profile) {
    SimProfile prof;
    for (const DevOp& d : plan.ops) {
      if (!d.conds.empty()) {
        k::cond_eval(ex.cbits.as<uint8_t>(), ex.ncs, 1, (const k::Cond*)(plan.base() + d.cond_off),
                     (int)d.conds.size(), ex.flags.as<int>(), st.s);
        int flag = 0;
        HIPCHK(hipMemcpyAsync(&flag, ex.flags.get(), sizeof(int), hipMemcpyDeviceToHost, st.s));
        st.sync();
        if (!flag) continue;
        DevOp plain = d;
        plain.conds.clear();
        Event a, b;
        HIPCHK(hipEventRecord(a.e, st.s));
        ex.run(plain, 1);
        HIPCHK(hipEventRecord(b.e, st.s));
        st.sync();
        double s = elapsed_ms(a, b) / 1e3;
        if (plain.category == 0) prof.oq_calls++, prof.oq_time += s;
        else if (plain.category == 1) prof.nq_calls++, prof.nq_time += s;
        else if (plain.category == 2) prof.mq_calls++, prof.mq_time += s;
        continue;
      }
      if (d.kind == D_NOOP) continue;
      Event a, b;
      HIPCHK(hipEventRecord(a.e, st.s));
      ex.run(d, 1);
      HIPCHK(hipEventRecord(b.e, st.s));
      st.sync();
      double s = elapsed_ms(a, b) / 1e3;
      if (d.category == 0) prof.oq_calls++, prof.oq_time += s;
      else if (d.category == 1) prof.nq_calls++, prof.nq_time += s;
      else if (d.category == 2) prof.mq_calls++, prof.mq_time += s;
    }
    ex.check_err();
    prof.total_time = now_s() - t0;
    res->has_profile = true;
    res->prof = prof;
  } else {
    for (const DevOp& d : plan.ops) ex.run(d, 1);
    ex.check_err();
  }
  res->cbits.assign(p.nc, 0);
  if (p.nc) HIPCHK(hipMemcpyAsync(res->cbits.data(), ex.cbits.get(), p.nc, hipMemcpyDeviceToHost, st.s));
  st.sync();
  ex.settle();
  res->state = std::move(ex.owned);
  return res;
}

// ---------------------------------------------------------------------------
// Allocation-free evolution into a caller's device state
// ---------------------------------------------------------------------------
// simulate()'s launch plan for a unitary circuit, with the plan constants
// copied into this GPU's static scratch array (noalloc.hip) and the state in
// the caller's buffer: no device allocation. Every check runs before any
// device call. Returns after the work on `stream` has finished.
void evolve_into(const Options& o, const char* json, size_t len, void* state, uint64_t state_len,
                 hipStream_t stream) {
  Parsed p = read_and_validate(json, len);
  p.ops = compile(p.c.instructions);
  for (const Op& op : p.ops)
    if (!op.is_unitary())
      throw ValueErr("the allocation-free evolution takes unitary circuits only (no measure, reset or conditional)");
  if (p.n >= 64 || state_len != (1ull << p.n))
    throw ValueErr("state buffer holds " + std::to_string(state_len) + " amplitudes; the circuit needs 2^" +
                   std::to_string(p.n));
  if (state == nullptr) throw ValueErr("state buffer is null");
  Plan plan = build_plan(p.ops, 0, p.ops.size(), o.prec, o.fusion, o.max_fused, (int)p.n, o.tile);
  uint64_t scratch_bytes = 0;
  void* scratch = noalloc_scratch(&scratch_bytes);
  if (plan.blob.size() > scratch_bytes)
    throw ValueErr("the circuit's plan constants need " + std::to_string(plan.blob.size()) +
                   " bytes; the allocation-free evolution holds at most " + std::to_string(scratch_bytes));
  if (!plan.blob.empty())
    HIPCHK(hipMemcpyAsync(scratch, plan.blob.data(), plan.blob.size(), hipMemcpyHostToDevice, stream));
  const uint64_t stride = 1ull << p.n;
  k::init_zero(o.prec, state, stride, 1, (int)p.n, stream);
  const uint8_t* cbase = (const uint8_t*)scratch;
  for (const DevOp& d : plan.ops) run_unitary(d, o.prec, state, stride, 1, (int)p.n, cbase, k::Flag{}, true, stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(stream));
}

// ---------------------------------------------------------------------------
// simulate_shots() / run_shots (shots.rs:58-260, simulator.rs:256-317)
// ---------------------------------------------------------------------------
Counts simulate_shots(const Options& o, const char* json, size_t len, uint64_t shots, bool profile_call) {
  const double total_t0 = now_s();
  const bool profile = profile_call || o.profile;
  Parsed p = read_and_validate(json, len);
  const double preprocessing_time = now_s() - total_t0;
  const double fusion_t0 = now_s();
  p.ops = compile(p.c.instructions);
  const double fusion_time = now_s() - fusion_t0;
  const uint64_t seed = o.has_seed ? o.seed : random_seed();
  const double shots_t0 = now_s();
  const int n = (int)p.n;
  const uint64_t nc = p.nc;

  std::map<std::string, uint64_t> counts;
  std::vector<double> shot_times;
  const char* strategy = "empty";
  uint64_t parallel_shots = 0, working_bytes = 0, prefix_ops = 0;

  auto finish = [&]() {
    Counts out(counts.begin(), counts.end());
    if (profile)
      write_shots_profile(preprocessing_time, fusion_time, now_s() - shots_t0, now_s() - total_t0, shots, shot_times,
                          strategy, parallel_shots, working_bytes, prefix_ops);
    return out;
  };
  if (shots == 0) return finish();

  Layout lay = state_layout(p.n, o.prec);
  if (lay.bytes > UINT64_MAX - nc) throw MemErr("Working memory size overflow");
  const uint64_t worker_bytes = lay.bytes + nc;
  check_budget(worker_bytes, o.budget);
  const size_t prefix_end = prefix_end_of(p.ops);
  bool terminal = true;
  for (size_t i = prefix_end; i < p.ops.size(); ++i)
    if (p.ops[i].kind != OpKind::Measure) terminal = false;

  if (o.sample_terminal && terminal) {
    std::map<uint64_t, uint64_t> mapping;  // cbit -> qubit, last write wins
    for (size_t i = prefix_end; i < p.ops.size(); ++i) mapping[p.ops[i].cbit] = p.ops[i].qubit;
    if (mapping.empty()) {
      strategy = "terminal_sampling";
      counts[std::string(nc, '0')] = shots;
      return finish();
    }
    std::vector<uint64_t> qubits;
    for (auto& kv : mapping) qubits.push_back(kv.second);
    std::sort(qubits.begin(), qubits.end());
    qubits.erase(std::unique(qubits.begin(), qubits.end()), qubits.end());
    std::reverse(qubits.begin(), qubits.end());
    const uint64_t m = qubits.size();
    std::vector<std::pair<uint64_t, uint64_t>> shifts;  // (cbit, shift)
    for (auto& kv : mapping) {
      size_t pos = std::find(qubits.begin(), qubits.end(), kv.second) - qubits.begin();
      shifts.push_back({kv.first, m - 1 - pos});
    }
    const uint64_t cb = cdf_bytes(m);
    if (worker_bytes <= UINT64_MAX - cb && worker_bytes + cb <= o.budget) {
      Stream st;
      const uint64_t used = worker_bytes + cb;
      Plan plan = build_plan(p.ops, 0, prefix_end, o.prec, o.fusion, o.max_fused, n, o.tile,
                             lay.bytes <= o.budget - used);
      plan.upload(st.s);
      Exec ex(o.prec, n, nc, 1, st.s);
      ex.prepare(plan);
      ex.start_zero(plan);
      k::rng_init(ex.rng.as<k::RngState>(), 1, seed, 0, st.s);
      for (const DevOp& d : plan.ops) ex.run(d, 1);
      std::vector<uint64_t> idx = sample_indices(ex, qubits, shots, seed);
      std::unordered_map<uint64_t, uint64_t> sampled;
      for (uint64_t x : idx) sampled[x]++;
      std::vector<uint8_t> cbits(nc, 0);
      for (auto& kv : sampled) {
        for (auto& cs : shifts) cbits[cs.first] = (uint8_t)((kv.first >> cs.second) & 1);
        counts[format_bits(cbits.data(), nc)] += kv.second;
      }
      strategy = "terminal_sampling";
      parallel_shots = 1;
      working_bytes = worker_bytes + cb;
      prefix_ops = prefix_end;
      return finish();
    }
    // Fall back to trajectories when the sampling table exceeds the budget.
  }

  // Trajectories: a batch of `workers` independent shots per launch sequence.
  const uint64_t budget = o.budget;
  const bool cache_prefix = shots > 1 && prefix_end > 0 && lay.bytes <= budget - worker_bytes;
  const uint64_t prefix_bytes = cache_prefix ? lay.bytes : 0;
  const uint64_t kMaxBatch = 1024;
  uint64_t workers = std::min<uint64_t>(shots, o.has_max_parallel_shots ? o.max_parallel_shots : kMaxBatch);
  workers = std::min<uint64_t>(workers, kMaxBatch);
  workers = std::min<uint64_t>(workers, (budget - prefix_bytes) / worker_bytes);
  workers = std::max<uint64_t>(workers, 1);

  Stream st;
  // With a cached prefix, the unitary prefix ops[0, prefix_end) has its own
  // plan, run once on one state (out of place when the budget holds a second
  // state while the prefix runs); the batched tail sweeps in place.
  Plan plan = build_plan(p.ops, cache_prefix ? prefix_end : 0, p.ops.size(), o.prec, o.fusion, o.max_fused, n,
                         o.tile);
  plan.upload(st.s);
  DeviceMem prefix;
  if (cache_prefix) {
    Plan pplan = build_plan(p.ops, 0, prefix_end, o.prec, o.fusion, o.max_fused, n, o.tile,
                            lay.bytes <= budget - worker_bytes);
    pplan.upload(st.s);
    Exec pex(o.prec, n, nc, 1, st.s);
    pex.prepare(pplan);
    pex.start_zero(pplan);
    for (const DevOp& d : pplan.ops) pex.run(d, 1);
    st.sync();
    pex.settle();
    prefix = std::move(pex.owned);
  }
  const size_t tail_begin = 0;
  Exec ex(o.prec, n, nc, (uint32_t)workers, st.s);
  ex.prepare(plan);
  std::vector<uint8_t> all(shots * ex.ncs);
  if (profile) shot_times.assign(shots, 0.0);
  for (uint64_t s0 = 0; s0 < shots; s0 += workers) {
    const uint32_t w = (uint32_t)std::min<uint64_t>(workers, shots - s0);
    Event a, b;
    if (profile) HIPCHK(hipEventRecord(a.e, st.s));
    if (cache_prefix) k::broadcast(o.prec, ex.sp, ex.stride, w, prefix.get(), n, st.s);
    else k::init_zero(o.prec, ex.sp, ex.stride, w, n, st.s);
    HIPCHK(hipMemsetAsync(ex.cbits.get(), 0, ex.ncs * w, st.s));
    k::rng_init(ex.rng.as<k::RngState>(), w, seed, s0, st.s);
    for (size_t i = tail_begin; i < plan.ops.size(); ++i) ex.run(plan.ops[i], w);
    if (profile) HIPCHK(hipEventRecord(b.e, st.s));
    HIPCHK(hipMemcpyAsync(all.data() + s0 * ex.ncs, ex.cbits.get(), ex.ncs * w, hipMemcpyDeviceToHost, st.s));
    if (profile) {
      st.sync();
      double t = elapsed_ms(a, b) / 1e3;
      for (uint32_t i = 0; i < w; ++i) shot_times[s0 + i] = t;
    }
  }
  ex.check_err();
  for (uint64_t s = 0; s < shots; ++s) counts[format_bits(all.data() + s * ex.ncs, nc)]++;
  strategy = "trajectories";
  parallel_shots = workers;
  working_bytes = prefix_bytes + workers * worker_bytes;
  prefix_ops = cache_prefix ? prefix_end : 0;
  return finish();
}

// ---------------------------------------------------------------------------
// SimulationResult queries (simulator.rs:110-191)
// ---------------------------------------------------------------------------
template <typename F>
static auto with_state(const Result& r, F fn) {
  Stream st;
  Exec ex(r.prec, (int)r.n, 0, 1, st.s, r.state.get());
  return fn(ex);
}

void result_probabilities(const Result& r, const std::vector<uint64_t>& qubits, double* out) {
  with_state(r, [&](Exec& ex) {
    const int m = (int)qubits.size();
    std::vector<uint8_t> map = marginal_map(ex.n, qubits);
    DeviceMem bins = reduce_virtual(ex, map, ex.n - m);
    const uint64_t count = 1ull << m;
    std::vector<double> tmp(2 * count);
    HIPCHK(hipMemcpyAsync(tmp.data(), bins.get(), 16 * count, hipMemcpyDeviceToHost, ex.st));
    HIPCHK(hipStreamSynchronize(ex.st));
    for (uint64_t i = 0; i < count; ++i) out[i] = tmp[2 * i];
    return 0;
  });
}

Counts result_counts(const Result& r, uint64_t shots, const std::vector<uint64_t>& qubits, uint64_t seed) {
  Counts out;
  if (shots == 0) return out;
  return with_state(r, [&](Exec& ex) {
    std::vector<uint64_t> idx = sample_indices(ex, qubits, shots, seed);
    std::unordered_map<uint64_t, uint64_t> sampled;
    for (uint64_t x : idx) sampled[x]++;
    const size_t w = qubits.size();
    std::map<std::string, uint64_t> c;
    for (auto& kv : sampled) {
      std::string bits(w, '0');
      for (size_t i = 0; i < w; ++i)
        if ((kv.first >> (w - 1 - i)) & 1) bits[i] = '1';
      c[bits] += kv.second;
    }
    return Counts(c.begin(), c.end());
  });
}

double result_fidelity(const Result& r, const void* other, uint64_t len) {
  if (len != (1ull << r.n)) return 0.0;
  return with_state(r, [&](Exec& ex) {
    DeviceMem o(16 * len);
    HIPCHK(hipMemcpyAsync(o.get(), other, 16 * len, hipMemcpyHostToDevice, ex.st));
    const int n = ex.n;
    const int seg1 = n < 11 ? n : 11;
    uint64_t count = 1ull << (n - seg1);
    DeviceMem a(16 * count), b(16 * count);
    k::seg_reduce_dot(ex.prec, ex.sp, n, o.get(), seg1, a.get(), ex.st);
    void* src = a.get();
    void* dst = b.get();
    while (count > 1) {
      int lg = 0;
      while ((1ull << lg) < count) ++lg;
      int s = lg < 11 ? lg : 11;
      k::seg_reduce_f64(src, 0, count, 1, s, dst, 0, ex.st);
      count >>= s;
      std::swap(src, dst);
    }
    double d[2];
    HIPCHK(hipMemcpyAsync(d, src, 16, hipMemcpyDeviceToHost, ex.st));
    HIPCHK(hipStreamSynchronize(ex.st));
    return d[0] * d[0] + d[1] * d[1];
  });
}

void result_statevector(const Result& r, void* out) {
  const uint64_t len = 1ull << r.n;
  Stream st;
  if (r.prec == 16) {
    HIPCHK(hipMemcpyAsync(out, r.state.get(), 16 * len, hipMemcpyDeviceToHost, st.s));
  } else {
    const uint64_t chunk = 1ull << 26;
    DeviceMem tmp(16 * std::min(len, chunk));
    for (uint64_t off = 0; off < len; off += chunk) {
      uint64_t c = std::min(chunk, len - off);
      k::to_c128(8, (const char*)r.state.get() + 8 * off, c, tmp.get(), st.s);
      HIPCHK(hipMemcpyAsync((char*)out + 16 * off, tmp.get(), 16 * c, hipMemcpyDeviceToHost, st.s));
    }
  }
  st.sync();
}

void result_amplitude(const Result& r, uint64_t index, double* out) {
  Stream st;
  if (r.prec == 16) {
    HIPCHK(hipMemcpyAsync(out, (const char*)r.state.get() + 16 * index, 16, hipMemcpyDeviceToHost, st.s));
    st.sync();
  } else {
    float f[2];
    HIPCHK(hipMemcpyAsync(f, (const char*)r.state.get() + 8 * index, 8, hipMemcpyDeviceToHost, st.s));
    st.sync();
    out[0] = f[0];
    out[1] = f[1];
  }
}

void result_compare(const Result& a, const Result& b, double* max_abs, double* l2) {
  if (a.n != b.n) throw ValueErr("results have different qubit counts");
  const uint64_t len = 1ull << a.n;
  Stream st;
  const uint64_t nb = k::diff_blocks(len);
  DeviceMem part(16 * nb);
  k::diff(a.prec, a.state.get(), b.prec, b.state.get(), len, part.get(), st.s);
  std::vector<double> h(2 * nb);
  HIPCHK(hipMemcpyAsync(h.data(), part.get(), 16 * nb, hipMemcpyDeviceToHost, st.s));
  st.sync();
  double m = 0, s = 0;
  for (uint64_t i = 0; i < nb; ++i) {
    m = std::max(m, h[2 * i]);
    s += h[2 * i + 1];
  }
  *max_abs = m;
  *l2 = std::sqrt(s);
}

void result_gather(const Result& r, const uint64_t* idx, uint64_t count, double* out) {
  const uint64_t len = 1ull << r.n;
  for (uint64_t i = 0; i < count; ++i)
    if (idx[i] >= len) throw ValueErr("amplitude index out of range");
  if (count == 0) return;
  Stream st;
  DeviceMem di(8 * count), dout(16 * count);
  HIPCHK(hipMemcpyAsync(di.get(), idx, 8 * count, hipMemcpyHostToDevice, st.s));
  k::gather(r.prec, r.state.get(), di.as<uint64_t>(), count, dout.get(), st.s);
  HIPCHK(hipMemcpyAsync(out, dout.get(), 16 * count, hipMemcpyDeviceToHost, st.s));
  st.sync();
}

// ---------------------------------------------------------------------------
// Timed evolution (benchmarks, planner training data)
// ---------------------------------------------------------------------------
EvolveStats evolve_timed(const Options& o, const char* json, size_t len, int repeats, int record,
                         std::unique_ptr<Result>* keep) {
  Parsed p = read_and_validate(json, len);
  Layout lay = state_layout(p.n, o.prec);
  check_budget(lay.bytes + p.nc, o.budget);
  p.ops = compile(p.c.instructions);
  const size_t pe = prefix_end_of(p.ops);
  const int n = (int)p.n;
  Stream st;
  Plan plan = build_plan(p.ops, 0, pe, o.prec, o.fusion, o.max_fused, n, o.tile,
                         lay.bytes <= o.budget - (lay.bytes + p.nc));
  plan.upload(st.s);
  Exec ex(o.prec, n, p.nc, 1, st.s);
  ex.prepare(plan);
  EvolveStats stats;
  stats.gates = count_gates(p.c.instructions);
  stats.launches = plan.ops.size();
  for (int rep = 0; rep < std::max(1, repeats); ++rep) {
    ex.start_zero(plan);
    const bool rec = record && rep == std::max(1, repeats) - 1;
    Event t0, t1;
    std::vector<Event> evs;
    if (rec) evs.resize(plan.ops.size() + 1);
    HIPCHK(hipEventRecord(t0.e, st.s));
    for (size_t i = 0; i < plan.ops.size(); ++i) {
      if (rec) HIPCHK(hipEventRecord(evs[i].e, st.s));
      ex.run(plan.ops[i], 1);
    }
    if (rec) HIPCHK(hipEventRecord(evs[plan.ops.size()].e, st.s));
    HIPCHK(hipEventRecord(t1.e, st.s));
    st.sync();
    stats.total_ms = elapsed_ms(t0, t1);
    if (rec) {
      stats.ops.clear();
      size_t sweep = 0;
      for (size_t i = 0; i < plan.ops.size(); ++i) {
        const DevOp& d = plan.ops[i];
        OpTiming t;
        t.kind = op_kind_code(d);
        t.k = d.kq;
        t.low = d.low;
        t.gates = d.gates;
        t.ms = elapsed_ms(evs[i], evs[i + 1]);
        t.bytes = op_bytes(d, n, o.prec);
        if (d.kind == D_XSWEEP) {
          // One state written, the part of the source that was read.
          t.bytes = (1.0 + sweep_read_fraction(ex.zero_started, ex.zero_masks, sweep)) * o.prec * std::ldexp(1.0, n);
          ++sweep;
        }
        std::copy(d.feat, d.feat + mlp::kFeatures, t.feat);
        stats.ops.push_back(t);
      }
    }
  }
  ex.check_err();
  if (keep) {
    auto res = std::make_unique<Result>();
    res->prec = o.prec;
    res->n = (uint32_t)n;
    res->nc = p.nc;
    res->cbits.assign(p.nc, 0);
    ex.settle();
    res->state = std::move(ex.owned);
    *keep = std::move(res);
  }
  return stats;
}

// ---------------------------------------------------------------------------
// Plan description (host only)
// ---------------------------------------------------------------------------
std::string plan_description(const char* json, size_t len, int prec, int fusion, int max_fused, int tile,
                             bool permute) {
  Parsed p = read_and_validate(json, len);
  p.ops = compile(p.c.instructions);
  const size_t pe = prefix_end_of(p.ops);
  Plan plan = build_plan(p.ops, 0, p.ops.size(), prec, fusion, max_fused, (int)p.n, tile, permute);
  static const char* names[] = {"one",     "x",     "swap", "diagonal", "dense", "fused",
                                "measure", "reset", "noop", "tile",     "tile"};
  std::ostringstream js;
  js << "{\"num_qubits\": " << p.n << ", \"num_cbits\": " << p.nc << ", \"compiled_ops\": " << p.ops.size()
     << ", \"prefix_ops\": " << pe << ", \"gates\": " << count_gates(p.c.instructions)
     << ", \"blob_bytes\": " << plan.blob.size() << ", \"blob_hash\": \"" << [&] {
          uint64_t h = 1469598103934665603ull;  // FNV-1a
          for (uint8_t c : plan.blob) h = (h ^ c) * 1099511628211ull;
          std::ostringstream hs;
          hs << std::hex << h;
          return hs.str();
        }() << "\", \"learned_choices\": [";
  for (size_t i = 0; i < plan.learned_choices.size(); ++i)
    js << (i ? ", " : "") << "{\"pinned\": " << (plan.learned_choices[i] >= 100 ? "true" : "false")
       << ", \"k\": " << plan.learned_choices[i] % 100 << "}";
  // A run from |0...0> (zero start): each out-of-place sweep's fractions of
  // the source read and of the tiles computed; bytes count what is read.
  std::vector<uint64_t> zmasks;
  const bool zstart = zero_start_masks(plan, &zmasks) && plan.uses_alt;
  if (!zstart) zmasks.clear();
  size_t sweep = 0;
  js << "], \"zero_start\": " << (zstart ? "true" : "false") << ", \"launches\": [";
  for (size_t i = 0; i < plan.ops.size(); ++i) {
    const DevOp& d = plan.ops[i];
    double bytes = op_bytes(d, (int)p.n, prec), readf = 1.0, workf = 1.0;
    if (d.kind == D_XSWEEP) {
      readf = sweep_read_fraction(zstart, zmasks, sweep);
      workf = sweep_work_fraction(zmasks, sweep);
      bytes = (1.0 + readf) * prec * std::ldexp(1.0, (int)p.n);
      ++sweep;
    }
    js << (i ? ", " : "") << "{\"kind\": \"" << names[d.kind] << "\", \"k\": " << d.kq
       << ", \"gates\": " << d.gates << ", \"conditional\": " << (d.conds.empty() ? "false" : "true")
       << ", \"mfma\": " << (d.mfma ? "true" : "false") << ", \"entries\": " << d.pat.size()
       << ", \"category\": " << d.category << ", \"bytes\": " << json_num(bytes)
       << ", \"read_fraction\": " << json_num(readf) << ", \"work_fraction\": " << json_num(workf)
       << ", \"qubits\": [";
    for (size_t t = 0; t < d.touched.size(); ++t) js << (t ? ", " : "") << d.touched[t];
    js << "], \"ops\": [";
    for (size_t t = 0; t < d.src.size(); ++t) js << (t ? ", " : "") << d.src[t];
    js << "], \"features\": [";
    for (int t = 0; t < mlp::kFeatures; ++t) js << (t ? ", " : "") << json_num(d.feat[t]);
    js << "]";
    if (d.kind == D_TILE) {
      js << ", \"exact\": " << (d.exact ? "true" : "false") << ", \"chains\": " << d.chains
         << ", \"fixed_chains\": " << d.fixed_chains;
      // Layout of the sweep: load and store bit position of each local bit
      // (local order), and the low positions that trade places with local
      // bits 8, 9 before the store.
      auto list = [&](const char* key, const uint32_t* v, uint32_t cnt) {
        js << ", \"" << key << "\": [";
        for (uint32_t t = 0; t < cnt; ++t) js << (t ? ", " : "") << v[t];
        js << "]";
      };
      list("local_qubits", d.tile_lq.data(), (uint32_t)d.tile_lq.size());
      list("load_positions", d.targs.spos, d.targs.L);
      list("store_positions", d.targs.opos, d.targs.L);
      list("swaps", d.targs.sbit, d.targs.nswap);
      js << ", \"blocks\": [";
      for (size_t b = 0; b < d.tile_blocks.size(); ++b) {
        js << (b ? ", " : "") << "{\"mfma\": " << (d.tile_mfma[b] ? "true" : "false")
           << ", \"split\": " << (d.tile_mfma[b] >= 4 ? "true" : "false") << ", \"qubits\": [";
        for (size_t t = 0; t < d.tile_blocks[b].size(); ++t) js << (t ? ", " : "") << d.tile_blocks[b][t];
        js << "]}";
      }
      js << "]";
    }
    if (d.kind == D_XSWEEP) {
      // Out-of-place sweep: read position of each local bit (contiguous:
      // local bit j at position j), destination position of each local bit
      // and of each tile-index bit, and the blocks.
      const k::XSweepArgs& a = d.xargs;
      auto list = [&](const char* key, const uint32_t* v, uint32_t cnt) {
        js << ", \"" << key << "\": [";
        for (uint32_t t = 0; t < cnt; ++t) js << (t ? ", " : "") << v[t];
        js << "]";
      };
      uint32_t store[kXTileMax] = {};
      for (uint32_t j = 0; j < a.L; ++j) store[a.wlocal[j]] = a.wpos[j];
      js << ", \"exact\": " << (d.exact ? "true" : "false") << ", \"out_of_place\": true, \"gather\": "
         << (a.gather ? "true" : "false")
         << ", \"persist\": " << (d.persist ? "true" : "false");
      list("local_qubits", d.tile_lq.data(), (uint32_t)d.tile_lq.size());
      list("load_positions", a.rpos, a.L);
      list("store_positions", store, a.L);
      list("tile_store_positions", a.wtile, a.ntb);
      js << ", \"blocks\": [";
      for (size_t b = 0; b < d.tile_blocks.size(); ++b) {
        js << (b ? ", " : "") << "{\"mfma\": " << (d.tile_mfma[b] ? "true" : "false")
           << ", \"split\": " << (d.tile_mfma[b] >= 4 ? "true" : "false") << ", \"qubits\": [";
        for (size_t t = 0; t < d.tile_blocks[b].size(); ++t) js << (t ? ", " : "") << d.tile_blocks[b][t];
        js << "]}";
      }
      js << "]";
    }
    js << "}";
  }
  js << "]}";
  return js.str();
}

// ---------------------------------------------------------------------------
// Batched small circuits
// ---------------------------------------------------------------------------
namespace {

struct BatchProgram {
  std::vector<bk::BatchCircuit> circ;
  std::vector<bk::BatchOp> ops;
  std::vector<float> mats;
  std::vector<uint32_t> order;
  uint64_t total_shots = 0;
  uint32_t max_n = 0;
};

struct ParsedBatch {
  std::vector<Parsed> p;
  std::vector<size_t> pe;
};

BatchProgram build_batch(const ParsedBatch& pb, const std::vector<uint64_t>& shots,
                         const std::vector<uint64_t>& seeds, int max_fused, const std::vector<uint32_t>& subset) {
  BatchProgram prog;
  std::vector<double> cost;
  for (uint32_t idx : subset) {
    const Parsed& p = pb.p[idx];
    bk::BatchCircuit c{};
    c.n = (uint32_t)p.n;
    c.op_off = prog.ops.size();
    c.seed = seeds[idx];
    c.shots = (uint32_t)shots[idx];
    c.out_off = prog.total_shots;
    prog.total_shots += shots[idx];
    std::vector<FusedBlock> blocks = greedy_fuse(p.ops, 0, pb.pe[idx], max_fused);
    for (FusedBlock& b : blocks) {
      build_block_matrix(b, p.ops);
      bk::BatchOp op{};
      op.k = (uint32_t)b.qubits.size();
      op.mat_off = (uint32_t)prog.mats.size();
      if (op.k > 5) throw ValueErr("batched blocks are limited to 5 qubits");
      for (uint32_t i = 0; i < op.k; ++i) op.q[i] = b.qubits[i];
      std::vector<uint32_t> srt(b.qubits.begin(), b.qubits.end());
      std::sort(srt.begin(), srt.end());
      for (uint32_t i = 0; i < op.k; ++i) op.s[i] = srt[i];
      for (const cplx& v : b.mat) {
        prog.mats.push_back((float)v.re);
        prog.mats.push_back((float)v.im);
      }
      prog.ops.push_back(op);
    }
    c.nops = (uint32_t)(prog.ops.size() - c.op_off);
    prog.circ.push_back(c);
    cost.push_back(std::ldexp((double)(c.nops + 2), (int)c.n));
    prog.max_n = std::max(prog.max_n, c.n);
  }
  prog.order.resize(prog.circ.size());
  for (uint32_t i = 0; i < prog.order.size(); ++i) prog.order[i] = i;
  std::stable_sort(prog.order.begin(), prog.order.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
  return prog;
}

int device_attr(hipDeviceAttribute_t a) {
  int dev = 0, v = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&v, a, dev));
  return v;
}

}  // namespace

void batch_device_info(int block, int* info) {
  if (block != 256) throw ValueErr("the single-launch batched kernel runs 256-thread workgroups");
  const int max_lds = device_attr(hipDeviceAttributeMaxSharedMemoryPerBlock);
  info[0] = max_lds;
  info[5] = bk::kBatchLdsMax;
  info[1] = bk::lds_bytes(bk::kBatchLdsMax) <= 65536 ? 0 : -1;  // no attribute call needed within 64 KiB
  info[2] = info[3] = info[4] = info[7] = -1;
  bk::func_info(&info[2], &info[3], &info[4], &info[7]);
  info[8] = device_attr(hipDeviceAttributeMaxSharedMemoryPerMultiprocessor);
  info[10] = -1;
  info[9] = bk::occupancy(bk::lds_bytes(bk::kBatchLdsMax), &info[10]);
}

namespace {
ParsedBatch parse_batch(const std::vector<std::pair<const char*, size_t>>& circuits,
                        const std::vector<uint64_t>& shots, BatchOutcome* out) {
  const size_t nc = circuits.size();
  ParsedBatch pb;
  pb.p.resize(nc);
  pb.pe.resize(nc);
  if (out) {
    out->nq.resize(nc);
    out->nc.resize(nc);
    out->cmap.resize(nc);
  }
  for (size_t i = 0; i < nc; ++i) {
    Parsed& p = pb.p[i];
    try {
      p = read_and_validate(circuits[i].first, circuits[i].second);
      p.ops = compile(p.c.instructions);
    } catch (const ValueErr& e) {
      throw ValueErr("Circuit " + std::to_string(i) + ": " + e.msg);
    }
    if (p.n > 30 || p.n < 1) throw ValueErr("Circuit " + std::to_string(i) + ": batched circuits need 1..30 qubits");
    if (shots[i] > 0xFFFFFFFFull) throw ValueErr("too many shots for one batched circuit");
    size_t pe = prefix_end_of(p.ops);
    std::map<uint64_t, uint64_t> mapping;
    for (size_t j = pe; j < p.ops.size(); ++j) {
      if (p.ops[j].kind != OpKind::Measure)
        throw ValueErr("Circuit " + std::to_string(i) + ": batched execution needs terminal measurements only");
      mapping[p.ops[j].cbit] = p.ops[j].qubit;
    }
    pb.pe[i] = pe;
    if (out) {
      out->nq[i] = p.n;
      out->nc[i] = p.nc;
      out->cmap[i].assign(mapping.begin(), mapping.end());
    }
  }
  return pb;
}
}  // namespace

std::string batch_program_text(const std::vector<std::pair<const char*, size_t>>& circuits,
                               const std::vector<uint64_t>& shots, const std::vector<uint64_t>& seeds, int max_fused,
                               int lmax) {
  const size_t nc = circuits.size();
  if (shots.size() != nc || seeds.size() != nc) throw ValueErr("shots and seeds need one entry per circuit");
  ParsedBatch pb = parse_batch(circuits, shots, nullptr);
  std::vector<uint32_t> all(nc);
  for (uint32_t i = 0; i < nc; ++i) all[i] = i;
  BatchProgram prog = build_batch(pb, shots, seeds, max_fused, all);
  uint64_t slot = 0;
  for (auto& c : prog.circ)
    if ((int)c.n > lmax) slot = std::max<uint64_t>(slot, 1ull << c.n);
  std::ostringstream t;
  t.precision(9);
  t << prog.circ.size() << " " << prog.ops.size() << " " << prog.mats.size() << " " << prog.total_shots << " "
    << slot << "\n";
  for (auto& c : prog.circ)
    t << c.n << " " << c.nops << " " << c.op_off << " " << c.seed << " " << c.shots << " " << c.out_off << "\n";
  for (uint32_t o : prog.order) t << o << " ";
  t << "\n";
  for (auto& op : prog.ops) {
    t << op.k << " " << op.mat_off;
    for (int i = 0; i < 5; ++i) t << " " << op.q[i];
    for (int i = 0; i < 5; ++i) t << " " << op.s[i];
    t << "\n";
  }
  for (float f : prog.mats) t << f << " ";
  t << "\n";
  return t.str();
}

namespace {
// Measured qubits of a terminal-measurement circuit, highest first, and for
// each written cbit the bit of the outcome index it reads (last write wins),
// as in simulate_shots' terminal sampling.
void terminal_map(const Parsed& p, size_t pe, std::vector<uint64_t>& qubits,
                  std::vector<std::pair<uint64_t, uint64_t>>& cmap) {
  std::map<uint64_t, uint64_t> mapping;  // cbit -> qubit
  for (size_t j = pe; j < p.ops.size(); ++j) mapping[p.ops[j].cbit] = p.ops[j].qubit;
  qubits.clear();
  for (auto& kv : mapping) qubits.push_back(kv.second);
  std::sort(qubits.begin(), qubits.end());
  qubits.erase(std::unique(qubits.begin(), qubits.end()), qubits.end());
  std::reverse(qubits.begin(), qubits.end());
  const uint64_t m = qubits.size();
  cmap.clear();
  for (auto& kv : mapping) {
    const size_t pos = std::find(qubits.begin(), qubits.end(), kv.second) - qubits.begin();
    cmap.push_back({kv.first, m - 1 - pos});
  }
}

struct BatchSetup {
  const ParsedBatch* pb;
  const std::vector<uint64_t>* shots;
  const std::vector<uint64_t>* seeds;
  std::vector<std::vector<uint64_t>> qubits;  // per circuit, highest first
  int prec = 8, tile = kTileDefault, device = 0;
};

// Run `set` (indices into the batch, in this order). Planner threads build
// the circuits' launch plans in order, at most kAhead ahead of execution;
// `workers` GPU threads, each with its own stream and device arena, take the
// circuits in the same order and run each on simulate_shots' terminal-sampling
// path (upload, evolve, sample on the device) with the circuit's own seed.
// With `out` null nothing is recorded (tuning). Returns the wall time in ms.
double run_stream_pool(const BatchSetup& su, const std::vector<uint32_t>& set, int workers, int fused,
                       BatchOutcome* out) {
  constexpr size_t kAhead = 64;
  const size_t total = set.size();
  const unsigned hw = std::max(2u, std::thread::hardware_concurrency());
  const int planners = (int)std::min<unsigned>(12u, std::max(1u, hw - (unsigned)std::min<int>(workers, (int)hw - 1)));
  std::vector<std::unique_ptr<Plan>> plans(total);
  std::vector<double> plan_us(total, 0.0);
  std::vector<char> ready(total, 0);
  std::mutex mu;
  std::condition_variable cv_ready, cv_space;
  size_t next_plan = 0;
  std::atomic<size_t> next_job{0};
  std::atomic<bool> stop{false};
  std::exception_ptr err;
  auto fail = [&](std::exception_ptr e) {
    {
      std::lock_guard<std::mutex> lk(mu);
      if (!err) err = e;
      stop.store(true);
    }
    cv_ready.notify_all();
    cv_space.notify_all();
  };
  auto plan_work = [&]() {
    try {
      for (;;) {
        size_t j;
        {
          std::unique_lock<std::mutex> lk(mu);
          cv_space.wait(lk, [&] { return stop.load() || next_plan >= total || next_plan < next_job.load() + kAhead; });
          if (stop.load() || next_plan >= total) return;
          j = next_plan++;
        }
        const uint32_t c = set[j];
        const double t0 = now_s();
        // Sweep out of place exactly when a simulate_shots call under the
        // default max_memory_mb would (its budget then holds the state, the
        // classical bits, the sampling table and a second state), so the
        // circuit's counts equal that call's.
        const Parsed& pc = su.pb->p[c];
        const uint64_t sb = state_layout(pc.n, su.prec).bytes, budget = Options{}.budget;
        const uint64_t used = sb + pc.nc + cdf_bytes(su.qubits[c].size());
        const bool permute = used <= budget && sb <= budget - used;
        auto plan = std::make_unique<Plan>(
            build_plan(pc.ops, 0, su.pb->pe[c], su.prec, FUSE_GREEDY, fused, (int)pc.n, su.tile, permute));
        {
          std::lock_guard<std::mutex> lk(mu);
          plans[j] = std::move(plan);
          plan_us[j] = (now_s() - t0) * 1e6;
          ready[j] = 1;
        }
        cv_ready.notify_all();
      }
    } catch (...) {
      fail(std::current_exception());
    }
  };
  auto gpu_work = [&]() {
    try {
      HIPCHK(hipSetDevice(su.device));
      DeviceArena arena;
      t_arena = &arena;
      try {
        Stream st;
        for (;;) {
          size_t j;
          {
            // Under the lock, so a planner checking for space cannot miss it.
            std::lock_guard<std::mutex> lk(mu);
            j = next_job.fetch_add(1);
          }
          cv_space.notify_all();
          if (j >= total) break;
          std::unique_ptr<Plan> plan;
          {
            std::unique_lock<std::mutex> lk(mu);
            cv_ready.wait(lk, [&] { return stop.load() || ready[j]; });
            if (stop.load()) break;
            plan = std::move(plans[j]);
          }
          const uint32_t c = set[j];
          const double c0 = now_s();
          const Parsed& p = su.pb->p[c];
          const int n = (int)p.n;
          const uint64_t shots = (*su.shots)[c], seed = (*su.seeds)[c];
          plan->upload(st.s);
          Exec ex(su.prec, n, p.nc, 1, st.s);
          ex.prepare(*plan);  // second state buffer (from the worker's arena) for out-of-place sweeps
          ex.start_zero(*plan);
          k::rng_init(ex.rng.as<k::RngState>(), 1, seed, 0, st.s);
          for (const DevOp& d : plan->ops) ex.run(d, 1);
          std::vector<uint64_t> ix;
          if (shots && !su.qubits[c].empty()) ix = sample_indices(ex, su.qubits[c], shots, seed);
          else ix.assign(shots, 0);
          if (out && !out->state_off.empty()) {
            const uint64_t len = 1ull << n;
            double* dst = out->states.data() + 2 * out->state_off[c];
            if (su.prec == 16) {
              HIPCHK(hipMemcpyAsync(dst, ex.sp, 16 * len, hipMemcpyDeviceToHost, st.s));
            } else {
              std::vector<float> tmp(2 * len);
              HIPCHK(hipMemcpyAsync(tmp.data(), ex.sp, 8 * len, hipMemcpyDeviceToHost, st.s));
              HIPCHK(hipStreamSynchronize(st.s));
              for (uint64_t q = 0; q < 2 * len; ++q) dst[q] = (double)tmp[q];
            }
          }
          HIPCHK(hipStreamSynchronize(st.s));
          if (out) {
            for (uint64_t s2 = 0; s2 < shots; ++s2) out->idx[out->offsets[c] + s2] = (uint32_t)ix[s2];
            out->latency_us[c] = plan_us[j] + (now_s() - c0) * 1e6;
          }
        }
        st.sync();
      } catch (...) {
        t_arena = nullptr;
        throw;
      }
      t_arena = nullptr;
    } catch (...) {
      fail(std::current_exception());
    }
  };
  const double t0 = now_s();
  std::vector<std::thread> pool;
  for (int i = 0; i < planners; ++i) pool.emplace_back(plan_work);
  for (int w = 0; w < workers; ++w) pool.emplace_back(gpu_work);
  for (auto& t : pool) t.join();
  if (err) std::rethrow_exception(err);
  return (now_s() - t0) * 1e3;
}
// The single launch: every circuit of `set` on one workgroup of
// k_batch_lds, captured in a hipGraph and launched on a stream of its own, so
// it runs while the stream pool works through the other circuits.
struct LdsBatch {
  Stream st;
  DeviceMem circ, list, ops, mats, outd, times, dump, dump_off;
  hipGraphExec_t exec = nullptr;
  Event ev0, ev1;
  std::vector<uint32_t> set;
  BatchProgram prog;
  std::vector<uint64_t> doff;

  // Returns the circuits of `s` it leaves to the pool: those with a fused
  // block wider than 4 qubits (a single compiled op can be, e.g. c4x).
  std::vector<uint32_t> start(const ParsedBatch& pb, const std::vector<uint64_t>& shots,
                              const std::vector<uint64_t>& seeds, const std::vector<uint32_t>& s, int fused,
                              BatchOutcome* out, int device) {
    HIPCHK(hipSetDevice(device));
    std::vector<uint32_t> rest;
    prog = build_batch(pb, shots, seeds, fused, s);
    for (size_t j = 0; j < s.size(); ++j) {
      bool ok = true;
      for (uint32_t o = 0; o < prog.circ[j].nops; ++o) ok = ok && prog.ops[prog.circ[j].op_off + o].k <= 4;
      (ok ? set : rest).push_back(s[j]);
    }
    if (!rest.empty()) prog = build_batch(pb, shots, seeds, fused, set);
    if (set.empty()) return rest;
    int lds_n = 1;
    for (size_t j = 0; j < set.size(); ++j) {
      bk::BatchCircuit& c = prog.circ[j];
      if ((int)c.n > bk::kBatchLdsMax) throw RuntimeErr("internal: circuit too large for the single launch");
      lds_n = std::max(lds_n, (int)c.n);
      c.out_off = out->offsets[set[j]];  // outcomes land in the batch-wide order
    }
    const bool keep = !out->state_off.empty();
    if (keep) {
      uint64_t acc = 0;
      for (size_t j = 0; j < set.size(); ++j) {
        doff.push_back(acc);
        acc += 1ull << prog.circ[j].n;
      }
      doff.push_back(acc);
    }
    auto up = [&](DeviceMem& m, const void* p, size_t bytes) {
      m.alloc(std::max<size_t>(bytes, 16));
      if (bytes) HIPCHK(hipMemcpyAsync(m.get(), p, bytes, hipMemcpyHostToDevice, st.s));
    };
    up(circ, prog.circ.data(), prog.circ.size() * sizeof(bk::BatchCircuit));
    up(list, prog.order.data(), prog.order.size() * sizeof(uint32_t));  // heaviest circuits first
    up(ops, prog.ops.data(), prog.ops.size() * sizeof(bk::BatchOp));
    up(mats, prog.mats.data(), prog.mats.size() * sizeof(float));
    outd.alloc(4 * std::max<uint64_t>(1, out->idx.size()));
    times.alloc(16 * set.size());
    if (keep) {
      up(dump_off, doff.data(), doff.size() * sizeof(uint64_t));
      dump.alloc(8 * std::max<uint64_t>(1, doff.back()));
    }
    bk::LdsLaunch a{};
    a.circ = circ.as<bk::BatchCircuit>();
    a.list = list.as<uint32_t>();
    a.count = (uint32_t)set.size();
    a.lds_n = lds_n;
    a.ops = ops.as<bk::BatchOp>();
    a.mats = mats.as<float>();
    a.out = outd.as<uint32_t>();
    a.times = times.as<uint64_t>();
    a.dump = keep ? dump.as<float2>() : nullptr;
    a.dump_off = keep ? dump_off.as<uint64_t>() : nullptr;
    // One kernel node; the uploads above precede the graph on the stream.
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(st.s, hipStreamCaptureModeThreadLocal));
    try {
      bk::launch_lds(a, st.s);
    } catch (...) {
      (void)hipStreamEndCapture(st.s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(st.s, &g));
    const hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(ie);
    HIPCHK(hipEventRecord(ev0.e, st.s));
    HIPCHK(hipGraphLaunch(exec, st.s));
    HIPCHK(hipEventRecord(ev1.e, st.s));
    return rest;
  }

  void finish(BatchOutcome* out) {
    if (set.empty()) return;
    HIPCHK(hipStreamSynchronize(st.s));
    HIPCHK(hipGetLastError());
    out->lds_ms = elapsed_ms(ev0, ev1);
    std::vector<uint32_t> got(out->idx.size());
    if (!got.empty()) HIPCHK(hipMemcpy(got.data(), outd.get(), 4 * got.size(), hipMemcpyDeviceToHost));
    std::vector<uint64_t> tm(2 * set.size());
    HIPCHK(hipMemcpy(tm.data(), times.get(), 16 * set.size(), hipMemcpyDeviceToHost));
    const double khz = (double)device_attr(hipDeviceAttributeWallClockRate);
    for (size_t j = 0; j < set.size(); ++j) {
      const uint32_t c = set[j];
      const uint64_t off = out->offsets[c], ns = prog.circ[j].shots;
      std::copy(got.begin() + off, got.begin() + off + ns, out->idx.begin() + off);
      out->latency_us[c] = (double)(tm[2 * j + 1] - tm[2 * j]) / khz * 1e3;
    }
    if (!doff.empty()) {
      std::vector<float> f(2 * doff.back());
      HIPCHK(hipMemcpy(f.data(), dump.get(), 8 * doff.back(), hipMemcpyDeviceToHost));
      for (size_t j = 0; j < set.size(); ++j) {
        const uint64_t len = 1ull << prog.circ[j].n;
        double* dst = out->states.data() + 2 * out->state_off[set[j]];
        for (uint64_t q = 0; q < 2 * len; ++q) dst[q] = (double)f[2 * doff[j] + q];
      }
    }
    HIPCHK(hipGraphExecDestroy(exec));
    exec = nullptr;
  }

  ~LdsBatch() {
    if (exec) {
      (void)hipStreamSynchronize(st.s);
      (void)hipGraphExecDestroy(exec);
    }
  }
};
}  // namespace

std::unique_ptr<BatchOutcome> batch_run(const std::vector<std::pair<const char*, size_t>>& circuits,
                                        const std::vector<uint64_t>& shots, const std::vector<uint64_t>& seeds,
                                        const BatchOptions& opt) {
  const size_t nc = circuits.size();
  if (shots.size() != nc || seeds.size() != nc) throw ValueErr("shots and seeds need one entry per circuit");
  if (opt.persistent)
    throw RuntimeErr("the persistent batched kernel is disabled in this build (it stalled on its first GPU run; "
                     "the cause is not identified yet); the single-launch kernel (lds_qubits) replaces it");
  if (opt.lds_qubits < 0 || opt.lds_qubits > bk::kBatchLdsMax)
    throw ValueErr("lds_qubits must be 0 (off) or 1.." + std::to_string(bk::kBatchLdsMax));
  if (opt.workers < -1 || opt.workers > 16) throw ValueErr("workers must be 0 (4), -1 (tuned) or 1..16");
  if (opt.max_fused < 0 || opt.max_fused > 4) throw ValueErr("batched fusion width must be 0 (tuned) or 1..4");
  if (opt.tile != -1 && opt.tile != 0 && (opt.tile < kTileMin || opt.tile > kTileMax))
    throw ValueErr("tile_qubits must be -1 (tuned), 0 or in 10..12");
  if (opt.prec != 8 && opt.prec != 16) throw ValueErr("precision must be complex64 or complex128");
  auto out = std::make_unique<BatchOutcome>();
  ParsedBatch pb = parse_batch(circuits, shots, out.get());
  BatchSetup su;
  su.pb = &pb;
  su.shots = &shots;
  su.seeds = &seeds;
  su.prec = opt.prec;
  const bool tune_tile = opt.tile == -1;
  su.tile = tune_tile ? kTileDefault : opt.tile;
  HIPCHK(hipGetDevice(&su.device));
  su.qubits.resize(nc);
  uint64_t total_shots = 0, total_amps = 0;
  out->offsets.resize(nc);
  for (size_t i = 0; i < nc; ++i) {
    terminal_map(pb.p[i], pb.pe[i], su.qubits[i], out->cmap[i]);
    out->offsets[i] = total_shots;
    total_shots += shots[i];
  }
  out->idx.assign(total_shots, 0);
  out->latency_us.assign(nc, 0.0);
  if (opt.dump_states) {
    out->state_off.resize(nc);
    for (size_t i = 0; i < nc; ++i) {
      out->state_off[i] = total_amps;
      total_amps += 1ull << pb.p[i].n;
    }
    // Kept states are complex128 host copies; refuse what would not fit.
    constexpr uint64_t kKeepMax = 64ull << 30;
    if (total_amps > kKeepMax / 16)
      throw MemErr("keeping the final states needs " + std::to_string(16 * total_amps >> 20) +
                   " MiB of host memory, above the 65536 MiB limit");
    out->states.assign(2 * total_amps, 0.0);
  }
  // Largest estimated work first (2^n amplitudes times compiled ops), so the
  // pool's tail holds small circuits.
  std::vector<uint32_t> order(nc);
  for (uint32_t i = 0; i < nc; ++i) order[i] = i;
  auto cost = [&](uint32_t i) { return std::ldexp((double)(pb.pe[i] + 8), (int)pb.p[i].n); };
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost(a) > cost(b); });

  // Circuits for the single launch: complex64, at most lds_qubits qubits,
  // every qubit measured (the full state's index is then the outcome), at
  // most 2^32 - 1 shots. The rest run on the stream pool.
  std::vector<uint32_t> lds_set, pool_set;
  for (uint32_t i : order) {
    const Parsed& p = pb.p[i];
    bool full = su.qubits[i].size() == p.n;
    for (size_t q = 0; full && q < su.qubits[i].size(); ++q) full = su.qubits[i][q] == p.n - 1 - q;
    if (opt.lds_qubits > 0 && opt.prec == 8 && (int)p.n <= opt.lds_qubits && full && shots[i] <= 0xFFFFFFFFull)
      lds_set.push_back(i);
    else
      pool_set.push_back(i);
  }
  const size_t npool = pool_set.size();

  int workers = opt.workers > 0 ? opt.workers : 4;
  int fused = opt.max_fused ? opt.max_fused : 3;
  if (opt.autotune && npool > 0 && (opt.workers < 0 || !opt.max_fused || tune_tile)) {
    // Calibration subset: evenly spaced circuits of the pool's cost order, run
    // once untimed (plans, arenas and code objects warm), then timed per
    // candidate: the fusion width at the default tile, then the tile shape
    // (LDS tile sweeps of 10, 11 or 12 qubits, or none) at the chosen width.
    // The worker count is tuned only when asked for (workers = -1): on a
    // 96-circuit subset the start-up of each worker weighs more than in the
    // whole call, and the subset's pick (2 workers) ran 4,096 circuits 1.34x
    // slower than the default 4.
    std::vector<uint32_t> sub;
    const size_t want = std::min<size_t>(npool, 96);
    for (size_t i = 0; i < want; ++i) sub.push_back(pool_set[(i * npool) / want]);
    run_stream_pool(su, sub, 4, fused, nullptr);
    double best = 1e300;
    auto name = [](int w, int f, int t) {
      return "workers=" + std::to_string(w) + ",k<=" + std::to_string(f) + ",tile=" + std::to_string(t);
    };
    for (int w : {2, 4, 8}) {
      if (opt.workers > 0 ? w != opt.workers : opt.workers == 0 && w != workers) continue;
      for (int f = 2; f <= 4; ++f) {
        if (opt.max_fused && f != opt.max_fused) continue;
        const double ms = run_stream_pool(su, sub, w, f, nullptr);
        out->tuning.push_back({name(w, f, su.tile), ms});
        if (ms < best) {
          best = ms;
          workers = w;
          fused = f;
        }
      }
    }
    if (tune_tile) {
      int best_tile = su.tile;
      for (int t : {0, 10, 11}) {
        su.tile = t;
        const double ms = run_stream_pool(su, sub, workers, fused, nullptr);
        out->tuning.push_back({name(workers, fused, t), ms});
        if (ms < best) {
          best = ms;
          best_tile = t;
        }
      }
      su.tile = best_tile;
    }
  }
  const double t0 = now_s();
  LdsBatch lds;
  if (!lds_set.empty()) {
    const std::vector<uint32_t> rest =
        lds.start(pb, shots, seeds, lds_set, std::max(1, std::min(fused, 4)), out.get(), su.device);
    pool_set.insert(pool_set.end(), rest.begin(), rest.end());
  }
  if (!pool_set.empty()) run_stream_pool(su, pool_set, workers, fused, out.get());
  lds.finish(out.get());
  out->total_ms = (now_s() - t0) * 1e3;
  out->workers = workers;
  out->fused = fused;
  out->tile = su.tile;
  out->streams = (uint32_t)workers + (lds.set.empty() ? 0u : 1u);
  out->lds_count = lds.set.size();
  return out;
}

Counts batch_counts(const BatchOutcome& b, uint64_t i) {
  std::map<std::string, uint64_t> c;
  const uint64_t nshots = (i + 1 < b.offsets.size() ? b.offsets[i + 1] : b.idx.size()) - b.offsets[i];
  const uint64_t ncb = b.nc[i];
  std::unordered_map<uint32_t, uint64_t> hist;
  for (uint64_t s = 0; s < nshots; ++s) hist[b.idx[b.offsets[i] + s]]++;
  std::vector<uint8_t> bits(ncb, 0);
  for (auto& kv : hist) {
    for (auto& m : b.cmap[i]) bits[m.first] = (uint8_t)((kv.first >> m.second) & 1);
    c[format_bits(bits.data(), ncb)] += kv.second;
  }
  return Counts(c.begin(), c.end());
}

// ---------------------------------------------------------------------------
// Learned planner hook
// ---------------------------------------------------------------------------
void set_planner_weights(const float* w, size_t count) {
  if (!w || count == 0) {
    g_planner.clear();
    return;
  }
  if (count != (size_t)mlp::kMlpParams)
    throw ValueErr("planner weights must have " + std::to_string(mlp::kMlpParams) + " entries");
  g_planner.assign(w, w + count);
}

float mlp_eval_host(const float* w, const float* x) {
  constexpr int F = mlp::kFeatures, H = mlp::kHidden;
  const float* W1 = w;
  const float* b1 = W1 + F * H;
  const float* W2 = b1 + H;
  const float* b2 = W2 + H * H;
  const float* w3 = b2 + H;
  const float b3 = w3[H];
  float h1[H], h2[H];
  for (int j = 0; j < H; ++j) {
    float a = b1[j];
    for (int f = 0; f < F; ++f) a = std::fma(x[f], W1[f * H + j], a);
    h1[j] = a > 0.f ? a : 0.f;
  }
  for (int j = 0; j < H; ++j) {
    float c = b2[j];
    for (int i = 0; i < H; ++i) c = std::fma(h1[i], W2[i * H + j], c);
    h2[j] = c > 0.f ? c : 0.f;
  }
  float o = b3;
  for (int i = 0; i < H; ++i) o = std::fma(h2[i], w3[i], o);
  return o;
}

namespace {
// Predicted time (relative units) of one launch for the block; the model
// predicts log(ms per 2^24 amplitudes).
double predicted_time(const std::vector<uint32_t>& q, const std::vector<int>& opidx, const std::vector<Op>& ops,
                      int n) {
  float f[mlp::kFeatures];
  block_features(q, opidx, ops, n, f);
  return std::exp((double)mlp_eval_host(g_planner.data(), f));
}

// Benefit of running blocks a and b as one launch instead of two.
double planner_score(const FusedBlock& a, const FusedBlock& b, const std::vector<Op>& ops, int n) {
  std::vector<uint32_t> u(a.qubits);
  for (uint32_t q : b.qubits)
    if (std::find(u.begin(), u.end(), q) == u.end()) u.push_back(q);
  std::vector<int> both(a.ops);
  both.insert(both.end(), b.ops.begin(), b.ops.end());
  // A merge saves a launch; it is skipped only when the model predicts the
  // merged block at least 5 % slower than its two parts.
  return 1.05 * (predicted_time(a.qubits, a.ops, ops, n) + predicted_time(b.qubits, b.ops, ops, n)) -
         predicted_time(u, both, ops, n);
}
}  // namespace

// ---------------------------------------------------------------------------
// MLP training on the GPU
// ---------------------------------------------------------------------------
namespace {
struct MlpDev {
  DeviceMem w, g, m, v, x, t, h1, h2, y, dy, d1, d2, loss;
  mlp::Scratch s{};
  explicit MlpDev(int nb) {
    const size_t P = mlp::kMlpParams, F = mlp::kFeatures, H = mlp::kHidden;
    w.alloc(4 * P);
    g.alloc(4 * P);
    m.alloc(4 * P);
    v.alloc(4 * P);
    x.alloc(4 * F * nb);
    t.alloc(4 * nb);
    h1.alloc(4 * H * nb);
    h2.alloc(4 * H * nb);
    y.alloc(4 * nb);
    dy.alloc(4 * nb);
    d1.alloc(4 * H * nb);
    d2.alloc(4 * H * nb);
    loss.alloc(4);
    s = {h1.as<float>(), h2.as<float>(), y.as<float>(), dy.as<float>(), d1.as<float>(), d2.as<float>()};
  }
};
}  // namespace

void mlp_grad(const float* w, const float* X, const float* T, int nb, float* g, float* loss) {
  if (nb <= 0) throw ValueErr("empty batch");
  Stream st;
  MlpDev d(nb);
  HIPCHK(hipMemcpyAsync(d.w.get(), w, 4 * mlp::kMlpParams, hipMemcpyHostToDevice, st.s));
  HIPCHK(hipMemcpyAsync(d.x.get(), X, 4 * (size_t)mlp::kFeatures * nb, hipMemcpyHostToDevice, st.s));
  HIPCHK(hipMemcpyAsync(d.t.get(), T, 4 * (size_t)nb, hipMemcpyHostToDevice, st.s));
  mlp::grad(d.w.as<float>(), d.x.as<float>(), d.t.as<float>(), nb, d.s, d.g.as<float>(), d.loss.as<float>(), st.s);
  HIPCHK(hipMemcpyAsync(g, d.g.get(), 4 * mlp::kMlpParams, hipMemcpyDeviceToHost, st.s));
  HIPCHK(hipMemcpyAsync(loss, d.loss.get(), 4, hipMemcpyDeviceToHost, st.s));
  st.sync();
}

void mlp_predict(const float* w, const float* X, uint64_t n, float* y) {
  Stream st;
  DeviceMem dw(4 * mlp::kMlpParams), dx(4 * mlp::kFeatures * std::max<uint64_t>(n, 1)), dyv(4 * std::max<uint64_t>(n, 1));
  HIPCHK(hipMemcpyAsync(dw.get(), w, 4 * mlp::kMlpParams, hipMemcpyHostToDevice, st.s));
  HIPCHK(hipMemcpyAsync(dx.get(), X, 4 * mlp::kFeatures * n, hipMemcpyHostToDevice, st.s));
  mlp::forward(dw.as<float>(), dx.as<float>(), (int)n, nullptr, nullptr, dyv.as<float>(), st.s);
  HIPCHK(hipMemcpyAsync(y, dyv.get(), 4 * n, hipMemcpyDeviceToHost, st.s));
  st.sync();
}

void mlp_train(float* w, const float* X, const float* T, uint64_t n, int epochs, int batch, float lr, uint64_t seed,
               float* loss_curve) {
  if (n == 0 || batch <= 0 || epochs <= 0) throw ValueErr("empty training set");
  const int nb = (int)std::min<uint64_t>((uint64_t)batch, n);
  Stream st;
  MlpDev d(nb);
  DeviceMem X_d(4 * mlp::kFeatures * n), T_d(4 * n), idx_d(4 * n);
  HIPCHK(hipMemcpyAsync(d.w.get(), w, 4 * mlp::kMlpParams, hipMemcpyHostToDevice, st.s));
  HIPCHK(hipMemsetAsync(d.m.get(), 0, 4 * mlp::kMlpParams, st.s));
  HIPCHK(hipMemsetAsync(d.v.get(), 0, 4 * mlp::kMlpParams, st.s));
  HIPCHK(hipMemcpyAsync(X_d.get(), X, 4 * mlp::kFeatures * n, hipMemcpyHostToDevice, st.s));
  HIPCHK(hipMemcpyAsync(T_d.get(), T, 4 * n, hipMemcpyHostToDevice, st.s));
  std::vector<uint32_t> order(n);
  for (uint64_t i = 0; i < n; ++i) order[i] = (uint32_t)i;
  int step = 0;
  const uint64_t per_epoch = n / (uint64_t)nb;
  for (int e = 0; e < epochs; ++e) {
    // Fisher-Yates with the Philox stream (seed, stream = epoch).
    for (uint64_t i = n - 1; i > 0; --i) {
      const uint64_t j = (uint64_t)(uniform53(seed, (uint64_t)e, i) * (double)(i + 1));
      std::swap(order[i], order[std::min(j, i)]);
    }
    HIPCHK(hipMemcpyAsync(idx_d.get(), order.data(), 4 * n, hipMemcpyHostToDevice, st.s));
    for (uint64_t b = 0; b < std::max<uint64_t>(1, per_epoch); ++b) {
      mlp::gather(X_d.as<float>(), T_d.as<float>(), idx_d.as<uint32_t>() + b * nb, nb, d.x.as<float>(),
                  d.t.as<float>(), st.s);
      mlp::grad(d.w.as<float>(), d.x.as<float>(), d.t.as<float>(), nb, d.s, d.g.as<float>(),
                b + 1 == std::max<uint64_t>(1, per_epoch) ? d.loss.as<float>() : nullptr, st.s);
      mlp::adam(d.w.as<float>(), d.g.as<float>(), d.m.as<float>(), d.v.as<float>(), lr, ++step, st.s);
    }
    if (loss_curve) {
      HIPCHK(hipMemcpyAsync(loss_curve + e, d.loss.get(), 4, hipMemcpyDeviceToHost, st.s));
    }
    // The host shuffle of the next epoch overwrites `order` only after the
    // copy of this epoch's indices has completed.
    st.sync();
  }
  HIPCHK(hipMemcpyAsync(w, d.w.get(), 4 * mlp::kMlpParams, hipMemcpyDeviceToHost, st.s));
  st.sync();
  HIPCHK(hipGetLastError());
}

void split_f16_host(const double* x, uint64_t n, uint16_t* hi, uint16_t* lo) {
  for (uint64_t i = 0; i < n; ++i) half_split(x[i], hi[i], lo[i]);
}

}  // namespace dq

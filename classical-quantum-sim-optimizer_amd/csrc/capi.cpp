// C ABI (include/dqsim_sv.h): error mapping, handles and thread-local errors.
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <random>
#include <new>
#include <string>
#include <vector>

#include "../../include/dqsim_sv.h"
#include "engine.hpp"
#include "pblock.hpp"
#include "planner.hpp"

struct dq_sv {
  dq::Options o;
};
struct dq_result {
  std::unique_ptr<dq::Result> r;
};
struct dq_batch {
  std::unique_ptr<dq::BatchOutcome> b;
};

namespace {
thread_local std::string g_err;

template <typename F>
int guard(F fn) {
  g_err.clear();
  try {
    fn();
    return DQ_OK;
  } catch (const dq::ValueErr& e) {
    g_err = e.msg;
    return DQ_EVALUE;
  } catch (const dq::MemErr& e) {
    g_err = e.msg;
    return DQ_EMEMORY;
  } catch (const dq::RuntimeErr& e) {
    g_err = e.msg;
    return DQ_ERUNTIME;
  } catch (const dq::DeviceErr& e) {
    g_err = e.msg;
    return DQ_EDEVICE;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return DQ_EMEMORY;
  } catch (const std::exception& e) {
    g_err = e.what();
    return DQ_ERUNTIME;
  }
}

int emit(const dq::Counts& c, dq_count_cb cb, void* user) {
  if (!cb) return DQ_OK;
  for (auto& kv : c)
    if (cb(user, kv.first.data(), kv.first.size(), kv.second)) break;
  return DQ_OK;
}

void check_query(const dq::Result& r, const uint64_t* qubits, uint64_t k) {
  for (uint64_t i = 0; i < k; ++i) {
    bool bad = qubits[i] >= r.n;
    for (uint64_t j = 0; j < i && !bad; ++j) bad = qubits[j] == qubits[i];
    if (bad) throw dq::ValueErr("Query qubits must be distinct and within the statevector's range");
  }
}
}  // namespace

// ---- P-block backend ----
struct dq_pb_result {
  std::unique_ptr<dq::pb::Result> r;
};

namespace {
dq::pb::Options pb_options(const dq_pb_opts* o) {
  dq::pb::Options p;
  if (o->max_memory_mb == 0 || o->max_memory_mb > (uint64_t)INT64_MAX / (1024 * 1024))
    throw dq::ValueErr("max_memory_mb must be positive and representable");
  p.memory = o->max_memory_mb * 1024 * 1024;
  p.has_max_block = o->has_max_block_qubits != 0;
  p.max_block_qubits = o->max_block_qubits;
  p.max_parallel_shots = o->max_parallel_shots;
  p.sample_terminal = o->sample_terminal != 0;
  p.split_separable = o->split_separable != 0;
  p.max_split_qubits = o->max_split_qubits;
  p.separation_tolerance = o->separation_tolerance;
  p.profile = o->profile != 0;
  p.has_seed = o->has_seed != 0;
  p.seed = o->seed;
  dq::pb::check_options(p);
  return p;
}
char* dup_string(const std::string& s) {
  char* p = (char*)std::malloc(s.size() + 1);
  if (!p) throw std::bad_alloc();
  std::memcpy(p, s.c_str(), s.size() + 1);
  return p;
}
}  // namespace

struct dq_counts {
  std::string keys;
  std::vector<uint64_t> offsets{0};
  std::vector<uint64_t> values;
  bool failed = false;
};

extern "C" {

dq_counts* dq_counts_new(void) { return new (std::nothrow) dq_counts; }

int dq_counts_append(void* table, const char* bits, size_t len, uint64_t count) {
  auto* t = static_cast<dq_counts*>(table);
  try {
    t->keys.append(bits, len);
    t->offsets.push_back(t->keys.size());
    t->values.push_back(count);
    return 0;
  } catch (...) {
    t->failed = true;
    return 1;
  }
}

int dq_counts_view(const dq_counts* t, uint64_t* n, const char** keys, const uint64_t** offsets,
                   const uint64_t** values) {
  return guard([&] {
    if (t->failed) throw std::bad_alloc();
    *n = t->values.size();
    *keys = t->keys.data();
    *offsets = t->offsets.data();
    *values = t->values.data();
  });
}

void dq_counts_free(dq_counts* t) { delete t; }

void dq_sv_opts_default(dq_sv_opts* o) {
  std::memset(o, 0, sizeof *o);
  o->max_memory_mb = 1024;
  o->sample_terminal = 1;
  o->precision = DQ_PREC_C128;
  o->fusion = DQ_FUSE_NONE;
  o->max_fused_qubits = 4;
  o->tile_qubits = dq::kTileDefault;
}

int dq_sv_create(const dq_sv_opts* o, dq_sv** out) {
  return guard([&] {
    if (!o || !out) throw dq::ValueErr("null argument");
    const uint64_t mib = 1024ull * 1024ull;
    if (o->max_memory_mb == 0 || o->max_memory_mb > (uint64_t)INT64_MAX / mib)
      throw dq::ValueErr("max_memory_mb must be positive and representable in bytes");
    if (o->has_max_parallel_shots && o->max_parallel_shots == 0)
      throw dq::ValueErr("max_parallel_shots must be positive or None");
    if (o->precision != DQ_PREC_C128 && o->precision != DQ_PREC_C64)
      throw dq::ValueErr("precision must be complex128 or complex64");
    if (o->fusion < DQ_FUSE_NONE || o->fusion > DQ_FUSE_LEARNED) throw dq::ValueErr("unknown fusion mode");
    if (o->max_fused_qubits < 1 || o->max_fused_qubits > 5) throw dq::ValueErr("max_fused_qubits must be in 1..5");
    if (o->tile_qubits != 0 && (o->tile_qubits < dq::kTileMin || o->tile_qubits > dq::kXTileMax))
      throw dq::ValueErr("tile_qubits must be 0 or in 10..13");
    auto* s = new dq_sv;
    s->o.has_seed = o->has_seed != 0;
    s->o.seed = o->seed;
    s->o.profile = o->profile != 0;
    s->o.budget = o->max_memory_mb * mib;
    s->o.has_max_parallel_shots = o->has_max_parallel_shots != 0;
    s->o.max_parallel_shots = o->max_parallel_shots;
    s->o.sample_terminal = o->sample_terminal != 0;
    s->o.prec = o->precision;
    s->o.fusion = o->fusion;
    s->o.max_fused = o->max_fused_qubits;
    s->o.tile = o->tile_qubits;
    *out = s;
  });
}

void dq_sv_free(dq_sv* sim) { delete sim; }

int dq_sv_simulate_shots(dq_sv* sim, const char* json, size_t len, uint64_t shots, int32_t profile, dq_count_cb cb,
                         void* user) {
  return guard([&] {
    dq::Counts c = dq::simulate_shots(sim->o, json, len, shots, profile != 0);
    emit(c, cb, user);
  });
}

int dq_sv_simulate(dq_sv* sim, const char* json, size_t len, dq_result** out) {
  return guard([&] {
    auto* r = new dq_result;
    try {
      r->r = dq::simulate(sim->o, json, len);
    } catch (...) {
      delete r;
      throw;
    }
    *out = r;
  });
}

int dq_sv_evolve_noalloc(dq_sv* sim, const char* json, size_t len, void* state, uint64_t state_len, void* stream) {
  return guard([&] { dq::evolve_into(sim->o, json, len, state, state_len, (hipStream_t)stream); });
}

uint32_t dq_result_num_qubits(const dq_result* r) { return r->r->n; }
uint64_t dq_result_num_cbits(const dq_result* r) { return r->r->nc; }

int dq_result_classical_bits(const dq_result* r, uint8_t* out, uint64_t len) {
  return guard([&] {
    if (len != r->r->nc) throw dq::ValueErr("length mismatch");
    if (len) std::memcpy(out, r->r->cbits.data(), len);
  });
}

int dq_result_profile(const dq_result* r, dq_sim_profile* out) {
  if (!r->r->has_profile) return 0;
  const dq::SimProfile& p = r->r->prof;
  out->apply_one_qubit_calls = p.oq_calls;
  out->apply_one_qubit_time = p.oq_time;
  out->apply_n_qubit_calls = p.nq_calls;
  out->apply_n_qubit_time = p.nq_time;
  out->measure_qubit_calls = p.mq_calls;
  out->measure_qubit_time = p.mq_time;
  out->total_time = p.total_time;
  return 1;
}

int dq_result_statevector_c128(const dq_result* r, void* out, uint64_t len) {
  return guard([&] {
    if (len != (1ull << r->r->n)) throw dq::ValueErr("length mismatch");
    dq::result_statevector(*r->r, out);
  });
}

int dq_result_probabilities(const dq_result* r, const uint64_t* qubits, uint64_t k, double* out) {
  return guard([&] {
    check_query(*r->r, qubits, k);
    dq::result_probabilities(*r->r, std::vector<uint64_t>(qubits, qubits + k), out);
  });
}

int dq_result_counts(const dq_result* r, uint64_t shots, const uint64_t* qubits, uint64_t k, int32_t has_seed,
                     uint64_t seed, dq_count_cb cb, void* user) {
  return guard([&] {
    check_query(*r->r, qubits, k);
    uint64_t s = seed;
    if (!has_seed) {
      std::random_device rd;
      s = ((uint64_t)rd() << 32) ^ rd();
    }
    dq::Counts c = dq::result_counts(*r->r, shots, std::vector<uint64_t>(qubits, qubits + k), s);
    emit(c, cb, user);
  });
}

int dq_result_fidelity(const dq_result* r, const void* other, uint64_t len, double* out) {
  return guard([&] { *out = dq::result_fidelity(*r->r, other, len); });
}

int dq_result_amplitude(const dq_result* r, uint64_t index, double* re_im) {
  return guard([&] {
    if (index >> r->r->n) throw dq::ValueErr("amplitude index out of range");
    dq::result_amplitude(*r->r, index, re_im);
  });
}

int dq_result_compare(const dq_result* a, const dq_result* b, double* max_abs, double* l2) {
  return guard([&] { dq::result_compare(*a->r, *b->r, max_abs, l2); });
}

int dq_result_gather(const dq_result* r, const uint64_t* indices, uint64_t count, double* out) {
  return guard([&] { dq::result_gather(*r->r, indices, count, out); });
}

void dq_result_free(dq_result* r) { delete r; }

const char* dq_last_error(void) { return g_err.c_str(); }

int dq_plan_json(const char* json, size_t len, int32_t precision, int32_t fusion, int32_t max_fused,
                 int32_t tile_qubits, int32_t out_of_place, char** out) {
  return guard([&] {
    if (tile_qubits != 0 && (tile_qubits < dq::kTileMin || tile_qubits > dq::kXTileMax))
      throw dq::ValueErr("tile_qubits must be 0 or in 10..13");
    std::string s = dq::plan_description(json, len, precision, fusion, max_fused, tile_qubits, out_of_place != 0);
    char* p = (char*)std::malloc(s.size() + 1);
    if (!p) throw std::bad_alloc();
    std::memcpy(p, s.c_str(), s.size() + 1);
    *out = p;
  });
}

int dq_batch_program_text(const char* const* jsons, const size_t* lens, uint64_t ncirc, const uint64_t* shots,
                          const uint64_t* seeds, int32_t max_fused, int32_t lds_qubits, char** out) {
  return guard([&] {
    std::vector<std::pair<const char*, size_t>> cs;
    for (uint64_t i = 0; i < ncirc; ++i) cs.push_back({jsons[i], lens[i]});
    std::string s = dq::batch_program_text(cs, std::vector<uint64_t>(shots, shots + ncirc),
                                           std::vector<uint64_t>(seeds, seeds + ncirc), max_fused, lds_qubits);
    char* p = (char*)std::malloc(s.size() + 1);
    if (!p) throw std::bad_alloc();
    std::memcpy(p, s.c_str(), s.size() + 1);
    *out = p;
  });
}

void dq_string_free(char* s) { std::free(s); }

int dq_device_count(void) { return dq::device_count(); }

int dq_set_sweep_relabel(int32_t on) { return dq::set_sweep_relabel(on != 0) ? 1 : 0; }
int dq_set_sweep_pinned(int32_t on) { return dq::set_sweep_pinned(on != 0) ? 1 : 0; }

int dq_set_sweep_mode(int32_t mode) { return dq::set_sweep_mode(mode); }

int dq_set_block_arith(int32_t mode) { return dq::set_block_arith(mode); }
int dq_set_wave_runs(int32_t mode) { return dq::set_wave_runs(mode); }

int dq_set_zero_start(int32_t on) { return dq::set_zero_start(on != 0) ? 1 : 0; }

int dq_split_f16(const double* x, uint64_t n, uint16_t* hi, uint16_t* lo) {
  return guard([&] {
    if (n && (!x || !hi || !lo)) throw std::invalid_argument("null argument");
    dq::split_f16_host(x, n, hi, lo);
  });
}

int dq_mfma_fp32_peak(int32_t reps, double* tflops, double* ms_per_launch) {
  return guard([&] { *tflops = dq::mfma_fp32_peak_tflops(reps, ms_per_launch); });
}

int dq_sv_evolve_timed(dq_sv* sim, const char* json, size_t len, int32_t repeats, dq_op_timing* out, uint64_t cap,
                       uint64_t* nops, uint64_t* gates, double* total_ms, dq_result** keep) {
  return guard([&] {
    std::unique_ptr<dq::Result> kept;
    dq::EvolveStats s = dq::evolve_timed(sim->o, json, len, repeats, out != nullptr, keep ? &kept : nullptr);
    for (uint64_t i = 0; i < s.ops.size() && i < cap && out; ++i) {
      out[i].kind = s.ops[i].kind;
      out[i].k = s.ops[i].k;
      out[i].low = s.ops[i].low;
      out[i].gates = s.ops[i].gates;
      out[i].ms = s.ops[i].ms;
      out[i].bytes = s.ops[i].bytes;
      static_assert(sizeof out[i].feat == sizeof s.ops[i].feat, "planner feature count matches the C ABI");
      std::memcpy(out[i].feat, s.ops[i].feat, sizeof out[i].feat);
    }
    if (nops) *nops = s.launches;
    if (gates) *gates = s.gates;
    if (total_ms) *total_ms = s.total_ms;
    if (keep) {
      auto* r = new dq_result;
      r->r = std::move(kept);
      *keep = r;
    }
  });
}

int dq_batch_run(const char* const* jsons, const size_t* lens, uint64_t ncirc, const uint64_t* shots,
                 const uint64_t* seeds, const dq_batch_opts* o, dq_batch** out) {
  return guard([&] {
    std::vector<std::pair<const char*, size_t>> cs;
    for (uint64_t i = 0; i < ncirc; ++i) cs.push_back({jsons[i], lens[i]});
    dq::BatchOptions opt;
    if (o) {
      opt.workers = o->workers;
      opt.max_fused = o->max_fused_qubits;
      opt.autotune = o->autotune != 0;
      opt.tile = o->tile_qubits;
      if (o->precision != DQ_PREC_C64 && o->precision != DQ_PREC_C128)
        throw dq::ValueErr("precision must be DQ_PREC_C64 or DQ_PREC_C128");
      opt.prec = o->precision == DQ_PREC_C128 ? 16 : 8;
      opt.dump_states = o->dump_states != 0;
      opt.persistent = o->persistent != 0;
      opt.lds_qubits = o->lds_qubits;
    }
    auto* b = new dq_batch;
    try {
      b->b = dq::batch_run(cs, std::vector<uint64_t>(shots, shots + ncirc), std::vector<uint64_t>(seeds, seeds + ncirc),
                           opt);
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
  });
}

int dq_batch_counts(const dq_batch* b, uint64_t i, dq_count_cb cb, void* user) {
  return guard([&] {
    if (i >= b->b->nq.size()) throw dq::ValueErr("circuit index out of range");
    emit(dq::batch_counts(*b->b, i), cb, user);
  });
}

int dq_batch_indices(const dq_batch* b, uint64_t i, uint32_t* out, uint64_t len) {
  return guard([&] {
    const auto& x = *b->b;
    if (i >= x.nq.size()) throw dq::ValueErr("circuit index out of range");
    const uint64_t n = (i + 1 < x.offsets.size() ? x.offsets[i + 1] : x.idx.size()) - x.offsets[i];
    if (len != n) throw dq::ValueErr("length mismatch");
    std::memcpy(out, x.idx.data() + x.offsets[i], 4 * n);
  });
}

int dq_batch_info(const dq_batch* b, double* total_ms, int32_t* workers, int32_t* max_fused, int32_t* tile_qubits,
                  uint32_t* streams, double* latency_us, uint64_t ncirc) {
  return guard([&] {
    const auto& x = *b->b;
    if (total_ms) *total_ms = x.total_ms;
    if (workers) *workers = x.workers;
    if (max_fused) *max_fused = x.fused;
    if (tile_qubits) *tile_qubits = x.tile;
    if (streams) *streams = x.streams;
    if (latency_us)
      for (uint64_t i = 0; i < ncirc && i < x.latency_us.size(); ++i) latency_us[i] = x.latency_us[i];
  });
}

int dq_batch_state_c128(const dq_batch* b, uint64_t i, void* out, uint64_t len) {
  return guard([&] {
    const auto& x = *b->b;
    if (x.state_off.empty()) throw dq::ValueErr("states were not kept (dump_states = 0)");
    if (i >= x.nq.size() || len != (1ull << x.nq[i])) throw dq::ValueErr("index or length mismatch");
    double* o = (double*)out;
    for (uint64_t k = 0; k < 2 * len; ++k) o[k] = x.states[2 * x.state_off[i] + k];
  });
}

uint64_t dq_batch_tuning_count(const dq_batch* b) { return b->b->tuning.size(); }

int dq_batch_tuning_entry(const dq_batch* b, uint64_t i, char* name, uint64_t cap, double* ms) {
  return guard([&] {
    const auto& t = b->b->tuning;
    if (i >= t.size()) throw dq::ValueErr("tuning entry out of range");
    if (cap == 0 || t[i].first.size() + 1 > cap) throw dq::ValueErr("name buffer too small");
    std::memcpy(name, t[i].first.c_str(), t[i].first.size() + 1);
    *ms = t[i].second;
  });
}

uint32_t dq_batch_num_qubits(const dq_batch* b, uint64_t i) {
  return i < b->b->nq.size() ? (uint32_t)b->b->nq[i] : 0;
}

void dq_batch_free(dq_batch* b) { delete b; }

int dq_batch_lds_info(const dq_batch* b, uint64_t* count, double* ms) {
  return guard([&] {
    if (!b || !count || !ms) throw std::invalid_argument("null argument");
    *count = b->b->lds_count;
    *ms = b->b->lds_ms;
  });
}

int dq_batch_device_info(int32_t block, int32_t* info11) {
  return guard([&] {
    if (block != 256) throw std::invalid_argument("block must be 256");
    static_assert(sizeof(int32_t) == sizeof(int), "int32_t is int");
    dq::batch_device_info(block, (int*)info11);
    info11[6] = dq::device_count();
  });
}

uint64_t dq_mlp_param_count(void) { return (uint64_t)dq::mlp::kMlpParams; }

int dq_set_planner_weights(const float* w, uint64_t count) {
  return guard([&] { dq::set_planner_weights(w, (size_t)count); });
}

int dq_mlp_train(float* w, const float* X, const float* y, uint64_t n, int32_t epochs, int32_t batch, float lr,
                 uint64_t seed, float* loss_curve) {
  return guard([&] { dq::mlp_train(w, X, y, n, epochs, batch, lr, seed, loss_curve); });
}

int dq_mlp_grad(const float* w, const float* X, const float* y, int32_t n, float* grad, float* loss) {
  return guard([&] { dq::mlp_grad(w, X, y, n, grad, loss); });
}

int dq_mlp_predict(const float* w, const float* X, uint64_t n, float* out) {
  return guard([&] { dq::mlp_predict(w, X, n, out); });
}


void dq_pb_opts_default(dq_pb_opts* o) {
  *o = dq_pb_opts{};
  o->max_memory_mb = 1024;
  o->max_parallel_shots = 1;
  o->sample_terminal = 1;
  o->max_split_qubits = 12;
}

int dq_pb_check_opts(const dq_pb_opts* o) {
  return guard([&] { (void)pb_options(o); });
}

int dq_pb_simulate(const dq_pb_opts* o, const char* json, size_t len, dq_pb_result** out) {
  return guard([&] {
    const dq::pb::Options p = pb_options(o);
    auto* r = new dq_pb_result;
    try {
      r->r = dq::pb::simulate(p, json, len);
    } catch (...) {
      delete r;
      throw;
    }
    *out = r;
  });
}

int dq_pb_simulate_shots(const dq_pb_opts* o, const char* json, size_t len, uint64_t shots, int32_t profile,
                         dq_count_cb cb, void* user) {
  return guard([&] {
    const dq::pb::Options p = pb_options(o);
    dq::Counts c = dq::pb::simulate_shots(p, json, len, shots, profile != 0);
    emit(c, cb, user);
  });
}

int dq_pb_last_shots_stats(uint64_t* out4) {
  return guard([&] {
    const dq::pb::ShotsStats st = dq::pb::last_shots_stats();
    out4[0] = st.lockstep_shots;
    out4[1] = st.lone_shots;
    out4[2] = st.peak_batch_bytes;
    out4[3] = st.batch_size;
  });
}

void dq_pb_result_free(dq_pb_result* r) { delete r; }

int dq_pb_result_json(const dq_pb_result* r, int32_t what, char** out) {
  return guard([&] {
    std::ostringstream js;
    const dq::pb::Result& res = *r->r;
    if (what == 0) {
      js << "[";
      for (size_t i = 0; i < res.physical.size(); ++i) js << (i ? ", " : "") << res.physical[i];
      js << "]";
    } else if (what == 1) {
      js << "[";
      for (size_t i = 0; i < res.cbits.size(); ++i)
        js << (i ? ", " : "") << "[" << res.cbits[i].first << ", " << res.cbits[i].second << "]";
      js << "]";
    } else if (what == 2) {
      js << "[";
      auto bq = dq::pb::block_qubits(res);
      for (size_t b = 0; b < bq.size(); ++b) {
        js << (b ? ", " : "") << "[";
        for (size_t i = 0; i < bq[b].size(); ++i) js << (i ? ", " : "") << bq[b][i];
        js << "]";
      }
      js << "]";
    } else if (what == 3) {
      js << dq::pb::diagnostics_json(res, false);
    } else if (what == 4) {
      js << (dq::pb::profiled(res) ? dq::pb::diagnostics_json(res, true) : std::string("null"));
    } else {
      throw dq::ValueErr("unknown result field");
    }
    *out = dup_string(js.str());
  });
}

int dq_pb_statevector_len(const dq_pb_result* r, uint64_t* len) {
  return guard([&] { *len = dq::pb::statevector_len(*r->r); });
}

int dq_pb_statevector(const dq_pb_result* r, void* out, uint64_t len) {
  return guard([&] {
    if (len != dq::pb::statevector_len(*r->r)) throw dq::ValueErr("statevector length mismatch");
    dq::pb::statevector(*r->r, out);
  });
}

int dq_pb_amplitude(const dq_pb_result* r, const char* bits, size_t len, double* out) {
  return guard([&] { dq::pb::amplitude(*r->r, std::string(bits, len), out); });
}

int dq_pb_probabilities(const dq_pb_result* r, const uint64_t* qubits, uint64_t k, int32_t all, char** out) {
  return guard([&] {
    std::vector<uint64_t> qs(qubits, qubits + k);
    auto p = dq::pb::probabilities(*r->r, all ? nullptr : &qs);
    std::ostringstream js;
    js << "[";
    char buf[40];
    for (size_t i = 0; i < p.size(); ++i) {
      std::snprintf(buf, sizeof buf, "%.17g", p[i].second);
      js << (i ? ", " : "") << "[" << p[i].first << ", " << buf << "]";
    }
    js << "]";
    *out = dup_string(js.str());
  });
}

int dq_pb_counts(const dq_pb_result* r, uint64_t shots, const uint64_t* qubits, uint64_t k, int32_t all, uint64_t seed,
                 dq_count_cb cb, void* user) {
  return guard([&] {
    std::vector<uint64_t> qs(qubits, qubits + k);
    dq::Counts c = dq::pb::counts(*r->r, shots, all ? nullptr : &qs, seed);
    emit(c, cb, user);
  });
}

int dq_pb_expectation(const dq_pb_result* r, const char* pauli, size_t len, double* out) {
  return guard([&] { *out = dq::pb::expectation_value(*r->r, std::string(pauli, len)); });
}

}  // extern "C"

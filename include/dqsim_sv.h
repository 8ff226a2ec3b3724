/*
 * dqsim_sv.h - C ABI of the MI355X dense statevector engine (libdqsim_sv.so).
 *
 * This is the drop-in boundary for the reference's statevector backend. The
 * reference exposes that backend only as the PyO3 module `dqsim._core`
 * (src/lib.rs:16-32); every entry point below names the reference item it
 * replaces, so a binding (PyO3 `extern "C"` block, ctypes, cffi) can map the
 * Python API onto it one to one.
 *
 * Conventions
 *  - Circuits are the JSON transport of src/types.rs:7-296, passed as UTF-8
 *    text with an explicit length (what `model_dump_json()` returns).
 *  - Amplitudes use little-endian wire order (qubit 0 = index bit 0). Query
 *    qubit lists put qubits[0] at the most significant output bit.
 *  - Every call returns DQ_OK or an error code; dq_last_error() then gives the
 *    message for the calling thread. Error codes map to Python exceptions:
 *      DQ_EVALUE  -> ValueError    DQ_EMEMORY -> MemoryError
 *      DQ_ETYPE   -> TypeError     DQ_ERUNTIME / DQ_EDEVICE -> RuntimeError
 *  - All calls are re-entrant: each call uses its own HIP stream, and a
 *    result may be queried from several threads at once.
 *  - Nothing in these signatures is a framework type: plain pointers and
 *    sizes only.
 */
#ifndef DQSIM_SV_H
#define DQSIM_SV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DQ_OK 0
#define DQ_EVALUE 1
#define DQ_EMEMORY 2
#define DQ_ETYPE 3
#define DQ_ERUNTIME 4
#define DQ_EDEVICE 5

/* Amplitude precision (bytes per amplitude). */
#define DQ_PREC_C128 16
#define DQ_PREC_C64 8

/* Gate fusion planner. */
#define DQ_FUSE_NONE 0
#define DQ_FUSE_GREEDY 1
#define DQ_FUSE_LEARNED 2

typedef struct dq_sv dq_sv;
typedef struct dq_result dq_result;

/* Constructor options of StatevectorSimulator
 * (src/monolithic/statevector/simulator.rs:225-254, docs/simulators.md:31-39).
 * precision / fusion / max_fused_qubits are additive: the defaults
 * (complex128, no block fusion) give the reference's FP64 results. */
typedef struct dq_sv_opts {
  uint64_t seed;
  int32_t has_seed;
  int32_t profile;
  uint64_t max_memory_mb;
  uint64_t max_parallel_shots;
  int32_t has_max_parallel_shots;
  int32_t sample_terminal;
  int32_t precision;        /* DQ_PREC_C128 (default) or DQ_PREC_C64 */
  int32_t fusion;           /* DQ_FUSE_NONE (default), _GREEDY, _LEARNED */
  int32_t max_fused_qubits; /* 1..5, default 4 */
  int32_t tile_qubits;      /* complex64 + fusion: fused blocks are grouped into
                               sweeps over 2^tile_qubits-amplitude LDS tiles
                               (10..13, default 12; 13 only for out-of-place
                               sweeps, in-place ones use 12); 0 = one launch
                               per block */
} dq_sv_opts;

/* SimulationProfile fields (simulator.rs:43-59). */
typedef struct dq_sim_profile {
  uint64_t apply_one_qubit_calls;
  double apply_one_qubit_time;
  uint64_t apply_n_qubit_calls;
  double apply_n_qubit_time;
  uint64_t measure_qubit_calls;
  double measure_qubit_time;
  double total_time;
} dq_sim_profile;

/* One (bitstring, count) entry; return nonzero to abort the iteration. */
typedef int (*dq_count_cb)(void* user, const char* bits, size_t len, uint64_t count);

/* Packed count table. No counterpart in the reference: a faster way to take
 * counts from any entry point with a dq_count_cb. Pass dq_counts_append as
 * `cb` and a table as `user`; the library appends every outcome itself, with
 * no call into the host language per outcome. dq_counts_view gives n
 * entries in emission order: key i is keys[offsets[i] .. offsets[i+1]), its
 * count values[i]; the pointers stay valid until the next append or
 * dq_counts_free. DQ_EMEMORY from the view if an append ran out of memory. */
typedef struct dq_counts dq_counts;
dq_counts* dq_counts_new(void);
int dq_counts_append(void* table, const char* bits, size_t len, uint64_t count);
int dq_counts_view(const dq_counts* t, uint64_t* n, const char** keys, const uint64_t** offsets,
                   const uint64_t** values);
void dq_counts_free(dq_counts* t);

/* Per-launch device timing (engine-recorded). */
typedef struct dq_op_timing {
  int32_t kind;  /* 0 one,1 x,2 swap,3 diag,4 dense,5 fused(VALU),6 measure,7 reset,8 fused(MFMA),
                    9 tile sweep (k = tile qubits, gates = ops applied),
                    10 exact complex128 tile sweep, 11 out-of-place tile sweep,
                    12 exact complex128 out-of-place tile sweep */
  int32_t k;     /* qubits touched */
  int32_t low;   /* lowest qubit touched */
  int32_t gates; /* compiled ops folded into the launch */
  double ms;
  double bytes;  /* algorithmic HBM bytes */
  float feat[16];/* fusion-planner features of the launch */
} dq_op_timing;

/* Fill `o` with the reference defaults (max_memory_mb = 1024,
 * sample_terminal = 1, no seed, complex128, no fusion). */
void dq_sv_opts_default(dq_sv_opts* o);

/* StatevectorSimulator::new (simulator.rs:225-254). DQ_EVALUE when
 * max_memory_mb is 0 or overflows bytes, or max_parallel_shots is 0. */
int dq_sv_create(const dq_sv_opts* o, dq_sv** out);
void dq_sv_free(dq_sv* sim);

/* StatevectorSimulator::simulate_shots (simulator.rs:256-317, shots.rs:58-260).
 * Calls `cb` once per distinct classical bitstring (highest cbit first). */
int dq_sv_simulate_shots(dq_sv* sim, const char* json, size_t len, uint64_t shots, int32_t profile, dq_count_cb cb,
                         void* user);

/* StatevectorSimulator::simulate (simulator.rs:319-371). */
int dq_sv_simulate(dq_sv* sim, const char* json, size_t len, dq_result** out);

/* Allocation-free variant of dq_sv_simulate. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Evolves a unitary circuit (no measure, reset or conditional: DQ_EVALUE)
 * from |0...0> into the caller's device buffer `state` of `state_len` = 2^n
 * amplitudes in the simulator's precision (complex128 or complex64,
 * interleaved re/im), with the launch plan and arithmetic of dq_sv_simulate
 * on the same options when that call keeps one state buffer (complex64 tile
 * sweeps in place: sweep mode 0, or a max_memory_mb that holds no second
 * state), so the state equals that call's statevector bit for bit; against
 * dq_sv_simulate's out-of-place sweeps it differs only by FP32 rounding.
 * Work runs on `stream` (a hipStream_t; NULL = the null stream) and the
 * call returns when it has finished. The only other device memory used is a
 * static per-GPU scratch array of DQ_NOALLOC_SCRATCH_BYTES for the plan's
 * constants; a circuit whose constants need more is rejected with DQ_EVALUE
 * before any device work. The array is shared by every call on a GPU and
 * keeps the last call's contents: calls that overlap on one GPU race.
 * max_memory_mb does not apply (the library allocates nothing). */
#define DQ_NOALLOC_SCRATCH_BYTES (16ull << 20)
int dq_sv_evolve_noalloc(dq_sv* sim, const char* json, size_t len, void* state, uint64_t state_len, void* stream);

/* SimulationResult (simulator.rs:85-191). */
uint32_t dq_result_num_qubits(const dq_result* r);
uint64_t dq_result_num_cbits(const dq_result* r);
/* classical_bits getter: out[i] = bit i, len = num_cbits. */
int dq_result_classical_bits(const dq_result* r, uint8_t* out, uint64_t len);
/* profile getter: returns 1 and fills `out` when profiling was on, else 0. */
int dq_result_profile(const dq_result* r, dq_sim_profile* out);
/* statevector getter: independent complex128 copy, len = 2^n elements. */
int dq_result_statevector_c128(const dq_result* r, void* out, uint64_t len);
/* probabilities(qubits): out has 2^k FP64 entries (all, including zeros). */
int dq_result_probabilities(const dq_result* r, const uint64_t* qubits, uint64_t k, double* out);
/* counts(shots, qubits, seed). has_seed = 0 draws a fresh seed. */
int dq_result_counts(const dq_result* r, uint64_t shots, const uint64_t* qubits, uint64_t k, int32_t has_seed,
                     uint64_t seed, dq_count_cb cb, void* user);
/* fidelity(other): |<self|other>|^2, 0.0 on a length mismatch. */
int dq_result_fidelity(const dq_result* r, const void* other_c128, uint64_t len, double* out);
/* Bounded single-amplitude query (the reference's MpsResult.amplitude,
 * docs/simulators.md:122-123): re, im of amplitude `index`. */
int dq_result_amplitude(const dq_result* r, uint64_t index, double* re_im);
/* Bulk form of dq_result_amplitude: out[2i], out[2i+1] = amplitude
 * indices[i] (no reference counterpart). */
int dq_result_gather(const dq_result* r, const uint64_t* indices, uint64_t count, double* out);
void dq_result_free(dq_result* r);

/* Accuracy report (no reference counterpart): max |a-b| and the L2 norm of
 * a-b over two results with the same qubit count (any precisions). */
int dq_result_compare(const dq_result* a, const dq_result* b, double* max_abs, double* l2);

/* Message of the calling thread's last error ("" if none). */
const char* dq_last_error(void);

/* ---- no reference counterpart: host-only inspection and measurement ---- */

/* Parse, validate, compile and fuse without touching the GPU; writes a JSON
 * description of the launch plan (launches, blob size and FNV-1a hash of the
 * plan's matrix blob, learned choices; per out-of-place sweep the zero-start
 * read_fraction and work_fraction). out_of_place != 0 describes the plan of a
 * call whose budget holds a second state (complex64 fused sweeps may then run
 * out of place, see dq_set_sweep_mode). Free it with dq_string_free. */
int dq_plan_json(const char* json, size_t len, int32_t precision, int32_t fusion, int32_t max_fused,
                 int32_t tile_qubits, int32_t out_of_place, char** out);
void dq_string_free(char* s);

/* ---- P-block backend: the state as a product of dense blocks ----
 * Replaces PBlockSimulator / PBlockResult (src/distributed/pblock/
 * simulator.rs:24-322, result.rs:21-293). Inputs are circuit JSON in the
 * transport schema; a distributed input is first merged into one circuit by
 * the host wrapper (input.rs:54-224), with one quantum register per
 * physical qubit. */
typedef struct dq_pb_result dq_pb_result;
typedef struct dq_pb_opts {
  uint64_t max_memory_mb;        /* default 1024 */
  int32_t has_max_block_qubits;  /* 0: no limit */
  uint64_t max_block_qubits;
  uint64_t max_parallel_shots;   /* default 1 */
  int32_t sample_terminal;       /* default 1 */
  int32_t split_separable;       /* default 0 */
  uint64_t max_split_qubits;     /* default 12 */
  double separation_tolerance;   /* default 0 */
  int32_t profile;
  int32_t has_seed;
  uint64_t seed;
} dq_pb_opts;
void dq_pb_opts_default(dq_pb_opts* o);
/* The constructor's checks (simulator.rs:58-95): DQ_EVALUE with its message. */
int dq_pb_check_opts(const dq_pb_opts* o);
int dq_pb_simulate(const dq_pb_opts* o, const char* json, size_t len, dq_pb_result** out);
int dq_pb_simulate_shots(const dq_pb_opts* o, const char* json, size_t len, uint64_t shots, int32_t profile,
                         dq_count_cb cb, void* user);
void dq_pb_result_free(dq_pb_result* r);
/* How the calling thread's last dq_pb_simulate_shots ran its trajectory
 * shots (no reference counterpart; diagnostics of the lockstep batches):
 * out4 = {shots completed in lockstep batches, shots run alone (evicted
 * from a batch, or all of them when batches are off), the largest batch's
 * accounted bytes (slots x a slot's peak bytes, plus a retained prefix;
 * never above max_memory_mb), the largest batch's slot count}. */
int dq_pb_last_shots_stats(uint64_t* out4);
/* JSON text of a result field; free with dq_string_free. what: 0 physical
 * qubits, 1 classical bits [[cbit, value]...], 2 block qubits (physical
 * labels), 3 diagnostics, 4 profile (diagnostics + total_time; "null" when
 * the simulator did not profile). */
int dq_pb_result_json(const dq_pb_result* r, int32_t what, char** out);
int dq_pb_statevector_len(const dq_pb_result* r, uint64_t* len);  /* DQ_EMEMORY beyond the budget */
int dq_pb_statevector(const dq_pb_result* r, void* out_c128, uint64_t len);
int dq_pb_amplitude(const dq_pb_result* r, const char* bits, size_t len, double* out_re_im);
/* [[index, probability]...] for probabilities > 0; all != 0: every qubit,
 * highest label first. */
int dq_pb_probabilities(const dq_pb_result* r, const uint64_t* qubits, uint64_t k, int32_t all, char** out_json);
int dq_pb_counts(const dq_pb_result* r, uint64_t shots, const uint64_t* qubits, uint64_t k, int32_t all,
                 uint64_t seed, dq_count_cb cb, void* user);
int dq_pb_expectation(const dq_pb_result* r, const char* pauli, size_t len, double* out);

/* Process-wide switch for relabelled tile sweeps (default off): a sweep may
 * store its tile with some bit positions exchanged so that the next sweep's
 * contiguous low positions hold the qubits it needs; every run of sweeps
 * ends in the identity layout, so results do not depend on the switch
 * (complex128 exact sweeps stay bit-identical). It cuts the sweeps of a
 * 32-qubit depth-40 random circuit by 9 % but each sweep costs more, and the
 * whole run measured about 5 % slower, hence off by default. No counterpart in
 * the reference: a tuning and A/B switch. Returns the previous setting. */
int dq_set_sweep_relabel(int32_t on);

/* Process-wide sweep mode of complex64 fused plans (no counterpart in the
 * reference: a tuning and A/B switch). When a call's max_memory_mb holds a
 * second state, its tile sweeps may run out of place: each sweep reads its
 * tiles from one state buffer as contiguous runs and writes them to the other
 * in the bit layout the next sweep reads contiguously; the last sweep of a
 * run restores the qubits' own positions. 0 = always in place; out of place
 * with 1 = persistent 256-thread workgroups, 2 = one 256-thread workgroup per
 * tile (default), 3 = one 512-thread workgroup per tile (tiles of 11-12
 * qubits; 256 threads otherwise), 4 = persistent 512-thread workgroups.
 * Results stay within complex64 rounding of each other (blocks may group
 * differently). Exact complex128 sweeps (complex128 without fusion) run out
 * of place, one 256-thread workgroup per tile, in every mode but 0, and stay
 * bit-identical to one launch per op (consecutive tiles share 3 qubits, so
 * stores move 128-byte runs). Returns the previous mode; -1 for a mode
 * outside 0..4. */
int dq_set_sweep_mode(int32_t mode);

/* Process-wide tile rule of out-of-place sweeps (no counterpart in the
 * reference: a tuning and A/B switch). on != 0 keeps qubits 0..3 in every
 * tile; 0 (default) needs them only in the first and last tile of a run of
 * sweeps and otherwise asks consecutive tiles to share 4 qubits, which leaves
 * more tile qubits for gates and so fewer sweeps. Results stay within
 * complex64 rounding of each other. Returns the previous setting. */
int dq_set_sweep_pinned(int32_t on);

/* Process-wide arithmetic of 3-qubit blocks in complex64 tile sweeps (no
 * counterpart in the reference: a tuning and A/B switch). 0 = FP32 matrix
 * cores (v_mfma_f32_16x16x4f32); 1 = split precision: every FP32 operand
 * (tile amplitude, embedding entry) as a sum of two FP16 parts on
 * v_mfma_f32_16x16x32_f16, all four partial products accumulated in FP32
 * (about 2^-22 relative per element, a quarter of the matrix-core cycles;
 * the default). Persistent out-of-place sweeps (sweep modes 1 and 4) stay on
 * FP32. Results
 * stay within complex64 rounding of each other. Returns the previous mode;
 * -1 for a mode outside 0..1. */
int dq_set_block_arith(int32_t mode);

/* Process-wide barrier placement between the blocks of 256-thread
 * out-of-place complex64 sweeps (no counterpart in the reference: a tuning
 * and A/B switch). 1 (default): consecutive blocks of at most 3 qubits that
 * leave two tile bits untouched form a run in which each wave owns the
 * amplitudes whose two such bits equal its index, with a workgroup barrier
 * only after a run's last block; 0: a barrier after every block. Results are
 * bit-identical either way. Returns the previous mode; -1 for a mode outside
 * 0..1. */
int dq_set_wave_runs(int32_t mode);

/* Process-wide zero start (no counterpart in the reference: a tuning and A/B
 * switch). on != 0 (default): a plan that opens with one-shot out-of-place
 * sweeps starts from a synthesised |0...0> (no initialisation, no read by
 * the first sweep) and, while some qubit has not yet been in a tile, its
 * sweeps store the tiles that are all zeros without reading them. 0:
 * initialise the state and read it. Results are identical bit for bit.
 * Returns the previous setting. */
int dq_set_zero_start(int32_t on);

/* The host-side FP16 split of split-precision block matrices (no reference
 * counterpart; exposed for tests): hi[i] = x[i] rounded to the nearest FP16
 * (ties to even), lo[i] = x[i] - value(hi[i]) rounded the same way, as
 * IEEE binary16 bit patterns. |x[i]| >= 65520 is DQ_ERUNTIME. */
int dq_split_f16(const double* x, uint64_t n, uint16_t* hi, uint16_t* lo);

/* Number of visible GPUs (0 on a host without one). */
int dq_device_count(void);

/* Measured FP32 matrix-core throughput: `reps` launches of a register-only
 * v_mfma_f32_16x16x4f32 loop (8 independent chains per wave, 8 waves per
 * SIMD) on the current device; TFLOP/s and milliseconds per launch. */
int dq_mfma_fp32_peak(int32_t reps, double* tflops, double* ms_per_launch);

/* Evolve the circuit's unitary prefix from |0...0> `repeats` times on the
 * simulator's settings; returns per-launch timings of the last repeat
 * (up to `cap` records), the launch count and the total device time. With
 * `keep` non-null the final state is returned as a result. */
int dq_sv_evolve_timed(dq_sv* sim, const char* json, size_t len, int32_t repeats, dq_op_timing* out, uint64_t cap,
                       uint64_t* nops, uint64_t* gates, double* total_ms, dq_result** keep);

/* ---- batched small circuits (no reference counterpart: the reference runs
 * such workloads as a loop of simulate_shots calls) ---- */
typedef struct dq_batch dq_batch;
typedef struct dq_batch_opts {
  int32_t workers;          /* host threads, one HIP stream each: 1..16; 0 = 4;
                               -1 = tuned with autotune (2, 4 or 8) */
  int32_t max_fused_qubits; /* 1..4; 0 = tuned */
  int32_t autotune;         /* time candidate settings on a 96-circuit sample */
  int32_t tile_qubits;      /* LDS tile sweeps of each circuit's plan: 0 = off, 10..12;
                               -1 = tuned with autotune (10, 11, 12 or off), else 12 */
  int32_t precision;        /* DQ_PREC_C64 or DQ_PREC_C128 */
  int32_t dump_states;      /* keep final states for dq_batch_state_c128 */
  int32_t persistent;       /* 1 = the persistent kernel (removed: DQ_ERUNTIME) */
  int32_t lds_qubits;       /* single launch for complex64 circuits of up to this
                               many qubits (0..12) that measure every qubit; 0 = off */
} dq_batch_opts;
/* Run `ncirc` circuits (terminal measurements only) with shots[i] shots and
 * seed seeds[i] each. A pool of host threads, each with its own HIP stream
 * and stream-ordered device arena, takes circuits largest first; each circuit
 * follows dq_sv_simulate_shots' terminal-sampling path (greedy fusion of
 * width max_fused_qubits, the given tile and precision, seed seeds[i]), so
 * its counts equal those of that call. Complex64 circuits of up to
 * lds_qubits qubits that measure every qubit run instead in one hipGraph
 * launch of the single-launch kernel (one workgroup per circuit, the state in
 * LDS, FP32 block arithmetic, the chunked FP64 sampler with Philox draws),
 * concurrently with the pool. */
int dq_batch_run(const char* const* jsons, const size_t* lens, uint64_t ncirc, const uint64_t* shots,
                 const uint64_t* seeds, const dq_batch_opts* o, dq_batch** out);
/* Counts of circuit i (same format as simulate_shots). */
int dq_batch_counts(const dq_batch* b, uint64_t i, dq_count_cb cb, void* user);
/* Circuits run by the single launch and its device milliseconds. */
int dq_batch_lds_info(const dq_batch* b, uint64_t* count, double* ms);
/* Sampled outcomes of circuit i over its measured qubits, the highest
 * measured qubit as the most significant bit (len = shots[i]). */
int dq_batch_indices(const dq_batch* b, uint64_t i, uint32_t* out, uint64_t len);
/* Wall ms of the run (tuning excluded), chosen workers / fusion width / tile
 * qubits / streams, and per-circuit latency in microseconds (ncirc entries:
 * plan, upload, evolution and sampling of that circuit). */
int dq_batch_info(const dq_batch* b, double* total_ms, int32_t* workers, int32_t* max_fused, int32_t* tile_qubits,
                  uint32_t* streams, double* latency_us, uint64_t ncirc);
/* Final state of circuit i as complex128 (needs dump_states). */
int dq_batch_state_c128(const dq_batch* b, uint64_t i, void* out, uint64_t len);
/* Autotuning record: the number of candidates timed, and candidate i's name
 * ("workers=W,k<=F,tile=T") and wall ms on the calibration sample. */
uint64_t dq_batch_tuning_count(const dq_batch* b);
int dq_batch_tuning_entry(const dq_batch* b, uint64_t i, char* name, uint64_t cap, double* ms);
/* Qubit count of circuit i (0 when i is out of range). */
uint32_t dq_batch_num_qubits(const dq_batch* b, uint64_t i);
void dq_batch_free(dq_batch* b);
/* Launch limits of the single-launch batched kernel (block must be 256; no
 * launch): info11 = {device max LDS per block, 0 when the largest LDS request
 * stays within the default 64 KiB, kernel max dynamic LDS, kernel static LDS,
 * kernel max threads, largest LDS-resident circuit (qubits), device count,
 * kernel private bytes per lane, device LDS per CU, occupancy query rc,
 * resident workgroups per CU at the largest LDS request}. */
int dq_batch_device_info(int32_t block, int32_t* info11);
/* Host-only text dump of the batched program for these circuits at fusion
 * width `max_fused` (header line: circuits, ops, matrix floats, total shots,
 * the largest state above `lds_qubits` qubits in amplitudes; then one line
 * per circuit, the launch order, one line per op, the FP32 matrices). Free it
 * with dq_string_free. */
int dq_batch_program_text(const char* const* jsons, const size_t* lens, uint64_t ncirc, const uint64_t* shots,
                          const uint64_t* seeds, int32_t max_fused, int32_t lds_qubits, char** out);

/* ---- learned fusion planner (no reference counterpart) ----
 * A 16 -> 64 -> 64 -> 1 FP32 MLP predicting log(ms per 2^24 amplitudes) of a
 * fused launch from its features. Packed weights: W1[16][64], b1[64],
 * W2[64][64], b2[64], w3[64], b3 (dq_mlp_param_count() entries). */
uint64_t dq_mlp_param_count(void);
/* Install weights for fusion = DQ_FUSE_LEARNED; count = 0 clears them. */
int dq_set_planner_weights(const float* w, uint64_t count);
/* Train on the GPU (Adam, mean squared error, Philox-shuffled mini-batches);
 * w is read and written; loss_curve gets one value per epoch (optional). */
int dq_mlp_train(float* w, const float* X, const float* y, uint64_t n, int32_t epochs, int32_t batch, float lr,
                 uint64_t seed, float* loss_curve);
/* One GPU forward + backward pass over a batch: gradients and loss. */
int dq_mlp_grad(const float* w, const float* X, const float* y, int32_t n, float* grad, float* loss);
/* GPU forward pass. */
int dq_mlp_predict(const float* w, const float* X, uint64_t n, float* out);

#ifdef __cplusplus
}
#endif

#endif /* DQSIM_SV_H */

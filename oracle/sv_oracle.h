/*
 * CPU FP64 oracle for the dense statevector path.  TEST INFRASTRUCTURE ONLY:
 * tests, the smoke check and the benchmark's CPU-baseline leg use it as the
 * checker.  The product (libdqsim_sv.so) never links or calls it.
 *
 * It restates the reference crate's algorithm (FelixVelasquez1025/
 * Classical-Quantum-Sim-Optimizer, dqsim 0.1.2) in plain C:
 *   gates ........... src/gates.rs:23-274, src/types.rs:303-344
 *   compile ......... src/monolithic/statevector/execution.rs:66-274,368-416
 *   execute ......... execution.rs:276-366
 *   kernels ......... src/engine.rs:64-453
 *   shots ........... src/monolithic/statevector/shots.rs:58-260
 *   result queries .. src/monolithic/statevector/simulator.rs:139-190, engine.rs:457-507
 * with this build's canonical choices where the reference leaves the order
 * of floating-point sums or the random stream to its libraries:
 *   - every FP64 sum of probabilities is the pairwise tree over the index
 *     range (engine.rs:379-391 uses a rayon reduction, shots.rs:139-142 a
 *     sequential sum);
 *   - CDFs are the engine's block scan (blocks of 2048 bins: per-thread
 *     sequential runs of 8, row-local Hillis-Steele steps, sequential row and
 *     wave offsets, block totals scanned recursively; see block_scan);
 *   - random draws are Philox4x32-10 (key = seed, counter = (draw, stream))
 *     instead of ChaCha8 (rand_chacha 0.3.1).
 * Complex arithmetic is written out in num-complex order and the file is
 * built with -ffp-contract=off, so FP64 results are reproducible bit for bit.
 *
 * Instructions arrive pre-parsed (oracle.py converts the JSON transport).
 */
#ifndef SV_ORACLE_H
#define SV_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Instruction kinds, in the order of src/types.rs:60-296. */
enum {
  OR_ID, OR_X, OR_Y, OR_Z, OR_H, OR_S, OR_SDG, OR_T, OR_TDG, OR_SX, OR_SXDG,
  OR_U3, OR_U2, OR_U1, OR_U, OR_P, OR_RX, OR_RY, OR_RZ, OR_U0,
  OR_CX, OR_CZ, OR_CY, OR_CH, OR_SWAP, OR_CSX,
  OR_CRX, OR_CRY, OR_CRZ, OR_CU1, OR_CP, OR_CU3, OR_CU, OR_RXX, OR_RZZ,
  OR_CCX, OR_CSWAP, OR_RCCX, OR_RC3X, OR_C3X, OR_C3SQRTX, OR_C4X,
  OR_GATE, OR_MEASURE, OR_RESET, OR_BARRIER, OR_KIND_COUNT
};

#define OR_MAX_OPERANDS 8
#define OR_MAX_COND 8

typedef struct or_inst {
  int32_t kind;
  int32_t nq;                        /* operands, Instruction::qubits() order */
  uint64_t q[OR_MAX_OPERANDS];
  double theta, phi, lam, gamma;
  double params[2];
  int32_t nparams;
  char name[32];                     /* generic gate name, lower case */
  uint64_t cbit;
  int32_t ncond;                     /* conditions wrapping this instruction, outermost first */
  uint64_t cbase[OR_MAX_COND], csize[OR_MAX_COND], cvalue[OR_MAX_COND];
} or_inst;

typedef struct or_shot_info {
  int32_t strategy;                  /* 0 empty, 1 terminal_sampling, 2 trajectories */
  uint64_t parallel_shots;
  uint64_t working_bytes;
  uint64_t prefix_ops;
} or_shot_info;

/* Philox4x32-10 draw used everywhere (see header comment). */
double or_uniform53(uint64_t seed, uint64_t stream, uint64_t draw);
void or_philox(uint32_t ctr[4], uint32_t k0, uint32_t k1);

/* Gate matrix of a 1-qubit instruction kind (row-major re,im x4); 0 if none. */
int or_gate_1q(int kind, double theta, double phi, double lam, double out[8]);

/* Number of compiled operations and their kinds (for compile tests):
 * kinds: 0 one,1 x,2 swap,3 diagonal,4 dense,5 measure,6 reset,7 conditional. */
int64_t or_compile_kinds(const or_inst* insts, int64_t count, int32_t* kinds, int64_t cap);
/* One instruction compiled alone (execution.rs rules, no 1q fusion), for
 * the P-block restatement: 1 compiled, 0 no-op, -1 error. */
typedef struct {
  int32_t kind; /* 0 one, 1 x, 2 swap, 3 diagonal, 4 dense, 5 measure, 6 reset, 7 no-op */
  uint32_t target, a, b, qubit;
  uint64_t cbit;
  int32_t nctl;
  uint32_t cq[5];
  uint8_t cv[5];
  int32_t k;
  uint32_t qubits[5];
  int32_t nmat;
} or_opdesc;
int or_compile_one(const or_inst* in, or_opdesc* d, double* mat, int64_t mat_cap);

/* simulate(): one trajectory from |0..0>, RNG stream 0.  state: 2^n complex128. */
int or_simulate(const or_inst* insts, int64_t count, int n, uint64_t nc, uint64_t seed, double* state,
                uint8_t* cbits);

/* run_shots(): bits_out has shots x max(nc,1) bytes, row per shot in shot
 * order (terminal sampling fills rows in shot order too). budget in bytes;
 * max_batch is the engine's concurrent-trajectory cap; prec = 16 or 8
 * bytes/amplitude for budget accounting (arithmetic is always FP64). */
int or_run_shots(const or_inst* insts, int64_t count, int n, uint64_t nc, uint64_t shots, uint64_t seed,
                 uint64_t budget, int has_mps, uint64_t mps, int sample_terminal, uint64_t max_batch, int prec,
                 uint8_t* bits_out, or_shot_info* info);

/* Kernels on a caller state (complex128 interleaved). */
void or_apply_one(double* s, int n, uint32_t target, const double u[8], const uint32_t* cq, const uint8_t* cv,
                  int nctl);
void or_apply_x(double* s, int n, uint32_t target, const uint32_t* cq, const uint8_t* cv, int nctl);
void or_apply_swap(double* s, int n, uint32_t a, uint32_t b, const uint32_t* cq, const uint8_t* cv, int nctl);
void or_apply_diagonal(double* s, int n, const uint32_t* qubits, int k, const double* diag);
void or_apply_dense(double* s, int n, const uint32_t* qubits, int k, const double* mat);
/* measure_qubit with an explicit uniform r in [0,1): returns the outcome. */
int or_measure(double* s, int n, uint32_t qubit, double r, double* p0, double* p1);
/* The same on a complex64 state held as complex128: kept amplitudes are
 * scaled in FP64 and rounded to FP32 (a complex64 collapse). */
int or_measure_f32(double* s, int n, uint32_t qubit, double r, double* p0, double* p1);
/* Pairwise-tree sum of a power-of-two array (in place on a copy). */
double or_ptree(const double* x, uint64_t len);
/* Marginal probabilities, qubits[0] = MSB of the output (2^k entries). */
void or_marginal(const double* s, int n, const uint64_t* qubits, int k, double* out);
/* Sample indices over a marginal exactly as the engine does. */
int or_sample_marginal(const double* s, int n, const uint64_t* qubits, int k, uint64_t shots, uint64_t seed,
                       uint64_t* out);
/* The canonical CDF (inclusive prefix) of x[0..count); returns its last entry. */
double or_cdf(const double* x, uint64_t count, double* out);
/* |<a|b>|^2 with the pairwise tree. */
double or_fidelity(const double* a, const double* b, int n);

/* Sampling of the batched-circuit kernel: `nch` = min(2^n, block) equal
 * chunks, sequential FP64 chunk sums and prefix, binary search then a
 * sequential in-chunk scan (see batch.hip). */
int or_batch_sample(const double* s, int n, uint64_t shots, uint64_t seed, int block, uint64_t* out);

/* Threads used by the parallel loops (OpenMP); 0 = library default. */
void or_set_threads(int t);

#ifdef __cplusplus
}
#endif
#endif

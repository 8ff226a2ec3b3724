// Host-side launch interface of the statevector kernels (kernels.hip).
//
// State layout in HBM: `batch` states back to back, each 2^n interleaved
// complex amplitudes ({re, im} as float2 for complex64 or double2 for
// complex128), little-endian qubit order (qubit 0 = index bit 0).
// Every launcher takes the element type as `prec` (8 or 16 bytes/amplitude).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace dq {
namespace k {

constexpr int kMaxExpand = 7;

// Bit-insertion description: a compact counter p is widened by inserting a
// zero at each position in `pos` (ascending), then `fixed` is OR-ed in.
struct Expand {
  uint32_t nbits = 0;
  uint32_t pos[kMaxExpand] = {};
  uint64_t fixed = 0;
};

// Per-trajectory measurement record (one per batch slot).
struct MeasState {
  double p0, p1, scale;
  int outcome;  // last outcome
  int active;   // 1 if the last measure ran (condition held, finite norm)
  int xflag;    // reset: apply X (active && outcome == 1)
  int pad;
};

// Per-trajectory RNG record.
struct RngState {
  uint64_t seed, stream, draw;
};

struct Cond {
  uint64_t base, size, value;
};

// Optional per-slot enable flag: the kernel for slot b runs only if
// flag[b * stride] != 0. A null pointer means unconditional.
struct Flag {
  const int* ptr = nullptr;
  uint32_t stride = 1;
};

// ---- gates (reference-semantics arithmetic, no FMA contraction) ----
// u: 4 complex FP64 values (re, im interleaved), row-major.
void one(int prec, void* s, uint64_t stride, uint32_t batch, int n, const Expand& e, uint32_t target,
         const double* u, Flag f, hipStream_t st);
void xgate(int prec, void* s, uint64_t stride, uint32_t batch, int n, const Expand& e, uint32_t target, Flag f,
           hipStream_t st);
void swap(int prec, void* s, uint64_t stride, uint32_t batch, int n, const Expand& e, uint32_t qa, uint32_t qb,
          Flag f, hipStream_t st);
// Sparse diagonal: entries with value != 1. qpos: the gate's qubits sorted
// ascending (k of them); patterns: for each entry, the index bits to set.
void diag(int prec, void* s, uint64_t stride, uint32_t batch, int n, const Expand& e, int nentries,
          const uint64_t* patterns, const double* values, Flag f, hipStream_t st);
// Dense k<=5 gather/matvec/scatter. mat_dev: device pointer to 4^k complex
// values (FP64 for prec 16, FP32 for prec 8), row-major with qubits[0] MSB.
// offsets_dev: device pointer to 2^k index offsets.
void dense(int prec, void* s, uint64_t stride, uint32_t batch, int n, int k, const Expand& e, const void* mat_dev,
           const uint64_t* offsets_dev, Flag f, hipStream_t st);
// Fused complex64 block (FMA arithmetic, matrix FP32 on device).
void fused_block(void* s, uint64_t stride, uint32_t batch, int n, int k, const Expand& e, const float* mat_dev,
                 const uint64_t* offsets_dev, hipStream_t st);
// Fused complex64 block on FP32 MFMA (k = 4 or 5). Returns false when the
// layout is not supported by the MFMA kernel.
bool fused_block_mfma(void* s, uint64_t stride, uint32_t batch, int n, int k, const Expand& e,
                      const float* embed_dev, const uint64_t* offsets_dev, hipStream_t st);

// ---- LDS-tile sweep (complex64) ----
// One sweep applies several fused blocks whose qubits all lie in a tile set S
// (|S| = L <= 12, S always holds the qubits at bit positions 0..3). Each workgroup loads one
// 2^L-amplitude tile into LDS, applies every block there, and writes it back.
// Blob of one sweep (copied to LDS by every workgroup): nblocks TileBlock
// descriptors, then the block matrices (16-byte aligned).
struct TileBlock {
  uint32_t k;        // 1..5 qubits
  uint32_t kind;     // 3: 8x8 real embedding on 4x4x1 (k <= 2); 1: permuted 16x16
                     // embedding (k = 3); 2: 2^(k+1) real embedding (k = 4, 5);
                     // 4: the k = 3 embedding split into FP16 parts for 16x16x32 f16;
                     // 5: the k = 4 embedding split into FP16 parts for 32x32x16 f16
  uint32_t mat_off;  // byte offset of the matrix in the blob
  uint32_t loc[5];   // local bit of each block qubit, loc[0] = MSB of the block index
  uint32_t srt[5];   // the same local bits sorted ascending
  uint32_t pad[3];
};
static_assert(sizeof(TileBlock) == 64, "TileBlock is 64 bytes");
// A sweep may also relabel the state's bit positions within its tile (the
// store writes the same addresses the load read, so the sweep stays in
// place): as a loaded tile goes into LDS, lane bit sbit[s] and register bit
// s (local bit 8 + s) exchange contents for s < nswap; the blocks address
// that swapped local order, and local bit i is stored at bit position
// opos[i]. Local bits 0..F-1 (F = 4 complex64, 3 complex128) are positions
// 0..F-1 both ways, so loads and stores move contiguous runs.
struct TileArgs {
  uint32_t L;         // tile qubits
  uint32_t nblocks;
  uint32_t blob16;    // blob size in 16-byte units
  uint32_t kmax;      // widest block of the sweep (3 or 4)
  uint32_t spos[12];  // bit position of load-order local bit i
  uint32_t opos[12];  // bit position of LDS local bit i (after the lane swaps) at the store
  uint32_t srt[12];   // the tile's bit positions sorted ascending
  uint32_t nswap;     // lane swaps between the load and LDS (0..2)
  uint32_t sbit[2];   // lane bit of swap s (< F)
  uint32_t split;     // 1: 3-qubit blocks in split precision (kind 4), the tile held scaled by 2^14 in LDS
};
// Exact complex128 sweep (no fusion): the blob holds nops TileOp descriptors,
// then each op's FP64 data (16-byte aligned). Ops are applied in program
// order with the arithmetic of the per-op kernels.
struct TileOp {
  uint32_t kind;      // 0: 1-qubit pair update, 1: pair swap (X / SWAP), 2: sparse diagonal, 3: dense
  uint32_t nbits;     // zero bits inserted into the group counter (sorted local positions in pos)
  uint32_t fixed;     // local bits ORed into every group index (control values)
  uint32_t abit, bbit;  // pair members: i | abit and i | bbit
  uint32_t k;         // dense width or diagonal entry count
  uint32_t data_off;  // byte offset of the op's data in the blob
  uint32_t qmask;     // local bits of all the op's qubits (targets and controls)
  uint32_t pos[8];    // sorted local positions of the op's qubits (targets and controls)
  uint32_t q[8];      // dense: local bit of each block qubit, q[0] = MSB of the block index;
                      // in a register chain: pair update or swap: q[0] = register pairs it
                      // applies to (bit p: pair (0,1) (2,3) (0,2) (1,3)); diagonal: q[0] =
                      // registers it multiplies (bit j), q[1] = byte offset of their
                      // (re, im) pairs in the op's data
};
static_assert(sizeof(TileOp) == 96, "TileOp is 96 bytes");
// A chain of consecutive ops applied by one thread to the 2^nq amplitudes of
// its group held in registers (nq = 1..2 chain qubits), or one op applied
// through LDS (nq = 0). Blob: nchains TileChain, then the TileOps, then data.
// Register-chain op letters of the compiled two-qubit chain signatures:
// update on register bit 0 (pairs (0,1),(2,3)) or bit 1 (pairs (0,2),(1,3)),
// swap of registers 1,3 or 2,3, multiplier on register 3.
enum ChainOp { CH_NONE = 0, CH_U1 = 1, CH_U2 = 2, CH_S13 = 3, CH_S23 = 4, CH_D3 = 5 };
constexpr uint32_t kChainSigCount = 12;
constexpr int kChainSigs[kChainSigCount][3] = {
    {0, 0, 0},                      // 0: not compiled
    {CH_U1, CH_U2, CH_S13}, {CH_U1, CH_U2, CH_S23}, {CH_U2, CH_U1, CH_S13}, {CH_U2, CH_U1, CH_S23},
    {CH_U1, CH_U2, CH_D3},  {CH_U2, CH_U1, CH_D3},  {CH_U1, CH_D3, 0},      {CH_U1, CH_S23, 0},
    {CH_U2, CH_D3, 0},      {CH_U2, CH_S13, 0},     {CH_D3, 0, 0}};
// A one-qubit chain holding a single update is compiled too (signature 1 in
// bits 8.. of nq = 1).
struct TileChain {
  uint32_t nq;     // chain qubits (0: single op through LDS); bits 8..: compiled signature
  uint32_t first;  // index of the chain's first op
  uint32_t count;  // ops in the chain
  uint32_t cbits;  // local bits of the chain qubits, ascending, 8 bits each
};
void tile_sweep_c128(void* s, uint64_t stride, uint32_t batch, int n, const TileArgs& a, const void* blob_dev,
                     hipStream_t st);

// Dynamic LDS of one workgroup: the tile plus the blob.
size_t tile_lds_bytes(const TileArgs& a);
void tile_sweep(void* s, uint64_t stride, uint32_t batch, int n, const TileArgs& a, const void* blob_dev,
                hipStream_t st);

// ---- out-of-place (permuting) tile sweep (complex64) ----
// Reads every tile from `src` and writes it to `dst` (distinct buffers), so
// the state may change its bit layout (qubit -> bit position) from sweep to
// sweep. Tile tau's local bit j is read from position j (contiguous mode: the
// tile is one contiguous run of 2^L amplitudes) or, in gather mode, from
// position rpos[j] with the tile index spread over the remaining positions in
// ascending order. The blocks (TileBlock array, as in tile_sweep) address
// local bits. The store walks the tile in output order: store index u (bit j
// of u < L) takes local index sum bit_j(u) << wlocal[j] and goes to position
// wpos[j]; tile-index bit k goes to position wtile[k]. With wpos ascending,
// consecutive lanes write consecutive addresses.
struct XSweepArgs {
  uint32_t L;          // tile qubits
  uint32_t nblocks;
  uint32_t blob16;     // blob size in 16-byte units
  uint32_t kmax;       // widest block (3 or 4)
  uint32_t gather;     // 0: contiguous read; 1: read positions rpos
  uint32_t ntb;        // tile-index bits (n - L)
  uint32_t threads;    // workgroup size: 256, or 512 for L >= 11 (L = 13: 512 only)
  uint32_t split;      // 1: 3-qubit blocks in split precision (kind 4), the tile held scaled by 2^14 in LDS
  uint32_t from_zero;  // 1: the source is |0...0> (one-shot launches): nothing is read, tile 0 starts as
                       // amplitude 1 at local index 0, every other tile stores zeros without block work
  uint32_t pad;
  uint32_t rpos[13];   // gather: source position of local bit j
  uint32_t rsrt[13];   // gather: rpos sorted ascending
  uint32_t wlocal[13]; // store index bit j -> local bit
  uint32_t wpos[13];   // store index bit j -> destination position
  uint32_t wtile[52];  // tile-index bit k -> destination position
  uint64_t zero_mask;  // one-shot launches: tiles whose index has one of these bits set hold zeros (their
                       // qubit is still 0 in every amplitude of the source): they store zeros without reading
};
size_t xsweep_lds_bytes(const XSweepArgs& a);
// persist: workgroups walk tiles with a one-tile prefetch (grid 64 per CU);
// otherwise one workgroup per tile.
void xsweep(const void* src, void* dst, uint64_t stride, uint32_t batch, int n, const XSweepArgs& a,
            const void* blob_dev, bool persist, hipStream_t st);
// Exact complex128 out-of-place sweep (k_xsweep128): one 256-thread
// workgroup per tile of 10 or 11 qubits; a.nblocks chains and the op records
// in the blob as for tile_sweep_c128; split, kmax and persistence unused.
void xsweep128(const void* src, void* dst, uint64_t stride, uint32_t batch, int n, const XSweepArgs& a,
               const void* blob_dev, hipStream_t st);

// ---- state set-up ----
void init_zero(int prec, void* s, uint64_t stride, uint32_t batch, int n, hipStream_t st);
// FP32 matrix-core peak microbenchmark (`out` holds mfma_peak_out_floats(grid)
// floats: one per thread).
void mfma_peak(float* out, uint32_t grid, int iters, hipStream_t st);
uint64_t mfma_peak_out_floats(uint32_t grid);
uint64_t mfma_peak_flops_per_launch(uint32_t grid, int iters);
void broadcast(int prec, void* dst, uint64_t stride, uint32_t batch, const void* src, int n, hipStream_t st);

// ---- deterministic FP64 reductions ----
// Channel pair (x, y) per element. Source kinds:
//   PROB:  x = |a|^2 if bit `branch` of the physical index is 0 (or branch<0),
//          y = |a|^2 if that bit is 1.
//   DOT:   (re, im) of conj(a) * other, other complex128 on device.
//   F64:   a double2 array.
// Virtual index v maps to physical index phys(v) = sum_j bit_j(v) << map[j]
// (identity when map == nullptr).
struct ProbSpec {
  int branch = -1;
  const uint8_t* map = nullptr;  // host pointer, n entries; copied into the launch
};
// Reduce `count` virtual elements in segments of 2^seg (seg <= 11) with the
// pairwise tree; writes count >> seg double2 per batch slot at out + b*out_stride.
void seg_reduce_prob(int prec, const void* s, uint64_t stride, uint32_t batch, int n, const ProbSpec& ps, int seg,
                     void* out, uint64_t out_stride, hipStream_t st);
void seg_reduce_dot(int prec, const void* s, int n, const void* other, int seg, void* out, hipStream_t st);
void seg_reduce_f64(const void* in, uint64_t in_stride, uint64_t count, uint32_t batch, int seg, void* out,
                    uint64_t out_stride, hipStream_t st);

// ---- measurement ----
void meas_decide(const void* totals, uint64_t tot_stride, uint32_t batch, MeasState* ms, RngState* rng,
                 uint8_t* cbits, uint64_t nc, int64_t cbit, Flag f, int* err, hipStream_t st);
void collapse(int prec, void* s, uint64_t stride, uint32_t batch, int n, uint32_t qubit, const MeasState* ms,
              hipStream_t st);
void cond_eval(const uint8_t* cbits, uint64_t nc, uint32_t batch, const Cond* conds_dev, int ncond, int* flags,
               hipStream_t st);
void rng_init(RngState* rng, uint32_t batch, uint64_t seed, uint64_t stream0, hipStream_t st);

// ---- CDF and sampling ----
// Canonical inclusive prefix of `count` bins (stride bstride doubles): blocks
// of 2048 bins scanned by one workgroup each (8 consecutive bins per thread
// summed in order, thread totals scanned within 16-lane rows, then rows,
// then waves, each by fixed additions), block totals scanned recursively and
// added as block offsets. Writes C (count entries) and total = C[count-1]
// into *total; group_tmp holds cdf_tmp_entries(count) doubles. With stride 1,
// C may be the bins themselves (each 2048-bin block is read whole before it
// is written).
uint64_t cdf_tmp_entries(uint64_t count);
void cdf_build(const double* bins, uint64_t bstride, uint64_t count, double* C, double* group_tmp, double* total,
               int* err, hipStream_t st);
// Materialised sampling: idx = first j with C[j]/total > r (last -> 1.0).
void sample_cdf(const double* C, uint64_t count, const double* total, uint64_t seed, uint64_t stream0,
                uint64_t shots, uint64_t* out, hipStream_t st);
// Blocked sampling: C over blocks of 2^11 bins, bins recomputed per shot as
// pairwise sums of 2^binseg amplitudes under `map`.
void sample_blocked(int prec, const void* s, int n, const uint8_t* map, int binseg, const double* C,
                    uint64_t nblocks, const double* total, uint64_t seed, uint64_t stream0, uint64_t shots,
                    uint64_t* out, hipStream_t st);

// ---- accuracy ----
// Per-block (max |a-b|, sum |a-b|^2) for two states; diff_blocks(count)
// double2 values are written.
uint64_t diff_blocks(uint64_t count);
void diff(int prec_a, const void* a, int prec_b, const void* b, uint64_t count, void* out, hipStream_t st);

// out[i] = state[idx[i]] as complex128.
void gather(int prec, const void* s, const uint64_t* idx, uint64_t count, void* out, hipStream_t st);

// ---- conversion ----
void to_c128(int prec, const void* s, uint64_t count, void* out, hipStream_t st);

// ---- P-block: product of dense complex128 blocks ----
// Merge: out[r * nl + j] = left[j] * right[r].
void pb_kron(void* out, const void* left, uint64_t nl, const void* right, uint64_t nr, hipStream_t st);
void pb_scale(void* s, uint64_t count, double re, double im, hipStream_t st);  // s[i] *= (re, im)
void pb_scale_real(void* s, uint64_t count, double f, hipStream_t st);         // s[i] *= f
// out[i] = in[i with a bit `bit` inserted at local position q], i < half.
void pb_extract(void* out, const void* in, uint64_t half, uint32_t q, uint32_t bit, hipStream_t st);
// Batched (slot b at b * its size): merge of each slot's pair of blocks;
// half of each slot's block at the slot's measured outcome (ms[b].outcome);
// each slot's one-qubit basis factor |outcome>.
void pb_kron_batch(void* out, const void* left, uint64_t nl, const void* right, uint64_t nr, uint32_t batch,
                   hipStream_t st);
void pb_extract_ms(void* out, const void* in, uint64_t half, uint32_t q, const MeasState* ms, uint32_t batch,
                   hipStream_t st);
void pb_basis_ms(void* out, const MeasState* ms, uint32_t batch, hipStream_t st);
void pb_norms(double* out, const void* s, uint64_t count, hipStream_t st);  // |s[i]|^2
// CDF tables of blocks of at most 1024 amplitudes in one launch: cdf[j] is
// the sequential inclusive prefix of |amps[i]|^2 (blocks this small keep the
// sequential order of the P-block restatement; larger blocks take
// cdf_build's block scan) and totals[slot] its last entry.
struct PbSmallTable {
  const void* amps;
  double* cdf;
  uint64_t len;
  uint64_t slot;
};
void pb_small_tables(const PbSmallTable* tabs_dev, uint32_t count, double* totals, hipStream_t st);
void pb_set2(void* s, double a_re, double a_im, double b_re, double b_im, hipStream_t st);  // s[0], s[1]
// Separation test of local qubit q (one workgroup); writes the remaining
// amplitudes to rem and res[10] = s0.re, s0.im, s1.re, s1.im, loss, total,
// kept, norm, pivot, bad-norm flag.
void pb_separate(const void* s, uint64_t half, uint32_t q, void* rem, double* res, hipStream_t st);
// One factor table of a product: index bit j of an output index b is bit
// shifts[shift_off + j] of b; len entries (double2 or double).
struct PbTable {
  const void* data;
  uint64_t len;
  uint32_t k, shift_off;
};
void pb_product_c128(void* out, uint64_t count, const PbTable* tabs_dev, uint32_t ntab, const uint32_t* shifts_dev,
                     hipStream_t st);
void pb_product_f64(double* out, uint64_t count, const PbTable* tabs_dev, uint32_t ntab, const uint32_t* shifts_dev,
                    hipStream_t st);
// Product sampling from per-table CDFs: out[s * ntab + t] = outcome index of
// table t in shot s (Philox stream stream0 + s, draw t).
void pb_sample(uint32_t* out, uint64_t shots, const PbTable* tabs_dev, uint32_t ntab, uint64_t seed, uint64_t stream0,
               hipStream_t st);
// Pauli expectation of one block: partial needs pb_pauli_blocks(size)
// double2, out receives the block's complex sum.
uint32_t pb_pauli_blocks(uint64_t size);
void pb_pauli(const void* s, uint64_t size, uint64_t xmask, uint64_t ymask, uint64_t zmask, uint32_t nq,
              void* partial, void* out, hipStream_t st);

}  // namespace k
}  // namespace dq

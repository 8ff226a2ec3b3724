// CDNA4 (gfx950) statevector kernels.
//
// The plain kernels restate the reference CPU loops of src/engine.rs with the
// same complex arithmetic (num-complex operation order, no FMA contraction:
// this file is built with -ffp-contract=off), so a complex128 run reproduces
// the FP64 results of the reference algorithm bit for bit:
//   one / xgate ...... for_each_pair + apply_one_qubit / apply_x (engine.rs:64-155)
//   diag ............. apply_diagonal (engine.rs:157-181), identity entries skipped
//   swap ............. apply_swap (engine.rs:183-241)
//   dense ............ DenseLayout::apply (engine.rs:253-348)
//   seg_reduce/meas .. measure_qubit (engine.rs:367-435), marginal_probs (:437-453)
//   cdf/sample ....... terminal sampling (shots.rs:137-158), sample_counts (engine.rs:457-507)
// Controlled gates enumerate only the amplitudes whose controls hold, and
// sparse diagonals only the non-identity entries, so a launch moves exactly
// the bytes the gate changes.
//
// Reductions use one canonical summation order, the pairwise tree over the
// index range (sum(x[0:2^k]) = sum(lower half) + sum(upper half)), which is
// what the butterfly over lanes plus the LDS tree computes; the CPU oracle
// uses the same tree, so FP64 probabilities agree bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>
#include <string>

#include "kernels.hpp"
#include "philox.hpp"

namespace dq {
namespace k {
namespace {

constexpr int TPB = 256;
constexpr int ITEMS = 4;

template <typename T>
struct V2;
template <>
struct V2<float> {
  using t = float2;
};
template <>
struct V2<double> {
  using t = double2;
};

template <typename T>
struct M2 {
  T v[8];  // u00r u00i u01r u01i u10r u10i u11r u11i
};

struct Offsets {
  uint64_t o[32];
};

__device__ __forceinline__ uint64_t expand(uint64_t p, const Expand& e) {
#pragma unroll
  for (int i = 0; i < kMaxExpand; ++i) {
    if ((uint32_t)i < e.nbits) {
      uint64_t low = (1ull << e.pos[i]) - 1;
      p = (p & low) | ((p & ~low) << 1);
    }
  }
  return p | e.fixed;
}

__device__ __forceinline__ bool flag_off(const Flag& f, uint32_t b) {
  return f.ptr != nullptr && f.ptr[(uint64_t)b * f.stride] == 0;
}

inline uint32_t blocks_for(uint64_t count, uint64_t per_block) {
  uint64_t b = (count + per_block - 1) / per_block;
  return (uint32_t)(b == 0 ? 1 : b);
}

// ---------------------------------------------------------------------------
// 1-qubit update on amplitude pairs (engine.rs:113-128)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(TPB) void k_one(typename V2<T>::t* __restrict__ s, uint64_t stride, uint64_t count,
                                             Expand e, uint64_t tbit, M2<T> u, Flag f) {
  using V = typename V2<T>::t;
  const uint32_t b = blockIdx.y;
  if (flag_off(f, b)) return;
  V* st = s + (uint64_t)b * stride;
  const uint64_t base = (uint64_t)blockIdx.x * (TPB * ITEMS) + threadIdx.x;
  V a[ITEMS], c[ITEMS];
  uint64_t idx[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    uint64_t p = base + (uint64_t)j * TPB;
    idx[j] = p < count ? expand(p, e) : ~0ull;
    if (p < count) {
      a[j] = st[idx[j]];
      c[j] = st[idx[j] | tbit];
    }
  }
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    if (idx[j] == ~0ull) continue;
    V x = a[j], y = c[j], na, nc;
    na.x = (u.v[0] * x.x - u.v[1] * x.y) + (u.v[2] * y.x - u.v[3] * y.y);
    na.y = (u.v[0] * x.y + u.v[1] * x.x) + (u.v[2] * y.y + u.v[3] * y.x);
    nc.x = (u.v[4] * x.x - u.v[5] * x.y) + (u.v[6] * y.x - u.v[7] * y.y);
    nc.y = (u.v[4] * x.y + u.v[5] * x.x) + (u.v[6] * y.y + u.v[7] * y.x);
    st[idx[j]] = na;
    st[idx[j] | tbit] = nc;
  }
}

// Controlled permutation: swap the pair members (engine.rs:152-155); also
// SWAP/CSWAP with tbit = a_bit and obit = b_bit (engine.rs:183-241).
template <typename T>
__global__ __launch_bounds__(TPB) void k_pairswap(typename V2<T>::t* __restrict__ s, uint64_t stride,
                                                  uint64_t count, Expand e, uint64_t abit, uint64_t bbit, Flag f) {
  using V = typename V2<T>::t;
  const uint32_t b = blockIdx.y;
  if (flag_off(f, b)) return;
  V* st = s + (uint64_t)b * stride;
  const uint64_t base = (uint64_t)blockIdx.x * (TPB * ITEMS) + threadIdx.x;
  V a[ITEMS], c[ITEMS];
  uint64_t idx[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    uint64_t p = base + (uint64_t)j * TPB;
    idx[j] = p < count ? expand(p, e) : ~0ull;
    if (p < count) {
      a[j] = st[idx[j] | abit];
      c[j] = st[idx[j] | bbit];
    }
  }
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    if (idx[j] == ~0ull) continue;
    st[idx[j] | abit] = c[j];
    st[idx[j] | bbit] = a[j];
  }
}

// Sparse diagonal (engine.rs:157-181): only entries != 1 are visited.
template <typename T>
struct DiagArgs {
  uint64_t pat[32];
  T val[64];
};

template <typename T>
__global__ __launch_bounds__(TPB) void k_diag(typename V2<T>::t* __restrict__ s, uint64_t stride, uint64_t count,
                                              Expand e, int nent, DiagArgs<T> d, Flag f) {
  using V = typename V2<T>::t;
  const uint32_t b = blockIdx.y;
  if (flag_off(f, b)) return;
  V* st = s + (uint64_t)b * stride;
  const uint64_t base = (uint64_t)blockIdx.x * (TPB * ITEMS) + threadIdx.x;
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    uint64_t p = base + (uint64_t)j * TPB;
    if (p >= count) continue;
    uint64_t i0 = expand(p, e);
    for (int t = 0; t < nent; ++t) {
      uint64_t i = i0 | d.pat[t];
      V a = st[i], r;
      T dr = d.val[2 * t], di = d.val[2 * t + 1];
      r.x = a.x * dr - a.y * di;
      r.y = a.x * di + a.y * dr;
      st[i] = r;
    }
  }
}

// Dense k<=5 gather / matvec / scatter (engine.rs:290-347).
template <typename T, int K>
__global__ __launch_bounds__(TPB) void k_dense(typename V2<T>::t* __restrict__ s, uint64_t stride, uint64_t count,
                                               Expand e, const T* __restrict__ m, Offsets off, Flag f) {
  using V = typename V2<T>::t;
  constexpr int D = 1 << K;
  const uint32_t b = blockIdx.y;
  if (flag_off(f, b)) return;
  V* st = s + (uint64_t)b * stride;
  const uint64_t p = (uint64_t)blockIdx.x * TPB + threadIdx.x;
  if (p >= count) return;
  const uint64_t base = expand(p, e);
  V v[D];
#pragma unroll
  for (int j = 0; j < D; ++j) v[j] = st[base | off.o[j]];
#pragma unroll
  for (int row = 0; row < D; ++row) {
    T ar = 0, ai = 0;
#pragma unroll
    for (int col = 0; col < D; ++col) {
      T mr = m[2 * (row * D + col)], mi = m[2 * (row * D + col) + 1];
      ar = ar + (mr * v[col].x - mi * v[col].y);
      ai = ai + (mr * v[col].y + mi * v[col].x);
    }
    V r;
    r.x = ar;
    r.y = ai;
    // Outputs are written after all inputs are read: v[] is private.
    st[base | off.o[row]] = r;
  }
}

// Fused complex64 block: FMA arithmetic, matrix in FP32.
template <int K>
__global__ __launch_bounds__(TPB) void k_fused(float2* __restrict__ s, uint64_t stride, uint64_t count, Expand e,
                                               const float* __restrict__ m, Offsets off) {
  constexpr int D = 1 << K;
  float2* st = s + (uint64_t)blockIdx.y * stride;
  const uint64_t p = (uint64_t)blockIdx.x * TPB + threadIdx.x;
  if (p >= count) return;
  const uint64_t base = expand(p, e);
  float2 v[D];
#pragma unroll
  for (int j = 0; j < D; ++j) v[j] = st[base | off.o[j]];
#pragma unroll
  for (int row = 0; row < D; ++row) {
    float ar = 0.f, ai = 0.f;
#pragma unroll
    for (int col = 0; col < D; ++col) {
      float mr = m[2 * (row * D + col)], mi = m[2 * (row * D + col) + 1];
      ar = __builtin_fmaf(mr, v[col].x, ar);
      ar = __builtin_fmaf(-mi, v[col].y, ar);
      ai = __builtin_fmaf(mr, v[col].y, ai);
      ai = __builtin_fmaf(mi, v[col].x, ai);
    }
    st[base | off.o[row]] = make_float2(ar, ai);
  }
}

// Fused complex64 block on FP32 MFMA (k = 4 or 5).
//
// A block of 2^k amplitudes (one "group", the amplitudes that differ only on
// the block's qubits) is the vector v; the block matrix U is applied as the
// real product [Re w; Im w] = E [Re v; Im v] with E = [[Ur, -Ui], [Ui, Ur]]
// (R = 2^(k+1) rows). A wave takes 32 groups as the 32 columns of B and runs
// v_mfma_f32_32x32x2f32 over K = R in steps of 2:
//   A fragment (step s, row tile t): lane l holds E[32t + l%32][2s + l/32]
//     -- the whole of E stays in registers for the kernel's lifetime;
//   B fragment: lane l holds component 2s + l/32 of group l%32. Components
//     < 2^k are real parts of input 2s + l/32, the rest imaginary parts, so
//     lane l only ever needs the inputs of parity l/32 of its group: it loads
//     those 2^(k-1) amplitudes once and feeds B from registers.
//   D: lane l, register r holds row 8(r/4) + 4(l/32) + r%4 of column l%32.
//     For k = 4 rows < 16 are real parts and rows >= 16 the matching
//     imaginary parts (registers r and r+8); for k = 5 tile 0 holds real
//     and tile 1 imaginary parts of the same 16 outputs. Each lane therefore
//     stores whole complex outputs of its own group.
typedef float floatx16 __attribute__((ext_vector_type(16)));

template <int K, int TILES>
__global__ __launch_bounds__(TPB) void k_fused_mfma(float2* __restrict__ s, uint64_t stride, uint64_t ngroups,
                                                    Expand e, const float* __restrict__ emb, Offsets off) {
  constexpr int D = 1 << K;
  constexpr int R = 2 * D;
  constexpr int T = R / 32;
  constexpr int STEPS = R / 2;
  constexpr int NIN = D / 2;
  float2* st = s + (uint64_t)blockIdx.y * stride;
  const int lane = threadIdx.x & 63;
  const int j = lane & 31, h = lane >> 5;
  float a[T][STEPS];
#pragma unroll
  for (int t = 0; t < T; ++t)
#pragma unroll
    for (int q = 0; q < STEPS; ++q) a[t][q] = emb[(32 * t + j) * R + 2 * q + h];
  uint64_t offs[NIN];
#pragma unroll
  for (int i = 0; i < NIN; ++i) offs[i] = off.o[2 * i + h];
  const uint64_t waves = (uint64_t)gridDim.x * (TPB / 64);
  const uint64_t ntiles = (ngroups + 31) / 32;
  // Each wave takes TILES consecutive tiles per iteration: all their loads are
  // issued before the first MFMA, which doubles the bytes in flight per wave.
  for (uint64_t tile0 = ((uint64_t)blockIdx.x * (TPB / 64) + (threadIdx.x >> 6)) * TILES; tile0 < ntiles;
       tile0 += waves * TILES) {
    uint64_t base[TILES];
    bool valid[TILES];
    float2 v[TILES][NIN];
#pragma unroll
    for (int u = 0; u < TILES; ++u) {
      const uint64_t g = (tile0 + u) * 32 + j;
      valid[u] = g < ngroups;
      base[u] = expand(valid[u] ? g : 0, e);
#pragma unroll
      for (int i = 0; i < NIN; ++i) v[u][i] = valid[u] ? st[base[u] | offs[i]] : make_float2(0.f, 0.f);
    }
    floatx16 acc[TILES][T];
#pragma unroll
    for (int u = 0; u < TILES; ++u)
#pragma unroll
      for (int t = 0; t < T; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[u][t][r] = 0.f;
#pragma unroll
    for (int q = 0; q < STEPS; ++q) {
#pragma unroll
      for (int u = 0; u < TILES; ++u) {
        const float b = q < STEPS / 2 ? v[u][q].x : v[u][q - STEPS / 2].y;
#pragma unroll
        for (int t = 0; t < T; ++t)
          acc[u][t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[t][q], b, acc[u][t], 0, 0, 0);
      }
    }
#pragma unroll
    for (int u = 0; u < TILES; ++u) {
      if (!valid[u]) continue;
      if constexpr (K == 4) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          const int o = 8 * (r / 4) + 4 * h + (r % 4);
          st[base[u] | off.o[o]] = make_float2(acc[u][0][r], acc[u][0][r + 8]);
        }
      } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int o = 8 * (r / 4) + 4 * h + (r % 4);
          st[base[u] | off.o[o]] = make_float2(acc[u][0][r], acc[u][1][r]);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// LDS-tile sweep: several fused blocks per HBM pass
// ---------------------------------------------------------------------------
// Local index t of a tile: bit i <-> bit position spos[i] of the state; the
// tile's other bits come from the tile number (zeros inserted at the tile's
// positions). spos[0..3] are positions 0..3, so 16 consecutive threads move
// one contiguous 128-byte run. A sweep may relabel positions within its
// tile: lane swaps on the loaded registers, then store positions opos.
// A workgroup first copies the sweep's blob (block descriptors, then their
// matrices) into LDS, then walks its tiles persistently: while the blocks of
// one tile are applied in LDS, the loads of its next tile are in flight (the
// block arithmetic reads only LDS, so it never waits on those loads).
// Block kinds (TileBlock.kind):
//   3  k = 2 (1-qubit blocks arrive padded to 2 qubits): 8x8 real embedding
//      on v_mfma_f32_4x4x1f32, each lane computing one whole group;
//   1  k = 3: permuted 16x16 real embedding on v_mfma_f32_16x16x4f32, each
//      lane holding whole complex inputs and outputs;
//   2  k = 4: 32x32 real embedding on v_mfma_f32_32x32x2f32 with the
//      operand layout of k_fused_mfma.
// 5-qubit blocks stay out of sweeps (ordinary launches): their operands would
// raise the kernel's register use past three workgroups per CU.
// MFMA matrices are stored fragment-major (the A value of K step q for lane l
// at [q][l]), so every operand read is one conflict-free 256-byte LDS row.

// Value held by lane (lane ^ (1 << j)) of the same 16-lane row, j < 4 and
// wave-uniform. Every DPP pattern used is its own inverse, so none depends on
// the direction convention of a shift: quad_perm [1,0,3,2] (xor 1),
// quad_perm [2,3,0,1] (xor 2), quad_perm [3,2,1,0] then row_half_mirror
// (xor 3, then xor 7: xor 4), row_ror:8 (xor 8).
__device__ __forceinline__ int lane_xor(int v, uint32_t j) {
  switch (j) {
    case 0: return __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false);
    case 1: return __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false);
    case 2: {
      const int t = __builtin_amdgcn_update_dpp(v, v, 0x1B, 0xF, 0xF, false);
      return __builtin_amdgcn_update_dpp(t, t, 0x141, 0xF, 0xF, false);
    }
    default: return __builtin_amdgcn_update_dpp(v, v, 0x128, 0xF, 0xF, false);
  }
}

// Exchange of lane bit j with a register bit: `a` holds register-bit value
// 0, `b` value 1; `h` is lane bit j of this lane. Afterwards the lane with
// bit j = 0 holds (a, partner's a) and the lane with bit j = 1 holds
// (partner's b, b): lane bit and register bit have traded places. Each lane
// sends the one value its partner needs (one DPP move per word).
__device__ __forceinline__ void lane_swap(float& ax, float& ay, float& bx, float& by, uint32_t j, bool h) {
  const float sx = h ? ax : bx, sy = h ? ay : by;
  const float rx = __int_as_float(lane_xor(__float_as_int(sx), j));
  const float ry = __int_as_float(lane_xor(__float_as_int(sy), j));
  if (h) {
    ax = rx;
    ay = ry;
  } else {
    bx = rx;
    by = ry;
  }
}
__device__ __forceinline__ double lane_xor_d(double v, uint32_t j) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)lane_xor((int)(uint32_t)u, j), hi = (uint32_t)lane_xor((int)(uint32_t)(u >> 32), j);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ void lane_swap(double& ax, double& ay, double& bx, double& by, uint32_t j, bool h) {
  const double sx = h ? ax : bx, sy = h ? ay : by;
  const double rx = lane_xor_d(sx, j), ry = lane_xor_d(sy, j);
  if (h) {
    ax = rx;
    ay = ry;
  } else {
    bx = rx;
    by = ry;
  }
}

// Lane swaps of a loaded tile (registers of local index tid + TPB i, real
// and imaginary parts in x[i], y[i]) before it goes into LDS: for s <
// nswap, lane bit sbit[s] trades places with register bit s. In-register
// (DPP), so the relabelling costs no LDS traffic; the sweep's blocks address
// the swapped local order.
template <int PER, typename T>
__device__ __forceinline__ void tile_lane_swaps(T (&x)[PER], T (&y)[PER], const TileArgs& a) {
  const uint32_t nswap = __builtin_amdgcn_readfirstlane(a.nswap);
  if (nswap == 0) return;
  auto pass = [&](int bit, uint32_t j) {
    const bool h = (threadIdx.x >> j) & 1;
#pragma unroll
    for (int i = 0; i < PER; ++i)
      if (!(i & bit)) lane_swap(x[i], y[i], x[i + bit], y[i + bit], j, h);
  };
  pass(1, __builtin_amdgcn_readfirstlane(a.sbit[0]));
  if (nswap == 2) pass(2, __builtin_amdgcn_readfirstlane(a.sbit[1]));
}
template <int PER>
__device__ __forceinline__ void tile_lane_swaps(float2 (&r)[PER], const TileArgs& a) {
  if (__builtin_amdgcn_readfirstlane(a.nswap) == 0) return;
  float x[PER], y[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    x[i] = r[i].x;
    y[i] = r[i].y;
  }
  tile_lane_swaps(x, y, a);
#pragma unroll
  for (int i = 0; i < PER; ++i) r[i] = make_float2(x[i], y[i]);
}

// Tiles of split-precision sweeps sit in LDS scaled by 2^14: amplitudes of a
// normalised state (|a| <= 1) then map into [2^-14, 2^14] for every |a| >=
// 2^-28, FP16's normal range, and stay below its 65504 limit with room for
// FP32 rounding of the norm. Powers of two scale exactly.
constexpr float kSplitScale = 16384.f, kSplitUnscale = 1.f / 16384.f;
__device__ __forceinline__ float2 split_scale(float2 v) { return make_float2(v.x * kSplitScale, v.y * kSplitScale); }

// Store of a tile from LDS: local bit i goes to bit position opos[i] (equal
// to spos[i] unless the sweep relabels).
template <int L, typename V, bool SPLIT = false>
__device__ __forceinline__ void tile_store(V* __restrict__ st, const V* tile, uint64_t base, const TileArgs& a) {
  constexpr int PER = (1 << L) / TPB;
  uint64_t olo = 0;
#pragma unroll
  for (int bit = 0; bit < 8; ++bit)
    if ((threadIdx.x >> bit) & 1) olo |= 1ull << a.opos[bit];
  // Opaque to the optimiser: otherwise the PER per-lane offsets olo | ohi(i)
  // are hoisted out of the tile loop into 2 * PER more VGPRs.
  asm volatile("" : "+v"(olo));
  uint64_t om[L - 8];
#pragma unroll
  for (int bit = 0; bit < L - 8; ++bit) om[bit] = 1ull << a.opos[8 + bit];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    uint64_t o = 0;
#pragma unroll
    for (int bit = 0; bit < L - 8; ++bit)
      if ((i >> bit) & 1) o |= om[bit];
    if constexpr (SPLIT) {
      const float2 v = tile[threadIdx.x + TPB * i];
      st[base | olo | o] = make_float2(v.x * kSplitUnscale, v.y * kSplitUnscale);
    } else {
      st[base | olo | o] = tile[threadIdx.x + TPB * i];
    }
  }
}

__device__ __forceinline__ uint32_t insert_zeros32(uint32_t p, const uint32_t* srt, int k) {
  for (int i = 0; i < k; ++i) {
    const uint32_t low = (1u << srt[i]) - 1;
    p = (p & low) | ((p & ~low) << 1);
  }
  return p;
}

template <int K>
__device__ __forceinline__ uint32_t local_offset(const TileBlock& b, int c) {
  uint32_t o = 0;
#pragma unroll
  for (int q = 0; q < K; ++q)
    if ((c >> (K - 1 - q)) & 1) o |= 1u << b.loc[q];
  return o;
}

// Group g of a block sits at local index IZ(g) (zeros inserted at the block's
// local bits). With g = u + v, u and v on disjoint bits, IZ(g) = IZ(u) | IZ(v):
// the lane-dependent part is fixed per block and the loop-dependent part is
// wave-uniform, so the address update per group runs on the scalar unit (the
// block's sorted bits are read into scalar registers first). An XOR swizzle of
// the LDS slots halved the bank conflicts but its address arithmetic cost more
// than it saved (1.66 s vs 1.51 s per 32-qubit step), so tiles are unswizzled.
template <int K>
__device__ __forceinline__ void uniform_bits(const TileBlock& b, uint32_t* srt) {
#pragma unroll
  for (int i = 0; i < K; ++i) srt[i] = __builtin_amdgcn_readfirstlane(b.srt[i]);
}

typedef float floatx4 __attribute__((ext_vector_type(4)));

// Groups of a block in one wave: local index = the lane's part | a
// wave-uniform part, on disjoint bits. Without WB the wave bits are the two
// group-index bits above the lane's (wave v takes every fourth row of
// groups, a different set for every block). With WB (256-thread workgroups
// of out-of-place sweeps) the block belongs to a run of blocks that leave two
// tile bits w0 < w1 untouched, and wave v owns the local indices whose bits
// w0, w1 equal v's two bits for every block of the run: consecutive blocks of
// a run need no workgroup barrier, since each wave reads back only what it
// wrote itself and a wave's LDS accesses complete in issue order. The
// descriptor carries pad[2] = w0 | w1 << 8 and pad[0] = the same two bits
// with the block's bits below each removed, so zeros inserted at pad[0]'s
// positions and then at the block's bits land at the block's bits and w0, w1.
// `lanes` groups are covered per step by the lane part (lane_grp < lanes).
template <int K, bool WB, int NTH>
struct GroupWalk {
  uint32_t lane_part, mask, step, base, wpart;
  __device__ __forceinline__ GroupWalk(const TileBlock& b, const uint32_t* srt, uint32_t lane_grp, uint32_t lanes) {
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    mask = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) mask |= 1u << srt[i];
    if constexpr (WB) {
      static_assert(NTH == 256, "wave-owned groups take two wave bits");
      const uint32_t wl = __builtin_amdgcn_readfirstlane(b.pad[2]), wc = __builtin_amdgcn_readfirstlane(b.pad[0]);
      const uint32_t c[2] = {wc & 255u, wc >> 8};
      const uint32_t w0 = wl & 255u, w1 = wl >> 8;
      lane_part = insert_zeros32(insert_zeros32(lane_grp, c, 2), srt, K);
      step = insert_zeros32(insert_zeros32(lanes, c, 2), srt, K);
      mask |= (1u << w0) | (1u << w1);
      base = 0;
      wpart = ((w & 1u) << w0) | ((w >> 1) << w1);
    } else {
      lane_part = insert_zeros32(lane_grp, srt, K);
      step = insert_zeros32(lanes * (NTH / 64), srt, K);
      base = insert_zeros32(w * lanes, srt, K);
      wpart = 0;
    }
  }
  // Wave-uniform part of the next group row (masked-carry stepping: with the
  // skipped bits set, adding `step` carries over them).
  __device__ __forceinline__ uint32_t next() {
    const uint32_t g = __builtin_amdgcn_readfirstlane(base) | wpart;
    base = ((base | mask) + step) & ~mask;
    return g;
  }
};

// k = 3 on v_mfma_f32_16x16x4f32 (A lane l: A[l%16][l/16]; B lane l:
// B[l/16][l%16]; D lane l, register r: D[4(l/16)+r][l%16]). The embedding
// e (16x16, row-major) is permuted so that
//   row 2c+p    = real (p=0) / imaginary (p=1) part of output c,
//   column 4s+u = input 2u + s/2, real (s even) / imaginary (s odd) part,
// so lane l (u = l/16) feeds the four K steps from inputs 2u and 2u+1 of
// group l%16 and receives outputs 2u and 2u+1 of that group whole.
// `e` holds the A fragments: e[64 s + l] = embedding[l % 16][4 s + l / 16].
template <int L, int NTH = TPB, bool WB = false>
__device__ __forceinline__ void tile_block_mfma16(float2* tile, const TileBlock& b, const float* e) {
  const int lane = threadIdx.x & 63;
  const int col = lane & 15, u = lane >> 4;
  float a[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) a[s] = e[64 * s + lane];
  uint32_t srt[3];
  uniform_bits<3>(b, srt);
  GroupWalk<3, WB, NTH> gw(b, srt, col, 16);
  const uint32_t lane_part = gw.lane_part;
  const uint32_t o0 = local_offset<3>(b, 2 * u) ^ lane_part, o1 = local_offset<3>(b, 2 * u + 1) ^ lane_part;
  // 16-group tiles per wave, U of them per iteration with independent MFMA
  // chains interleaved.
  constexpr int NT = (1 << L) / 8 / 16;  // 16-group tiles of the block
  constexpr int PERW = NT / (NTH / 64);   // L >= 10: at least 2 per wave
  constexpr int U = PERW >= 4 ? 4 : PERW;
  static_assert(NT % (NTH / 64) == 0 && PERW % U == 0, "16-group tiles split evenly over the waves");
  // Byte addresses: the lane's part plus a wave-uniform group part (scalar) on
  // disjoint bits.
  unsigned char* p0 = (unsigned char*)(tile + o0);
  unsigned char* p1 = (unsigned char*)(tile + o1);
#pragma unroll
  for (int it = 0; it < PERW; it += U) {
    uint32_t bs[U];
    float2 x0[U], x1[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      bs[j] = gw.next() * 8u;
      x0[j] = *(const float2*)(p0 + bs[j]);
      x1[j] = *(const float2*)(p1 + bs[j]);
    }
    floatx4 acc[U];
#pragma unroll
    for (int j = 0; j < U; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < U; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], x0[j].x, acc[j], 0, 0, 0);
#pragma unroll
    for (int j = 0; j < U; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1], x0[j].y, acc[j], 0, 0, 0);
#pragma unroll
    for (int j = 0; j < U; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2], x1[j].x, acc[j], 0, 0, 0);
#pragma unroll
    for (int j = 0; j < U; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[3], x1[j].y, acc[j], 0, 0, 0);
#pragma unroll
    for (int j = 0; j < U; ++j) {
      *(float2*)(p0 + bs[j]) = make_float2(acc[j][0], acc[j][1]);
      *(float2*)(p1 + bs[j]) = make_float2(acc[j][2], acc[j][3]);
    }
  }
}

typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));

// Two FP32 values split into FP16 parts, a = ha + la and b = hb + lb to about
// 2^-23 relative: h packs (ha, hb) rounded to nearest even, l packs the
// residuals a - ha, b - hb, each formed exactly by one fused multiply-add
// (a * one - ha with the FP16 operand widened exactly) and rounded to
// nearest even into the low or high half of l: three instructions per pair
// (v_cvt_pk_f16_f32, v_fma_mixlo_f16, v_fma_mixhi_f16). `one` is 1.0 hidden
// from constant folding (see split_one), which keeps each residual a single
// mixed-precision multiply-add; written in C, the compiler also inserts the
// wait states its writes need before the matrix core reads them. Inputs are
// tile values held scaled by kSplitScale, so they stay in FP16's normal range.
__device__ __forceinline__ void split_f16(float a, float b, float one, uint32_t& h, uint32_t& l) {
#pragma clang fp contract(fast)
  const halfx2 hv = __builtin_convertvector((floatx2){a, b}, halfx2);
  const _Float16 la = (_Float16)(a * one - (float)hv[0]);
  const _Float16 lb = (_Float16)(b * one - (float)hv[1]);
  h = __builtin_bit_cast(uint32_t, hv);
  l = __builtin_bit_cast(uint32_t, (halfx2){la, lb});
}
// 1.0 in a scalar register, opaque to constant folding (an empty asm
// statement: no instruction is emitted).
__device__ __forceinline__ float split_one() {
  float one = 1.0f;
  asm volatile("" : "+s"(one));
  return one;
}

// k = 3 in split precision on v_mfma_f32_16x16x32_f16 (A lane l: A[l%16][8u
// + j], B lane l: B[8u + j][l%16], u = l/16, j = 0..7; D as for 16x16x4).
// Lane l holds inputs 2u and 2u+1 of group l%16 whole, as for the FP32 path,
// and feeds its 8 K slots with their FP16 parts: slots 0-1 high parts of
// input 2u (real, imaginary), 2-3 its low parts, 4-7 the same for input
// 2u+1. The embedding E (rows and columns as for the FP32 path) is split
// the same way, E = Eh + El, and two MFMAs accumulate in FP32
//   D = [Eh Eh] [xh; xl] + [El El] [xh; xl] = (Eh + El)(xh + xl),
// all four partial products, so the result matches an FP32 contraction to
// about 2^-22 relative per element: a quarter of the 16x16x4 path's matrix
// cycles for six VALU operations per lane and group tile. `e` holds per lane
// l four packed FP16 pairs: e[64 s + l], s = 0, 1: Eh of inputs 2u + s (real
// part column, imaginary part column) for row l%16; s = 2, 3: the same of El.
template <int L, int NTH = TPB, bool WB = false>
__device__ __forceinline__ void tile_block_mfma16h(float2* tile, const TileBlock& b, const uint32_t* e) {
  const int lane = threadIdx.x & 63;
  const int col = lane & 15, u = lane >> 4;
  uint32_t f[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) f[s] = e[64 * s + lane];
  const uint32_t fh[4] = {f[0], f[0], f[1], f[1]}, fl[4] = {f[2], f[2], f[3], f[3]};
  const halfx8 ah = __builtin_bit_cast(halfx8, fh), al = __builtin_bit_cast(halfx8, fl);
  const float one = split_one();
  uint32_t srt[3];
  uniform_bits<3>(b, srt);
  GroupWalk<3, WB, NTH> gw(b, srt, col, 16);
  const uint32_t lane_part = gw.lane_part;
  const uint32_t o0 = local_offset<3>(b, 2 * u) ^ lane_part, o1 = local_offset<3>(b, 2 * u + 1) ^ lane_part;
  constexpr int NT = (1 << L) / 8 / 16;  // 16-group tiles of the block
  constexpr int PERW = NT / (NTH / 64);
  constexpr int U = PERW >= 4 ? 4 : PERW;
  static_assert(NT % (NTH / 64) == 0 && PERW % U == 0, "16-group tiles split evenly over the waves");
  // Byte addresses: the lane's part (o0, o1 on the block's bits and the lane's
  // group bits) plus a wave-uniform group part on the other bits (scalar);
  // the parts share no bit, so adding them equals OR-ing the indices.
  unsigned char* p0 = (unsigned char*)(tile + o0);
  unsigned char* p1 = (unsigned char*)(tile + o1);
#pragma unroll
  for (int it = 0; it < PERW; it += U) {
    uint32_t bs[U];
    float2 x0[U], x1[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      bs[j] = gw.next() * 8u;
      x0[j] = *(const float2*)(p0 + bs[j]);
      x1[j] = *(const float2*)(p1 + bs[j]);
    }
    floatx4 acc[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      uint32_t wv[4];
      split_f16(x0[j].x, x0[j].y, one, wv[0], wv[1]);
      split_f16(x1[j].x, x1[j].y, one, wv[2], wv[3]);
      const halfx8 bv = __builtin_bit_cast(halfx8, wv);
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bv, floatx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bv, acc[j], 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
      *(float2*)(p0 + bs[j]) = make_float2(acc[j][0], acc[j][1]);
      *(float2*)(p1 + bs[j]) = make_float2(acc[j][2], acc[j][3]);
    }
  }
}

// k = 2 (1-qubit blocks arrive padded to 2) on v_mfma_f32_4x4x1f32: 16
// blocks b of 4x4x1 per instruction (A lane 4b+i: A_b[i][0]; B lane 4b+j:
// B_b[0][j]; D lane 4b+j, register i: D_b[i][j]). Every lane owns one group
// of 4 amplitudes and feeds its 8 real inputs, one per K step, as column j of
// block b = lane/4; its A operand at step kk is E[4h + lane%4][kk] of the 8x8
// real embedding E (the same in every block), so D_b[i][j] accumulates output
// row 4h + i of the lane's own group. Two accumulators (h = 0, 1) give the
// lane its group's 4 outputs: 0.5 MFMA cycles per amplitude, half of a
// 3-qubit block's. `e` holds the A fragments: e[16 i + 8 h + kk] = E[4h+i][kk].
template <int L, int NTH = TPB, bool WB = false>
__device__ __forceinline__ void tile_block_mfma4(float2* tile, const TileBlock& b, const float* e) {
  const int lane = threadIdx.x & 63;
  float a[16];
  const float4* e4 = (const float4*)(e + 16 * (lane & 3));
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float4 v = e4[q];
    a[4 * q] = v.x;
    a[4 * q + 1] = v.y;
    a[4 * q + 2] = v.z;
    a[4 * q + 3] = v.w;
  }
  uint32_t srt[2];
  uniform_bits<2>(b, srt);
  GroupWalk<2, WB, NTH> gw(b, srt, lane, 64);
  const uint32_t lane_part = gw.lane_part;
  uint32_t off[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) off[j] = local_offset<2>(b, j) ^ lane_part;
  // 64-group rows per wave, U of them per iteration with independent chains.
  constexpr int NR = (1 << L) / 4 / 64;  // 64-group rows of the block
  constexpr int PERW = NR / (NTH / 64);  // L >= 10: at least 1 per wave
  constexpr int U = 1;
  static_assert(NR % (NTH / 64) == 0 && PERW % U == 0, "64-group rows split evenly over the waves");
  // Byte addresses: the lane's part plus a wave-uniform group part (scalar) on
  // disjoint bits.
  unsigned char* pj[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) pj[j] = (unsigned char*)(tile + off[j]);
#pragma unroll
  for (int it = 0; it < PERW; it += U) {
    uint32_t bs[U];
    float2 x[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      bs[u] = gw.next() * 8u;
#pragma unroll
      for (int j = 0; j < 4; ++j) x[u][j] = *(const float2*)(pj[j] + bs[u]);
    }
    floatx4 lo[U], hi[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      lo[u] = floatx4{0.f, 0.f, 0.f, 0.f};
      hi[u] = floatx4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float bv = (kk & 1) ? x[u][kk >> 1].y : x[u][kk >> 1].x;
        lo[u] = __builtin_amdgcn_mfma_f32_4x4x1f32(a[kk], bv, lo[u], 0, 0, 0);
        hi[u] = __builtin_amdgcn_mfma_f32_4x4x1f32(a[8 + kk], bv, hi[u], 0, 0, 0);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      *(float2*)(pj[0] + bs[u]) = make_float2(lo[u][0], lo[u][1]);
      *(float2*)(pj[1] + bs[u]) = make_float2(lo[u][2], lo[u][3]);
      *(float2*)(pj[2] + bs[u]) = make_float2(hi[u][0], hi[u][1]);
      *(float2*)(pj[3] + bs[u]) = make_float2(hi[u][2], hi[u][3]);
    }
  }
}

// k = 4 on v_mfma_f32_32x32x2f32 with the operand layout of k_fused_mfma.
// `emb` holds the A fragments: emb[64 q + l] = E[l % 32][2 q + l / 32].
template <int L, int NTH = TPB>
__device__ __forceinline__ void tile_block_mfma32(float2* tile, const TileBlock& b, const float* emb) {
  constexpr int K = 4, STEPS = 16, NIN = 8, NOUT = 8;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 31, h = lane >> 5;
  uint32_t srt[K];
  uniform_bits<K>(b, srt);
  const uint32_t lane_part = insert_zeros32(j, srt, K);
  uint32_t offs[NIN], outs[NOUT];
#pragma unroll
  for (int i = 0; i < NIN; ++i) offs[i] = local_offset<K>(b, 2 * i + h) ^ lane_part;
#pragma unroll
  for (int r = 0; r < NOUT; ++r) outs[r] = local_offset<K>(b, 8 * (r / 4) + 4 * h + (r % 4)) ^ lane_part;
  // 32-group tiles per wave (at L = 10 only waves 0 and 1 have one), U per
  // iteration with independent MFMA chains sharing each A operand read.
  constexpr int NT = (1 << L) / 16 / 32;
  constexpr int PERW = (NT + NTH / 64 - 1) / (NTH / 64);
  constexpr int U = PERW >= 2 ? 2 : 1;
  static_assert(PERW % U == 0, "32-group tiles per wave split evenly");
  const uint32_t mask = (1u << srt[0]) | (1u << srt[1]) | (1u << srt[2]) | (1u << srt[3]);
  const uint32_t w = __builtin_amdgcn_readfirstlane(wave);
  const uint32_t step = insert_zeros32(32 * (NTH / 64), srt, K);
  uint32_t base = insert_zeros32(w * 32, srt, K);
  if (NT < NTH / 64 && (int)w >= NT) return;  // wave-uniform: no 32-group tile for this wave
#pragma unroll
  for (int it = 0; it < PERW; it += U) {
    uint32_t bs[U];
    float2 v[U][NIN];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      bs[u] = base;
      base = ((base | mask) + step) & ~mask;
#pragma unroll
      for (int i = 0; i < NIN; ++i) v[u][i] = tile[bs[u] ^ offs[i]];
    }
    floatx16 acc[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[u][r] = 0.f;
#pragma unroll
    for (int q = 0; q < STEPS; ++q) {
      const float av = emb[q * 64 + lane];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float bv = q < STEPS / 2 ? v[u][q].x : v[u][q - STEPS / 2].y;
        acc[u] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc[u], 0, 0, 0);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int r = 0; r < NOUT; ++r) tile[bs[u] ^ outs[r]] = make_float2(acc[u][r], acc[u][r + 8]);
  }
}

// k = 4 in split precision on v_mfma_f32_32x32x16_f16 (A lane l: A[l%32][8h
// + j], B lane l: B[8h + j][l%32], h = l/32, j = 0..7; D as for 32x32x2).
// Group tiles, inputs, outputs and the embedding's rows and columns are those
// of tile_block_mfma32: lane half h holds inputs 2i + h (i = 0..7) of group
// l%32 whole. Instruction t (0..3) takes the lane's inputs i = 2t and 2t+1 as
// slots 0-1 high parts (real, imaginary) of input 2(2t) + h, 2-3 its low
// parts, 4-7 the same for input 2(2t+1) + h; per t the A fragments repeat the
// embedding columns of those inputs over their high and low slots, once with
// Eh and once with El, so the eight MFMAs accumulate all four partial
// products (Eh + El)(xh + xl) in FP32. `e` holds per (t, lane) four dwords
// e[4 (64 t + l) + w]: w = 0, 1 the packed Eh pairs (real-, imaginary-part
// column) of the two inputs, w = 2, 3 the El pairs. The tile is held scaled by
// kSplitScale.
template <int L, int NTH = TPB>
__device__ __forceinline__ void tile_block_mfma32h(float2* tile, const TileBlock& b, const uint32_t* e) {
  constexpr int K = 4, NIN = 8, NOUT = 8;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 31, h = lane >> 5;
  uint32_t srt[K];
  uniform_bits<K>(b, srt);
  const uint32_t lane_part = insert_zeros32(j, srt, K);
  uint32_t offs[NIN], outs[NOUT];
#pragma unroll
  for (int i = 0; i < NIN; ++i) offs[i] = local_offset<K>(b, 2 * i + h) ^ lane_part;
#pragma unroll
  for (int r = 0; r < NOUT; ++r) outs[r] = local_offset<K>(b, 8 * (r / 4) + 4 * h + (r % 4)) ^ lane_part;
  constexpr int NT = (1 << L) / 16 / 32;
  constexpr int PERW = (NT + NTH / 64 - 1) / (NTH / 64);
  constexpr int U = PERW >= 2 ? 2 : 1;
  static_assert(PERW % U == 0, "32-group tiles per wave split evenly");
  const uint32_t mask = (1u << srt[0]) | (1u << srt[1]) | (1u << srt[2]) | (1u << srt[3]);
  const uint32_t w = __builtin_amdgcn_readfirstlane(wave);
  const uint32_t step = insert_zeros32(32 * (NTH / 64), srt, K);
  uint32_t base = insert_zeros32(w * 32, srt, K);
  if (NT < NTH / 64 && (int)w >= NT) return;  // wave-uniform: no 32-group tile for this wave
  const float one = split_one();
  const uint4* e4 = (const uint4*)e;
#pragma unroll
  for (int it = 0; it < PERW; it += U) {
    uint32_t bs[U];
    float2 v[U][NIN];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      bs[u] = base;
      base = ((base | mask) + step) & ~mask;
#pragma unroll
      for (int i = 0; i < NIN; ++i) v[u][i] = tile[bs[u] ^ offs[i]];
    }
    floatx16 acc[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[u][r] = 0.f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const uint4 f = e4[64 * t + lane];
      const uint32_t fh[4] = {f.x, f.x, f.y, f.y}, fl[4] = {f.z, f.z, f.w, f.w};
      const halfx8 ah = __builtin_bit_cast(halfx8, fh), al = __builtin_bit_cast(halfx8, fl);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        uint32_t wv[4];
        split_f16(v[u][2 * t].x, v[u][2 * t].y, one, wv[0], wv[1]);
        split_f16(v[u][2 * t + 1].x, v[u][2 * t + 1].y, one, wv[2], wv[3]);
        const halfx8 bv = __builtin_bit_cast(halfx8, wv);
        acc[u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bv, acc[u], 0, 0, 0);
        acc[u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bv, acc[u], 0, 0, 0);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int r = 0; r < NOUT; ++r) tile[bs[u] ^ outs[r]] = make_float2(acc[u][r], acc[u][r + 8]);
  }
}

// K4: the sweep holds 4-qubit blocks (the 32x32x2 path is compiled in only
// then, keeping the register count of 3-qubit sweeps at four workgroups/CU).
// SPLIT: 3-qubit blocks run in split precision (kind 4) on a tile held scaled
// by kSplitScale in LDS.
template <int L, bool K4, bool SPLIT>
__global__ __launch_bounds__(TPB) void k_tile(float2* __restrict__ s, uint64_t stride, uint64_t ntiles, TileArgs a,
                                              const float4* __restrict__ blob) {
  extern __shared__ float4 lds4[];
  constexpr int PER = (1 << L) / TPB;
  float2* tile = (float2*)lds4;
  const unsigned char* meta = (const unsigned char*)(lds4 + (1 << L) / 2);
  for (uint32_t i = threadIdx.x; i < a.blob16; i += TPB) lds4[(1 << L) / 2 + i] = blob[i];
  const TileBlock* blocks = (const TileBlock*)meta;
  float2* st = s + (uint64_t)blockIdx.y * stride;
  // Global offset of local index t = tid + TPB * i: low part from the thread,
  // high part from i (disjoint bit sets, so the offsets OR together).
  uint64_t lo = 0;
#pragma unroll
  for (int bit = 0; bit < 8; ++bit)
    if ((threadIdx.x >> bit) & 1) lo |= 1ull << a.spos[bit];
  asm volatile("" : "+v"(lo));  // keeps the PER offsets lo | hi(i) out of VGPRs
  uint64_t hm[L - 8];
#pragma unroll
  for (int bit = 0; bit < L - 8; ++bit) hm[bit] = 1ull << a.spos[8 + bit];
  auto hi = [&](int i) {
    uint64_t o = 0;
#pragma unroll
    for (int bit = 0; bit < L - 8; ++bit)
      if ((i >> bit) & 1) o |= hm[bit];
    return o;
  };
  auto tile_base = [&](uint64_t tau) {
    uint64_t base = tau;
#pragma unroll
    for (int i = 0; i < L; ++i) {
      const uint64_t low = (1ull << a.srt[i]) - 1;
      base = (base & low) | ((base & ~low) << 1);
    }
    return base;
  };
  // Pipeline per tile: issue the loads of the next tile, apply the blocks to
  // this tile in LDS, store it, then move the loaded registers into LDS. The
  // next tile's loads are consumed in the iteration that issued them, after
  // this tile's stores, so the wait for them lets the stores stay in flight
  // (vector-memory completions count in issue order, stores included).
  float2 r[PER];
  uint64_t tau = blockIdx.x;
  if (tau < ntiles) {
    const uint64_t b0 = tile_base(tau);
#pragma unroll
    for (int i = 0; i < PER; ++i) r[i] = st[b0 | lo | hi(i)];
    tile_lane_swaps(r, a);
#pragma unroll
    for (int i = 0; i < PER; ++i) tile[threadIdx.x + TPB * i] = SPLIT ? split_scale(r[i]) : r[i];
  }
  __syncthreads();  // the first tile and the blob are in LDS
  for (; tau < ntiles; tau += gridDim.x) {
    const uint64_t base = tile_base(tau);
    const uint64_t nxt = tau + gridDim.x;
    const bool more = nxt < ntiles;
    if (more) {
      const uint64_t nb = tile_base(nxt);
#pragma unroll
      for (int i = 0; i < PER; ++i) r[i] = st[nb | lo | hi(i)];
    }
    for (uint32_t bi = 0; bi < a.nblocks; ++bi) {
      const TileBlock& b = blocks[bi];
      const float* m = (const float*)(meta + b.mat_off);
      if (b.kind == 3) tile_block_mfma4<L>(tile, b, m);
      else if (SPLIT && b.kind == 4) tile_block_mfma16h<L>(tile, b, (const uint32_t*)m);
      else if (K4 && SPLIT && b.kind == 5) tile_block_mfma32h<L>(tile, b, (const uint32_t*)m);
      else if (!K4 || b.kind == 1) tile_block_mfma16<L>(tile, b, m);
      else tile_block_mfma32<L>(tile, b, m);
      __syncthreads();
    }
    tile_store<L, float2, SPLIT>(st, tile, base, a);
    if (more) {
      tile_lane_swaps(r, a);
      __syncthreads();  // every wave has read this tile out of LDS
#pragma unroll
      for (int i = 0; i < PER; ++i) tile[threadIdx.x + TPB * i] = SPLIT ? split_scale(r[i]) : r[i];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Out-of-place (permuting) tile sweep, complex64 (XSweepArgs in kernels.hpp).
// ---------------------------------------------------------------------------
// Bit scatter of the low `nb` bits of v to positions pos[0..nb).
__device__ __forceinline__ uint64_t scatter_bits(uint64_t v, const uint32_t* pos, uint32_t nb) {
  uint64_t o = 0;
  for (uint32_t k = 0; k < nb; ++k) o |= ((v >> k) & 1ull) << pos[k];
  return o;
}

// The blocks of one sweep applied to the LDS tile. Without WB a barrier
// follows every block; with WB (wave-owned groups, see GroupWalk) only the
// last block of each run, as its descriptor's pad[1] bit 0 says (blocks of 4
// qubits keep the per-block wave split and form runs of their own).
template <int L, bool K4, int NTH, bool SPLIT, bool WB = false>
__device__ __forceinline__ void tile_apply_blocks(float2* tile, const TileBlock* blocks, const unsigned char* meta,
                                                  uint32_t nblocks) {
  for (uint32_t bi = 0; bi < nblocks; ++bi) {
    const TileBlock& b = blocks[bi];
    const float* m = (const float*)(meta + b.mat_off);
    if (b.kind == 3) tile_block_mfma4<L, NTH, WB>(tile, b, m);
    else if (SPLIT && b.kind == 4) tile_block_mfma16h<L, NTH, WB>(tile, b, (const uint32_t*)m);
    else if (K4 && SPLIT && b.kind == 5) tile_block_mfma32h<L, NTH>(tile, b, (const uint32_t*)m);
    else if (!K4 || b.kind == 1) tile_block_mfma16<L, NTH, WB>(tile, b, m);
    else tile_block_mfma32<L, NTH>(tile, b, m);
    if (!WB || (__builtin_amdgcn_readfirstlane(b.pad[1]) & 1u)) __syncthreads();
  }
}

// Store of a tile from LDS in output order: store index u = tid + TPB * i
// reads local index lo_l | hi_l(i) and writes destination offset lo_w | hi_w(i)
// (plus the tile's base), so each wave's stores cover runs of consecutive
// addresses as long as the low destination positions are tile bits. T is
// float2 (complex64 tiles) or double2 (exact complex128 tiles, SPLIT false).
template <int L, int NTH, bool SPLIT, bool ZERO = false, typename T = float2>
__device__ __forceinline__ void xsweep_store(T* __restrict__ dst, const T* tile, uint64_t wbase,
                                             const XSweepArgs& a) {
  constexpr int PER = (1 << L) / NTH;
  constexpr int LB = NTH == 512 ? 9 : 8;  // thread-index bits
  uint32_t lo_l = 0;
  uint64_t lo_w = 0;
#pragma unroll
  for (int bit = 0; bit < LB; ++bit)
    if ((threadIdx.x >> bit) & 1) {
      lo_l |= 1u << a.wlocal[bit];
      lo_w |= 1ull << a.wpos[bit];
    }
  // Opaque to the optimiser, so the PER per-store offsets are formed per use
  // instead of being hoisted out of the tile loop into 3 * PER VGPRs.
  asm volatile("" : "+v"(lo_l), "+v"(lo_w));
  uint32_t hl[L - LB];
  uint64_t hw[L - LB];
#pragma unroll
  for (int bit = 0; bit < L - LB; ++bit) {
    hl[bit] = 1u << a.wlocal[LB + bit];
    hw[bit] = 1ull << a.wpos[LB + bit];
  }
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    uint32_t l = lo_l;
    uint64_t w = wbase | lo_w;
#pragma unroll
    for (int bit = 0; bit < L - LB; ++bit)
      if ((i >> bit) & 1) {
        l |= hl[bit];
        w |= hw[bit];
      }
    if constexpr (ZERO) {
      if constexpr (sizeof(T) == 16) dst[w] = make_double2(0.0, 0.0);
      else dst[w] = make_float2(0.f, 0.f);
    } else if constexpr (SPLIT) {
      const float2 v = tile[l];
      dst[w] = make_float2(v.x * kSplitUnscale, v.y * kSplitUnscale);
    } else {
      dst[w] = tile[l];
    }
  }
}

// Source offsets of local index tid + TPB * i: contiguous (the tile is one
// run starting at tau << L) or gathered from positions rpos.
template <int L, int NTH>
struct XSweepRead {
  static constexpr int LB = NTH == 512 ? 9 : 8;  // thread-index bits
  uint64_t lo;
  uint64_t hm[L - LB];
  __device__ __forceinline__ explicit XSweepRead(const XSweepArgs& a) {
    lo = 0;
    if (a.gather) {
#pragma unroll
      for (int bit = 0; bit < LB; ++bit)
        if ((threadIdx.x >> bit) & 1) lo |= 1ull << a.rpos[bit];
#pragma unroll
      for (int bit = 0; bit < L - LB; ++bit) hm[bit] = 1ull << a.rpos[LB + bit];
    } else {
      lo = threadIdx.x;
#pragma unroll
      for (int bit = 0; bit < L - LB; ++bit) hm[bit] = (uint64_t)NTH << bit;
    }
    asm volatile("" : "+v"(lo));
  }
  __device__ __forceinline__ uint64_t base(uint64_t tau, const XSweepArgs& a) const {
    if (!a.gather) return tau << L;
    uint64_t b = tau;
#pragma unroll
    for (int i = 0; i < L; ++i) {
      const uint64_t low = (1ull << a.rsrt[i]) - 1;
      b = (b & low) | ((b & ~low) << 1);
    }
    return b;
  }
  __device__ __forceinline__ uint64_t off(int i) const {
    uint64_t o = lo;
#pragma unroll
    for (int bit = 0; bit < L - LB; ++bit)
      if ((i >> bit) & 1) o |= hm[bit];
    return o;
  }
};

// GBLOB: the blob (block descriptors and matrices) is read from global
// memory, where every workgroup of the launch finds it in cache, instead of
// being copied into LDS behind the tile; used for blobs large enough to cost
// resident workgroups (see xsweep_blob_in_lds).
template <int L, bool K4, bool PERSIST, int NTH, bool GBLOB, bool SPLIT>
__global__ __launch_bounds__(NTH) void k_xsweep(const float2* __restrict__ src, float2* __restrict__ dst,
                                                uint64_t stride, uint64_t ntiles, XSweepArgs a,
                                                const float4* __restrict__ blob) {
  extern __shared__ float4 lds4[];
  constexpr int PER = (1 << L) / NTH;
  float2* tile = (float2*)lds4;
  if (!PERSIST && (blockIdx.x & a.zero_mask) != 0) {
    // A tile of zeros (a qubit fixed to 1 by its index is 0 in every
    // amplitude of the source, e.g. every tile but tile 0 of |0...0>): every
    // block maps it to zeros, so it only stores zeros (no read, no blob copy,
    // no block work).
    xsweep_store<L, NTH, SPLIT, true>(dst + (uint64_t)blockIdx.y * stride, tile,
                                      scatter_bits(blockIdx.x, a.wtile, a.ntb), a);
    return;
  }
  const unsigned char* meta;
  if (GBLOB) {
    meta = (const unsigned char*)blob;
  } else {
    meta = (const unsigned char*)(lds4 + (1 << L) / 2);
    for (uint32_t i = threadIdx.x; i < a.blob16; i += NTH) lds4[(1 << L) / 2 + i] = blob[i];
  }
  const TileBlock* blocks = (const TileBlock*)meta;
  const float2* sb = src + (uint64_t)blockIdx.y * stride;
  float2* db = dst + (uint64_t)blockIdx.y * stride;
  const XSweepRead<L, NTH> rd(a);
  float2 r[PER];
  uint64_t tau = blockIdx.x;
  if (!PERSIST) {
    if (a.from_zero) {
#pragma unroll
      for (int i = 0; i < PER; ++i)
        tile[threadIdx.x + NTH * i] =
            make_float2(threadIdx.x == 0 && i == 0 ? (SPLIT ? kSplitScale : 1.f) : 0.f, 0.f);
    } else {
      const uint64_t b0 = rd.base(tau, a);
#pragma unroll
      for (int i = 0; i < PER; ++i) r[i] = sb[b0 | rd.off(i)];
#pragma unroll
      for (int i = 0; i < PER; ++i) tile[threadIdx.x + NTH * i] = SPLIT ? split_scale(r[i]) : r[i];
    }
    __syncthreads();  // the tile and the blob are in LDS
    tile_apply_blocks<L, K4, NTH, SPLIT, NTH == 256>(tile, blocks, meta, a.nblocks);
    xsweep_store<L, NTH, SPLIT>(db, tile, scatter_bits(tau, a.wtile, a.ntb), a);
    return;
  }
  // Persistent: the next tile's loads are in flight while this tile's blocks
  // run in LDS (as in k_tile).
  if (tau < ntiles) {
    const uint64_t b0 = rd.base(tau, a);
#pragma unroll
    for (int i = 0; i < PER; ++i) r[i] = sb[b0 | rd.off(i)];
#pragma unroll
    for (int i = 0; i < PER; ++i) tile[threadIdx.x + NTH * i] = SPLIT ? split_scale(r[i]) : r[i];
  }
  __syncthreads();
  for (; tau < ntiles; tau += gridDim.x) {
    const uint64_t nxt = tau + gridDim.x;
    const bool more = nxt < ntiles;
    if (more) {
      const uint64_t nb = rd.base(nxt, a);
#pragma unroll
      for (int i = 0; i < PER; ++i) r[i] = sb[nb | rd.off(i)];
    }
    tile_apply_blocks<L, K4, NTH, SPLIT, NTH == 256>(tile, blocks, meta, a.nblocks);
    xsweep_store<L, NTH, SPLIT>(db, tile, scatter_bits(tau, a.wtile, a.ntb), a);
    if (more) {
      __syncthreads();  // every wave has read this tile out of LDS
#pragma unroll
      for (int i = 0; i < PER; ++i) tile[threadIdx.x + NTH * i] = SPLIT ? split_scale(r[i]) : r[i];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Exact complex128 sweep: consecutive compiled ops applied to 2^L-amplitude
// double2 tiles in LDS, each with the arithmetic and index set of its
// per-op kernel (k_one, k_pairswap, k_diag, k_dense), in program order, so
// every amplitude sees the same sequence of operations as one launch per op.
// ---------------------------------------------------------------------------
// Group index p -> local index: zeros inserted at the op's sorted positions
// (held in registers; loops fully unrolled so nothing is indexed at run time).
// Op descriptors sit in LDS and are the same for every lane: reading a field
// through readfirstlane keeps it in a scalar register, so decoding it runs on
// the scalar unit instead of costing vector instructions per lane and group.
__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

struct TileGroup {
  uint32_t low[8];  // (1 << pos[i]) - 1 for i < nb
  uint32_t nb, fixed;
  __device__ __forceinline__ explicit TileGroup(const TileOp& op) {
    nb = uni(op.nbits);
#pragma unroll
    for (int i = 0; i < 8; ++i) low[i] = (uint32_t)i < nb ? (1u << uni(op.pos[i])) - 1 : 0xFFFFFFFFu;
    fixed = uni(op.fixed);
  }
  __device__ __forceinline__ uint32_t operator()(uint32_t p) const {
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if ((uint32_t)i < nb) p = (p & low[i]) | ((p & ~low[i]) << 1);  // wave-uniform branch
    return p | fixed;
  }
};

template <int L>
__device__ __forceinline__ void tile128_op(double2* tile, const TileOp& op, const unsigned char* meta) {
  const uint32_t count = (1u << L) >> uni(op.nbits);
  const double* d = (const double*)(meta + uni(op.data_off));
  const TileGroup tile_group(op);
  const uint32_t kind = uni(op.kind), abit = uni(op.abit), bbit = uni(op.bbit);
  if (kind == 0) {
    double u[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) u[i] = d[i];
    for (uint32_t p = threadIdx.x; p < count; p += TPB) {
      const uint32_t i = tile_group(p);
      const double2 x = tile[i], y = tile[i | bbit];
      double2 na, nc;
      na.x = (u[0] * x.x - u[1] * x.y) + (u[2] * y.x - u[3] * y.y);
      na.y = (u[0] * x.y + u[1] * x.x) + (u[2] * y.y + u[3] * y.x);
      nc.x = (u[4] * x.x - u[5] * x.y) + (u[6] * y.x - u[7] * y.y);
      nc.y = (u[4] * x.y + u[5] * x.x) + (u[6] * y.y + u[7] * y.x);
      tile[i] = na;
      tile[i | bbit] = nc;
    }
  } else if (kind == 1) {
    for (uint32_t p = threadIdx.x; p < count; p += TPB) {
      const uint32_t i = tile_group(p);
      const double2 a = tile[i | abit], c = tile[i | bbit];
      tile[i | abit] = c;
      tile[i | bbit] = a;
    }
  } else if (kind == 2) {
    const uint32_t* pat = (const uint32_t*)d;
    const double* val = d + 16;  // patterns occupy the first 128 bytes
    const uint32_t nk = uni(op.k);
    for (uint32_t p = threadIdx.x; p < count; p += TPB) {
      const uint32_t i0 = tile_group(p);
      for (uint32_t t = 0; t < nk; ++t) {
        const uint32_t i = i0 | uni(pat[t]);
        const double2 a = tile[i];
        const double dr = val[2 * t], di = val[2 * t + 1];
        double2 r;
        r.x = a.x * dr - a.y * di;
        r.y = a.x * di + a.y * dr;
        tile[i] = r;
      }
    }
  } else {
    // Dense, k <= 3 in sweeps.
    const uint32_t kk = uni(op.k), D = 1u << kk;
    uint32_t qb[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) qb[q] = (uint32_t)q < kk ? 1u << uni(op.q[q]) : 0u;
    uint32_t off[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      uint32_t o = 0;
#pragma unroll
      for (int q = 0; q < 3; ++q)
        if ((uint32_t)q < kk && ((j >> (kk - 1 - q)) & 1)) o |= qb[q];
      off[j] = o;
    }
    for (uint32_t p = threadIdx.x; p < count; p += TPB) {
      const uint32_t base = tile_group(p);
      double2 v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if ((uint32_t)j < D) v[j] = tile[base | off[j]];
#pragma unroll
      for (int row = 0; row < 8; ++row) {
        if ((uint32_t)row >= D) break;
        double ar = 0, ai = 0;
#pragma unroll
        for (int col = 0; col < 8; ++col) {
          if ((uint32_t)col >= D) break;
          const double mr = d[2 * (row * D + col)], mi = d[2 * (row * D + col) + 1];
          ar = ar + (mr * v[col].x - mi * v[col].y);
          ai = ai + (mr * v[col].y + mi * v[col].x);
        }
        double2 r;
        r.x = ar;
        r.y = ai;
        tile[base | off[row]] = r;
      }
    }
  }
}

// Register chains: consecutive 1-qubit, X and diagonal ops on <= 2 qubits in
// total, applied by one thread to the 2^NQ amplitudes of its group held in
// registers (register j holds local index base | combo(j), bit i of j <->
// chain qubit c[i]), op after op with each op's own expression, so every
// amplitude goes through the same arithmetic as with one LDS pass per op.
// The host resolves each op to the register pairs (or, for a diagonal, the
// registers and multipliers) it touches, so the thread runs straight-line
// updates selected by wave-uniform mask bits.
// Register pairs of a chain: bit p of an op's pair mask selects pair
// (kPairA[p], kPairB[p]) of the 2^NQ chain registers.
__device__ constexpr int kPairA[4] = {0, 2, 0, 1};
__device__ constexpr int kPairB[4] = {1, 3, 2, 3};

template <int L, int NQ>
__device__ __forceinline__ void tile128_chain(double2* tile, const TileChain& ch, const TileOp* ops,
                                              const unsigned char* meta) {
  constexpr int D = 1 << NQ;
  uint32_t c[3], low[3];
  const uint32_t cbits = uni(ch.cbits), first = uni(ch.first), count = uni(ch.count);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    c[i] = (cbits >> (8 * i)) & 0xFF;
    low[i] = (1u << c[i]) - 1;
  }
  uint32_t off[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < NQ; ++i)
      if ((j >> i) & 1) o |= 1u << c[i];
    off[j] = o;
  }
  const uint32_t groups = (1u << L) >> NQ;
  for (uint32_t g = threadIdx.x; g < groups; g += TPB) {
    uint32_t base = g;
#pragma unroll
    for (int i = 0; i < NQ; ++i) base = (base & low[i]) | ((base & ~low[i]) << 1);
    double2 v[D];
#pragma unroll
    for (int j = 0; j < D; ++j) v[j] = tile[base | off[j]];
    for (uint32_t o = 0; o < count; ++o) {
      const TileOp& op = ops[first + o];
      const double* d = (const double*)(meta + uni(op.data_off));
      const uint32_t kind = uni(op.kind);
      if (kind == 2) {
        // Per-register multipliers resolved on the host: bit j of q[0] marks
        // the registers a diagonal entry applies to, values at byte q[1].
        const uint32_t m = uni(op.q[0]);
        const double* dl = d + uni(op.q[1]) / sizeof(double);
        double dv[2 * D];
#pragma unroll
        for (int i = 0; i < 2 * D; ++i) dv[i] = dl[i];
#pragma unroll
        for (int j = 0; j < D; ++j)
          if ((m >> j) & 1) {
            const double2 a = v[j];
            const double dr = dv[2 * j], di = dv[2 * j + 1];
            double2 r;
            r.x = a.x * dr - a.y * di;
            r.y = a.x * di + a.y * dr;
            v[j] = r;
          }
        continue;
      }
      // Pair update (kind 0) or swap (kind 1) on the pairs the host selected
      // (target bit and control values resolved per register pair).
      const uint32_t pm = uni(op.q[0]);
      constexpr int NP = NQ == 1 ? 1 : 4;
      if (kind == 0) {
        double u[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) u[i] = d[i];
#pragma unroll
        for (int pp = 0; pp < NP; ++pp)
          if ((pm >> pp) & 1) {
            const double2 x = v[kPairA[pp]], y = v[kPairB[pp]];
            double2 na, nc;
            na.x = (u[0] * x.x - u[1] * x.y) + (u[2] * y.x - u[3] * y.y);
            na.y = (u[0] * x.y + u[1] * x.x) + (u[2] * y.y + u[3] * y.x);
            nc.x = (u[4] * x.x - u[5] * x.y) + (u[6] * y.x - u[7] * y.y);
            nc.y = (u[4] * x.y + u[5] * x.x) + (u[6] * y.y + u[7] * y.x);
            v[kPairA[pp]] = na;
            v[kPairB[pp]] = nc;
          }
      } else {
#pragma unroll
        for (int pp = 0; pp < NP; ++pp)
          if ((pm >> pp) & 1) {
            const double2 t = v[kPairA[pp]];
            v[kPairA[pp]] = v[kPairB[pp]];
            v[kPairB[pp]] = t;
          }
      }
    }
#pragma unroll
    for (int j = 0; j < D; ++j) tile[base | off[j]] = v[j];
  }
}

// One op of a compiled two-qubit chain signature: every register index is
// fixed at compile time (a swap is a renaming), the arithmetic is the generic
// chain's expression for the same op.
template <int OP>
__device__ __forceinline__ void chain_fixed_op(double2* v, const TileOp& op, const unsigned char* meta) {
  if constexpr (OP == CH_U1 || OP == CH_U2) {
    constexpr int P0A = OP == CH_U1 ? 0 : 0, P0B = OP == CH_U1 ? 1 : 2;
    constexpr int P1A = OP == CH_U1 ? 2 : 1, P1B = 3;
    const double* d = (const double*)(meta + uni(op.data_off));
    double u[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) u[i] = d[i];
    auto upd = [&](int a, int b) {
      const double2 x = v[a], y = v[b];
      double2 na, nc;
      na.x = (u[0] * x.x - u[1] * x.y) + (u[2] * y.x - u[3] * y.y);
      na.y = (u[0] * x.y + u[1] * x.x) + (u[2] * y.y + u[3] * y.x);
      nc.x = (u[4] * x.x - u[5] * x.y) + (u[6] * y.x - u[7] * y.y);
      nc.y = (u[4] * x.y + u[5] * x.x) + (u[6] * y.y + u[7] * y.x);
      v[a] = na;
      v[b] = nc;
    };
    upd(P0A, P0B);
    upd(P1A, P1B);
  } else if constexpr (OP == CH_S13 || OP == CH_S23) {
    constexpr int A = OP == CH_S13 ? 1 : 2;
    const double2 t = v[A];
    v[A] = v[3];
    v[3] = t;
  } else if constexpr (OP == CH_D3) {
    const double* dl = (const double*)(meta + uni(op.data_off) + uni(op.q[1]));
    const double dr = dl[6], di = dl[7];
    const double2 a = v[3];
    double2 r;
    r.x = a.x * dr - a.y * di;
    r.y = a.x * di + a.y * dr;
    v[3] = r;
  }
}

template <int L, int O0, int O1, int O2>
__device__ __forceinline__ void tile128_chain_fixed(double2* tile, const TileChain& ch, const TileOp* ops,
                                                    const unsigned char* meta) {
  const uint32_t cbits = uni(ch.cbits), first = uni(ch.first);
  const uint32_t c0 = cbits & 0xFF, c1 = (cbits >> 8) & 0xFF;
  const uint32_t low0 = (1u << c0) - 1, low1 = (1u << c1) - 1;
  const uint32_t off[4] = {0u, 1u << c0, 1u << c1, (1u << c0) | (1u << c1)};
  constexpr uint32_t groups = (1u << L) >> 2;
  for (uint32_t g = threadIdx.x; g < groups; g += TPB) {
    uint32_t base = g;
    base = (base & low0) | ((base & ~low0) << 1);
    base = (base & low1) | ((base & ~low1) << 1);
    double2 v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = tile[base | off[j]];
    chain_fixed_op<O0>(v, ops[first], meta);
    if constexpr (O1 != CH_NONE) chain_fixed_op<O1>(v, ops[first + 1], meta);
    if constexpr (O2 != CH_NONE) chain_fixed_op<O2>(v, ops[first + 2], meta);
#pragma unroll
    for (int j = 0; j < 4; ++j) tile[base | off[j]] = v[j];
  }
}

// One-qubit chain with one update: register pair (0,1) of each group.
template <int L>
__device__ __forceinline__ void tile128_chain1_fixed(double2* tile, const TileChain& ch, const TileOp* ops,
                                                     const unsigned char* meta) {
  const uint32_t c0 = uni(ch.cbits) & 0xFF, first = uni(ch.first);
  const uint32_t low0 = (1u << c0) - 1, off1 = 1u << c0;
  const TileOp& op = ops[first];
  const double* d = (const double*)(meta + uni(op.data_off));
  constexpr uint32_t groups = (1u << L) >> 1;
  for (uint32_t g = threadIdx.x; g < groups; g += TPB) {
    const uint32_t base = (g & low0) | ((g & ~low0) << 1);
    double u[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) u[i] = d[i];
    const double2 x = tile[base], y = tile[base | off1];
    double2 na, nc;
    na.x = (u[0] * x.x - u[1] * x.y) + (u[2] * y.x - u[3] * y.y);
    na.y = (u[0] * x.y + u[1] * x.x) + (u[2] * y.y + u[3] * y.x);
    nc.x = (u[4] * x.x - u[5] * x.y) + (u[6] * y.x - u[7] * y.y);
    nc.y = (u[4] * x.y + u[5] * x.x) + (u[6] * y.y + u[7] * y.x);
    tile[base] = na;
    tile[base | off1] = nc;
  }
}

template <int L>
__device__ __forceinline__ void tile128_chain_sig(double2* tile, const TileChain& ch, uint32_t sig,
                                                  const TileOp* ops, const unsigned char* meta) {
#define SIG(N)                                                                                           \
  case N:                                                                                                \
    tile128_chain_fixed<L, kChainSigs[N][0], kChainSigs[N][1], kChainSigs[N][2]>(tile, ch, ops, meta); \
    break
  switch (sig) {
    SIG(1); SIG(2); SIG(3); SIG(4); SIG(5); SIG(6); SIG(7); SIG(8); SIG(9); SIG(10); SIG(11);
    default: break;
  }
#undef SIG
}
static_assert(kChainSigCount == 12, "one case per compiled chain signature");

template <int L>
__global__ __launch_bounds__(TPB, 3) void k_tile128(double2* __restrict__ s, uint64_t stride, uint64_t ntiles,
                                                 TileArgs a, const float4* __restrict__ blob) {
  extern __shared__ float4 lds4[];
  constexpr int PER = (1 << L) / TPB;
  double2* tile = (double2*)lds4;
  const unsigned char* meta = (const unsigned char*)(lds4 + (1 << L));
  for (uint32_t i = threadIdx.x; i < a.blob16; i += TPB) lds4[(1 << L) + i] = blob[i];
  const TileChain* chains = (const TileChain*)meta;
  const TileOp* ops = (const TileOp*)(meta + sizeof(TileChain) * a.nblocks);
  double2* st = s + (uint64_t)blockIdx.y * stride;
  uint64_t lo = 0;
#pragma unroll
  for (int bit = 0; bit < 8; ++bit)
    if ((threadIdx.x >> bit) & 1) lo |= 1ull << a.spos[bit];
  asm volatile("" : "+v"(lo));  // keeps the PER offsets lo | hi(i) out of VGPRs
  uint64_t hm[L - 8];
#pragma unroll
  for (int bit = 0; bit < L - 8; ++bit) hm[bit] = 1ull << a.spos[8 + bit];
  auto hi = [&](int i) {
    uint64_t o = 0;
#pragma unroll
    for (int bit = 0; bit < L - 8; ++bit)
      if ((i >> bit) & 1) o |= hm[bit];
    return o;
  };
  auto tile_base = [&](uint64_t tau) {
    uint64_t base = tau;
#pragma unroll
    for (int i = 0; i < L; ++i) {
      const uint64_t low = (1ull << a.srt[i]) - 1;
      base = (base & low) | ((base & ~low) << 1);
    }
    return base;
  };
  // Same pipeline as k_tile: next tile's loads in flight while this tile's
  // ops run; the loaded registers move to LDS after this tile's stores.
  // Real and imaginary parts in separate arrays: an array of double2 was
  // kept in scratch memory instead of registers.
  double rx[PER], ry[PER];
  uint64_t tau = blockIdx.x;
  if (tau < ntiles) {
    const uint64_t b0 = tile_base(tau);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const double2 v = st[b0 | lo | hi(i)];
      rx[i] = v.x;
      ry[i] = v.y;
    }
    tile_lane_swaps(rx, ry, a);
#pragma unroll
    for (int i = 0; i < PER; ++i) tile[threadIdx.x + TPB * i] = make_double2(rx[i], ry[i]);
  }
  __syncthreads();
  for (; tau < ntiles; tau += gridDim.x) {
    const uint64_t base = tile_base(tau);
    const uint64_t nxt = tau + gridDim.x;
    const bool more = nxt < ntiles;
    if (more) {
      const uint64_t nb = tile_base(nxt);
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const double2 v = st[nb | lo | hi(i)];
        rx[i] = v.x;
        ry[i] = v.y;
      }
    }
    for (uint32_t ci = 0; ci < a.nblocks; ++ci) {
      const TileChain& ch = chains[ci];
      const uint32_t nq = uni(ch.nq);
      if ((nq >> 8) && (nq & 0xFF) == 1) tile128_chain1_fixed<L>(tile, ch, ops, meta);
      else if (nq >> 8) tile128_chain_sig<L>(tile, ch, nq >> 8, ops, meta);
      else if (nq == 0) tile128_op<L>(tile, ops[uni(ch.first)], meta);
      else if (nq == 1) tile128_chain<L, 1>(tile, ch, ops, meta);
      else tile128_chain<L, 2>(tile, ch, ops, meta);
      __syncthreads();
    }
    tile_store<L>(st, tile, base, a);
    if (more) {
      tile_lane_swaps(rx, ry, a);
      __syncthreads();
#pragma unroll
      for (int i = 0; i < PER; ++i) tile[threadIdx.x + TPB * i] = make_double2(rx[i], ry[i]);
    }
    __syncthreads();
  }
}

// Exact complex128 sweep out of place: one workgroup per 2^L-amplitude tile,
// read as one contiguous run (or gathered from positions rpos at a run's
// first sweep), the tile's compiled ops applied in program order with
// k_tile128's arithmetic (the same chains and per-op code), and the tile
// stored through the next sweep's layout into the other state buffer. Zero
// start as in k_xsweep: from_zero synthesises |0...0> (nothing is read);
// tiles whose index meets zero_mask hold zeros and only store zeros.
template <int L>
__global__ __launch_bounds__(TPB, 3) void k_xsweep128(const double2* __restrict__ src, double2* __restrict__ dst,
                                                   uint64_t stride, XSweepArgs a, const float4* __restrict__ blob) {
  extern __shared__ float4 lds4[];
  constexpr int PER = (1 << L) / TPB;
  double2* tile = (double2*)lds4;
  double2* db = dst + (uint64_t)blockIdx.y * stride;
  const uint64_t wbase = scatter_bits(blockIdx.x, a.wtile, a.ntb);
  if ((blockIdx.x & a.zero_mask) != 0) {
    xsweep_store<L, TPB, false, true, double2>(db, tile, wbase, a);
    return;
  }
  const unsigned char* meta = (const unsigned char*)(lds4 + (1 << L));
  for (uint32_t i = threadIdx.x; i < a.blob16; i += TPB) lds4[(1 << L) + i] = blob[i];
  const TileChain* chains = (const TileChain*)meta;
  const TileOp* ops = (const TileOp*)(meta + sizeof(TileChain) * a.nblocks);
  if (a.from_zero) {
#pragma unroll
    for (int i = 0; i < PER; ++i)
      tile[threadIdx.x + TPB * i] = make_double2(threadIdx.x == 0 && i == 0 ? 1.0 : 0.0, 0.0);
  } else {
    const double2* sb = src + (uint64_t)blockIdx.y * stride;
    const XSweepRead<L, TPB> rd(a);
    const uint64_t b0 = rd.base(blockIdx.x, a);
    // Real and imaginary parts in separate register arrays (as in k_tile128).
    double rx[PER], ry[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const double2 v = sb[b0 | rd.off(i)];
      rx[i] = v.x;
      ry[i] = v.y;
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) tile[threadIdx.x + TPB * i] = make_double2(rx[i], ry[i]);
  }
  __syncthreads();  // the tile and the blob are in LDS
  for (uint32_t ci = 0; ci < a.nblocks; ++ci) {
    const TileChain& ch = chains[ci];
    const uint32_t nq = uni(ch.nq);
    if ((nq >> 8) && (nq & 0xFF) == 1) tile128_chain1_fixed<L>(tile, ch, ops, meta);
    else if (nq >> 8) tile128_chain_sig<L>(tile, ch, nq >> 8, ops, meta);
    else if (nq == 0) tile128_op<L>(tile, ops[uni(ch.first)], meta);
    else if (nq == 1) tile128_chain<L, 1>(tile, ch, ops, meta);
    else tile128_chain<L, 2>(tile, ch, ops, meta);
    __syncthreads();
  }
  xsweep_store<L, TPB, false, false, double2>(db, tile, wbase, a);
}

// ---------------------------------------------------------------------------
// State set-up
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(TPB) void k_init_zero(typename V2<T>::t* __restrict__ s, uint64_t stride,
                                                   uint64_t len) {
  using V = typename V2<T>::t;
  V* st = s + (uint64_t)blockIdx.y * stride;
  const uint64_t base = (uint64_t)blockIdx.x * (TPB * ITEMS) + threadIdx.x;
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    uint64_t i = base + (uint64_t)j * TPB;
    if (i < len) {
      V z;
      z.x = i == 0 ? T(1) : T(0);
      z.y = T(0);
      st[i] = z;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_broadcast(typename V2<T>::t* __restrict__ d, uint64_t stride,
                                                   const typename V2<T>::t* __restrict__ src, uint64_t len) {
  using V = typename V2<T>::t;
  V* st = d + (uint64_t)blockIdx.y * stride;
  const uint64_t base = (uint64_t)blockIdx.x * (TPB * ITEMS) + threadIdx.x;
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    uint64_t i = base + (uint64_t)j * TPB;
    if (i < len) st[i] = src[i];
  }
}

// ---------------------------------------------------------------------------
// Pairwise-tree segmented reductions (FP64, two channels)
// ---------------------------------------------------------------------------
struct Map64 {
  uint8_t q[64];
};

template <typename T>
struct ProbSrc {
  const typename V2<T>::t* s;
  uint64_t stride;
  int branch;
  int identity;
  int n;
  Map64 map;
  __device__ __forceinline__ uint64_t phys(uint64_t v) const {
    if (identity) return v;
    uint64_t p = 0;
    for (int j = 0; j < n; ++j) p |= ((v >> j) & 1ull) << map.q[j];
    return p;
  }
  __device__ __forceinline__ double2 operator()(uint32_t b, uint64_t v) const {
    uint64_t i = phys(v);
    typename V2<T>::t a = s[(uint64_t)b * stride + i];
    double re = (double)a.x, im = (double)a.y;
    double pr = re * re + im * im;
    double2 r;
    if (branch < 0) {
      r.x = pr;
      r.y = 0.0;
    } else if ((i >> branch) & 1ull) {
      r.x = 0.0;
      r.y = pr;
    } else {
      r.x = pr;
      r.y = 0.0;
    }
    return r;
  }
};

template <typename T>
struct DotSrc {
  const typename V2<T>::t* s;
  const double2* other;
  __device__ __forceinline__ double2 operator()(uint32_t, uint64_t v) const {
    typename V2<T>::t a = s[v];
    double2 o = other[v];
    double ar = (double)a.x, ai = (double)a.y;
    double2 r;
    // conj(a) * o = (ar*or - (-ai)*oi, ar*oi + (-ai)*or)
    r.x = ar * o.x - (-ai) * o.y;
    r.y = ar * o.y + (-ai) * o.x;
    return r;
  }
};

struct F64Src {
  const double2* p;
  uint64_t stride;
  __device__ __forceinline__ double2 operator()(uint32_t b, uint64_t v) const { return p[(uint64_t)b * stride + v]; }
};

__device__ __forceinline__ double2 add2(double2 a, double2 b) {
  double2 r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  return r;
}

__device__ __forceinline__ double2 shfl_xor2(double2 v, int m) {
  double2 r;
  r.x = __shfl_xor(v.x, m, 64);
  r.y = __shfl_xor(v.y, m, 64);
  return r;
}

// Block = 2048 consecutive virtual elements; element j*256 + t is loaded by
// thread t (coalesced). Lane butterflies give the pairwise sums of aligned
// runs of up to 64 elements; the 32 run sums per block are combined in LDS.
// The indexing below (8 loads of 256 threads per 2048-element block, 4 waves
// of 64 lanes, run j * 4 + wave) holds only for 256-thread workgroups.
static_assert(TPB == 256, "k_seg_reduce assumes 256 threads: 4 waves x 8 loads per 2048-element block");
template <class Src>
__global__ __launch_bounds__(TPB) void k_seg_reduce(Src src, uint64_t count, int seg, double2* __restrict__ out,
                                                    uint64_t out_stride) {
  __shared__ double2 runs[32];
  const uint32_t b = blockIdx.y;
  const uint32_t t = threadIdx.x;
  const uint32_t lane = t & 63, wave = t >> 6;
  const uint64_t blk0 = (uint64_t)blockIdx.x * 2048;
  double2 val[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    uint64_t v = blk0 + (uint64_t)j * 256 + t;
    if (v < count) {
      val[j] = src(b, v);
    } else {
      val[j].x = 0.0;
      val[j].y = 0.0;
    }
  }
  const int s1 = seg < 6 ? seg : 6;
  for (int l = 0; l < s1; ++l) {
#pragma unroll
    for (int j = 0; j < 8; ++j) val[j] = add2(val[j], shfl_xor2(val[j], 1 << l));
  }
  double2* o = out + (uint64_t)b * out_stride;
  if (seg <= 6) {
    const uint64_t mask = (1ull << seg) - 1;
    if ((lane & mask) == 0) {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        uint64_t v = blk0 + (uint64_t)j * 256 + t;
        if (v < count) o[v >> seg] = val[j];
      }
    }
    return;
  }
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < 8; ++j) runs[4 * j + wave] = val[j];
  }
  __syncthreads();
  const int width = 1 << (seg - 6);  // runs per output segment
  const int nout = 32 / width;
  if ((int)t < nout) {
    double2* r = runs + t * width;
    for (int w = 1; w < width; w *= 2)
      for (int k2 = 0; k2 < width; k2 += 2 * w) r[k2] = add2(r[k2], r[k2 + w]);
    uint64_t v = blk0 + ((uint64_t)t << seg);
    if (v < count) o[v >> seg] = r[0];
  }
}

// ---------------------------------------------------------------------------
// Measurement (engine.rs:367-435)
// ---------------------------------------------------------------------------
__global__ void k_meas_decide(const double2* __restrict__ tot, uint64_t tot_stride, uint32_t batch,
                              MeasState* __restrict__ ms, RngState* __restrict__ rng, uint8_t* __restrict__ cbits,
                              uint64_t nc, int64_t cbit, Flag f, int* err) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  MeasState m = ms[b];
  m.active = 0;
  m.xflag = 0;
  if (!flag_off(f, b)) {
    double2 p = tot[(uint64_t)b * tot_stride];
    double total = p.x + p.y;
    if (!(isfinite(total) && total > 0.0)) {
      atomicOr(err, 1);
    } else {
      RngState r = rng[b];
      double u = uniform53(r.seed, r.stream, r.draw);
      rng[b].draw = r.draw + 1;
      int outcome = u < p.y / total ? 1 : 0;
      double po = outcome ? p.y : p.x;
      m.p0 = p.x;
      m.p1 = p.y;
      m.scale = 1.0 / sqrt(po);
      m.outcome = outcome;
      m.active = 1;
      m.xflag = outcome;
      if (cbit >= 0) cbits[(uint64_t)b * nc + (uint64_t)cbit] = (uint8_t)outcome;
    }
  }
  ms[b] = m;
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_collapse(typename V2<T>::t* __restrict__ s, uint64_t stride, uint64_t len,
                                                  uint32_t q, const MeasState* __restrict__ ms) {
  using V = typename V2<T>::t;
  const uint32_t b = blockIdx.y;
  const MeasState m = ms[b];
  if (!m.active) return;
  V* st = s + (uint64_t)b * stride;
  const uint64_t base = (uint64_t)blockIdx.x * (TPB * ITEMS) + threadIdx.x;
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    uint64_t i = base + (uint64_t)j * TPB;
    if (i >= len) continue;
    V a = st[i];
    if ((int)((i >> q) & 1ull) == m.outcome) {
      a.x = (T)((double)a.x * m.scale);
      a.y = (T)((double)a.y * m.scale);
    } else {
      a.x = T(0);
      a.y = T(0);
    }
    st[i] = a;
  }
}

// Conditional evaluation (execution.rs:286-295): the tested bits are read as
// a little-endian integer and compared with the value; nested conditions AND.
__global__ void k_cond_eval(const uint8_t* __restrict__ cbits, uint64_t nc, uint32_t batch,
                            const Cond* __restrict__ conds, int ncond, int* __restrict__ flags) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  int ok = 1;
  for (int c = 0; c < ncond; ++c) {
    uint64_t actual = 0;
    for (uint64_t bit = 0; bit < conds[c].size; ++bit)
      actual |= (uint64_t)cbits[(uint64_t)b * nc + conds[c].base + bit] << bit;
    ok = ok && (actual == conds[c].value);
  }
  flags[b] = ok;
}

__global__ void k_rng_init(RngState* rng, uint32_t batch, uint64_t seed, uint64_t stream0) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  rng[b].seed = seed;
  rng[b].stream = stream0 + b;
  rng[b].draw = 0;
}

// ---------------------------------------------------------------------------
// CDF (block scan) and sampling
// ---------------------------------------------------------------------------
// Canonical inclusive prefix P of bins x[0..N): blocks of kScanBlock bins,
// each scanned by one workgroup (block_scan below); above one block, the
// block totals T[g] (each block's last output) are scanned the same way,
// recursively, and P[j] = O[g] + L[j] with O[g] = 0 for g = 0, else the scan
// of T at g - 1. block_scan of y[0..m), thread t owning y[8t .. 8t+8):
//   p[e]   sequential: p[0] = y[8t], p[e] = p[e-1] + y[8t+e] (0 past m);
//   R      the thread totals p[7] scanned in rows of 16 lanes (Hillis-Steele:
//          for d = 1, 2, 4, 8 lane i adds lane i-d of its row, DPP row_shr);
//   S[i]   = RO[row] + R[i], RO = 0, R[15], RO[1] + R[31], RO[2] + R[47];
//   TO[t]  = WO[w] + E[t], E = S of the previous lane (0 for lane 0), WO the
//          sequential prefix of the four wave totals S[63] (WO[0] = 0);
//   out    L[8t+e] = TO[t] + p[e].
// Every addition of zero is exact, so the bins' order of addition is fixed
// and the CPU restatement reproduces P bit for bit.
constexpr uint32_t kScanBlock = 2048, kScanPer = 8;

// LDS slot of bin i of a block: one pad slot per 8 bins, so a thread's 8
// consecutive bins sit 9 slots after its neighbour's (conflict-free).
__device__ __forceinline__ uint32_t scan_slot(uint32_t i) { return i + (i >> 3); }

// Value held by lane i-d of the same 16-lane row (0.0 for the row's first d
// lanes): DPP row_shr:d on both 32-bit halves.
template <int D>
__device__ __forceinline__ double row_shr_or_zero(double v) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, 0x110 + D, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), 0x110 + D, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}
__device__ __forceinline__ double lane_value(double v, int lane) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// One workgroup per block: x[(lo + i) * xstride] for i < m -> L[lo + i];
// bt[block] = L[lo + m - 1].
__global__ __launch_bounds__(256) void k_scan_block(const double* __restrict__ x, uint64_t xstride, uint64_t count,
                                                    double* __restrict__ L, double* __restrict__ bt) {
  __shared__ double v[kScanBlock + kScanBlock / 8];
  __shared__ double wt[4];
  const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
  const uint64_t lo = (uint64_t)blockIdx.x * kScanBlock;
  const uint32_t m = (uint32_t)(count - lo < kScanBlock ? count - lo : kScanBlock);
  for (uint32_t i = t; i < kScanBlock; i += 256) v[scan_slot(i)] = i < m ? x[(lo + i) * xstride] : 0.0;
  __syncthreads();
  double p[kScanPer];
  double acc = v[scan_slot(kScanPer * t)];
  p[0] = acc;
#pragma unroll
  for (uint32_t e = 1; e < kScanPer; ++e) {
    acc = acc + v[scan_slot(kScanPer * t + e)];
    p[e] = acc;
  }
  double r = acc;
  r = r + row_shr_or_zero<1>(r);
  r = r + row_shr_or_zero<2>(r);
  r = r + row_shr_or_zero<4>(r);
  r = r + row_shr_or_zero<8>(r);
  const double rt0 = lane_value(r, 15), rt1 = lane_value(r, 31), rt2 = lane_value(r, 47);
  const double ro1 = rt0, ro2 = ro1 + rt1, ro3 = ro2 + rt2;
  const uint32_t row = lane >> 4;
  const double sv = (row == 0 ? 0.0 : row == 1 ? ro1 : row == 2 ? ro2 : ro3) + r;
  double ex = __shfl_up(sv, 1, 64);
  if (lane == 0) ex = 0.0;
  if (lane == 63) wt[w] = sv;
  __syncthreads();
  const double wo1 = wt[0], wo2 = wo1 + wt[1], wo3 = wo2 + wt[2];
  const double to = (w == 0 ? 0.0 : w == 1 ? wo1 : w == 2 ? wo2 : wo3) + ex;
#pragma unroll
  for (uint32_t e = 0; e < kScanPer; ++e) v[scan_slot(kScanPer * t + e)] = to + p[e];
  __syncthreads();
  for (uint32_t i = t; i < m; i += 256) L[lo + i] = v[scan_slot(i)];
  if (t == 0) bt[blockIdx.x] = v[scan_slot(m - 1)];
}

// P[j] = O[g] + L[j] for the blocks g >= 1 (block 0's offset is 0).
__global__ __launch_bounds__(TPB) void k_scan_offsets(double* __restrict__ P, uint64_t count,
                                                      const double* __restrict__ q) {
  const uint64_t j = kScanBlock + (uint64_t)blockIdx.x * TPB + threadIdx.x;
  if (j >= count) return;
  P[j] = q[j / kScanBlock - 1] + P[j];
}

__global__ void k_scan_total(const double* __restrict__ P, uint64_t count, double* __restrict__ total,
                             int* __restrict__ err) {
  const double last = P[count - 1];
  *total = last;
  if (!(isfinite(last) && last > 0.0)) atomicOr(err, 2);
}

__device__ __forceinline__ double cdf_u(const double* C, uint64_t j, uint64_t count, double total) {
  return j + 1 == count ? 1.0 : C[j] / total;
}

// First j with U(j) > r (partition_point of p <= r), clamped to count-1.
__device__ __forceinline__ uint64_t partition(const double* C, uint64_t count, double total, double r) {
  uint64_t lo = 0, hi = count;
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo) / 2;
    if (cdf_u(C, mid, count, total) <= r) lo = mid + 1;
    else hi = mid;
  }
  return lo < count ? lo : count - 1;
}

__global__ __launch_bounds__(TPB) void k_sample_cdf(const double* __restrict__ C, uint64_t count,
                                                    const double* __restrict__ total, uint64_t seed,
                                                    uint64_t stream0, uint64_t shots, uint64_t* __restrict__ out) {
  const uint64_t shot = (uint64_t)blockIdx.x * TPB + threadIdx.x;
  if (shot >= shots) return;
  double r = uniform53(seed, stream0 + shot, 0);
  out[shot] = partition(C, count, *total, r);
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_sample_blocked(ProbSrc<T> src, int binseg, const double* __restrict__ C,
                                                        uint64_t nblocks, const double* __restrict__ total,
                                                        uint64_t seed, uint64_t stream0, uint64_t shots,
                                                        uint64_t* __restrict__ out) {
  const uint64_t shot = (uint64_t)blockIdx.x * TPB + threadIdx.x;
  if (shot >= shots) return;
  const double tot = *total;
  double r = uniform53(seed, stream0 + shot, 0);
  uint64_t kb = partition(C, nblocks, tot, r);
  double acc = kb == 0 ? 0.0 : C[kb - 1];
  uint64_t found = ~0ull, last_pos = kb * 2048;
  for (uint64_t j = 0; j < 2048; ++j) {
    uint64_t bin = kb * 2048 + j;
    // Pairwise sum of the bin's 2^binseg amplitudes (binary-counter stack).
    double stack[12];
    int level[12];
    int sp = 0;
    uint64_t nb = 1ull << binseg;
    for (uint64_t e = 0; e < nb; ++e) {
      double x = src(0, (bin << binseg) | e).x;
      int lv = 0;
      while (sp > 0 && level[sp - 1] == lv) {
        x = stack[sp - 1] + x;
        --sp;
        ++lv;
      }
      stack[sp] = x;
      level[sp] = lv;
      ++sp;
    }
    double p = stack[0];
    if (p > 0.0) last_pos = bin;
    acc = acc + p;
    if (acc / tot > r) {
      found = bin;
      break;
    }
  }
  out[shot] = found != ~0ull ? found : last_pos;
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_to_c128(const typename V2<T>::t* __restrict__ s, uint64_t count,
                                                 double2* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * TPB + threadIdx.x;
  if (i >= count) return;
  typename V2<T>::t a = s[i];
  double2 r;
  r.x = (double)a.x;
  r.y = (double)a.y;
  out[i] = r;
}

// Per-block max |a-b| and pairwise sum of |a-b|^2 between two states of
// possibly different precision (accuracy reports).
template <typename Ta, typename Tb>
__global__ __launch_bounds__(TPB) void k_diff(const typename V2<Ta>::t* __restrict__ a,
                                              const typename V2<Tb>::t* __restrict__ b, uint64_t count,
                                              double2* __restrict__ out) {
  __shared__ double mx[TPB], sq[TPB];
  const uint64_t base = (uint64_t)blockIdx.x * (TPB * ITEMS);
  double m = 0.0, acc[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    uint64_t i = base + (uint64_t)threadIdx.x * ITEMS + j;
    acc[j] = 0.0;
    if (i < count) {
      double dr = (double)a[i].x - (double)b[i].x, di = (double)a[i].y - (double)b[i].y;
      double d2 = dr * dr + di * di;
      acc[j] = d2;
      m = fmax(m, sqrt(d2));
    }
  }
  double t = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  mx[threadIdx.x] = m;
  sq[threadIdx.x] = t;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w /= 2) {
    if ((int)threadIdx.x < w) {
      mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + w]);
      sq[threadIdx.x] = sq[threadIdx.x] + sq[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = make_double2(mx[0], sq[0]);
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_gather(const typename V2<T>::t* __restrict__ s,
                                                const uint64_t* __restrict__ idx, uint64_t count,
                                                double2* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * TPB + threadIdx.x;
  if (i >= count) return;
  const typename V2<T>::t a = s[idx[i]];
  out[i] = make_double2((double)a.x, (double)a.y);
}

template <typename T>
M2<T> to_m2(const double* u) {
  M2<T> m;
  for (int i = 0; i < 8; ++i) m.v[i] = (T)u[i];
  return m;
}

template <typename T>
ProbSrc<T> make_prob(const void* s, uint64_t stride, int n, const ProbSpec& ps) {
  ProbSrc<T> src;
  src.s = (const typename V2<T>::t*)s;
  src.stride = stride;
  src.branch = ps.branch;
  src.n = n;
  src.identity = 1;
  for (int j = 0; j < 64; ++j) src.map.q[j] = (uint8_t)j;
  if (ps.map) {
    for (int j = 0; j < n; ++j) {
      src.map.q[j] = ps.map[j];
      if (ps.map[j] != j) src.identity = 0;
    }
  }
  return src;
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
#define DQ_DISPATCH(prec, F) \
  do {                       \
    if ((prec) == 16) {      \
      F(double);             \
    } else {                 \
      F(float);              \
    }                        \
  } while (0)

void one(int prec, void* s, uint64_t stride, uint32_t batch, int n, const Expand& e, uint32_t target,
         const double* u, Flag f, hipStream_t st) {
  const uint64_t count = (1ull << n) >> e.nbits;
  dim3 grid(blocks_for(count, TPB * ITEMS), batch);
#define L(T)                                                                                        \
  hipLaunchKernelGGL(k_one<T>, grid, dim3(TPB), 0, st, (typename V2<T>::t*)s, stride, count, e, \
                     1ull << target, to_m2<T>(u), f)
  DQ_DISPATCH(prec, L);
#undef L
}

void xgate(int prec, void* s, uint64_t stride, uint32_t batch, int n, const Expand& e, uint32_t target, Flag f,
           hipStream_t st) {
  const uint64_t count = (1ull << n) >> e.nbits;
  dim3 grid(blocks_for(count, TPB * ITEMS), batch);
#define L(T)                                                                                                \
  hipLaunchKernelGGL(k_pairswap<T>, grid, dim3(TPB), 0, st, (typename V2<T>::t*)s, stride, count, e, 0ull, \
                     1ull << target, f)
  DQ_DISPATCH(prec, L);
#undef L
}

void swap(int prec, void* s, uint64_t stride, uint32_t batch, int n, const Expand& e, uint32_t qa, uint32_t qb,
          Flag f, hipStream_t st) {
  const uint64_t count = (1ull << n) >> e.nbits;
  dim3 grid(blocks_for(count, TPB * ITEMS), batch);
#define L(T)                                                                                                  \
  hipLaunchKernelGGL(k_pairswap<T>, grid, dim3(TPB), 0, st, (typename V2<T>::t*)s, stride, count, e, 1ull << qa, \
                     1ull << qb, f)
  DQ_DISPATCH(prec, L);
#undef L
}

void diag(int prec, void* s, uint64_t stride, uint32_t batch, int n, const Expand& e, int nentries,
          const uint64_t* patterns, const double* values, Flag f, hipStream_t st) {
  if (nentries == 0) return;
  const uint64_t count = (1ull << n) >> e.nbits;
  dim3 grid(blocks_for(count, TPB * ITEMS), batch);
#define L(T)                                                                                                  \
  {                                                                                                           \
    DiagArgs<T> d;                                                                                            \
    for (int i = 0; i < nentries; ++i) {                                                                      \
      d.pat[i] = patterns[i];                                                                                 \
      d.val[2 * i] = (T)values[2 * i];                                                                        \
      d.val[2 * i + 1] = (T)values[2 * i + 1];                                                                \
    }                                                                                                         \
    hipLaunchKernelGGL(k_diag<T>, grid, dim3(TPB), 0, st, (typename V2<T>::t*)s, stride, count, e, nentries, d, f); \
  }
  DQ_DISPATCH(prec, L);
#undef L
}

void dense(int prec, void* s, uint64_t stride, uint32_t batch, int n, int kq, const Expand& e, const void* mat_dev,
           const uint64_t* offsets, Flag f, hipStream_t st) {
  const uint64_t count = (1ull << n) >> kq;
  dim3 grid(blocks_for(count, TPB), batch);
  Offsets off;
  for (int j = 0; j < 32; ++j) off.o[j] = j < (1 << kq) ? offsets[j] : 0;
#define L(T, KK)                                                                                               \
  hipLaunchKernelGGL((k_dense<T, KK>), grid, dim3(TPB), 0, st, (typename V2<T>::t*)s, stride, count, e,      \
                     (const T*)mat_dev, off, f)
#define LK(T)                  \
  switch (kq) {                \
    case 0: L(T, 0); break;    \
    case 1: L(T, 1); break;    \
    case 2: L(T, 2); break;    \
    case 3: L(T, 3); break;    \
    case 4: L(T, 4); break;    \
    default: L(T, 5); break;   \
  }
  DQ_DISPATCH(prec, LK);
#undef LK
#undef L
}

void fused_block(void* s, uint64_t stride, uint32_t batch, int n, int kq, const Expand& e, const float* mat_dev,
                 const uint64_t* offsets, hipStream_t st) {
  const uint64_t count = (1ull << n) >> kq;
  dim3 grid(blocks_for(count, TPB), batch);
  Offsets off;
  for (int j = 0; j < 32; ++j) off.o[j] = j < (1 << kq) ? offsets[j] : 0;
#define L(KK) hipLaunchKernelGGL(k_fused<KK>, grid, dim3(TPB), 0, st, (float2*)s, stride, count, e, mat_dev, off)
  switch (kq) {
    case 0: L(0); break;
    case 1: L(1); break;
    case 2: L(2); break;
    case 3: L(3); break;
    case 4: L(4); break;
    default: L(5); break;
  }
#undef L
}

bool fused_block_mfma(void* s, uint64_t stride, uint32_t batch, int n, int kq, const Expand& e, const float* emb,
                      const uint64_t* offsets, hipStream_t st) {
  if (kq != 4 && kq != 5) return false;
  const uint64_t ngroups = (1ull << n) >> kq;
  const uint64_t ntiles = (ngroups + 31) / 32;
  const uint64_t waves_wanted = ntiles;
  uint64_t blocks = (waves_wanted + 3) / 4;
  const uint64_t cap = 8192;
  if (blocks > cap) blocks = cap;
  dim3 grid((uint32_t)(blocks == 0 ? 1 : blocks), batch);
  Offsets off;
  for (int j = 0; j < 32; ++j) off.o[j] = j < (1 << kq) ? offsets[j] : 0;
  if (kq == 4)
    hipLaunchKernelGGL((k_fused_mfma<4, 2>), grid, dim3(TPB), 0, st, (float2*)s, stride, ngroups, e, emb, off);
  else
    hipLaunchKernelGGL((k_fused_mfma<5, 1>), grid, dim3(TPB), 0, st, (float2*)s, stride, ngroups, e, emb, off);
  return true;
}

#define HIPCHK_K(x)                                                                             \
  do {                                                                                          \
    const hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------
// FP32 matrix-core peak microbenchmark: every wave runs `iters` rounds of 8
// independent v_mfma_f32_16x16x4f32 chains on register operands (no memory
// traffic in the loop); each lane stores one sum so the chains stay live.
// ---------------------------------------------------------------------------
namespace {
__global__ __launch_bounds__(TPB) void k_mfma_peak(float* __restrict__ out, int iters) {
  floatx4 acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
  const float a = 1e-3f * (float)(threadIdx.x & 63), b = 1.0f + 1e-6f * (float)blockIdx.x;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[j], 0, 0, 0);
  }
  float sum = 0.f;
#pragma unroll
  for (int j = 0; j < 8; ++j) sum += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  out[(uint64_t)blockIdx.x * TPB + threadIdx.x] = sum;
}
}  // namespace

uint64_t mfma_peak_out_floats(uint32_t grid) { return (uint64_t)grid * TPB; }

uint64_t mfma_peak_flops_per_launch(uint32_t grid, int iters) {
  // 2 * 16 * 16 * 4 flops per instruction, 8 per round, one wave per 64 threads.
  return (uint64_t)grid * (TPB / 64) * (uint64_t)iters * 8ull * 2048ull;
}

void mfma_peak(float* out, uint32_t grid, int iters, hipStream_t st) {
  hipLaunchKernelGGL(k_mfma_peak, dim3(grid), dim3(TPB), 0, st, out, iters);
}

size_t tile_lds_bytes(const TileArgs& a) { return ((size_t)8 << a.L) + (size_t)16 * a.blob16; }
static size_t tile128_lds_bytes(const TileArgs& a) { return ((size_t)16 << a.L) + (size_t)16 * a.blob16; }

// Grid: 64 workgroups per CU (each walks ntiles / grid tiles with its
// prefetch), capped by the tile count. A grid of exactly the resident
// workgroup count measured 1.4-1.8x slower outside a profiler: it leaves the
// kernel's end to the slowest of a few long-lived workgroups.
static uint64_t tile_grid(uint64_t ntiles) {
  static thread_local std::map<int, int> cus_of;  // device -> CU count
  int dev = 0;
  HIPCHK_K(hipGetDevice(&dev));
  auto it = cus_of.find(dev);
  if (it == cus_of.end()) {
    int cus = 0;
    HIPCHK_K(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    it = cus_of.emplace(dev, std::max(1, cus)).first;
  }
  return std::min<uint64_t>(ntiles, (uint64_t)it->second * 64);
}

void tile_sweep(void* s, uint64_t stride, uint32_t batch, int n, const TileArgs& a, const void* blob,
                hipStream_t st) {
  const uint64_t ntiles = 1ull << (n - (int)a.L);
  const size_t lds = tile_lds_bytes(a);
  dim3 g((uint32_t)tile_grid(ntiles), batch);
#define LS(LL, K4, S) \
  hipLaunchKernelGGL((k_tile<LL, K4, S>), g, dim3(TPB), lds, st, (float2*)s, stride, ntiles, a, (const float4*)blob)
#define L(LL, K4) \
  if (a.split) LS(LL, K4, true); else LS(LL, K4, false)
  const bool k4 = a.kmax >= 4;
  switch (a.L) {
    case 10: if (k4) L(10, true); else L(10, false); break;
    case 11: if (k4) L(11, true); else L(11, false); break;
    default: if (k4) L(12, true); else L(12, false); break;
  }
#undef L
#undef LS
}

// The blob goes into LDS while it adds at most a quarter of the tile's bytes
// (8 KiB for 12-qubit tiles: four 256-thread workgroups still fit a CU's
// 160 KiB); a larger one would cost resident workgroups and is read from
// global memory. Measured on 32-qubit sweeps: LDS is faster for the few
// blocks of a random-circuit sweep (17.4 vs 18.1 ms), global for sweeps of
// 10+ blocks (41.8 vs 49.8 ms).
// 13-qubit tiles take the whole default 64 KiB of LDS: their blob stays in
// global memory.
bool xsweep_blob_in_lds(const XSweepArgs& a) {
  return a.L <= 12 && (size_t)16 * a.blob16 <= ((size_t)8 << a.L) / 4;
}
size_t xsweep_lds_bytes(const XSweepArgs& a) {
  return ((size_t)8 << a.L) + (xsweep_blob_in_lds(a) ? (size_t)16 * a.blob16 : 0);
}

void xsweep(const void* src, void* dst, uint64_t stride, uint32_t batch, int n, const XSweepArgs& a,
            const void* blob, bool persist, hipStream_t st) {
  if (a.L < 10 || a.L > 13 || (int)a.L > n || a.ntb != (uint32_t)(n - (int)a.L))
    throw std::runtime_error("xsweep: tile size does not match the state");
  if (a.L == 13 ? (a.threads != 512 || persist) : (a.threads != 256 && !(a.threads == 512 && a.L >= 11)))
    throw std::runtime_error("xsweep: 256 threads, or 512 for tiles of 11-12 qubits; 13-qubit tiles: 512 "
                             "threads, one workgroup per tile");
  const uint64_t ntiles = 1ull << (n - (int)a.L);
  const size_t lds = xsweep_lds_bytes(a);
  dim3 g((uint32_t)(persist ? tile_grid(ntiles) : ntiles), batch);
  if ((a.split || a.from_zero) && persist)
    throw std::runtime_error("xsweep: split-precision sweeps and sweeps from |0...0> run one workgroup per tile");
  const bool gb = !xsweep_blob_in_lds(a);
#define LGS(LL, K4, P, NT, G, S)                                                                                        \
  hipLaunchKernelGGL((k_xsweep<LL, K4, P, NT, G, S>), g, dim3(NT), lds, st, (const float2*)src, (float2*)dst, stride, \
                     ntiles, a, (const float4*)blob)
#define LG(LL, K4, P, NT, G) \
  if (!P && a.split) LGS(LL, K4, false, NT, G, true); else LGS(LL, K4, P, NT, G, false)
#define L(LL, K4, P, NT) \
  if (gb) LG(LL, K4, P, NT, true); else LG(LL, K4, P, NT, false)
#define LP(LL, K4, NT) \
  if (persist) { L(LL, K4, true, NT); } else { L(LL, K4, false, NT); }
#define LT(LL, K4) \
  if (a.threads == 512) { LP(LL, K4, 512); } else { LP(LL, K4, 256); }
  const bool k4 = a.kmax >= 4;
  switch (a.L) {
    case 10: if (k4) { LP(10, true, 256); } else { LP(10, false, 256); } break;
    case 11: if (k4) { LT(11, true); } else { LT(11, false); } break;
    case 12: if (k4) { LT(12, true); } else { LT(12, false); } break;
    default: if (k4) { LG(13, true, false, 512, true); } else { LG(13, false, false, 512, true); } break;
  }
#undef LT
#undef LP
#undef L
#undef LG
#undef LGS
}

void tile_sweep_c128(void* s, uint64_t stride, uint32_t batch, int n, const TileArgs& a, const void* blob,
                     hipStream_t st) {
  const uint64_t ntiles = 1ull << (n - (int)a.L);
  const size_t lds = tile128_lds_bytes(a);
  dim3 g((uint32_t)tile_grid(ntiles), batch);
  if (a.L == 10)
    hipLaunchKernelGGL(k_tile128<10>, g, dim3(TPB), lds, st, (double2*)s, stride, ntiles, a, (const float4*)blob);
  else
    hipLaunchKernelGGL(k_tile128<11>, g, dim3(TPB), lds, st, (double2*)s, stride, ntiles, a, (const float4*)blob);
}

void xsweep128(const void* src, void* dst, uint64_t stride, uint32_t batch, int n, const XSweepArgs& a,
               const void* blob, hipStream_t st) {
  if ((a.L != 10 && a.L != 11) || (int)a.L > n || a.ntb != (uint32_t)(n - (int)a.L) || a.threads != TPB)
    throw std::runtime_error("xsweep128: 10- or 11-qubit tiles of 256 threads matching the state");
  const size_t lds = ((size_t)16 << a.L) + (size_t)16 * a.blob16;
  if (lds > 65536) throw std::runtime_error("xsweep128: tile and blob exceed 64 KiB of LDS");
  dim3 g((uint32_t)(1ull << (n - (int)a.L)), batch);
  if (a.L == 10)
    hipLaunchKernelGGL(k_xsweep128<10>, g, dim3(TPB), lds, st, (const double2*)src, (double2*)dst, stride, a,
                       (const float4*)blob);
  else
    hipLaunchKernelGGL(k_xsweep128<11>, g, dim3(TPB), lds, st, (const double2*)src, (double2*)dst, stride, a,
                       (const float4*)blob);
}

void init_zero(int prec, void* s, uint64_t stride, uint32_t batch, int n, hipStream_t st) {
  const uint64_t len = 1ull << n;
  dim3 grid(blocks_for(len, TPB * ITEMS), batch);
#define L(T) hipLaunchKernelGGL(k_init_zero<T>, grid, dim3(TPB), 0, st, (typename V2<T>::t*)s, stride, len)
  DQ_DISPATCH(prec, L);
#undef L
}

void broadcast(int prec, void* d, uint64_t stride, uint32_t batch, const void* src, int n, hipStream_t st) {
  const uint64_t len = 1ull << n;
  dim3 grid(blocks_for(len, TPB * ITEMS), batch);
#define L(T)                                                                                   \
  hipLaunchKernelGGL(k_broadcast<T>, grid, dim3(TPB), 0, st, (typename V2<T>::t*)d, stride, \
                     (const typename V2<T>::t*)src, len)
  DQ_DISPATCH(prec, L);
#undef L
}

void seg_reduce_prob(int prec, const void* s, uint64_t stride, uint32_t batch, int n, const ProbSpec& ps, int seg,
                     void* out, uint64_t out_stride, hipStream_t st) {
  const uint64_t count = 1ull << n;
  dim3 grid(blocks_for(count, 2048), batch);
#define L(T)                                                                                          \
  hipLaunchKernelGGL(k_seg_reduce<ProbSrc<T>>, grid, dim3(TPB), 0, st, make_prob<T>(s, stride, n, ps), \
                     count, seg, (double2*)out, out_stride)
  DQ_DISPATCH(prec, L);
#undef L
}

void seg_reduce_dot(int prec, const void* s, int n, const void* other, int seg, void* out, hipStream_t st) {
  const uint64_t count = 1ull << n;
  dim3 grid(blocks_for(count, 2048), 1);
#define L(T)                                                                                                    \
  {                                                                                                             \
    DotSrc<T> src;                                                                                              \
    src.s = (const typename V2<T>::t*)s;                                                                        \
    src.other = (const double2*)other;                                                                          \
    hipLaunchKernelGGL(k_seg_reduce<DotSrc<T>>, grid, dim3(TPB), 0, st, src, count, seg, (double2*)out, 0ull); \
  }
  DQ_DISPATCH(prec, L);
#undef L
}

void seg_reduce_f64(const void* in, uint64_t in_stride, uint64_t count, uint32_t batch, int seg, void* out,
                    uint64_t out_stride, hipStream_t st) {
  dim3 grid(blocks_for(count, 2048), batch);
  F64Src src{(const double2*)in, in_stride};
  hipLaunchKernelGGL(k_seg_reduce<F64Src>, grid, dim3(TPB), 0, st, src, count, seg, (double2*)out, out_stride);
}

void meas_decide(const void* totals, uint64_t tot_stride, uint32_t batch, MeasState* ms, RngState* rng,
                 uint8_t* cbits, uint64_t nc, int64_t cbit, Flag f, int* err, hipStream_t st) {
  hipLaunchKernelGGL(k_meas_decide, dim3(blocks_for(batch, 64)), dim3(64), 0, st, (const double2*)totals, tot_stride,
                     batch, ms, rng, cbits, nc, cbit, f, err);
}

void collapse(int prec, void* s, uint64_t stride, uint32_t batch, int n, uint32_t qubit, const MeasState* ms,
              hipStream_t st) {
  const uint64_t len = 1ull << n;
  dim3 grid(blocks_for(len, TPB * ITEMS), batch);
#define L(T) \
  hipLaunchKernelGGL(k_collapse<T>, grid, dim3(TPB), 0, st, (typename V2<T>::t*)s, stride, len, qubit, ms)
  DQ_DISPATCH(prec, L);
#undef L
}

void cond_eval(const uint8_t* cbits, uint64_t nc, uint32_t batch, const Cond* conds, int ncond, int* flags,
               hipStream_t st) {
  hipLaunchKernelGGL(k_cond_eval, dim3(blocks_for(batch, 64)), dim3(64), 0, st, cbits, nc, batch, conds, ncond,
                     flags);
}

void rng_init(RngState* rng, uint32_t batch, uint64_t seed, uint64_t stream0, hipStream_t st) {
  hipLaunchKernelGGL(k_rng_init, dim3(blocks_for(batch, 64)), dim3(64), 0, st, rng, batch, seed, stream0);
}

// Canonical scan of `count` bins (stride bstride doubles) into C; tmp holds
// the block totals of every level (cdf_tmp_entries(count) doubles).
static void scan_into(const double* x, uint64_t xstride, uint64_t count, double* P, double* tmp, hipStream_t st) {
  const uint64_t nb = (count + kScanBlock - 1) / kScanBlock;
  hipLaunchKernelGGL(k_scan_block, dim3((uint32_t)nb), dim3(256), 0, st, x, xstride, count, P, tmp);
  if (nb == 1) return;
  // Block totals scanned in place (their own totals after them), then the
  // offsets added to every block after the first.
  scan_into(tmp, 1, nb, tmp, tmp + nb, st);
  hipLaunchKernelGGL(k_scan_offsets, dim3(blocks_for(count - kScanBlock, TPB)), dim3(TPB), 0, st, P, count,
                     (const double*)tmp);
}

uint64_t cdf_tmp_entries(uint64_t count) {
  uint64_t total = 0;
  for (uint64_t c = count; c > 1;) {
    c = (c + kScanBlock - 1) / kScanBlock;
    total += c;
  }
  return total + 1;
}

void cdf_build(const double* bins, uint64_t bstride, uint64_t count, double* C, double* gt, double* total, int* err,
               hipStream_t st) {
  scan_into(bins, bstride, count, C, gt, st);
  hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(1), 0, st, (const double*)C, count, total, err);
}

void sample_cdf(const double* C, uint64_t count, const double* total, uint64_t seed, uint64_t stream0,
                uint64_t shots, uint64_t* out, hipStream_t st) {
  if (shots == 0) return;
  hipLaunchKernelGGL(k_sample_cdf, dim3(blocks_for(shots, TPB)), dim3(TPB), 0, st, C, count, total, seed, stream0,
                     shots, out);
}

void sample_blocked(int prec, const void* s, int n, const uint8_t* map, int binseg, const double* C,
                    uint64_t nblocks, const double* total, uint64_t seed, uint64_t stream0, uint64_t shots,
                    uint64_t* out, hipStream_t st) {
  if (shots == 0) return;
  ProbSpec ps;
  ps.map = map;
#define L(T)                                                                                                   \
  hipLaunchKernelGGL(k_sample_blocked<T>, dim3(blocks_for(shots, TPB)), dim3(TPB), 0, st,                    \
                     make_prob<T>(s, 0, n, ps), binseg, C, nblocks, total, seed, stream0, shots, out)
  DQ_DISPATCH(prec, L);
#undef L
}

uint64_t diff_blocks(uint64_t count) { return blocks_for(count, TPB * ITEMS); }

void diff(int prec_a, const void* a, int prec_b, const void* b, uint64_t count, void* out, hipStream_t st) {
  dim3 grid(blocks_for(count, TPB * ITEMS));
#define L2(TA, TB)                                                                                       \
  hipLaunchKernelGGL((k_diff<TA, TB>), grid, dim3(TPB), 0, st, (const typename V2<TA>::t*)a,         \
                     (const typename V2<TB>::t*)b, count, (double2*)out)
  if (prec_a == 16 && prec_b == 16) L2(double, double);
  else if (prec_a == 16) L2(double, float);
  else if (prec_b == 16) L2(float, double);
  else L2(float, float);
#undef L2
}

void gather(int prec, const void* s, const uint64_t* idx, uint64_t count, void* out, hipStream_t st) {
  if (count == 0) return;
#define L(T)                                                                                       \
  hipLaunchKernelGGL(k_gather<T>, dim3(blocks_for(count, TPB)), dim3(TPB), 0, st,              \
                     (const typename V2<T>::t*)s, idx, count, (double2*)out)
  DQ_DISPATCH(prec, L);
#undef L
}

void to_c128(int prec, const void* s, uint64_t count, void* out, hipStream_t st) {
#define L(T)                                                                                      \
  hipLaunchKernelGGL(k_to_c128<T>, dim3(blocks_for(count, TPB)), dim3(TPB), 0, st,             \
                     (const typename V2<T>::t*)s, count, (double2*)out)
  DQ_DISPATCH(prec, L);
#undef L
}

// ---------------------------------------------------------------------------
// P-block (product of dense blocks): complex128 block buffers. Complex
// products and sums are written out as in the reference (no contraction).
// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ double2 pb_mul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double pb_nsq(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ uint64_t pb_insert(uint64_t i, uint32_t q, uint64_t bit) {
  const uint64_t low = (1ull << q) - 1;
  return (i & low) | ((i & ~low) << 1) | (bit << q);
}

// Merge (model.rs:199-244): out[r * nl + j] = left[j] * right[r].
__global__ __launch_bounds__(TPB) void k_pb_kron(double2* __restrict__ out, const double2* __restrict__ left,
                                                 uint64_t nl, const double2* __restrict__ right, uint64_t total) {
  const uint64_t i = (uint64_t)blockIdx.x * TPB + threadIdx.x;
  if (i >= total) return;
  const uint64_t r = i / nl, j = i - r * nl;
  out[i] = pb_mul(left[j], right[r]);
}

__global__ __launch_bounds__(TPB) void k_pb_scale(double2* __restrict__ s, uint64_t count, double2 c) {
  const uint64_t i = (uint64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < count) s[i] = pb_mul(s[i], c);
}

__global__ __launch_bounds__(TPB) void k_pb_scale_real(double2* __restrict__ s, uint64_t count, double f) {
  const uint64_t i = (uint64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < count) s[i] = make_double2(s[i].x * f, s[i].y * f);
}

// Amplitudes with local bit `q` equal to `bit`, in order (model.rs:486-489).
__global__ __launch_bounds__(TPB) void k_pb_extract(double2* __restrict__ out, const double2* __restrict__ in,
                                                    uint64_t half, uint32_t q, uint64_t bit) {
  const uint64_t i = (uint64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < half) out[i] = in[pb_insert(i, q, bit)];
}

__global__ void k_pb_set2(double2* __restrict__ s, double2 a, double2 b) {
  if (threadIdx.x == 0) {
    s[0] = a;
    s[1] = b;
  }
}

// Batched forms (slot b at b * stride): merge, and the measured half and
// basis factor of each slot's outcome (read from its MeasState).
__global__ __launch_bounds__(TPB) void k_pb_kron_b(double2* __restrict__ out, const double2* __restrict__ left,
                                                   uint64_t nl, const double2* __restrict__ right, uint64_t total,
                                                   uint64_t nr) {
  const uint64_t i = (uint64_t)blockIdx.x * TPB + threadIdx.x;
  if (i >= total) return;
  const uint64_t b = blockIdx.y;
  const uint64_t r = i / nl, j = i - r * nl;
  out[b * total + i] = pb_mul(left[b * nl + j], right[b * nr + r]);
}

__global__ __launch_bounds__(TPB) void k_pb_extract_ms(double2* __restrict__ out, const double2* __restrict__ in,
                                                       uint64_t half, uint32_t q, const MeasState* __restrict__ ms) {
  const uint64_t i = (uint64_t)blockIdx.x * TPB + threadIdx.x;
  const uint64_t b = blockIdx.y;
  if (i < half) out[b * half + i] = in[2 * half * b + pb_insert(i, q, (uint64_t)ms[b].outcome)];
}

__global__ __launch_bounds__(TPB) void k_pb_basis_ms(double2* __restrict__ out, const MeasState* __restrict__ ms,
                                                     uint32_t batch) {
  const uint32_t b = blockIdx.x * TPB + threadIdx.x;
  if (b >= batch) return;
  const int v = ms[b].outcome;
  out[2 * (uint64_t)b] = make_double2(v == 0 ? 1.0 : 0.0, 0.0);
  out[2 * (uint64_t)b + 1] = make_double2(v == 1 ? 1.0 : 0.0, 0.0);
}

__global__ __launch_bounds__(64) void k_pb_small_tables(const PbSmallTable* __restrict__ t, uint32_t count,
                                                       double* __restrict__ totals) {
  const uint32_t i = blockIdx.x * 64 + threadIdx.x;
  if (i >= count) return;
  const PbSmallTable d = t[i];
  const double2* a = (const double2*)d.amps;
  double acc = 0.0;
  for (uint64_t j = 0; j < d.len; ++j) {
    acc = acc + pb_nsq(a[j]);
    d.cdf[j] = acc;
  }
  totals[d.slot] = acc;
}

__global__ __launch_bounds__(TPB) void k_pb_norms(double* __restrict__ out, const double2* __restrict__ s,
                                                  uint64_t count) {
  const uint64_t i = (uint64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < count) out[i] = pb_nsq(s[i]);
}

// Separation test of local qubit q (model.rs:497-556), one workgroup: the
// pivot pair of largest weight |x0|^2 + |x1|^2 (the last one among equals),
// the normalised single-qubit factor s, the remaining amplitudes
// conj(s0) x0 + conj(s1) x1 and the sums loss = sum |x - s value|^2,
// total = sum |x|^2, kept = sum |value|^2 (pairwise, not sequential).
// res: s0.re, s0.im, s1.re, s1.im, loss, total, kept, norm, pivot, bad.
__global__ __launch_bounds__(TPB) void k_pb_separate(const double2* __restrict__ s, uint64_t half, uint32_t q,
                                                     double2* __restrict__ rem, double* __restrict__ res) {
  __shared__ double sw[TPB];
  __shared__ uint64_t si[TPB];
  __shared__ double2 single[2];
  __shared__ double sums[3][TPB];
  double bw = -1.0;
  uint64_t bi = 0;
  for (uint64_t i = threadIdx.x; i < half; i += TPB) {
    const double w = pb_nsq(s[pb_insert(i, q, 0)]) + pb_nsq(s[pb_insert(i, q, 1)]);
    if (w >= bw) {
      bw = w;
      bi = i;
    }
  }
  sw[threadIdx.x] = bw;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int h = TPB / 2; h > 0; h >>= 1) {
    if (threadIdx.x < (unsigned)h) {
      const double w2 = sw[threadIdx.x + h];
      const uint64_t i2 = si[threadIdx.x + h];
      if (w2 > sw[threadIdx.x] || (w2 == sw[threadIdx.x] && i2 > si[threadIdx.x])) {
        sw[threadIdx.x] = w2;
        si[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const uint64_t p = si[0];
    double2 a = s[pb_insert(p, q, 0)], b = s[pb_insert(p, q, 1)];
    const double norm = sqrt(pb_nsq(a) + pb_nsq(b));
    a = make_double2(a.x / norm, a.y / norm);
    b = make_double2(b.x / norm, b.y / norm);
    single[0] = a;
    single[1] = b;
    res[0] = a.x;
    res[1] = a.y;
    res[2] = b.x;
    res[3] = b.y;
    res[7] = norm;
    res[8] = (double)p;
    res[9] = (norm == 0.0 || !isfinite(norm)) ? 1.0 : 0.0;
  }
  __syncthreads();
  const double2 c0 = make_double2(single[0].x, -single[0].y), c1 = make_double2(single[1].x, -single[1].y);
  double loss = 0.0, total = 0.0, kept = 0.0;
  for (uint64_t i = threadIdx.x; i < half; i += TPB) {
    const double2 x0 = s[pb_insert(i, q, 0)], x1 = s[pb_insert(i, q, 1)];
    const double2 p0 = pb_mul(c0, x0), p1 = pb_mul(c1, x1);
    const double2 v = make_double2(p0.x + p1.x, p0.y + p1.y);
    rem[i] = v;
    const double2 m0 = pb_mul(single[0], v), m1 = pb_mul(single[1], v);
    loss = loss + pb_nsq(make_double2(x0.x - m0.x, x0.y - m0.y));
    total = total + pb_nsq(x0);
    loss = loss + pb_nsq(make_double2(x1.x - m1.x, x1.y - m1.y));
    total = total + pb_nsq(x1);
    kept = kept + pb_nsq(v);
  }
  sums[0][threadIdx.x] = loss;
  sums[1][threadIdx.x] = total;
  sums[2][threadIdx.x] = kept;
  __syncthreads();
  for (int h = TPB / 2; h > 0; h >>= 1) {
    if (threadIdx.x < (unsigned)h)
      for (int c = 0; c < 3; ++c) sums[c][threadIdx.x] = sums[c][threadIdx.x] + sums[c][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    res[4] = sums[0][0];
    res[5] = sums[1][0];
    res[6] = sums[2][0];
  }
}

__device__ __forceinline__ uint64_t pb_index(uint64_t b, const PbTable& t, const uint32_t* __restrict__ shifts) {
  uint64_t idx = 0;
  for (uint32_t j = 0; j < t.k; ++j) idx |= ((b >> shifts[t.shift_off + j]) & 1ull) << j;
  return idx;
}

// Product over blocks of one value each (result.rs:118-140, 179-195):
// out[b] = prod_t data_t[index_t(b)], folded from 1 in table order.
__global__ __launch_bounds__(TPB) void k_pb_product_c128(double2* __restrict__ out, uint64_t count,
                                                         const PbTable* __restrict__ tabs, uint32_t ntab,
                                                         const uint32_t* __restrict__ shifts) {
  const uint64_t b = (uint64_t)blockIdx.x * TPB + threadIdx.x;
  if (b >= count) return;
  double2 acc = make_double2(1.0, 0.0);
  for (uint32_t t = 0; t < ntab; ++t) acc = pb_mul(acc, ((const double2*)tabs[t].data)[pb_index(b, tabs[t], shifts)]);
  out[b] = acc;
}
__global__ __launch_bounds__(TPB) void k_pb_product_f64(double* __restrict__ out, uint64_t count,
                                                        const PbTable* __restrict__ tabs, uint32_t ntab,
                                                        const uint32_t* __restrict__ shifts) {
  const uint64_t b = (uint64_t)blockIdx.x * TPB + threadIdx.x;
  if (b >= count) return;
  double acc = 1.0;
  for (uint32_t t = 0; t < ntab; ++t) acc = acc * ((const double*)tabs[t].data)[pb_index(b, tabs[t], shifts)];
  out[b] = acc;
}

// Product sampling (model.rs:606-617): one Philox draw per block per shot,
// draw = u * cdf[last], outcome = first j with cdf[j] > draw (clamped).
__global__ __launch_bounds__(TPB) void k_pb_sample(uint32_t* __restrict__ out, uint64_t shots,
                                                   const PbTable* __restrict__ tabs, uint32_t ntab, uint64_t seed,
                                                   uint64_t stream0) {
  const uint64_t sh = (uint64_t)blockIdx.x * TPB + threadIdx.x;
  if (sh >= shots) return;
  for (uint32_t t = 0; t < ntab; ++t) {
    const double* cdf = (const double*)tabs[t].data;
    const uint64_t len = tabs[t].len;
    const double draw = uniform53(seed, stream0 + sh, t) * cdf[len - 1];
    uint64_t lo = 0, hi = len;
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if (cdf[mid] <= draw) lo = mid + 1;
      else hi = mid;
    }
    out[sh * ntab + t] = (uint32_t)(lo < len ? lo : len - 1);
  }
}

// Pauli expectation of one block (result.rs:253-293): terms
// conj(s[b ^ flip]) * phase(b) * s[b], partial sums per workgroup.
__global__ __launch_bounds__(TPB) void k_pb_pauli(const double2* __restrict__ s, uint64_t size, uint64_t xmask,
                                                  uint64_t ymask, uint64_t zmask, uint32_t nq,
                                                  double2* __restrict__ partial) {
  __shared__ double2 red[TPB];
  double2 acc = make_double2(0.0, 0.0);
  for (uint64_t b = (uint64_t)blockIdx.x * TPB + threadIdx.x; b < size; b += (uint64_t)gridDim.x * TPB) {
    double2 phase = make_double2(1.0, 0.0);
    for (uint32_t i = 0; i < nq; ++i) {
      const uint64_t bit = (b >> i) & 1;
      if ((ymask >> i) & 1) phase = pb_mul(phase, make_double2(0.0, bit == 0 ? 1.0 : -1.0));
      else if (((zmask >> i) & 1) && bit == 1) phase = make_double2(-phase.x, -phase.y);
    }
    const double2 o = s[b ^ (xmask | ymask)];
    const double2 t = pb_mul(pb_mul(make_double2(o.x, -o.y), phase), s[b]);
    acc = make_double2(acc.x + t.x, acc.y + t.y);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = TPB / 2; h > 0; h >>= 1) {
    if (threadIdx.x < (unsigned)h)
      red[threadIdx.x] = make_double2(red[threadIdx.x].x + red[threadIdx.x + h].x,
                                      red[threadIdx.x].y + red[threadIdx.x + h].y);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(TPB) void k_pb_sum2(const double2* __restrict__ in, uint64_t count,
                                                 double2* __restrict__ out) {
  __shared__ double2 red[TPB];
  double2 acc = make_double2(0.0, 0.0);
  for (uint64_t i = threadIdx.x; i < count; i += TPB) acc = make_double2(acc.x + in[i].x, acc.y + in[i].y);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = TPB / 2; h > 0; h >>= 1) {
    if (threadIdx.x < (unsigned)h)
      red[threadIdx.x] = make_double2(red[threadIdx.x].x + red[threadIdx.x + h].x,
                                      red[threadIdx.x].y + red[threadIdx.x + h].y);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

}  // namespace

void pb_kron(void* out, const void* left, uint64_t nl, const void* right, uint64_t nr, hipStream_t st) {
  const uint64_t total = nl * nr;
  hipLaunchKernelGGL(k_pb_kron, dim3(blocks_for(total, TPB)), dim3(TPB), 0, st, (double2*)out,
                     (const double2*)left, nl, (const double2*)right, total);
}
void pb_scale(void* s, uint64_t count, double re, double im, hipStream_t st) {
  hipLaunchKernelGGL(k_pb_scale, dim3(blocks_for(count, TPB)), dim3(TPB), 0, st, (double2*)s, count,
                     make_double2(re, im));
}
void pb_scale_real(void* s, uint64_t count, double f, hipStream_t st) {
  hipLaunchKernelGGL(k_pb_scale_real, dim3(blocks_for(count, TPB)), dim3(TPB), 0, st, (double2*)s, count, f);
}
void pb_extract(void* out, const void* in, uint64_t half, uint32_t q, uint32_t bit, hipStream_t st) {
  hipLaunchKernelGGL(k_pb_extract, dim3(blocks_for(half, TPB)), dim3(TPB), 0, st, (double2*)out,
                     (const double2*)in, half, q, (uint64_t)bit);
}
void pb_kron_batch(void* out, const void* left, uint64_t nl, const void* right, uint64_t nr, uint32_t batch,
                   hipStream_t st) {
  const uint64_t total = nl * nr;
  hipLaunchKernelGGL(k_pb_kron_b, dim3(blocks_for(total, TPB), batch), dim3(TPB), 0, st, (double2*)out,
                     (const double2*)left, nl, (const double2*)right, total, nr);
}
void pb_extract_ms(void* out, const void* in, uint64_t half, uint32_t q, const MeasState* ms, uint32_t batch,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_pb_extract_ms, dim3(blocks_for(half, TPB), batch), dim3(TPB), 0, st, (double2*)out,
                     (const double2*)in, half, q, ms);
}
void pb_basis_ms(void* out, const MeasState* ms, uint32_t batch, hipStream_t st) {
  hipLaunchKernelGGL(k_pb_basis_ms, dim3(blocks_for(batch, TPB)), dim3(TPB), 0, st, (double2*)out, ms, batch);
}
void pb_set2(void* s, double a_re, double a_im, double b_re, double b_im, hipStream_t st) {
  hipLaunchKernelGGL(k_pb_set2, dim3(1), dim3(64), 0, st, (double2*)s, make_double2(a_re, a_im),
                     make_double2(b_re, b_im));
}
void pb_small_tables(const PbSmallTable* tabs_dev, uint32_t count, double* totals, hipStream_t st) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_pb_small_tables, dim3(blocks_for(count, 64)), dim3(64), 0, st, tabs_dev, count, totals);
}
void pb_norms(double* out, const void* s, uint64_t count, hipStream_t st) {
  hipLaunchKernelGGL(k_pb_norms, dim3(blocks_for(count, TPB)), dim3(TPB), 0, st, out, (const double2*)s, count);
}
void pb_separate(const void* s, uint64_t half, uint32_t q, void* rem, double* res, hipStream_t st) {
  hipLaunchKernelGGL(k_pb_separate, dim3(1), dim3(TPB), 0, st, (const double2*)s, half, q, (double2*)rem, res);
}
void pb_product_c128(void* out, uint64_t count, const PbTable* tabs, uint32_t ntab, const uint32_t* shifts,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_pb_product_c128, dim3(blocks_for(count, TPB)), dim3(TPB), 0, st, (double2*)out, count, tabs,
                     ntab, shifts);
}
void pb_product_f64(double* out, uint64_t count, const PbTable* tabs, uint32_t ntab, const uint32_t* shifts,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_pb_product_f64, dim3(blocks_for(count, TPB)), dim3(TPB), 0, st, out, count, tabs, ntab,
                     shifts);
}
void pb_sample(uint32_t* out, uint64_t shots, const PbTable* tabs, uint32_t ntab, uint64_t seed, uint64_t stream0,
               hipStream_t st) {
  if (shots == 0) return;
  hipLaunchKernelGGL(k_pb_sample, dim3(blocks_for(shots, TPB)), dim3(TPB), 0, st, out, shots, tabs, ntab, seed,
                     stream0);
}
uint32_t pb_pauli_blocks(uint64_t size) { return (uint32_t)std::min<uint64_t>(blocks_for(size, TPB), 1024); }
void pb_pauli(const void* s, uint64_t size, uint64_t xmask, uint64_t ymask, uint64_t zmask, uint32_t nq,
              void* partial, void* out, hipStream_t st) {
  const uint32_t g = pb_pauli_blocks(size);
  hipLaunchKernelGGL(k_pb_pauli, dim3(g), dim3(TPB), 0, st, (const double2*)s, size, xmask, ymask, zmask, nq,
                     (double2*)partial);
  hipLaunchKernelGGL(k_pb_sum2, dim3(1), dim3(TPB), 0, st, (const double2*)partial, (uint64_t)g, (double2*)out);
}

}  // namespace k
}  // namespace dq

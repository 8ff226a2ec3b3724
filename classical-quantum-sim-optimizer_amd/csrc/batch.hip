// Single-launch batched-circuit kernel (complex64): thousands of small,
// uneven circuits in one launch (captured in a hipGraph by the host).
//
// One 256-thread workgroup per circuit; the circuit's whole state lives in
// LDS (2^n amplitudes, n <= kBatchLdsMax), so a circuit costs one workgroup's
// time and no launch of its own. Each workgroup:
//   1. writes |0...0> into LDS;
//   2. applies every fused block of its circuit (k <= 4 qubits, FP32 complex
//      matrix) as a gather / matvec / scatter over the block's groups, a
//      workgroup barrier between blocks;
//   3. optionally copies the final state out (tests keep final states);
//   4. samples its shots on chip: the state is cut into `nch` equal chunks
//      (one per thread), chunk probability sums are sequential FP64 sums,
//      thread 0 forms their sequential prefix, and each shot draws
//      r = Philox(seed, stream = shot, draw 0), binary-searches the chunk
//      prefix (last entry taken as 1.0) and scans its chunk sequentially; if
//      rounding leaves the scan short it takes the chunk's last non-zero
//      outcome. The CPU oracle restates this sampler (or_batch_sample).
// Per-circuit start and finish come from the device wall clock.
//
// There is no work queue, no atomic, no device-side function call and no
// private stack: blockIdx.x names the circuit, every helper is inlined into
// the kernel, and the LDS request stays within the default 64 KiB.
#include <hip/hip_runtime.h>

#include "batch.hpp"
#include "philox.hpp"

namespace dq {
namespace bk {
namespace {

constexpr int kBS = 256;  // threads per workgroup

template <int K>
__device__ __forceinline__ void apply_block(float2* st, uint32_t n, const BatchOp& op, const float* __restrict__ m) {
  // Batched circuits here have at most kBatchLdsMax qubits: 32-bit indices.
  constexpr int D = 1 << K;
  uint32_t off[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < K; ++i)
      if ((j >> (K - 1 - i)) & 1) o |= 1u << op.q[i];
    off[j] = o;
  }
  const uint32_t groups = (1u << n) >> K;
  for (uint32_t g = threadIdx.x; g < groups; g += kBS) {
    uint32_t base = g;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const uint32_t low = (1u << op.s[i]) - 1;
      base = (base & low) | ((base & ~low) << 1);
    }
    float2 v[D];
#pragma unroll
    for (int j = 0; j < D; ++j) v[j] = st[base | off[j]];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      float ar = 0.f, ai = 0.f;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        const float mr = m[2 * (r * D + c)], mi = m[2 * (r * D + c) + 1];
        ar = __builtin_fmaf(mr, v[c].x, ar);
        ar = __builtin_fmaf(-mi, v[c].y, ar);
        ai = __builtin_fmaf(mr, v[c].y, ai);
        ai = __builtin_fmaf(mi, v[c].x, ai);
      }
      st[base | off[r]] = make_float2(ar, ai);
    }
  }
}

__device__ __forceinline__ double prob(const float2* st, uint32_t i) {
  const float2 a = st[i];
  const double re = (double)a.x, im = (double)a.y;
  return re * re + im * im;
}

__global__ __launch_bounds__(kBS) void k_batch_lds(const BatchCircuit* __restrict__ circ,
                                                  const uint32_t* __restrict__ list, const BatchOp* __restrict__ ops,
                                                  const float* __restrict__ mats, uint32_t* __restrict__ out,
                                                  uint64_t* __restrict__ times, float2* __restrict__ dump,
                                                  const uint64_t* __restrict__ dump_off) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t idx = list[blockIdx.x];
  const BatchCircuit c = circ[idx];
  const uint32_t len = 1u << c.n;
  float2* st = (float2*)smem;
  double* part = (double*)(smem + ((size_t)8 << c.n));
  const uint64_t t0 = wall_clock64();
  for (uint32_t i = threadIdx.x; i < len; i += kBS) st[i] = make_float2(i == 0 ? 1.f : 0.f, 0.f);
  __syncthreads();
  for (uint32_t o = 0; o < c.nops; ++o) {
    const BatchOp op = ops[c.op_off + o];
    const float* m = mats + op.mat_off;
    switch (op.k) {
      case 1: apply_block<1>(st, c.n, op, m); break;
      case 2: apply_block<2>(st, c.n, op, m); break;
      case 3: apply_block<3>(st, c.n, op, m); break;
      default: apply_block<4>(st, c.n, op, m); break;
    }
    __syncthreads();
  }
  if (dump)
    for (uint32_t i = threadIdx.x; i < len; i += kBS) dump[dump_off[idx] + i] = st[i];
  // Chunked FP64 CDF and per-shot search (restated by or_batch_sample).
  const uint32_t nch = len < (uint32_t)kBS ? len : (uint32_t)kBS;
  const uint32_t cl = len / nch;
  if (threadIdx.x < nch) {
    double acc = 0.0;
    for (uint32_t i = threadIdx.x * cl; i < (threadIdx.x + 1) * cl; ++i) acc = acc + prob(st, i);
    part[threadIdx.x] = acc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double run = 0.0;
    for (uint32_t t = 0; t < nch; ++t) {
      run = run + part[t];
      part[t] = run;
    }
  }
  __syncthreads();
  const double total = part[nch - 1];
  for (uint32_t s = threadIdx.x; s < c.shots; s += kBS) {
    const double r = uniform53(c.seed, s, 0);
    uint32_t lo = 0, hi = nch;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) / 2;
      const double u = mid + 1 == nch ? 1.0 : part[mid] / total;
      if (u <= r) lo = mid + 1;
      else hi = mid;
    }
    const uint32_t t = lo < nch ? lo : nch - 1;
    double acc = t ? part[t - 1] : 0.0;
    uint32_t found = 0xFFFFFFFFu, lastpos = t * cl;
    for (uint32_t i = t * cl; i < (t + 1) * cl; ++i) {
      const double p = prob(st, i);
      if (p > 0.0) lastpos = i;
      acc = acc + p;
      if (acc / total > r) {
        found = i;
        break;
      }
    }
    out[c.out_off + s] = found != 0xFFFFFFFFu ? found : lastpos;
  }
  if (threadIdx.x == 0) {
    times[2 * idx] = t0;
    times[2 * idx + 1] = wall_clock64();
  }
}

}  // namespace

size_t lds_bytes(int n) { return ((size_t)8 << n) + (size_t)8 * kBS; }

int func_info(int* max_dynamic_lds, int* static_lds, int* max_threads, int* private_bytes) {
  hipFuncAttributes fa{};
  hipError_t e = hipFuncGetAttributes(&fa, (const void*)k_batch_lds);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return (int)e;
  }
  *max_dynamic_lds = fa.maxDynamicSharedSizeBytes;
  *static_lds = (int)fa.sharedSizeBytes;
  *max_threads = fa.maxThreadsPerBlock;
  *private_bytes = (int)fa.localSizeBytes;
  return 0;
}

int occupancy(size_t lds, int* blocks_per_cu) {
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void*)k_batch_lds, kBS, lds);
  if (e != hipSuccess) (void)hipGetLastError();
  return (int)e;
}

void launch_lds(const LdsLaunch& a, hipStream_t st) {
  if (a.count == 0) return;
  if (a.lds_n < 0 || a.lds_n > kBatchLdsMax) throw std::runtime_error("batch: LDS state size out of range");
  hipLaunchKernelGGL(k_batch_lds, dim3(a.count), dim3(kBS), lds_bytes(a.lds_n), st, a.circ, a.list, a.ops, a.mats,
                     a.out, a.times, a.dump, a.dump_off);
}

}  // namespace bk
}  // namespace dq

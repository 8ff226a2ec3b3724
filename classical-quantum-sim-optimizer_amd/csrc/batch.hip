// Persistent batched-circuit kernel (complex64): thousands of small, uneven
// circuits in one launch. DISABLED: it stalled on its first GPU run and no
// cause has been shown, so nothing launches it; the kernel stays for its
// launch-limit probe (func_info, set_lds_limit) and the host-only program dump
// that a CPU emulation of its logic runs on. Batched circuits run on the
// stream-pool runner (engine.cpp, run_stream_pool).
//
// Each workgroup pulls circuits from an atomic work queue (host order:
// largest estimated cost first) and runs a whole circuit by itself:
//   1. |0...0> in LDS when 2^n amplitudes fit (n <= lmax), else in the
//      workgroup's own HBM scratch slot;
//   2. every fused block of the circuit (k <= 4 qubits, FP32 matrix), one
//      pass over the state per block with a workgroup barrier between blocks;
//   3. terminal sampling on chip: the state is cut into `nch` equal chunks
//      (one per thread), chunk probability sums are sequential FP64 sums,
//      thread 0 forms their sequential prefix, and each shot draws
//      r = Philox(seed, stream = shot, draw 0), binary-searches the chunk
//      prefix (last entry taken as 1.0) and scans its chunk sequentially; if
//      rounding leaves the scan short it takes the chunk's last non-zero
//      outcome. The CPU oracle restates this (or_batch_sample).
// Per-circuit start/finish come from the device wall clock.
#include <hip/hip_runtime.h>

#include "batch.hpp"
#include "philox.hpp"

namespace dq {
namespace bk {
namespace {

template <int K, typename P>
__device__ __forceinline__ void apply_block(P st, uint32_t n, const BatchOp& op, const float* __restrict__ m, int bs) {
  // Batched circuits have at most 30 qubits: 32-bit indices throughout.
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
  for (uint32_t g = threadIdx.x; g < groups; g += bs) {
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

template <typename P>
__device__ void run_circuit(P st, const BatchCircuit& c, const BatchOp* __restrict__ ops,
                            const float* __restrict__ mats, int bs) {
  const uint64_t len = 1ull << c.n;
  for (uint64_t i = threadIdx.x; i < len; i += bs) st[i] = make_float2(i == 0 ? 1.f : 0.f, 0.f);
  __syncthreads();
  for (uint32_t o = 0; o < c.nops; ++o) {
    const BatchOp op = ops[c.op_off + o];
    const float* m = mats + op.mat_off;
    switch (op.k) {
      case 1: apply_block<1>(st, c.n, op, m, bs); break;
      case 2: apply_block<2>(st, c.n, op, m, bs); break;
      case 3: apply_block<3>(st, c.n, op, m, bs); break;
      case 4: apply_block<4>(st, c.n, op, m, bs); break;
      default: apply_block<5>(st, c.n, op, m, bs); break;
    }
    __syncthreads();
  }
}

template <typename P>
__device__ __forceinline__ double prob(P st, uint64_t i) {
  const float2 a = st[i];
  const double re = (double)a.x, im = (double)a.y;
  return re * re + im * im;
}

template <typename P>
__device__ void sample(P st, const BatchCircuit& c, double* part, uint32_t* __restrict__ out, int bs) {
  const uint64_t len = 1ull << c.n;
  const uint64_t nch = len < (uint64_t)bs ? len : (uint64_t)bs;
  const uint64_t cl = len / nch;
  if (threadIdx.x < nch) {
    double acc = 0.0;
    for (uint64_t i = threadIdx.x * cl; i < (threadIdx.x + 1) * cl; ++i) acc = acc + prob(st, i);
    part[threadIdx.x] = acc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double run = 0.0;
    for (uint64_t t = 0; t < nch; ++t) {
      run = run + part[t];
      part[t] = run;
    }
  }
  __syncthreads();
  const double total = part[nch - 1];
  for (uint64_t s = threadIdx.x; s < c.shots; s += bs) {
    const double r = uniform53(c.seed, s, 0);
    uint64_t lo = 0, hi = nch;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) / 2;
      const double u = mid + 1 == nch ? 1.0 : part[mid] / total;
      if (u <= r) lo = mid + 1;
      else hi = mid;
    }
    const uint64_t t = lo < nch ? lo : nch - 1;
    double acc = t ? part[t - 1] : 0.0;
    uint64_t found = ~0ull, lastpos = t * cl;
    for (uint64_t i = t * cl; i < (t + 1) * cl; ++i) {
      const double p = prob(st, i);
      if (p > 0.0) lastpos = i;
      acc = acc + p;
      if (acc / total > r) {
        found = i;
        break;
      }
    }
    out[c.out_off + s] = (uint32_t)(found != ~0ull ? found : lastpos);
  }
  __syncthreads();
}

template <int BS>
__global__ __launch_bounds__(BS) void k_batch(const BatchCircuit* __restrict__ circ, uint32_t ncirc,
                                              const uint32_t* __restrict__ order, const BatchOp* __restrict__ ops,
                                              const float* __restrict__ mats, float2* __restrict__ scratch,
                                              uint64_t slot_amps, int lmax, int* __restrict__ queue,
                                              uint32_t* __restrict__ out, uint64_t* __restrict__ times,
                                              float2* __restrict__ dump, const uint64_t* __restrict__ dump_off) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int cur;
  float2* lds = (float2*)smem;
  double* part = (double*)(smem + ((size_t)8 << lmax));
  float2* slot = scratch + (uint64_t)blockIdx.x * slot_amps;
  for (;;) {
    if (threadIdx.x == 0) cur = atomicAdd(queue, 1);
    __syncthreads();
    const int ci = __builtin_amdgcn_readfirstlane(cur);
    __syncthreads();
    if (ci >= (int)ncirc) break;
    const uint32_t idx = order[ci];
    const BatchCircuit c = circ[idx];
    const uint64_t t0 = wall_clock64();
    const uint64_t len = 1ull << c.n;
    if ((int)c.n <= lmax) {
      run_circuit(lds, c, ops, mats, BS);
      if (dump)
        for (uint64_t i = threadIdx.x; i < len; i += BS) dump[dump_off[idx] + i] = lds[i];
      sample(lds, c, part, out, BS);
    } else {
      run_circuit(slot, c, ops, mats, BS);
      if (dump)
        for (uint64_t i = threadIdx.x; i < len; i += BS) dump[dump_off[idx] + i] = slot[i];
      sample(slot, c, part, out, BS);
    }
    if (threadIdx.x == 0) {
      times[2 * idx] = t0;
      times[2 * idx + 1] = wall_clock64();
    }
  }
}

}  // namespace

size_t lds_bytes(int lmax, int block) { return ((size_t)8 << lmax) + (size_t)8 * block; }

static const void* kernel_for(int block) {
  if (block == 256) return (const void*)k_batch<256>;
  if (block == 512) return (const void*)k_batch<512>;
  return (const void*)k_batch<1024>;
}

int set_lds_limit(int block, size_t bytes) {
  hipError_t e = hipFuncSetAttribute(kernel_for(block), hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) (void)hipGetLastError();
  return (int)e;
}

int func_info(int block, int* max_dynamic_lds, int* static_lds, int* max_threads, int* private_bytes) {
  hipFuncAttributes fa{};
  hipError_t e = hipFuncGetAttributes(&fa, kernel_for(block));
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

int occupancy(int block, size_t lds, int* blocks_per_cu) {
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, kernel_for(block), block, lds);
  if (e != hipSuccess) (void)hipGetLastError();
  return (int)e;
}

}  // namespace bk
}  // namespace dq

// Deterministic grid reduction: wave shuffle -> LDS -> one FP64 partial per
// workgroup -> the last workgroup to arrive (ticket counter) sums the
// partials in a fixed order and writes the result.  The result therefore
// does not depend on scheduling, and no host round trip is needed.
//
// Publication without a device-scope fence: on CDNA3/4 each XCD has its own
// L2, and an agent-scope release fence writes back every dirty line of the
// XCD's L2 - in a kernel that streams hundreds of MB of stores that stalls
// each workgroup for the whole write-back.  Instead each partial is stored
// with an agent-scope atomic store (coherent across XCDs, written through
// the L2), the wave waits for the store to complete (s_waitcnt), and only
// then takes its ticket with a relaxed agent-scope atomic; the last
// workgroup reads the partials with agent-scope atomic loads.
//
// The ordering of the partial stores before the ticket rests on the
// hardware: s_waitcnt vmcnt(0) returns only once the write-through stores
// are acknowledged.  The language does not order relaxed atomics to
// different addresses, so compiler-only fences on both sides of the wait
// keep the compiler from moving a store past it or the ticket before it.
#pragma once
#include <hip/hip_runtime.h>

namespace yh {

// every store issued so far has completed before anything after this
__device__ __forceinline__ void stores_complete() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// NR values per thread; partials holds nblk * NR doubles.
// Returns true in the (whole) last block after out[] has been written.
template <int NR, int NT>
__device__ bool block_reduce_finish(double (&v)[NR], double* partials, int nblk, int blk, unsigned* ticket,
                                    double* out) {
  __shared__ double sm[NR][NT / 64];
  __shared__ int is_last;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < NR; r++) {
    double s = wave_sum(v[r]);
    if (lane == 0) sm[r][w] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; r++) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < NT / 64; q++) s += sm[r][q];
      __hip_atomic_store(&partials[(size_t)r * nblk + blk], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    stores_complete();
    unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    is_last = (t == (unsigned)(nblk - 1));
  }
  __syncthreads();
  if (!is_last) return false;
#pragma unroll
  for (int r = 0; r < NR; r++) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nblk; i += NT) s += __hip_atomic_load(&partials[(size_t)r * nblk + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s = wave_sum(s);
    __syncthreads();
    if (lane == 0) sm[r][w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double tot = 0.0;
#pragma unroll
      for (int q = 0; q < NT / 64; q++) tot += sm[r][q];
      out[r] = tot;
    }
  }
  if (threadIdx.x == 0) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  return true;
}

}  // namespace yh

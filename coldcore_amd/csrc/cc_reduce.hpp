// cc_reduce.hpp -- whole-field reductions and queries.
//
// Semantics follow reference src/coldcore/protocol.py:181-219 and
// src/plugins/covering/dctcore_fast.cuh:199-327:
//   count_eq(t)   sum [cnt == t]                   (int64)
//   deficit       sum max(0, mu - cnt)             (int64)
//   map_max       max(0, max A0)
//   map_hist      positive v -> bin (v-1)*nbins/vmax, clipped to nbins-1
//   map_extract   (idx, val) of A0 >= thr; found counts every hit
// The extraction here is a deterministic compaction: blocks own contiguous
// chunks, a scan of the per-chunk counts gives each chunk its output offset,
// and hits are written in ascending index order.  That is a strict
// refinement of the reference's unordered contract and equals the host
// backends' order (protocol.py:201-208), which keeps search trajectories
// identical across engines.
#pragma once
#include "cc_common.hpp"
#include "cc_extract.hpp"

__global__ void k_fill32(int32_t *p, int32_t v, long long n) {
  const long long gs = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gs)
    p[i] = v;
}

__global__ void k_fill16(uint16_t *p, uint16_t v, long long n) {
  const long long gs = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gs)
    p[i] = v;
}

__global__ void k_fill_u64(unsigned long long *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// A0[idx[i]] += 1 : the code multiset indicator
__global__ void k_scatter_code(int32_t *A0, const long long *idx, int M) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < M) atomicAdd(&A0[idx[i]], 1);
}

// MODE CC_EQ: sum [cnt == target];  CC_DEF: sum max(0, mu - cnt)
template <int MODE>
__global__ void __launch_bounds__(CC_BLOCK)
    k_cnt_reduce(const uint16_t *__restrict__ cnt, long long N, int target,
                 int mu, unsigned long long *out) {
  const long long nvec = N >> 3;
  const long long gs = (long long)gridDim.x * CC_BLOCK;
  const uint4 *v = (const uint4 *)cnt;
  long long acc = 0;
  for (long long i = (long long)blockIdx.x * CC_BLOCK + threadIdx.x; i < nvec;
       i += gs) {
    const uint4 q = v[i];
    const uint32_t wv[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int h = 0; h < 8; h++) {
      const int c = (int)((wv[h >> 1] >> (16 * (h & 1))) & 0xFFFFu);
      if (MODE == CC_EQ) acc += (c == target);
      else { const int d = mu - c; acc += d > 0 ? d : 0; }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < (int)(N & 7)) {
    const int c = (int)cnt[nvec * 8 + threadIdx.x];
    if (MODE == CC_EQ) acc += (c == target);
    else { const int d = mu - c; acc += d > 0 ? d : 0; }
  }
  acc = cc_block_sum(acc);
  if (threadIdx.x == 0 && acc) atomicAdd(out, (unsigned long long)acc);
}

__global__ void __launch_bounds__(CC_BLOCK)
    k_max_i32(const int32_t *__restrict__ A, long long N, int *out) {
  const long long nvec = N >> 2;
  const long long gs = (long long)gridDim.x * CC_BLOCK;
  const int4 *v = (const int4 *)A;
  int m = 0;
  for (long long i = (long long)blockIdx.x * CC_BLOCK + threadIdx.x; i < nvec;
       i += gs) {
    const int4 q = v[i];
    m = max(m, max(max(q.x, q.y), max(q.z, q.w)));
  }
  if (blockIdx.x == 0 && threadIdx.x < (int)(N & 3))
    m = max(m, A[nvec * 4 + threadIdx.x]);
  m = cc_wave_max(m);
  if ((threadIdx.x & 63) == 0 && m > 0) atomicMax(out, m);
}

__device__ __forceinline__ void cc_hist_one(int32_t v, int nbins, int32_t vmax,
                                            int32_t *sh) {
  if (v > 0) {
    long long b = (long long)(v - 1) * nbins / vmax;
    if (b >= nbins) b = nbins - 1;
    atomicAdd(&sh[b], 1);
  }
}

__global__ void __launch_bounds__(CC_BLOCK)
    k_hist(const int32_t *__restrict__ A, long long N, int nbins, int32_t vmax,
           int32_t *hist) {
  extern __shared__ int32_t sh[];
  for (int i = threadIdx.x; i < nbins; i += CC_BLOCK) sh[i] = 0;
  __syncthreads();
  const long long nvec = N >> 2;
  const long long gs = (long long)gridDim.x * CC_BLOCK;
  const int4 *v = (const int4 *)A;
  for (long long i = (long long)blockIdx.x * CC_BLOCK + threadIdx.x; i < nvec;
       i += gs) {
    const int4 q = v[i];
    cc_hist_one(q.x, nbins, vmax, sh);
    cc_hist_one(q.y, nbins, vmax, sh);
    cc_hist_one(q.z, nbins, vmax, sh);
    cc_hist_one(q.w, nbins, vmax, sh);
  }
  if (blockIdx.x == 0 && threadIdx.x < (int)(N & 3))
    cc_hist_one(A[nvec * 4 + threadIdx.x], nbins, vmax, sh);
  __syncthreads();
  for (int i = threadIdx.x; i < nbins; i += CC_BLOCK)
    if (sh[i]) atomicAdd(&hist[i], sh[i]);
}

__global__ void k_read_at(const int32_t *A, const long long *idx, int k,
                          int32_t *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < k) out[i] = A[idx[i]];
}

// out[i * L + d] = layer d at word idx[i] (after a layers transform); LT is
// the plane width the transform used
template <typename LT>
__global__ void k_layers_read_at(const SweepArgs a, int L, const long long *idx,
                                 int k, int32_t *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  const long long w = idx[i];
  for (int d = 0; d < L; d++) {
    if constexpr (sizeof(LT) == 2) out[(long long)i * L + d] = (int32_t)a.H[d][w];
    else out[(long long)i * L + d] = a.A[d][w];
  }
}

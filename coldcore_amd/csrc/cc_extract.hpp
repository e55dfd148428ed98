// cc_extract.hpp -- ordered extraction of the map entries >= thr.
//
// A deterministic compaction: blocks own contiguous chunks, a scan of the
// per-chunk counts gives each chunk its output offset, and hits are written
// in ascending index order (the host backends' order, reference
// src/coldcore/protocol.py:201-208; the reference's GPU order is
// unspecified).  The kernels have internal linkage: the engine and the
// allocation-free entry point each compile their own copy.
#pragma once
#include "cc_common.hpp"

// ---- ordered extraction: chunk counts -> exclusive scan -> ordered write
// Chunk b covers [b*CH, min(N, (b+1)*CH)); CH is a multiple of 4*CC_BLOCK.
__device__ __forceinline__ int cc_hits4(const int32_t *A, long long N,
                                        long long i, int32_t thr,
                                        int32_t e[4]) {
  if (i + 3 < N) {
    const int4 q = *(const int4 *)(A + i);
    e[0] = q.x; e[1] = q.y; e[2] = q.z; e[3] = q.w;
  } else {
#pragma unroll
    for (int j = 0; j < 4; j++) e[j] = (i + j < N) ? A[i + j] : INT32_MIN;
  }
  int m = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) m |= (e[j] >= thr && i + j < N) ? (1 << j) : 0;
  return m;
}

static __global__ void __launch_bounds__(CC_BLOCK)
    k_extract_count(const int32_t *__restrict__ A, long long N, int32_t thr,
                    long long CH, long long *bcnt) {
  const long long c0 = (long long)blockIdx.x * CH;
  const long long c1 = min(N, c0 + CH);
  long long acc = 0;
  for (long long i = c0 + 4LL * threadIdx.x; i < c1; i += 4LL * CC_BLOCK) {
    int32_t e[4];
    acc += __popc(cc_hits4(A, c1, i, thr, e));
  }
  acc = cc_block_sum(acc);
  if (threadIdx.x == 0) bcnt[blockIdx.x] = acc;
}

// exclusive scan of nb chunk counts by one workgroup of 1024 threads;
// boff[nb] = total
static __global__ void __launch_bounds__(1024)
    k_scan_offsets(const long long *bcnt, long long nb, long long *boff) {
  __shared__ long long part[1024];
  const int t = threadIdx.x;
  const long long per = (nb + 1023) / 1024;
  const long long lo = t * per, hi = min(nb, lo + per);
  long long s = 0;
  for (long long i = lo; i < hi; i++) s += bcnt[i];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    long long run = 0;
    for (int i = 0; i < 1024; i++) {
      const long long v = part[i];
      part[i] = run;
      run += v;
    }
    boff[nb] = run;
  }
  __syncthreads();
  long long run = part[t];
  for (long long i = lo; i < hi; i++) {
    boff[i] = run;
    run += bcnt[i];
  }
}

template <typename IDX>
static __global__ void __launch_bounds__(CC_BLOCK)
    k_extract_write(const int32_t *__restrict__ A, long long N, int32_t thr,
                    long long CH, const long long *boff, long long cap,
                    IDX *oidx, int32_t *oval) {
  __shared__ int wtot[CC_BLOCK / 64];
  const long long c0 = (long long)blockIdx.x * CH;
  const long long c1 = min(N, c0 + CH);
  long long run = boff[blockIdx.x];
  if (run >= cap) return;  // block-uniform: nothing of this chunk is stored
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (long long b = c0; b < c1; b += 4LL * CC_BLOCK) {
    const long long i = b + 4LL * threadIdx.x;
    int32_t e[4] = {0, 0, 0, 0};
    const int m = (i < c1) ? cc_hits4(A, c1, i, thr, e) : 0;
    const int c = __popc(m);
    const int incl = cc_wave_incl_scan(c);
    if (lane == 63) wtot[wid] = incl;
    __syncthreads();
    int before = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < CC_BLOCK / 64; k++) {
      before += (k < wid) ? wtot[k] : 0;
      tot += wtot[k];
    }
    long long pos = run + before + incl - c;
#pragma unroll
    for (int j = 0; j < 4; j++)
      if ((m >> j) & 1) {
        if (pos < cap) {
          oidx[pos] = (IDX)(i + j);
          oval[pos] = e[j];
        }
        pos++;
      }
    run += tot;
    __syncthreads();
  }
}


// cc_reduce.hpp -- whole-field reductions and queries.
//
// Semantics follow reference src/coldcore/protocol.py:181-219 (the counts,
// the binning rule and the ordered extraction the host backends define):
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

// zero the CC_NSLOT census slots of one region
__global__ void k_clear_slots(unsigned long long *slots, int stride) {
  if (threadIdx.x < CC_NSLOT) slots[(size_t)threadIdx.x * stride] = 0ull;
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

// ---- field scans -----------------------------------------------------
// A workgroup owns one contiguous slab of the plane (a multiple of 8 x
// CC_BLOCK words, so every slab starts 16-byte aligned), each thread folds
// 16-byte loads of the slab into a 32-bit partial, the wave combines the
// partials with DPP row shifts and row broadcasts (the total lands in lane
// 63), the waves' totals meet in LDS and one atomic per workgroup lands in
// one of CC_NSLOT census slots (cc_census_flush).  The host sizes the slabs
// so no 32-bit partial can wrap (at most 2^15 words per thread).

// the words of slab b: [b * slab, min(N, (b + 1) * slab))
__device__ __forceinline__ void cc_slab(long long N, long long slab,
                                        long long *w0, long long *w1) {
  *w0 = (long long)blockIdx.x * slab;
  *w1 = min(N, *w0 + slab);
}

// the predicate count of the two uint16 counters of dword x: EQ counts
// halves equal to the target (pat = target in both halves: a zero half of
// x ^ pat), DEF sums max(0, mu - c) over both halves
template <int MODE>
__device__ __forceinline__ uint32_t cc_census2(uint32_t x, uint32_t pat, int mu) {
  if (MODE == CC_EQ) {
    const uint32_t y = x ^ pat;
    return (uint32_t)((y & 0xFFFFu) == 0u) + (uint32_t)((y >> 16) == 0u);
  } else {
    const int lo = mu - (int)(x & 0xFFFFu), hi = mu - (int)(x >> 16);
    return (uint32_t)(lo > 0 ? lo : 0) + (uint32_t)(hi > 0 ? hi : 0);
  }
}

// MODE CC_EQ: sum [cnt == target];  CC_DEF: sum max(0, mu - cnt)
template <int MODE>
__global__ void __launch_bounds__(CC_BLOCK)
    k_cnt_census(const uint16_t *__restrict__ cnt, long long N, long long slab,
                 int target, int mu, unsigned long long *slots, int stride) {
  long long w0, w1;
  cc_slab(N, slab, &w0, &w1);
  const uint32_t pat = ((uint32_t)target & 0xFFFFu) * 0x00010001u;
  const long long v1 = w0 + ((w1 - w0) & ~7LL);     // end of the whole 8-word runs
  uint32_t part = 0;
  for (long long i = w0 + 8LL * threadIdx.x; i < v1; i += 8LL * CC_BLOCK) {
    const uint4 q = *(const uint4 *)(cnt + i);
    part += cc_census2<MODE>(q.x, pat, mu) + cc_census2<MODE>(q.y, pat, mu) +
            cc_census2<MODE>(q.z, pat, mu) + cc_census2<MODE>(q.w, pat, mu);
  }
  const long long j = v1 + threadIdx.x;             // the slab's last < 8 words
  if (j < w1) {
    const int c = (int)cnt[j];
    if (MODE == CC_EQ) part += (uint32_t)(c == target);
    else part += (uint32_t)(mu > c ? mu - c : 0);
  }
  cc_census_flush<false>(part, slots, stride);
}

// max(0, max of the map plane): DPP wave maxima, one atomicMax per workgroup
__global__ void __launch_bounds__(CC_BLOCK)
    k_map_peak(const int32_t *__restrict__ A, long long N, long long slab,
               unsigned long long *slots, int stride) {
  long long w0, w1;
  cc_slab(N, slab, &w0, &w1);
  const long long v1 = w0 + ((w1 - w0) & ~3LL);
  int m = 0;
  for (long long i = w0 + 4LL * threadIdx.x; i < v1; i += 4LL * CC_BLOCK) {
    const int4 q = *(const int4 *)(A + i);
    m = max(m, max(max(q.x, q.y), max(q.z, q.w)));
  }
  const long long j = v1 + threadIdx.x;
  if (j < w1) m = max(m, A[j]);
  cc_census_flush<true>((uint32_t)max(m, 0), slots, stride);
}

// bin of a positive map value: floor((v - 1) * nbins / vmax), clipped to the
// last bin.  (v - 1) * nbins < 2^43 and the quotient is below 4096, so the
// correctly rounded double quotient never reaches the next integer (its gap
// to it is at least 1 / vmax >= 2^-31, far above the rounding error) and the
// truncation is exact.
__device__ __forceinline__ int cc_bin(int32_t v, int nbins, int32_t vmax) {
  const double q = (double)((long long)(v - 1) * nbins) / (double)vmax;
  const int b = (int)q;
  return b < nbins ? b : nbins - 1;
}

// histogram of the positive map values: a workgroup bins its slab into a
// private LDS histogram (nbins <= 4096 counters), then adds the nonzero
// counters to the global one
__global__ void __launch_bounds__(CC_BLOCK)
    k_map_hist(const int32_t *__restrict__ A, long long N, long long slab,
               int nbins, int32_t vmax, int32_t *hist) {
  extern __shared__ int32_t bins[];
  for (int i = threadIdx.x; i < nbins; i += CC_BLOCK) bins[i] = 0;
  __syncthreads();
  long long w0, w1;
  cc_slab(N, slab, &w0, &w1);
  const long long v1 = w0 + ((w1 - w0) & ~3LL);
  for (long long i = w0 + 4LL * threadIdx.x; i < v1; i += 4LL * CC_BLOCK) {
    const int4 q = *(const int4 *)(A + i);
    const int32_t e[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (e[k] > 0) atomicAdd(&bins[cc_bin(e[k], nbins, vmax)], 1);
  }
  const long long j = v1 + threadIdx.x;
  if (j < w1 && A[j] > 0) atomicAdd(&bins[cc_bin(A[j], nbins, vmax)], 1);
  __syncthreads();
  for (int i = threadIdx.x; i < nbins; i += CC_BLOCK)
    if (bins[i]) atomicAdd(&hist[i], bins[i]);
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

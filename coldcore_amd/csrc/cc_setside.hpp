// cc_setside.hpp -- set-side exact gathers and the swap-evaluation fast path.
//
// A gather counts, for each candidate word c, the words x of its ball B(c)
// whose cnt satisfies a predicate.  The walk reads |B| cnt entries per
// candidate.  Near a cover the predicate set Z = {x : pred(cnt[x])} is small
// (the uncovered words), and the same count is #{z in Z : dist(c, z) <= R}
// whenever the support table is exactly the metric ball (the host proves
// that once).  Z is listed once per cnt state as packed digits, and every
// candidate is compared with every z by digit arithmetic
// (reference src/plugins/covering/dctcore_fast.cuh:87-189, 456-488 and
// dct_hc.cu:590-754 give the behaviour; the code here is its own design).
//
// MI355X layout: a digit takes b = bits(max axis - 1) bits (1 for binary
// cells, so n <= 64 binary axes qualify, 4 for axes up to 16), packed
// little-endian into one uint64.  The gather kernel holds 32 candidates per
// workgroup in LDS and streams a slice of Z once, each z compared with all
// 32 candidates into register counters (Z read once per 32 candidates);
// the Hamming compare is SWAR: xor, fold each b-bit field onto its low bit,
// popcount.
//
// Swap evaluation (reference dct_hc.cu:868-977): for drop d and candidate c,
// the gain of c after removing d equals the gain of c now plus the words of
// B(c) in L_d = {x in B(d) : cnt[x] == mu} (they become deficient).  One walk
// per drop gives its loss and lists L_d as (packed digits, drop index); the
// candidates' gains come from one gather, and a correction kernel adds the
// matches in their drop's list.  cnt is never modified.
#pragma once
#include "cc_common.hpp"
#include "cc_walk.hpp"

#define CCS_CB 32   // candidates per gather workgroup

struct Dig {
  int ax[CC_MAXN];
  int n, b, R, mu;
  uint64_t lsb;     // bit j * b of every digit j (the Hamming fold mask)
};

__device__ __forceinline__ uint64_t cc_pack(const Dig &g, long long w) {
  uint64_t p = 0;
  if (w < (1LL << 32)) {
    uint32_t r = (uint32_t)w;
    for (int j = 0; j < g.n; j++) {
      const uint32_t a = (uint32_t)g.ax[j], q = r / a;
      p |= (uint64_t)(r - q * a) << (g.b * j);
      r = q;
    }
  } else {
    unsigned long long r = (unsigned long long)w;
    for (int j = 0; j < g.n; j++) {
      const unsigned long long a = (unsigned long long)g.ax[j], q = r / a;
      p |= (uint64_t)(r - q * a) << (g.b * j);
      r = q;
    }
  }
  return p;
}

// is z within distance R of c (both packed)?  Hamming: differing digits;
// Lee: sum of circular digit distances; torus: every circular distance;
// open grid: every plain distance.
template <int PROB>
__device__ __forceinline__ bool cc_within(const Dig &g, uint64_t c, uint64_t z) {
  if (PROB == CC_HAMMING) {
    const uint64_t x = c ^ z;
    uint64_t y = x;
    for (int s = 1; s < g.b; s++) y |= x >> s;
    return __popcll(y & g.lsb) <= g.R;
  }
  const uint32_t fm = (1u << g.b) - 1u;
  int acc = 0;
  for (int j = 0; j < g.n; j++) {
    const int a = (int)((c >> (g.b * j)) & fm), e = (int)((z >> (g.b * j)) & fm);
    int d = a > e ? a - e : e - a;
    if (PROB == CC_GRID) {
      if (d > g.R) return false;
    } else {
      d = min(d, g.ax[j] - d);
      if (PROB == CC_TORUS) {
        if (d > g.R) return false;
      } else {
        acc += d;
        if (acc > g.R) return false;
      }
    }
  }
  return true;
}

// claim the next slots of a shared list for the lanes with hit set (one
// atomic per wave); returns this lane's slot (meaningful where hit)
__device__ __forceinline__ unsigned long long cc_claim(bool hit,
                                                       unsigned long long *count) {
  const unsigned long long m = __ballot(hit);
  const int lane = threadIdx.x & 63;
  unsigned long long base = 0;
  if (m) {
    if (lane == __ffsll((long long)m) - 1)
      base = atomicAdd(count, (unsigned long long)__popcll(m));
    base = __shfl(base, __ffsll((long long)m) - 1, 64);
  }
  return base + (unsigned long long)__popcll(m & ((1ull << lane) - 1ull));
}

// Z = {x : pred(cnt[x])} as packed digits; *zcount = |Z| (may exceed zcap;
// only zcap entries are stored).  Every lane of a wave runs the same trip
// count, so the ballots are well defined.
template <int MODE>
__global__ void __launch_bounds__(CC_BLOCK)
    k_zlist(const uint16_t *__restrict__ cnt, long long X, int target, Dig g,
            uint64_t *z, long long zcap, unsigned long long *zcount) {
  const long long gs = (long long)gridDim.x * CC_BLOCK;
  const long long trips = (X + gs - 1) / gs;
  const long long i0 = (long long)blockIdx.x * CC_BLOCK + threadIdx.x;
  for (long long k = 0; k < trips; k++) {
    const long long i = i0 + k * gs;
    const bool hit = i < X && cc_pred(MODE, (int)cnt[i], target, g.mu);
    if (!__any(hit)) continue;
    const unsigned long long s = cc_claim(hit, zcount);
    if (hit && (long long)s < zcap) z[s] = cc_pack(g, i);
  }
}

// out[c] += #{z in Z[slice] : dist(words[c], z) <= R}; blockIdx.y = a tile
// of CCS_CB candidates, blockIdx.x = a slice of zper entries of Z; out must
// be zeroed first
template <int PROB>
__global__ void __launch_bounds__(CC_BLOCK)
    k_gather_set(const long long *__restrict__ words, int nb,
                 const uint64_t *__restrict__ z, long long zn, long long zper,
                 Dig g, int32_t *out) {
  __shared__ uint64_t cd[CCS_CB];
  __shared__ int32_t red[CC_BLOCK / 64][CCS_CB];
  const int c0 = blockIdx.y * CCS_CB;
  const int nc = min(CCS_CB, nb - c0);
  if (threadIdx.x < CCS_CB)
    cd[threadIdx.x] = (int)threadIdx.x < nc ? cc_pack(g, words[c0 + threadIdx.x])
                                            : 0ull;
  __syncthreads();
  // the candidates are uniform across the workgroup: scalar registers
  uint64_t cr[CCS_CB];
  int32_t acc[CCS_CB];
#pragma unroll
  for (int c = 0; c < CCS_CB; c++) {
    const uint64_t v = cd[c];
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    cr[c] = ((uint64_t)hi << 32) | lo;
    acc[c] = 0;
  }
  const long long t0 = (long long)blockIdx.x * zper;
  const long long t1 = min(zn, t0 + zper);
  for (long long t = t0 + threadIdx.x; t < t1; t += CC_BLOCK) {
    const uint64_t zz = z[t];
#pragma unroll
    for (int c = 0; c < CCS_CB; c++) acc[c] += cc_within<PROB>(g, cr[c], zz) ? 1 : 0;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < CCS_CB; c++) {
    int32_t v = acc[c];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[wid][c] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < nc) {
    int32_t s = 0;
#pragma unroll
    for (int w = 0; w < CC_BLOCK / 64; w++) s += red[w][threadIdx.x];
    if (s) atomicAdd(&out[c0 + threadIdx.x], s);
  }
}

// per drop word (blockIdx = drop * split + segment): loss[d] += #{x in B(d) :
// 1 <= cnt <= mu} and the list L_d = {x in B(d) : cnt == mu} appended as
// (packed digits, d); *lcount = the true total (may exceed lcap)
template <bool GRID, int FMT>
__global__ void __launch_bounds__(CC_BLOCK)
    k_drop_walk(Geom g, Dig dg, const long long *__restrict__ drops, int split,
                const uint16_t *__restrict__ cnt, const uint64_t *__restrict__ pat,
                long long blen, int32_t *loss, uint64_t *ld, int32_t *ldrop,
                long long lcap, unsigned long long *lcount) {
  __shared__ long long disp[CC_MAXN * CC_CSTRIDE];
  __shared__ unsigned char okf[CC_MAXN * CC_CSTRIDE];
  __shared__ int digs[CC_MAXN], sax[CC_MAXN];
  __shared__ long long spw[CC_MAXN];
  const int di = blockIdx.x / split, seg = blockIdx.x - di * split;
  const long long w = drops[di];
  cc_walk_disp(g, w, disp, okf, digs, sax, spw);
  __syncthreads();
  long long acc = 0;
  const long long step = (long long)split * CC_BLOCK;
  for (long long base = (long long)seg * CC_BLOCK; base < blen; base += step) {
    const long long t = base + threadIdx.x;
    bool ok = false;
    long long x = 0;
    if (t < blen)
      x = cc_walk_word<GRID, FMT>(pat[t], w, disp, okf, &ok, digs, sax, spw, g.R);
    const int c = ok ? (int)cnt[x] : -1;
    acc += (c >= 1 && c <= g.mu) ? 1 : 0;
    const bool hit = c == g.mu;
    if (!__any(hit)) continue;
    const unsigned long long s = cc_claim(hit, lcount);
    if (hit && (long long)s < lcap) {
      ld[s] = cc_pack(dg, x);
      ldrop[s] = di;
    }
  }
  acc = cc_block_sum(acc);
  if (threadIdx.x == 0 && acc) atomicAdd(&loss[di], (int32_t)acc);
}

// gain[j] += #{(x, d) in L : d == cdrop[j], dist(cand[j], x) <= R}; skipped
// when the list overflowed (the host then takes the exact fallback)
template <int PROB>
__global__ void __launch_bounds__(CC_BLOCK)
    k_swap_corr(const long long *__restrict__ cand, const int32_t *__restrict__ cdrop,
                int split, const uint64_t *__restrict__ ld,
                const int32_t *__restrict__ ldrop, long long lcap,
                const unsigned long long *__restrict__ lcount, Dig g,
                int32_t *gain) {
  const long long nl = (long long)*lcount;
  if (nl > lcap) return;
  __shared__ uint64_t cdig;
  const int j = blockIdx.x / split, seg = blockIdx.x - j * split;
  if (threadIdx.x == 0) cdig = cc_pack(g, cand[j]);
  __syncthreads();
  const uint64_t c = cdig;
  const int d = cdrop[j];
  long long acc = 0;
  const long long step = (long long)split * CC_BLOCK;
  for (long long t = (long long)seg * CC_BLOCK + threadIdx.x; t < nl; t += step)
    acc += (ldrop[t] == d && cc_within<PROB>(g, c, ld[t])) ? 1 : 0;
  acc = cc_block_sum(acc);
  if (threadIdx.x == 0 && acc) atomicAdd(&gain[j], (int32_t)acc);
}

// cc_walk.hpp -- support-set walks: exact incremental updates and exact
// per-candidate gathers over one word's ball.
//
// The ball of centre w is w + D(p) for every pattern p, where D sums the
// per-(position, delta) board displacement ((dig_pos + delta) mod ax_pos -
// dig_pos) * stride_pos (reference src/plugins/covering/dctcore_core.cuh:
// 158-212).  The open grid resolves the same circular delta against the
// centre digit and drops a pattern whose displacement leaves the L-inf
// radius, which is a wrap past the board edge (dctcore_grid.cuh:239-286).
// Three packed formats are decoded: v1 (8-bit pos<<4|delta fields, pos < 16,
// the reference's wire format), wide (10-bit fields, pos < 64, at most 6
// per pattern) for the Hamming / Lee cells with more than 16 axes, and
// long-axis (16-bit pos<<10|delta fields, at most 4) for the L-inf cells
// with axes longer than 16, whose displacement is computed per field from
// the centre digit (a table of every (position, delta) would not fit).
//
// One workgroup walks one contiguous segment of one word's table
// (blockIdx = word * split + segment).  The displacement table lives in LDS
// (16 x 15 int64); each digit is computed by the thread that needs it, so
// no thread serialises the mixed-radix decomposition.  Updates are packed
// uint16 +-1 on the aligned 32-bit word (the plane is padded to an even
// length); gathers reduce over the wave with DPP shuffles.
#pragma once
#include "cc_common.hpp"

struct Geom {
  long long pw[CC_MAXN];
  int ax[CC_MAXN];
  int n, R, mu;
};

#define CC_WALK_UPDATE 0  // cnt += delta over the ball
#define CC_WALK_TRANS 1   // ... and count deficit transitions (update_delta)
#define CC_WALK_GATHER 2  // count predicate(cnt) over the ball

// the displacement (and, for the open grid, in-board flag) of every
// (position, delta) slot for centre w: the centre's digits first (one
// division per axis, 32-bit when the index allows), then the table; the
// caller synchronises after it
__device__ __forceinline__ void cc_walk_disp(const Geom &g, long long w,
                                             long long *disp,
                                             unsigned char *okf, int *digs,
                                             int *sax = nullptr,
                                             long long *spw = nullptr) {
  for (int p = threadIdx.x; p < g.n; p += blockDim.x) {
    if (sax) {
      sax[p] = g.ax[p];
      spw[p] = g.pw[p];
    }
    int d;
    if (w < (1LL << 32)) {
      const uint32_t w32 = (uint32_t)w;
      d = g.pw[p] > (long long)w32 ? 0
                                   : (int)((w32 / (uint32_t)g.pw[p]) % (uint32_t)g.ax[p]);
    } else {
      d = (int)((w / g.pw[p]) % g.ax[p]);
    }
    digs[p] = d;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < CC_MAXN * CC_CSTRIDE; t += blockDim.x) {
    const int p = t / CC_CSTRIDE, dl = t - p * CC_CSTRIDE + 1;
    long long off = 0;
    unsigned char ok = 0;
    if (p < g.n && dl < g.ax[p]) {
      const int a = g.ax[p];
      const int dig = digs[p];
      const int sh = (dig + dl < a) ? dl : dl - a;
      off = (long long)sh * g.pw[p];
      ok = (sh <= g.R && sh >= -g.R) ? 1 : 0;
    }
    disp[t] = off;
    okf[t] = ok;
  }
}

// the word pattern pt reaches from centre w; *ok = false when an open-grid
// move leaves the board
template <bool GRID, int FMT>
__device__ __forceinline__ long long cc_walk_word(uint64_t pt, long long w,
                                                  const long long *disp,
                                                  const unsigned char *okf,
                                                  bool *ok,
                                                  const int *digs = nullptr,
                                                  const int *sax = nullptr,
                                                  const long long *spw = nullptr,
                                                  int R = 0) {
  long long x = w;
  bool good = true;
  if constexpr (FMT == CC_FMT_X16) {
#pragma unroll
    for (int k = 0; k < CC_MAXPAT_X; k++) {
      const unsigned b = (unsigned)(pt >> (16 * k)) & 0xFFFFu;
      if (!b) break;
      const int pos = (int)(b >> 10), dl = (int)(b & 1023u);
      const int a = sax[pos];
      const int sh = digs[pos] + dl < a ? dl : dl - a;
      x += (long long)sh * spw[pos];
      if (GRID) good = good && sh <= R && sh >= -R;
    }
    *ok = good;
    return x;
  }
  constexpr int FB = FMT == CC_FMT_V1 ? 8 : 10;
  constexpr int NF = FMT == CC_FMT_V1 ? CC_MAXPAT : CC_MAXPAT_W;
  constexpr unsigned FM = (1u << FB) - 1u;
#pragma unroll
  for (int k = 0; k < NF; k++) {
    const unsigned b = (unsigned)(pt >> (FB * k)) & FM;
    if (!b) break;
    const int e = (int)(b >> 4) * CC_CSTRIDE + (int)(b & 15u) - 1;
    x += disp[e];
    if (GRID) good = good && okf[e];
  }
  *ok = good;
  return x;
}

template <int KIND, bool GRID, int FMT>
__global__ void __launch_bounds__(CC_BLOCK)
    k_walk(Geom g, WordsArg wa, const long long *__restrict__ wptr, int split,
           int delta, uint16_t *cnt, const uint64_t *__restrict__ pat,
           long long blen, int pmode, int target, int32_t *out,
           unsigned long long *trans) {
  __shared__ long long disp[CC_MAXN * CC_CSTRIDE];
  __shared__ unsigned char okf[CC_MAXN * CC_CSTRIDE];
  __shared__ int digs[CC_MAXN], sax[CC_MAXN];
  __shared__ long long spw[CC_MAXN];
  const int wi = blockIdx.x / split, seg = blockIdx.x - wi * split;
  const long long w = wptr ? wptr[wi] : wa.w[wi];
  cc_walk_disp(g, w, disp, okf, digs, sax, spw);
  __syncthreads();
  long long acc = 0;
  const long long step = (long long)split * CC_BLOCK;
  for (long long t = (long long)seg * CC_BLOCK + threadIdx.x; t < blen;
       t += step) {
    bool ok;
    const long long x = cc_walk_word<GRID, FMT>(pat[t], w, disp, okf, &ok, digs,
                                                sax, spw, g.R);
    if (GRID && !ok) continue;
    if (KIND == CC_WALK_GATHER) {
      acc += cc_pred(pmode, (int)cnt[x], target, g.mu) ? 1 : 0;
    } else {
      uint32_t *cell = (uint32_t *)(cnt + (x & ~1LL));
      const int sh = (int)(x & 1) * 16;
      const uint32_t add = 1u << sh;
      const uint32_t old = delta > 0 ? atomicAdd(cell, add) : atomicSub(cell, add);
      if (KIND == CC_WALK_TRANS) {
        const int c = (int)((old >> sh) & 0xFFFFu);
        acc += delta > 0 ? (c < g.mu) : (c <= g.mu);
      }
    }
  }
  if (KIND != CC_WALK_UPDATE) {
    acc = cc_block_sum(acc);
    if (threadIdx.x == 0) {
      if (KIND == CC_WALK_GATHER) {
        if (split == 1) out[wi] = (int32_t)acc;
        else if (acc) atomicAdd(&out[wi], (int32_t)acc);
      } else if (acc) {
        atomicAdd(trans, (unsigned long long)acc);
      }
    }
  }
}

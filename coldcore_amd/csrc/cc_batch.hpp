// cc_batch.hpp -- many uneven cells on one GPU: the transform of every cell
// by one persistent launch per sweep level, batched walks and reductions,
// and the device side of a batched LNS round (ruin k, greedy recreate,
// accept / revert) whose fixed launch sequence is captured once into a HIP
// graph.
//
// Batched transform: a cell's sweeps run in order (sweep s + 1 reads what
// sweep s wrote), so the work is split by sweep level: level l holds the
// tiles of every cell's l-th sweep.  One persistent launch per level deals
// the (sweep, tile) items out round-robin: workgroup g takes items g,
// g + G, g + 2G, ... of a list ordered largest tile first.  Items of one
// level never depend on each other, so no workgroup ever waits on another
// (the kernel boundary orders the levels, and the L2 write-back /
// invalidate between XCDs happens there once, not per tile).  A shared
// atomic work counter was measured slower (64 cells: 139 -> 132 us per
// transform without it): each claim is a device-scope atomic whose round
// trip crosses the XCDs, once per tile.
#pragma once
#include "cc_common.hpp"
#include "cc_sweep.hpp"

#define CCB_KMAX 8       // ruin size per LNS round
#define CCB_WSPLIT 8     // workgroups per walked ball
// LDS layout of the batch's generic tiles: unpadded (measured faster than
// the singleton's one-pad-word-per-32 layout on the batched workloads)
#ifndef CCB_PAD
#define CCB_PAD false
#endif

struct BCell {
  long long pw[CC_MAXN + 1];
  int ax[CC_MAXN];
  int prob, n, R, L, mu, fmt;
  long long X;
  int32_t *A[CC_MAXL];   // A[0]: map / code indicator; A[1..L-1]: layers
  uint16_t *H;           // l16: L uint16 layer planes of X words (inter-sweep)
  int l16;               // every layer value fits 16 bits (ball, code <= 65,535)
  uint16_t *cnt;         // padded to an even length
  const uint64_t *pat;   // Hamming / Lee: support table (v1 or wide)
  long long blen;
  long long *code;       // current code, M words (capacity cap)
  long long *best;       // best feasible code found by the LNS (bestM words)
  int cap;
  // LNS state (control kernels only)
  int M, uncov, bestM, add_budget, pre_uncov, nrem, nadd;
  int snap;              // the current code is the new best: copy it
  long long rem[CCB_KMAX], add[CCB_KMAX];
  unsigned long long rng;
  long long rounds, accepted, reverted;
  // reductions of the current step
  unsigned long long amax;   // (gain << 32) | (0xFFFFFFFF - index)
  unsigned long long nzero;  // words with cnt == 0
};

// pitch (words) of the 16-bit layer planes: a multiple of 8, so every
// plane starts 16-byte aligned
inline long long ccb_hpitch(long long X) { return (X + 7) & ~7LL; }

struct BItem {
  int cell, delta;   // delta 0: no-op slot
  long long w;
};

// ---------------------------------------------------------------- transform
template <int PROB>
__device__ __forceinline__ void ccb_tile_linf(const SweepArgs &a, long long tile,
                                              int32_t *__restrict__ S) {
  cc_sweep_tile<PROB, 1, int32_t, CCB_PAD, false>(a, tile, S);
}


// LT: inter-sweep layer planes int32 (A) or uint16 (H, cells whose layer
// values provably fit 16 bits; never with the runtime-length window path)
template <typename LT>
__device__ __forceinline__ void ccb_tile_t(const SweepArgs &a, int prob, int L,
                                           long long tile, int32_t *S) {
  if (prob == CC_HAMMING) {
    if (L == 2) cc_sweep_tile<CC_HAMMING, 2, LT, CCB_PAD, false>(a, tile, S);
    else if (L == 3) cc_sweep_tile<CC_HAMMING, 3, LT, CCB_PAD, false>(a, tile, S);
    else cc_sweep_tile<CC_HAMMING, 4, LT, CCB_PAD, false>(a, tile, S);
  } else if (prob == CC_LEE) {
    if (L == 2) cc_sweep_tile<CC_LEE, 2, LT, CCB_PAD, false>(a, tile, S);
    else if (L == 3) cc_sweep_tile<CC_LEE, 3, LT, CCB_PAD, false>(a, tile, S);
    else cc_sweep_tile<CC_LEE, 4, LT, CCB_PAD, false>(a, tile, S);
  } else if constexpr (sizeof(LT) == 2) {
    if (prob == CC_TORUS) cc_sweep_tile<CC_TORUS, 1, LT, CCB_PAD, false>(a, tile, S);
    else cc_sweep_tile<CC_GRID, 1, LT, CCB_PAD, false>(a, tile, S);
  } else if (prob == CC_TORUS) {
    ccb_tile_linf<CC_TORUS>(a, tile, S);
  } else {
    ccb_tile_linf<CC_GRID>(a, tile, S);
  }
}

// homogeneous Hamming groups of 16-bit cells (axis size 2 or 3, at least
// rr axes): register rounds of rr axes per LDS round trip.  Cubes stay at
// 8 or 9 words (rr = 3 binary, 2 ternary): the 16-word cubes of the
// singleton's round kernel would set this whole kernel's register count
// (187 VGPRs, 2 waves per SIMD, against 112 and 4), and every tile kind of
// a level shares it (64 cells: 133 -> 121 us per transform, the grid boxes
// 105 -> 78 us, with the smaller cubes).
template <int R1>
__device__ __forceinline__ void ccb_tile_round(const SweepArgs &a, long long tile,
                                               int32_t *S) {
  if (a.Q[0] == 2) cc_round_tile<CC_HAMMING, R1, 2, 3, uint16_t>(a, tile, S);
  else cc_round_tile<CC_HAMMING, R1, 3, 2, uint16_t>(a, tile, S);
}

__device__ __forceinline__ void ccb_tile(const SweepArgs &a, const BCell &c,
                                         long long tile, int32_t *S) {
  if (a.rr) {
    if (c.L == 2) ccb_tile_round<2>(a, tile, S);
    else if (c.L == 3) ccb_tile_round<3>(a, tile, S);
    else ccb_tile_round<4>(a, tile, S);
  } else if (c.l16) {
    ccb_tile_t<uint16_t>(a, c.prob, c.L, tile, S);
  } else {
    ccb_tile_t<int32_t>(a, c.prob, c.L, tile, S);
  }
}

// items[i] = (sweep argument index, tile), all of one sweep level; NT
// threads per workgroup work on one tile
template <int NT>
__global__ void __launch_bounds__(NT)
    k_ccb_sweeps(const SweepArgs *__restrict__ args, const int *__restrict__ argcell,
                 const BCell *__restrict__ cells, const int2 *__restrict__ items,
                 int nitems) {
  extern __shared__ int32_t S[];
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    // uniform item: its arguments load through the scalar unit
    const int ai = __builtin_amdgcn_readfirstlane(items[it].x);
    const long long tile = (long long)__builtin_amdgcn_readfirstlane(items[it].y);
    const SweepArgs &a = args[ai];
    const BCell &c = cells[argcell[ai]];
    ccb_tile(a, c, tile, S);
  }
}

// ---------------------------------------------------------------- seeds
// A0 := 0 then A0[code] += 1, every cell (blockIdx.y = cell)
__global__ void k_ccb_zero_a0(BCell *cells) {
  const BCell &c = cells[blockIdx.y];
  const long long gs = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < c.X; i += gs)
    c.A[0][i] = 0;
}

__global__ void k_ccb_scatter(BCell *cells) {
  const BCell &c = cells[blockIdx.y];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < c.M; i += gridDim.x * blockDim.x)
    atomicAdd(&c.A[0][c.code[i]], 1);
}

// ---------------------------------------------------------------- walks
// split workgroups per item (blockIdx = item * split + segment), each taking
// every split-th block of CC_BLOCK ball elements: cnt += delta over the ball
// of w in its cell.  Hamming / Lee walk their support table; torus / grid
// enumerate the (2R+1)^n L-inf offsets (the open grid drops the ones off the
// board).  Cell sizes stay below 2^26 words and L-inf balls below 7^8
// offsets, so all index arithmetic is 32-bit and the offset digits come
// from high-multiply divisions (exact: offset * (2R + 1) < 2^32).
__global__ void __launch_bounds__(CC_BLOCK)
    k_ccb_walk(const BCell *__restrict__ cells, const BItem *__restrict__ items,
               int split, int ipc, int istride) {
  // the launch walks the first ipc of every cell's istride item slots
  const int q = blockIdx.x / split, seg = blockIdx.x - q * split;
  const int cell = q / ipc, ii = cell * istride + (q - cell * ipc);
  const BItem itm = items[ii];
  if (itm.delta == 0) return;
  const BCell &c = cells[itm.cell];
  const int n = c.n, R = c.R, prob = c.prob;
  uint16_t *const cnt = c.cnt;
  const int w = (int)itm.w;
  __shared__ long long disp[CC_MAXN * CC_CSTRIDE];
  __shared__ int dig[CC_MAXN], sax[CC_MAXN], spw[CC_MAXN];
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    const int pw = (int)c.pw[p], a = c.ax[p];
    dig[p] = (w / pw) % a;
    sax[p] = a;
    spw[p] = pw;
  }
  __syncthreads();
  const bool add = itm.delta > 0;
  auto bump = [&](int x) {
    uint32_t *cell32 = (uint32_t *)(cnt + (x & ~1));
    const uint32_t v = 1u << ((x & 1) * 16);
    if (add) atomicAdd(cell32, v);
    else atomicSub(cell32, v);
  };
  const int t0 = seg * CC_BLOCK + (int)threadIdx.x, step = split * CC_BLOCK;
  if (prob == CC_TORUS || prob == CC_GRID) {
    const int side = 2 * R + 1;
    const uint32_t ms = cc_magic((uint32_t)side);
    int tot = 1;
    for (int p = 0; p < n; p++) tot *= side;
    const bool torus = prob == CC_TORUS;
    for (int t = t0; t < tot; t += step) {
      int x = 0, r = t;
      bool ok = true;
      for (int p = 0; p < n; p++) {
        const int q = cc_div(r, ms);
        const int a = sax[p];
        int d = dig[p] + (r - q * side) - R;
        r = q;
        if (torus) d += d < 0 ? a : (d >= a ? -a : 0);
        else ok = ok && d >= 0 && d < a;
        x += d * spw[p];
      }
      if (ok) bump(x);
    }
    return;
  }
  for (int t = threadIdx.x; t < CC_MAXN * CC_CSTRIDE; t += blockDim.x) {
    const int p = t / CC_CSTRIDE, dl = t - p * CC_CSTRIDE + 1;
    long long off = 0;
    if (p < n && dl < sax[p]) {
      const int a = sax[p];
      const int sh = (dig[p] + dl < a) ? dl : dl - a;
      off = (long long)sh * spw[p];
    }
    disp[t] = off;
  }
  __syncthreads();
  const int FB = c.fmt == CC_FMT_V1 ? 8 : 10;
  const int NF = c.fmt == CC_FMT_V1 ? CC_MAXPAT : CC_MAXPAT_W;
  const unsigned FM = (1u << FB) - 1u;
  const uint64_t *__restrict__ pat = c.pat;
  const int blen = (int)c.blen;
  for (int t = t0; t < blen; t += step) {
    const uint64_t pt = pat[t];
    long long x = w;
    for (int k = 0; k < NF; k++) {
      const unsigned b = (unsigned)(pt >> (FB * k)) & FM;
      if (!b) break;
      x += disp[(int)(b >> 4) * CC_CSTRIDE + (int)(b & 15u) - 1];
    }
    bump((int)x);
  }
}

// ---------------------------------------------------------------- reductions
__global__ void k_ccb_reset_red(BCell *cells, int ncells) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ncells) { cells[i].amax = 0ull; cells[i].nzero = 0ull; }
}

// one block per chunk (cell, first word; chunks of CCB_CHUNK words): number
// of uncovered words; with MAP also the largest map value and its smallest
// index.  The chunk table sizes the grid by the words, not by the largest
// cell times the cell count.
#define CCB_CHUNK 16384
__device__ __forceinline__ void ccb_key_max(unsigned long long &best, int32_t g,
                                            long long i) {
  if (g > 0) {
    const unsigned long long key = ((unsigned long long)(uint32_t)g << 32) |
                                   (0xFFFFFFFFull - (unsigned long long)i);
    best = key > best ? key : best;
  }
}

__global__ void __launch_bounds__(CC_BLOCK)
    k_ccb_reduce(BCell *cells, const long long *__restrict__ chunks, int with_map) {
  const long long ch = chunks[blockIdx.x];
  const int ci = (int)(ch >> 40);
  const long long w0 = ch & ((1LL << 40) - 1);
  BCell &c = cells[ci];
  const long long w1 = min(c.X, w0 + CCB_CHUNK);
  const uint16_t *__restrict__ cnt = c.cnt;
  const int32_t *__restrict__ map = c.A[0];
  unsigned long long z = 0, best = 0;
  // groups of 8 words: one 16-byte cnt load and two 16-byte map loads (w0 is
  // a multiple of CCB_CHUNK, so every group is 16-byte aligned)
  const int ng = (int)((w1 - w0) >> 3);
  for (int g = threadIdx.x; g < ng; g += CC_BLOCK) {
    const long long i = w0 + 8LL * g;
    const uint4 cv = *(const uint4 *)(cnt + i);
    const uint32_t cw[4] = {cv.x, cv.y, cv.z, cv.w};
#pragma unroll
    for (int k = 0; k < 4; k++)
      z += ((cw[k] & 0xFFFFu) == 0u) + ((cw[k] >> 16) == 0u);
    if (with_map) {
      const int4 a0 = *(const int4 *)(map + i), a1 = *(const int4 *)(map + i + 4);
      const int32_t gv[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
#pragma unroll
      for (int k = 0; k < 8; k++) ccb_key_max(best, gv[k], i + k);
    }
  }
  for (long long i = w0 + 8LL * ng + threadIdx.x; i < w1; i += CC_BLOCK) {
    z += cnt[i] == 0;
    if (with_map) ccb_key_max(best, map[i], i);
  }
  const long long zs = cc_block_sum((long long)z);
  cc_key_flush(best, &c.amax);
  if (threadIdx.x == 0 && zs) atomicAdd(&c.nzero, (unsigned long long)zs);
}

// best := code for the cells flagged by the control kernels (blockIdx.y =
// cell), all threads copying
__global__ void k_ccb_snapshot(BCell *cells) {
  BCell &c = cells[blockIdx.y];
  if (!c.snap) return;
  const int gs = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < c.M; i += gs)
    c.best[i] = c.code[i];
}

// ---------------------------------------------------------------- LNS control
__device__ __forceinline__ unsigned long long ccb_mix(unsigned long long x) {
  // splitmix64 finaliser
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ __forceinline__ unsigned long long ccb_next(BCell &c) {
  c.rng = ccb_mix(c.rng);
  return c.rng;
}

// ruin: remove k distinct random words (a partial Fisher-Yates from the
// tail); a feasible cell will re-add one fewer (notch descent).  Ruin and
// add leave amax / nzero cleared, so the reductions of a round need no
// separate reset launch.
__global__ void k_ccb_ruin(BCell *__restrict__ cells, int ncells, int k,
                           BItem *__restrict__ items) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= ncells) return;
  BCell &c = cells[ci];
  // the cell's state in registers: one round trip for the fields, one per
  // removed word for its two code slots
  const int M = c.M, uncov = c.uncov;
  long long *__restrict__ code = c.code;
  unsigned long long rng = c.rng;
  const int kk = min(k, M - 1);
  const int nrem = kk > 0 ? kk : 0;
  c.pre_uncov = uncov;
  c.nrem = nrem;
  c.nadd = 0;
  c.add_budget = kk <= 0 ? 0 : (uncov == 0 ? kk - 1 : kk);
  for (int j = 0; j < CCB_KMAX; j++) {
    BItem it{ci, 0, 0};
    if (j < nrem) {
      const int top = M - 1 - j;
      rng = ccb_mix(rng);
      const int r = (int)(rng % (unsigned long long)(top + 1));
      const long long t = code[r], u = code[top];
      code[r] = u;
      code[top] = t;
      c.rem[j] = t;
      it.delta = -1;
      it.w = t;
    }
    items[ci * CCB_KMAX + j] = it;
  }
  c.rng = rng;
  c.M = M - nrem;
  c.amax = 0ull;   // clear the accumulators for the round's first reduction
  c.nzero = 0ull;
}

// recreate step j: add the word of largest gain (smallest index on ties).
// j < 0: a greedy step instead (add while the cell has uncovered words; the
// added words are not recorded for a revert)
__global__ void k_ccb_add(BCell *__restrict__ cells, int ncells, int j,
                          BItem *__restrict__ items) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= ncells) return;
  BCell &c = cells[ci];
  BItem it{ci, 0, 0};
  // greedy (j < 0): a cell with an uncovered word has a positive gain
  // somewhere (mu = 1: the word itself), so amax != 0 below is the test
  const bool go = j < 0 ? true : j < c.add_budget;
  if (go && c.amax != 0ull && c.M < c.cap) {
    const long long w = (long long)(0xFFFFFFFFull - (c.amax & 0xFFFFFFFFull));
    c.code[c.M++] = w;
    if (j >= 0) c.add[c.nadd++] = w;
    it.delta = 1;
    it.w = w;
  }
  items[ci] = it;
  c.amax = 0ull;   // read: clear for the next reduction
  c.nzero = 0ull;
}

// accept / revert: a round that uncovers more words is undone with
// probability 0.9 (added words out, removed words back); a new best
// covering code is flagged for k_ccb_snapshot
__global__ void k_ccb_decide(BCell *__restrict__ cells, int ncells,
                             BItem *__restrict__ items) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= ncells) return;
  BCell &c = cells[ci];
  const int now = (int)c.nzero;
  bool revert = false;
  int snap = 0;
  if (c.nrem > 0 && now > c.pre_uncov) {
    const double u = (double)(ccb_next(c) >> 11) * (1.0 / 9007199254740992.0);
    revert = u > 0.10;
  }
  for (int j = 0; j < 2 * CCB_KMAX; j++) items[ci * 2 * CCB_KMAX + j] = BItem{ci, 0, 0};
  if (revert) {
    for (int j = 0; j < c.nadd; j++)
      items[ci * 2 * CCB_KMAX + j] = BItem{ci, -1, c.add[j]};
    c.M -= c.nadd;
    // the removed words follow the added ones, so a round of ruin size k
    // fills at most the first 2k slots (the revert walk launches only those)
    for (int j = 0; j < c.nrem; j++) {
      items[ci * 2 * CCB_KMAX + c.nadd + j] = BItem{ci, 1, c.rem[j]};
      c.code[c.M++] = c.rem[j];
    }
    c.uncov = c.pre_uncov;
    c.reverted++;
  } else {
    c.uncov = now;
    if (c.nrem > 0) c.accepted++;
    if (now == 0 && c.M < c.bestM) {
      c.bestM = c.M;
      snap = 1;
    }
  }
  c.snap = snap;
  c.rounds++;
}

// initial state after a recount: uncovered count, best = the start if feasible
__global__ void k_ccb_lns_init(BCell *cells, int ncells, unsigned long long seed) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= ncells) return;
  BCell &c = cells[ci];
  c.uncov = (int)c.nzero;
  c.rng = ccb_mix(seed * 0x100000001B3ull + (unsigned long long)ci);
  c.rounds = c.accepted = c.reverted = 0;
  c.bestM = 0x7FFFFFFF;
  c.snap = 0;
  if (c.uncov == 0) {
    c.bestM = c.M;
    c.snap = 1;
  }
}

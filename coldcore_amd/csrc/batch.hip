// batch.hip -- the batched multi-cell engine (additive C ABI, `ccb_*`).
//
// Holds many uneven cells at once (Hamming, Lee, king torus, open grid; any
// mix of axis sizes and radii R <= 3), all in one device arena beside the
// singleton `dct_*` cell.  A transform of every cell is one persistent
// launch per sweep level (cc_batch.hpp); a batched LNS round (ruin k random
// words per cell,
// greedy recreate by exact gain, accept or revert) is a fixed sequence of
// launches, captured once into a HIP graph and replayed.  No counterpart in
// the reference (one cell per process there, reference docs/PLUGIN.md:73-76).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/coldcore_amd/batch.h"
#include "cc_batch.hpp"

#define CKB(x)                                                              \
  do {                                                                      \
    hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) {                                                 \
      fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_),     \
              __FILE__, __LINE__);                                          \
      return -(100 + (int)e_);                                              \
    }                                                                       \
  } while (0)

namespace {

constexpr int CCB_LDS = 32 * 1024;   // default LDS budget of a sweep tile

struct Batch {
  bool live = false;
  int nc = 0;
  std::vector<BCell> h;               // host copies (geometry and pointers)
  BCell *d = nullptr;                 // device cell table
  std::vector<void *> allocs;
  std::vector<SweepArgs> args[3];     // 0 gain, 1 loss, 2 recount
  SweepArgs *dargs[3] = {};
  std::vector<int> argcell, need;
  std::vector<int2> items;
  std::vector<int> level_off;         // items of level l: [off[l], off[l+1])
  int *d_argcell = nullptr;
  int2 *d_items = nullptr;
  // the items of the cells a greedy run still covers (ccb_greedy): level l
  // is [open_off[l], open_off[l + 1]) of d_open
  int2 *d_open = nullptr;
  std::vector<int> open_off;
  int nargs = 0, nitems = 0, nlevels = 0;
  size_t lds = 0;
  int per_cu = 1, ncu = 1;
  int threads = 512;                  // threads per level tile (CCB_THREADS,
                                      // autotuned: 256, 512 or 1024)
  int budget = CCB_LDS;               // LDS budget the plan was built for
  std::vector<void *> plan_allocs;    // freed when the plan is rebuilt
  long long maxX = 0;
  BItem *d_ruin = nullptr, *d_add = nullptr, *d_rev = nullptr;
  std::vector<long long> chunks;      // reduction chunks: cell << 40 | word
  long long *d_chunks = nullptr;
  int snap_grid = 1;                  // snapshot blocks per cell
  hipStream_t st = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_k = -1;
};

Batch B;

template <typename T>
int balloc(T **p, size_t bytes) {
  if (hipMalloc((void **)p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return -1;
  }
  B.allocs.push_back((void *)*p);
  return 0;
}

void bfree_all() {
  if (B.st) (void)hipStreamSynchronize(B.st);
  if (B.exec) { (void)hipGraphExecDestroy(B.exec); B.exec = nullptr; }
  if (B.graph) { (void)hipGraphDestroy(B.graph); B.graph = nullptr; }
  for (void *p : B.allocs) (void)hipFree(p);
  B.allocs.clear();
  for (void *p : B.plan_allocs) (void)hipFree(p);
  B.plan_allocs.clear();
  if (B.st) { (void)hipStreamDestroy(B.st); B.st = nullptr; }
  B = Batch();
}

// largest partial distance one axis adds (Hamming 1, Lee floor(a/2))
int live_after(const BCell &c, int upto) {
  if (c.L == 1) return 1;
  long long w = 0;
  for (int j = 0; j < upto; j++) w += c.prob == CC_LEE ? c.ax[j] / 2 : 1;
  return (int)std::min<long long>(w, c.R) + 1;
}

// the cell's sweeps: groups of consecutive axes whose tile (L live planes
// x P words x TL columns, int32) fits the LDS budget; the first group is
// contiguous (TL = 1), later ones stage up to 32 columns
// axes per register round for a homogeneous Hamming group of axis size q
// (cubes of 8 or 9 words), 0 where the LDS tile runs
int round_axes_b(const BCell &c, int q) {
  if (c.prob != CC_HAMMING || !c.l16 || getenv("CCB_NO_ROUND")) return 0;
  return q == 2 ? 3 : (q == 3 ? 2 : 0);
}

// LDS bytes of a round tile of pt words: layer pairs packed per dword, one
// pad word per 16
long long round_lds(const BCell &c, long long pt) {
  return (long long)((c.L + 1) / 2) * (pt + (pt >> 4) + 1) * 4;
}

void plan_cell(int ci, const BCell &c) {
  int j = 0;
  while (j < c.n) {
    const long long s = c.pw[j];
    const int TL = j == 0 ? 1 : (int)std::min<long long>(32, s);
    // a homogeneous Hamming group of 16-bit layers is sized by the round
    // tile's packed footprint; binary groups stop at 4,096 words
    // (four rounds of three axes)
    const int q = c.ax[j], rrq = round_axes_b(c, q);
    bool homog = true;
    long long P = 1;
    int k = j;
    while (k < c.n) {
      const long long P2 = P * c.ax[k];
      const bool hk = homog && c.ax[k] == q;
      const bool packed = rrq > 0 && hk;
      const long long need = packed ? round_lds(c, P2 * TL) : (long long)c.L * P2 * TL * 4;
      const bool cap = packed && q != 3 && P2 * TL > 4096;
      if (k > j && (need > B.budget || k - j >= CC_MAXSW || cap)) break;
      P = P2;
      homog = hk;
      k++;
    }
    // strided groups: widen the tile to fill the budget (a group of one
    // short axis would otherwise make thousands of 96-word tiles)
    int TLw = TL;
    if (j > 0) {
      long long tl = std::min<long long>(s, B.budget / (4LL * c.L * P));
      if (tl >= 64) tl = tl / 32 * 32;
      TLw = (int)std::max<long long>(TL, tl);
    }
    const int rr = homog && k - j >= rrq ? rrq : 0;
    // contiguous groups that stopped short of the budget carry F "passenger"
    // rows of the axes above (F divides the words above the group; the tile
    // is F consecutive groups, transformed along the group axes only), so a
    // tile is not a few hundred words
    if (j == 0 && k < c.n && getenv("CCB_NO_ROWS") == nullptr) {
      const long long rest = c.X / P;
      long long F = 1;
      for (long long f = 2; f <= rest && P * f <= 16384; f++) {
        if (rest % f) continue;
        const long long need = rr ? (k - j > rr ? round_lds(c, P * f) : 0)
                                  : (long long)c.L * P * f * 4;
        if (need <= B.budget) F = f;
      }
      P *= F;
    }
    SweepArgs a;
    memset(&a, 0, sizeof(a));
    for (int d = 0; d < c.L; d++) {
      a.A[d] = c.A[d];
      a.H[d] = c.l16 ? c.H + (size_t)d * ccb_hpitch(c.X) : nullptr;
    }
    a.cnt = c.cnt;
    a.s = s;
    a.P = (int)P;
    a.TL = TLw;
    a.ncolb = (s + TLw - 1) / TLw;
    a.ntiles = (c.X / (s * P)) * a.ncolb;
    a.nax = k - j;
    a.c0 = j;
    a.R = c.R;
    a.mu = c.mu;
    long long sk = 1;
    for (int i = 0; i < a.nax; i++) {
      a.Q[i] = c.ax[j + i];
      a.sk[i] = (int)sk;
      sk *= a.Q[i];
    }
    for (int i = 0; i <= a.nax; i++) a.live[i] = live_after(c, j + i);
    a.rr = rr;
    const bool first = j == 0, last = k == c.n;
    for (int kind = 0; kind < 3; kind++) {
      SweepArgs b = a;
      b.in_kind = first ? (kind == 2 ? CC_IN_CODE : CC_IN_PRED) : CC_IN_LAYERS;
      b.out_kind = last ? (kind == 2 ? CC_OUT_CNT : CC_OUT_MAP) : CC_OUT_LAYERS;
      b.pmode = kind == 1 ? CC_CRIT : CC_DEF;
      // uint16 planes keep layer 0 in H[0] too: the first sweep writes it
      // whatever seeded it (A[0] holds the code indicator as int32)
      b.write_l0 = c.L == 1 || (first && (kind != 2 || c.l16));
      // the gain map's last sweep folds the cell's argmax into its BCell
      b.amax = kind == 0 && last ? &B.d[ci].amax : nullptr;
      B.args[kind].push_back(b);
    }
    B.argcell.push_back(ci);
    B.need.push_back((int)a.ntiles);
    if (a.rr)   // one round: registers only
      B.lds = std::max(B.lds, a.nax > a.rr ? (size_t)round_lds(c, P * TLw) : (size_t)0);
    else
      B.lds = std::max(B.lds, (size_t)c.L * cc_gplane<CCB_PAD>((int)(P * TLw)) * 4);
    j = k;
  }
}

// the level kernel for the batch's threads per tile
const void *sweeps_kernel() {
  if (B.threads == 1024) return (const void *)k_ccb_sweeps<1024>;
  if (B.threads == 512) return (const void *)k_ccb_sweeps<512>;
  return (const void *)k_ccb_sweeps<256>;
}

// The sweep plan of every cell for an LDS budget: sweeps, work items by
// level, their device copies and the persistent grid.  Rebuilding drops a
// captured graph (its launches hold the old plan).
int build_plan(int budget) {
  if (B.st) CKB(hipStreamSynchronize(B.st));
  if (B.exec) { CKB(hipGraphExecDestroy(B.exec)); B.exec = nullptr; }
  if (B.graph) { CKB(hipGraphDestroy(B.graph)); B.graph = nullptr; }
  B.graph_k = -1;
  for (void *p : B.plan_allocs) CKB(hipFree(p));
  B.plan_allocs.clear();
  for (auto &v : B.args) v.clear();
  B.argcell.clear();
  B.need.clear();
  B.items.clear();
  B.level_off.clear();
  B.lds = 0;
  B.budget = budget;
  const int ncells = B.nc;
  std::vector<std::vector<int>> cell_args(ncells);
  for (int i = 0; i < ncells; i++) {
    const int before = (int)B.argcell.size();
    plan_cell(i, B.h[i]);
    for (int a = before; a < (int)B.argcell.size(); a++) cell_args[i].push_back(a);
  }
  B.nargs = (int)B.argcell.size();
  for (int level = 0;; level++) {
    bool any = false;
    const int off = (int)B.items.size();
    for (int i = 0; i < ncells; i++) {
      if (level >= (int)cell_args[i].size()) continue;
      any = true;
      const int a = cell_args[i][level];
      for (int t = 0; t < B.need[a]; t++) B.items.push_back(make_int2(a, t));
    }
    if (!any) break;
    // largest tiles first: dealt round-robin, a descending list gives every
    // workgroup about the same mix of tile sizes
    if (!getenv("CCB_NO_LPT")) {
      auto cost = [&](const int2 &it) {
        const SweepArgs &sa = B.args[0][it.x];
        return (long long)sa.P * sa.TL * B.h[B.argcell[it.x]].L * (sa.nax + 2);
      };
      std::stable_sort(B.items.begin() + off, B.items.end(),
                       [&](const int2 &x, const int2 &y) { return cost(x) > cost(y); });
    }
    B.level_off.push_back(off);
  }
  B.nlevels = (int)B.level_off.size();
  B.nitems = (int)B.items.size();
  B.level_off.push_back(B.nitems);
  auto palloc = [](void **p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) { (void)hipGetLastError(); return -1; }
    B.plan_allocs.push_back(*p);
    return 0;
  };
  for (int k = 0; k < 3; k++)
    if (palloc((void **)&B.dargs[k], sizeof(SweepArgs) * B.nargs) ||
        hipMemcpy(B.dargs[k], B.args[k].data(), sizeof(SweepArgs) * B.nargs,
                  hipMemcpyHostToDevice) != hipSuccess)
      return -3;
  if (palloc((void **)&B.d_argcell, 4 * B.nargs) ||
      palloc((void **)&B.d_items, 8 * B.nitems) ||
      palloc((void **)&B.d_open, 8 * B.nitems) ||
      hipMemcpy(B.d_argcell, B.argcell.data(), 4 * B.nargs, hipMemcpyHostToDevice) ||
      hipMemcpy(B.d_items, B.items.data(), 8 * B.nitems, hipMemcpyHostToDevice))
    return -3;
  const void *kf = sweeps_kernel();
  if (B.lds > 48 * 1024)   // beyond the default dynamic LDS of a launch
    CKB(hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)B.lds + 1024));
  // persistent grids: what the CUs hold at once
  int per_cu = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kf, B.threads, B.lds) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  B.per_cu = per_cu;
  return 0;
}

int reduce_launch(int with_map, bool reset);

// open_only: the levels of d_open (cells a greedy run still works on)
int transform_launch(int kind, bool open_only = false) {
  if (kind == 2) {
    const int gx = (int)std::min<long long>((B.maxX + 255) / 256, 1024);
    hipLaunchKernelGGL(k_ccb_zero_a0, dim3(gx, B.nc), dim3(256), 0, B.st, B.d);
    CKB(hipGetLastError());
    hipLaunchKernelGGL(k_ccb_scatter, dim3(64, B.nc), dim3(256), 0, B.st, B.d);
    CKB(hipGetLastError());
  }
  // one persistent launch per sweep level: at most what the CUs hold at
  // once, at most one workgroup per item
  const std::vector<int> &off = open_only ? B.open_off : B.level_off;
  int2 *items = open_only ? B.d_open : B.d_items;
  for (int l = 0; l + 1 < (int)off.size(); l++) {
    const int n = off[l + 1] - off[l];
    if (n == 0) continue;
    const int grid = std::max(1, std::min(B.per_cu * B.ncu, n));
    if (B.threads == 1024)
      hipLaunchKernelGGL(k_ccb_sweeps<1024>, dim3(grid), dim3(1024), B.lds, B.st,
                         B.dargs[kind], B.d_argcell, B.d, items + off[l], n);
    else if (B.threads == 512)
      hipLaunchKernelGGL(k_ccb_sweeps<512>, dim3(grid), dim3(512), B.lds, B.st,
                         B.dargs[kind], B.d_argcell, B.d, items + off[l], n);
    else
      hipLaunchKernelGGL(k_ccb_sweeps<256>, dim3(grid), dim3(256), B.lds, B.st,
                         B.dargs[kind], B.d_argcell, B.d, items + off[l], n);
    CKB(hipGetLastError());
  }
  return 0;
}

// d_open / open_off: the items of the cells that still have an uncovered
// word (from a fresh reduction), in the level order of the full list
int build_open_items(std::vector<char> &open) {
  if (int rc = reduce_launch(0, true)) return rc;
  CKB(hipStreamSynchronize(B.st));
  std::vector<BCell> h(B.nc);
  CKB(hipMemcpy(h.data(), B.d, sizeof(BCell) * B.nc, hipMemcpyDeviceToHost));
  open.assign(B.nc, 0);
  // (nzero counts cnt == 0: with mu > 1 a cell stays open while any word
  // may still be deficient)
  for (int i = 0; i < B.nc; i++) open[i] = h[i].nzero != 0ull || h[i].mu > 1;
  std::vector<int2> sel;
  B.open_off.assign(1, 0);
  for (int l = 0; l < B.nlevels; l++) {
    for (int i = B.level_off[l]; i < B.level_off[l + 1]; i++)
      if (open[B.argcell[B.items[i].x]]) sel.push_back(B.items[i]);
    B.open_off.push_back((int)sel.size());
  }
  if (!sel.empty())
    CKB(hipMemcpy(B.d_open, sel.data(), 8 * sel.size(), hipMemcpyHostToDevice));
  return 0;
}

// reset = false when the previous control kernel (k_ccb_ruin, k_ccb_add)
// has already cleared every cell's accumulators
int reduce_launch(int with_map, bool reset) {
  if (reset) {
    hipLaunchKernelGGL(k_ccb_reset_red, dim3((B.nc + 63) / 64), dim3(64), 0, B.st,
                       B.d, B.nc);
    CKB(hipGetLastError());
  }
  hipLaunchKernelGGL(k_ccb_reduce, dim3((unsigned)B.chunks.size()), dim3(CC_BLOCK),
                     0, B.st, B.d, B.d_chunks, with_map);
  CKB(hipGetLastError());
  return 0;
}

int snapshot_launch() {
  hipLaunchKernelGGL(k_ccb_snapshot, dim3(B.snap_grid, B.nc), dim3(CC_BLOCK), 0,
                     B.st, B.d);
  CKB(hipGetLastError());
  return 0;
}

// one batched LNS round (fixed launch sequence)
int round_launch(int k) {
  const dim3 cg((B.nc + 63) / 64), cb(64);
  hipLaunchKernelGGL(k_ccb_ruin, cg, cb, 0, B.st, B.d, B.nc, k, B.d_ruin);
  CKB(hipGetLastError());
  // only the first k of a cell's CCB_KMAX ruin slots can hold a word
  hipLaunchKernelGGL(k_ccb_walk, dim3(B.nc * k * CCB_WSPLIT), dim3(CC_BLOCK),
                     0, B.st, B.d, B.d_ruin, CCB_WSPLIT, k, CCB_KMAX);
  CKB(hipGetLastError());
  for (int j = 0; j < k; j++) {
    // the gain transform leaves each cell's argmax in amax (its map sweep
    // folds it in; ruin / add cleared it), so no reduction runs here
    if (int rc = transform_launch(0)) return rc;
    hipLaunchKernelGGL(k_ccb_add, cg, cb, 0, B.st, B.d, B.nc, j, B.d_add);
    CKB(hipGetLastError());
    hipLaunchKernelGGL(k_ccb_walk, dim3(B.nc * CCB_WSPLIT), dim3(CC_BLOCK), 0, B.st,
                       B.d, B.d_add, CCB_WSPLIT, 1, 1);
    CKB(hipGetLastError());
  }
  if (int rc = reduce_launch(0, false)) return rc;
  hipLaunchKernelGGL(k_ccb_decide, cg, cb, 0, B.st, B.d, B.nc, B.d_rev);
  CKB(hipGetLastError());
  if (int rc = snapshot_launch()) return rc;
  hipLaunchKernelGGL(k_ccb_walk, dim3(B.nc * 2 * k * CCB_WSPLIT),
                     dim3(CC_BLOCK), 0, B.st, B.d, B.d_rev, CCB_WSPLIT,
                     2 * k, 2 * CCB_KMAX);
  CKB(hipGetLastError());
  return 0;
}

int sync_batch() {
  CKB(hipStreamSynchronize(B.st));
  return 0;
}

}  // namespace

#pragma GCC visibility push(default)
extern "C" {

// ccb_init: ncells cells; cell i has problem[i] (plugin ids: 0 Hamming,
// 1 torus L-inf, 2 Lee, 3 open grid), n[i] axes (axes concatenated), radius
// R[i] <= 3, and for Hamming / Lee its support table (npats[i] patterns of
// format fmt[i]: 0 v1, 1 wide; concatenated in pats).  cap: code words per
// cell.  Returns 0, -1 (invalid), -3 (allocation).
int ccb_init(int ncells, const int *problem, const int *n, const int *axes,
             const int *R, const uint64_t *pats, const long long *npats,
             const int *fmt, int cap) {
  if (ncells < 1 || ncells > 4096 || !problem || !n || !axes || !R || !npats ||
      !fmt || cap < 2 || cap > 65535)
    return -1;
  if (B.live) bfree_all();
  B.nc = ncells;
  B.h.assign(ncells, BCell());
  int aoff = 0;
  long long poff = 0;
  for (int i = 0; i < ncells; i++) {
    BCell &c = B.h[i];
    memset(&c, 0, sizeof(c));
    c.prob = problem[i];
    c.n = n[i];
    c.R = R[i];
    c.mu = 1;
    c.fmt = fmt[i];
    if (c.n < 1 || c.n > CC_MAXN || c.R < 1 || c.R > 3 || c.prob < 0 || c.prob > 3)
      { bfree_all(); return -1; }
    const bool linf = c.prob == CC_TORUS || c.prob == CC_GRID;
    c.L = linf ? 1 : c.R + 1;
    c.pw[0] = 1;
    for (int j = 0; j < c.n; j++) {
      c.ax[j] = axes[aoff + j];
      const int lim = linf ? CC_MAXAX_X : (c.prob == CC_HAMMING ? 10 : 16);
      if (c.ax[j] < 2 || c.ax[j] > lim || (c.prob == CC_TORUS && c.ax[j] < 2 * c.R + 1))
        { bfree_all(); return -1; }
      c.pw[j + 1] = c.pw[j] * c.ax[j];
      if (c.pw[j + 1] > (1LL << 26)) { bfree_all(); return -1; }
    }
    if ((linf && c.n > 8) || (c.fmt == CC_FMT_V1 && c.n > CC_MAXN_V1) ||
        (c.fmt != CC_FMT_V1 && c.fmt != CC_FMT_WIDE))
      { bfree_all(); return -1; }
    aoff += c.n;
    c.X = c.pw[c.n];
    c.blen = linf ? 0 : npats[i];
    if (!linf && (c.blen < 1 || c.blen > INT32_MAX || !pats)) { bfree_all(); return -1; }
    // 16-bit inter-sweep layers when every value fits: a layer value is a
    // count of seed words at one partial distance, at most the ball size for
    // the gain / loss indicators and the code size (<= cap) for a recount;
    // the L-inf cells also need every axis on the unrolled window path
    {
      double ball = 1.0;
      if (linf) {
        for (int j = 0; j < c.n; j++) ball *= std::min(c.ax[j], 2 * c.R + 1);
      } else {
        std::vector<double> poly(c.R + 1, 0.0);
        poly[0] = 1.0;
        for (int j = 0; j < c.n; j++) {
          std::vector<double> nx(c.R + 1, 0.0);
          for (int w = 0; w <= c.R; w++) {
            if (poly[w] == 0) continue;
            nx[w] += poly[w];
            for (int dl = 1; dl < c.ax[j]; dl++) {
              const int e = c.prob == CC_LEE ? std::min(dl, c.ax[j] - dl) : 1;
              if (w + e <= c.R) nx[w + e] += poly[w];
            }
          }
          poly = nx;
        }
        ball = 0;
        for (double v : poly) ball += v;
      }
      bool small_axes = true;
      for (int j = 0; j < c.n; j++) small_axes = small_axes && c.ax[j] <= 16;
      c.l16 = ball <= 65535.0 && (!linf || small_axes) && !getenv("CCB_INT32");
    }
    // device planes
    for (int d = 0; d < c.L; d++)
      if (balloc(&c.A[d], (size_t)c.X * 4)) { bfree_all(); return -3; }
    if (c.l16 && balloc(&c.H, (size_t)c.L * ccb_hpitch(c.X) * 2)) { bfree_all(); return -3; }
    if (balloc(&c.cnt, (size_t)((c.X + 1) & ~1LL) * 2) ||
        balloc(&c.code, (size_t)cap * 8) || balloc(&c.best, (size_t)cap * 8))
      { bfree_all(); return -3; }
    c.cap = cap;
    if (!linf) {
      uint64_t *dp = nullptr;
      if (balloc(&dp, (size_t)c.blen * 8)) { bfree_all(); return -3; }
      if (hipMemcpy(dp, pats + poff, (size_t)c.blen * 8, hipMemcpyHostToDevice) !=
          hipSuccess) { bfree_all(); return -3; }
      c.pat = dp;
      poff += c.blen;
    }
    if (hipMemset(c.cnt, 0, (size_t)((c.X + 1) & ~1LL) * 2) != hipSuccess)
      { bfree_all(); return -3; }
    B.maxX = std::max(B.maxX, c.X);
  }
  if (hipStreamCreateWithFlags(&B.st, hipStreamNonBlocking) != hipSuccess)
    { bfree_all(); return -3; }
  for (int i = 0; i < ncells; i++)
    for (long long w = 0; w < B.h[i].X; w += CCB_CHUNK)
      B.chunks.push_back(((long long)i << 40) | w);
  B.snap_grid = (cap + 4 * CC_BLOCK - 1) / (4 * CC_BLOCK);
  if (balloc(&B.d, sizeof(BCell) * ncells) ||
      balloc(&B.d_chunks, 8 * B.chunks.size()) ||
      balloc(&B.d_ruin, sizeof(BItem) * ncells * CCB_KMAX) ||
      balloc(&B.d_add, sizeof(BItem) * ncells) ||
      balloc(&B.d_rev, sizeof(BItem) * ncells * 2 * CCB_KMAX))
    { bfree_all(); return -3; }
  if (hipMemcpy(B.d, B.h.data(), sizeof(BCell) * ncells, hipMemcpyHostToDevice) ||
      hipMemcpy(B.d_chunks, B.chunks.data(), 8 * B.chunks.size(),
                hipMemcpyHostToDevice))
    { bfree_all(); return -3; }
  {
    int dev = 0, ncu = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    B.ncu = std::max(ncu, 1);
  }
  if (const char *nt = getenv("CCB_THREADS")) {
    const int t = atoi(nt);
    B.threads = t == 1024 ? 1024 : (t == 256 ? 256 : 512);
  }
  // CCB_LDS (bytes) overrides the initial LDS budget per tile, within
  // [4 KiB, 64 KiB]; the autotuner may replace it
  int budget0 = CCB_LDS;
  if (const char *lb = getenv("CCB_LDS"))
    budget0 = std::min(128 * 1024, std::max(4096, atoi(lb)));
  if (build_plan(budget0)) { bfree_all(); return -3; }
  // the set-up copies and clears ran outside the batch's (non-blocking)
  // stream: finish them before its first launch
  if (hipDeviceSynchronize() != hipSuccess) { bfree_all(); return -3; }
  B.live = true;
  return 0;
}

int ccb_free(void) {
  bfree_all();
  return 0;
}

// the code of one cell (a set or multiset of word indices, M <= cap)
int ccb_set_code(int cell, const long long *idx, int M) {
  if (!B.live || cell < 0 || cell >= B.nc || M < 0 || M > B.h[cell].cap ||
      (M && !idx))
    return -1;
  for (int i = 0; i < M; i++)
    if (idx[i] < 0 || idx[i] >= B.h[cell].X) return -1;
  CKB(hipStreamSynchronize(B.st));
  if (M) CKB(hipMemcpy(B.h[cell].code, idx, (size_t)M * 8, hipMemcpyHostToDevice));
  CKB(hipMemcpy(&B.d[cell].M, &M, 4, hipMemcpyHostToDevice));
  return 0;
}

// the transform of every cell in one persistent launch: init_mode / out_mode
// as dct_transform ((1,1) gain, (2,1) loss, (0,2) recount)
int ccb_transform(int init_mode, int out_mode) {
  if (!B.live) return -1;
  int kind;
  if (init_mode == 1 && out_mode == 1) kind = 0;
  else if (init_mode == 2 && out_mode == 1) kind = 1;
  else if (init_mode == 0 && out_mode == 2) kind = 2;
  else return -1;
  if (int rc = transform_launch(kind)) return rc;
  return sync_batch();
}

int ccb_read_cnt(int cell, uint16_t *out) {
  if (!B.live || cell < 0 || cell >= B.nc || !out) return -1;
  CKB(hipStreamSynchronize(B.st));
  CKB(hipMemcpy(out, B.h[cell].cnt, (size_t)B.h[cell].X * 2, hipMemcpyDeviceToHost));
  return 0;
}

int ccb_read_map(int cell, int32_t *out) {
  if (!B.live || cell < 0 || cell >= B.nc || !out) return -1;
  CKB(hipStreamSynchronize(B.st));
  CKB(hipMemcpy(out, B.h[cell].A[0], (size_t)B.h[cell].X * 4, hipMemcpyDeviceToHost));
  return 0;
}

// LNS state from the current cnt of every cell (call after a recount):
// uncovered counts, per-cell random streams from seed, best = start if
// feasible
int ccb_lns_init(unsigned long long seed) {
  if (!B.live) return -1;
  if (int rc = reduce_launch(0, true)) return rc;
  hipLaunchKernelGGL(k_ccb_lns_init, dim3((B.nc + 63) / 64), dim3(64), 0, B.st,
                     B.d, B.nc, seed);
  CKB(hipGetLastError());
  if (int rc = snapshot_launch()) return rc;
  CKB(hipStreamSynchronize(B.st));
  return 0;
}

// rounds batched LNS rounds, k words ruined per cell per round (k <=
// CCB_KMAX); graph != 0 replays one captured round, else launches the same
// sequence directly.  *ms: wall milliseconds of the rounds.
int ccb_lns_run(int rounds, int k, int graph, float *ms) {
  if (!B.live || rounds < 0 || k < 1 || k > CCB_KMAX) return -1;
  hipEvent_t e0, e1;
  CKB(hipEventCreate(&e0));
  CKB(hipEventCreate(&e1));
  if (graph && (!B.exec || B.graph_k != k)) {
    if (B.exec) { CKB(hipGraphExecDestroy(B.exec)); B.exec = nullptr; }
    if (B.graph) { CKB(hipGraphDestroy(B.graph)); B.graph = nullptr; }
    CKB(hipStreamBeginCapture(B.st, hipStreamCaptureModeThreadLocal));
    const int rc = round_launch(k);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(B.st, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    CKB(ec);
    B.graph = g;
    CKB(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
    B.graph_k = k;
  }
  CKB(hipEventRecord(e0, B.st));
  for (int r = 0; r < rounds; r++) {
    if (graph) CKB(hipGraphLaunch(B.exec, B.st));
    else if (int rc = round_launch(k)) return rc;
  }
  CKB(hipEventRecord(e1, B.st));
  CKB(hipEventSynchronize(e1));
  float t = 0;
  CKB(hipEventElapsedTime(&t, e0, e1));
  if (ms) *ms = t;
  CKB(hipEventDestroy(e0));
  CKB(hipEventDestroy(e1));
  return 0;
}

// steps greedy steps on every cell: gain transform, argmax, add (cells with
// no uncovered word left are unchanged), cnt kept exact by walks.  Start
// from a recounted code (ccb_transform(0, 2)); returns the number of cells
// still uncovered after the steps (>= 0).
int ccb_greedy(int steps) {
  if (!B.live || steps < 0) return -1;
  // only the cells that still have an uncovered word are transformed (a
  // covered cell has no positive gain, so its step would add nothing)
  std::vector<char> todo;
  if (int rc = build_open_items(todo)) return rc;
  // each step's gain transform folds the argmax into amax (cleared here
  // once, then by every k_ccb_add), so a step launches no reduction
  hipLaunchKernelGGL(k_ccb_reset_red, dim3((B.nc + 63) / 64), dim3(64), 0, B.st,
                     B.d, B.nc);
  CKB(hipGetLastError());
  for (int s = 0; s < steps; s++) {
    if (int rc = transform_launch(0, true)) return rc;
    hipLaunchKernelGGL(k_ccb_add, dim3((B.nc + 63) / 64), dim3(64), 0, B.st, B.d,
                       B.nc, -1, B.d_add);
    CKB(hipGetLastError());
    hipLaunchKernelGGL(k_ccb_walk, dim3(B.nc * CCB_WSPLIT), dim3(CC_BLOCK), 0, B.st,
                       B.d, B.d_add, CCB_WSPLIT, 1, 1);
    CKB(hipGetLastError());
  }
  if (int rc = reduce_launch(0, true)) return rc;
  if (int rc = sync_batch()) return rc;
  std::vector<BCell> h(B.nc);
  CKB(hipMemcpy(h.data(), B.d, sizeof(BCell) * B.nc, hipMemcpyDeviceToHost));
  int open = 0;
  for (const BCell &c : h) open += c.nzero != 0ull;
  return open;
}

// which 0: the current code, 1: the best feasible code; returns its size
int ccb_get_code(int cell, int which, long long *out, int cap) {
  if (!B.live || cell < 0 || cell >= B.nc || which < 0 || which > 1) return -1;
  CKB(hipStreamSynchronize(B.st));
  BCell c;
  CKB(hipMemcpy(&c, &B.d[cell], sizeof(BCell), hipMemcpyDeviceToHost));
  const int m = which ? (c.bestM == 0x7FFFFFFF ? 0 : c.bestM) : c.M;
  const int mc = std::min(m, std::max(cap, 0));
  if (mc && !out) return -1;
  if (mc) CKB(hipMemcpy(out, which ? c.best : c.code, (size_t)mc * 8,
                        hipMemcpyDeviceToHost));
  return m;
}

// out[0..5] = M, uncovered, best feasible size (-1: none), rounds,
// accepted, reverted
int ccb_stats(int cell, long long *out) {
  if (!B.live || cell < 0 || cell >= B.nc || !out) return -1;
  CKB(hipStreamSynchronize(B.st));
  BCell c;
  CKB(hipMemcpy(&c, &B.d[cell], sizeof(BCell), hipMemcpyDeviceToHost));
  out[0] = c.M;
  out[1] = c.uncov;
  out[2] = c.bestM == 0x7FFFFFFF ? -1 : c.bestM;
  out[3] = c.rounds;
  out[4] = c.accepted;
  out[5] = c.reverted;
  return 0;
}

// schedule of the batch: out[0..5] = cells, sweeps, work items, sweep
// levels (persistent launches per transform), workgroups per launch at
// most, LDS budget of the plan (bytes per tile)
int ccb_info(long long *out) {
  if (!B.live || !out) return -1;
  out[0] = B.nc;
  out[1] = B.nargs;
  out[2] = B.nitems;
  out[3] = B.nlevels;
  out[4] = (long long)B.per_cu * B.ncu;
  out[5] = B.budget;
  out[6] = B.threads;
  return 0;
}

// Time the batched gain transform under each LDS budget (8 to 128 KiB per
// tile: the axis grouping, tile widths and persistent grid follow from it)
// and each workgroup size of the level launches (256, 512, 1,024 threads
// on one tile: more threads shorten the longest tiles, which set a level's
// length) and keep the fastest plan.  reps timed transforms per plan after
// one warm-up; the map planes are overwritten, cnt is not.  Returns the
// chosen budget (bytes); *best_us = its time per transform.
int ccb_autotune(int reps, float *best_us) {
  if (!B.live || reps < 1) return -1;
  hipEvent_t e0, e1;
  CKB(hipEventCreate(&e0));
  CKB(hipEventCreate(&e1));
  float best = 1e30f;
  int win = B.budget, wthreads = B.threads;
  int dev = 0, lmax = 64 * 1024;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  for (int threads : {256, 512, 1024})
  for (int budget : {8192, 16384, 32768, 65536, 131072}) {
    if (budget + 1024 > lmax) continue;
    B.threads = threads;
    if (int rc = build_plan(budget)) return rc;
    if (int rc = transform_launch(0)) return rc;
    CKB(hipEventRecord(e0, B.st));
    for (int r = 0; r < reps; r++)
      if (int rc = transform_launch(0)) return rc;
    CKB(hipEventRecord(e1, B.st));
    CKB(hipEventSynchronize(e1));
    float ms = 0;
    CKB(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) { best = ms; win = budget; wthreads = threads; }
  }
  CKB(hipEventDestroy(e0));
  CKB(hipEventDestroy(e1));
  B.threads = wthreads;
  if (int rc = build_plan(win)) return rc;
  if (best_us) *best_us = best * 1000.0f / reps;
  return win;
}

}  // extern "C"
#pragma GCC visibility pop

// dct_engine.hip -- the plugin ABI v1 (`dct_*`) over HIP on MI355X (gfx950).
//
// Drop-in for the reference's libcoldcore_covering.so
// (reference src/plugins/covering/dct_hc.cu, contract src/include/coldcore/
// plugin.h:62-138 and docs/PLUGIN.md:71-160): same flat symbols, argument
// meaning, return codes (0 ok, -1 invalid / rejected, -2 space > 2^40,
// -3/-4/-5 allocation, -(100 + hipError_t) device error with a message on
// stderr) and validation-before-mutation rules.  One cell per process
// (singleton); every device operation runs in call order on one stream
// owned by the engine; calls that return data synchronise once; small
// ball_update batches return without waiting.
//
// Everything is HBM-resident (288 GB per MI355X): the layers_mode/cnt_mode
// placement arguments are accepted for ABI compatibility and all planes go
// to device memory, so dct_host_bytes() is 0.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "cc_common.hpp"
#include "cc_reduce.hpp"
#include "cc_launch.hpp"
#include "cc_walk.hpp"
#include "cc_psweep.hpp"
#include "cc_setside.hpp"

#define CK(x)                                                               \
  do {                                                                      \
    hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) {                                                 \
      fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_),     \
              __FILE__, __LINE__);                                          \
      return -(100 + (int)e_);                                              \
    }                                                                       \
  } while (0)

extern "C" int dct_autotune(int reps, float *best_us);

namespace {

constexpr int MAX_CODE = 65535;        // live occurrences (uint16 exactness)
constexpr int CODECAP = 1 << 16;       // staging words per chunk
constexpr int ASYNC_UPDATE_MAX = 1024; // larger update batches synchronise
constexpr int NV_WORDS = 128;          // words per by-value launch
constexpr long long EXTRACT_CH = 1 << 14;  // words per extraction chunk
// swap evaluation staging: candidates and drops (<= CODECAP each) with the
// candidate -> drop map in, gains, losses and the list count out
constexpr size_t SWAP_IN_BYTES = (size_t)CODECAP * (8 + 8 + 4);
constexpr size_t SWAP_RES_BYTES = (size_t)CODECAP * (4 + 4) + 16;

struct Sweep {
  int j0, nax, c0, P, TL;
  long long s, ncolb, ntiles;
  int lmax;
  int rq;       // round kernel: axis size Q of the homogeneous group, else 0
  int threads;  // round kernel: workgroup size (one cube per thread)
  size_t lds;   // LDS bytes with int32 layers
  size_t lds16; // LDS bytes in 16-bit mode (round kernel packs layer pairs)
};

// One sweep of the packed-layer path (cc_psweep.hpp).
struct PSweep {
  int j0, nax, RR, W, ND, P, TL, threads;
  int kern;  // 0 k_psweep, 1 / 2 k_pcol (lanes per cube), 3 k_pfirst,
             // 4 k_pfirstp, 5 k_pfirstw, 6 k_pstrip
  long long s, ncolb, ntiles;
  int nrounds;
  int ca[CC_PK_MAXRD], amask[CC_PK_MAXRD], qa[CC_PK_MAXRD];
  int in_nd, out_nd;  // compact planes read / written (0: seed in, sum out)
  int in_w8, out_w8;  // compact format is one byte per layer
  int in_nar, out_nar;  // narrow compact word (1: 16 bits, 2: 24 bits)
  int lin[CC_PK_MAXRD], kp[CC_PK_MAXRD];  // linear LDS addressing per round
  int idw[CC_MAXL], ioff[CC_MAXL], ibits[CC_MAXL];
  int odw[CC_MAXL], ooff[CC_MAXL], obits[CC_MAXL];
  size_t lds;
  int in_bytes, out_bytes;  // algorithmic bytes per word
  double cost;              // estimated relative time per word
  int na;                   // k_pfirstw: axes of the widening last round
  int mid_w8;               // k_pstrip: LDS exchange format (one dword)
  int lcmp;                 // k_psweep, two rounds: LDS words in that format
  int moff[CC_MAXL], mbits[CC_MAXL];
};

struct PPlan {
  bool ok = false;
  std::vector<PSweep> sw;
  long long bytes = 0;  // algorithmic bytes per word of one transform
};

// A code multiset kept as its distinct words ascending with their
// multiplicities (what a sparse seed stages) and, built on first use, a hash
// map (what update validation looks words up in).  Changes go through the
// map, which then becomes the only valid form.
struct Multiset {
  std::vector<long long> w;
  std::vector<int> m;
  std::unordered_map<long long, int> map;
  bool vec_ok = true, map_ok = true;
  int maxm = 0;   // largest multiplicity (valid with vec_ok)

  void clear() {
    w.clear(); m.clear(); map.clear();
    vec_ok = map_ok = true;
    maxm = 0;
  }
  // from an unsorted word list with repeats
  void assign(const long long *idx, int M) {
    w.assign(idx, idx + M);
    std::sort(w.begin(), w.end());
    m.clear();
    int k = 0;
    for (int i = 0; i < M; i++) {
      if (k && w[k - 1] == w[i]) { m[k - 1]++; continue; }
      w[k++] = w[i];
      m.push_back(1);
    }
    w.resize(k);
    maxm = 0;
    for (int c : m) maxm = std::max(maxm, c);
    map.clear();
    vec_ok = true;
    map_ok = false;
  }
  const std::unordered_map<long long, int> &get() {
    if (!map_ok) {
      map.clear();
      map.reserve(w.size());
      for (size_t i = 0; i < w.size(); i++) map.emplace(w[i], m[i]);
      map_ok = true;
    }
    return map;
  }
  std::unordered_map<long long, int> &mut() {
    get();
    vec_ok = false;
    return map;
  }
};

struct Engine {
  bool live = false;
  int prob = 0, n = 0, R = 0, L = 1, mu = 1;
  int ax[CC_MAXN] = {};
  long long pw[CC_MAXN + 1] = {};
  long long X = 0;
  int32_t *A[CC_MAXL] = {};   // A[0] map plane; A[1..R] views of `lay`
  uint16_t *H[CC_MAXL] = {};  // uint16 views of `lay` (16-bit mode)
  int32_t *lay = nullptr;     // one block for layers 1..R (pitch-aligned)
  long long pitch = 0;        // plane pitch in words (multiple of 64)
  int l16_field = 0;          // gain / loss run 16-bit inter-sweep planes
  int l16_count = 0;          // recount runs 16-bit inter-sweep planes
  int layers_live = 0;        // a layers transform left exact layers (1: int32
                              // planes A, 2: uint16 planes H)
  uint16_t *cnt = nullptr;
  long long cnt_len = 0;
  uint64_t *pat = nullptr, *pat2 = nullptr;
  long long blen = 0;
  int tfmt = CC_FMT_V1;  // packed format of the init table
  int fmt = 0, use_owner = 0;
  long long *d_code = nullptr;
  int codeM = 0;
  // sparse code seeds (recount without a code plane): the pending code's
  // distinct words ascending, their multiplicities, and the entry offsets
  // of every first-sweep tile
  long long *d_cw = nullptr;   // CODECAP words, then CODECAP multiplicities
  int32_t *d_cm = nullptr;     // the multiplicities of the last staging (in d_cw)
  int32_t *d_toff = nullptr;
  long long toff_cap = 0;
  int32_t *toff_cur = nullptr;   // the tile offsets of the last staging
  int sparse = 1;            // CC_SPARSE_CODE=0: code seeds through the A0 plane
  long long *d_words = nullptr;
  int32_t *d_out = nullptr;
  unsigned long long *d_scal = nullptr;  // [0] reductions [1] transitions
  int32_t *d_hist = nullptr;
  int64_t *d_eidx64 = nullptr;
  uint32_t *d_eidx32 = nullptr;
  int32_t *d_eval = nullptr;
  int ecap = 0;
  long long *d_bcnt = nullptr, *d_boff = nullptr;
  long long nchunks = 0;
  long long *h_words = nullptr;
  int32_t *h_out = nullptr;
  int32_t *h_out_dev = nullptr;  // h_out as the device addresses it
  unsigned long long *h_scal = nullptr;
  // census slot regions (cc_census_flush): CEN_* regions of CC_NSLOT
  // uint64 slots, cen_stride words apart (CC_SLOT_STRIDE), read back into
  // h_cen
  unsigned long long *d_cen = nullptr;
  unsigned long long *h_cen = nullptr;
  int cen_stride = 32;
  hipStream_t st = nullptr;
  long long bytes_dev = 0;
  long long stats[8] = {};
  unsigned long long trans_last = 0;
  // the pending code (set_code) and the represented one (cnt's code)
  Multiset uploaded, active;
  std::vector<long long> code_raw;   // the pending list as given (repeats kept)
  bool code_dev_ok = false;          // d_code holds code_raw
  // after a recount the represented multiset equals the pending one: the
  // flag stands for that instead of a copy (active_map / active_mut)
  bool active_shared = false;
  int active_size = 0;
  // the code seed's packed plan, kept while its inputs are unchanged
  bool cp_ok = false;
  double cp_s0 = 0, cp_cap = 0;
  int cp_opts[8] = {};
  PPlan cp_plan;
  std::vector<Sweep> plan;
  double ball = 0;           // exact metric ball size (closed form)
  int lds_budget = 64 * 1024;
  int lds_max = 64 * 1024;   // device limit per workgroup
  int ncu = 256;             // compute units
  int grid = -1;             // CC_SWEEP_GRID: workgroups per CU (0 = per tile)
  int generic_only = 0;      // CC_SWEEP_KERNEL=generic: no round kernel
  int packed = 1;            // CC_PACKED=0: never take the packed path
  int pk_col = 2;            // CC_PK_COL: column kernel 0 off, 1 whole, 2 split
  int autotune = 1;          // CC_AUTOTUNE=0: never time schedules by itself
  int tuned = 0;             // the schedules of this cell were timed
  int pk_narrow = 0;         // CC_PK_NARROW=1: 16 / 24-bit compact words (fewer
                             // bytes, but the sub-dword accesses measured
                             // slower on K_2(24,3): off by default)
  int compact = 1;           // CC_COMPACT=0: unpacked sweeps exchange int32 /
                             // uint16 layer planes instead of compact words
  int pk_nt = ~1;            // CC_PK_NT: bit i = sweep i stores non-temporal
  int pk_wide = 0;           // CC_PK_WIDE: the planner may take k_pfirstw /
                             // k_pstrip sweeps
  int pk_lcmp = 0;           // CC_PK_LCMP=1: two-round LDS sweeps exchange one
                             // compact dword per word (when it holds one;
                             // half the LDS, but the packing measured slower)
  int pk_first = 1;          // CC_PK_FIRST: first sweep 0 k_psweep, 1 persistent
                             // with prefetch, 2 LDS-free (cross-lane)
  PPlan pfield;              // packed plan of gain / loss (predicate seeds)

  // set-side gathers: exact when the walked table is exactly the metric
  // ball (proven at init / dct_set_tbl2); Z cached per cnt state
  int set_policy = 1;        // CC_SET_POLICY: 0 never, 1 auto, 2 always
  int set_ratio = 2;         // CC_SET_RATIO: list Z only while |Z| <= ratio |B|
  long long set_min = -1;    // CC_SET_MIN: walk reads from which the set side
                             // is tried (-1: max(|X| / 8, 2^21))
  int set_ok1 = 0, set_ok2 = 0;  // init table / ox table proven to be the ball
  int dbits = 4;             // bits per packed digit
  int region_walk = 0;       // region layers by ball walks: the table is the
                             // Hamming ball and digits pack into 64 bits
  int region_walk_mode = 1;  // CC_REGION_WALK: 0 never, 1 when cheaper, 2 always
  uint64_t *zbuf = nullptr;
  long long zcap = 0, zn = 0;
  int z_mode = -1, z_target = 0, z_mu = 0;
  unsigned long long z_ver = 0;
  int o_mode = -1, o_target = 0, o_mu = 0;   // last listing that overflowed
  long long o_F = 0, o_plus = 0;
  unsigned long long cnt_ver = 1;            // bumped by every cnt change
  // census folded into the last packed sweep (dct_set_fold, CC_FOLD=1 at
  // init): region CEN_FMAP holds max(0, max A0) while map_fold, CEN_FDEF
  // the cnt deficit of cnt state fold_cnt_ver
  int fold = 0;
  int scan_blocks = 2048;    // CC_SCAN_BLOCKS: workgroups of a field scan
                             // (2,048 measured 2.7 us faster per scan than
                             // 8,192 on 2^24 words; 1,024 no better)
  bool map_fold = false;
  bool cnt_fold_pending = false;
  unsigned long long fold_cnt_ver = 0;
  long long plus_words = 0;                  // words placed by +1 updates
  // swap evaluation fast path: drop lists (packed digits, drop index)
  uint64_t *lbuf = nullptr;
  int32_t *ldrop = nullptr;
  long long lcap = 0;
  // one packed upload [cand nc x 8][drops k x 8][cand -> drop nc x 4] and
  // one result region [gain nc x 4][loss k x 4][pad][list count 8]
  char *d_swap = nullptr, *h_swap = nullptr, *d_sres = nullptr, *h_sres = nullptr;
  // optional per-sweep timing of the last transform (additive ABI)
  int prof = 0;
  hipEvent_t ev[2 * CC_MAXN] = {};
  float sweep_ms[CC_MAXN] = {};
  int nswept = 0;
};

Engine E;

// the represented multiset, for lookups / for changes (materialised from
// the pending one on the first change after a recount)
const std::unordered_map<long long, int> &active_map() {
  return E.active_shared ? E.uploaded.get() : E.active.get();
}
std::unordered_map<long long, int> &active_mut() {
  if (E.active_shared) {
    E.active = E.uploaded;
    E.active_shared = false;
  }
  return E.active.mut();
}

// Forced sweep boundaries of the packed schedules (dct_sweep_force): bit k
// set means a sweep ends after axis k - 1 (0 < k < n).  0: the planner
// chooses.  Cleared by init and free.
uint64_t g_cuts = 0;

bool set_ok() { return E.fmt == 1 ? E.set_ok2 : E.set_ok1; }

Dig make_dig() {
  Dig d;
  memset(&d, 0, sizeof(d));
  for (int i = 0; i < E.n; i++) d.ax[i] = E.ax[i];
  d.n = E.n;
  d.b = E.dbits;
  d.R = E.R;
  d.mu = E.mu;
  d.lsb = 0;
  for (int i = 0; i < E.n; i++) d.lsb |= 1ull << (E.dbits * i);
  return d;
}

Geom make_geom() {
  Geom g;
  memset(&g, 0, sizeof(g));
  for (int i = 0; i < E.n; i++) {
    g.pw[i] = E.pw[i];
    g.ax[i] = E.ax[i];
  }
  g.n = E.n;
  g.R = E.R;
  g.mu = E.mu;
  return g;
}

bool valid_words(const long long *w, int n) {
  if (n < 0 || n > MAX_CODE || (n && !w)) return false;
  for (int i = 0; i < n; i++)
    if (w[i] < 0 || w[i] >= E.X) return false;
  return true;
}

// Slabs of a field scan (k_cnt_census, k_map_peak, k_map_hist): whole
// multiples of 8 x CC_BLOCK words, at most 2^17 words (512 per thread, so a
// wave's 32-bit deficit total stays below 2^32), at least N / maxb words so
// the grid stays within maxb workgroups where that bound allows (maxb <= 0:
// the engine's scan_blocks).
struct Slabs { unsigned blocks; long long slab; };
Slabs scan_slabs(long long N, long long maxb = 0, long long cap = 1LL << 17) {
  if (maxb <= 0) maxb = E.scan_blocks;
  const long long unit = 8LL * CC_BLOCK;
  long long slab = ((N + maxb - 1) / maxb + unit - 1) / unit * unit;
  slab = std::max(unit, std::min(slab, cap));
  return Slabs{(unsigned)std::max(1LL, (N + slab - 1) / slab), slab};
}

int pass_blocks(long long items) {
  long long b = (items + CC_BLOCK * 8LL - 1) / (CC_BLOCK * 8LL);
  return (int)std::max(1LL, std::min(b, 8192LL));
}

// Split blocks per word so a single-word walk still spreads over the chip
// (reference dct_hc.cu:523-531): min(8192/nwords, ceil(|B|/256), 65535).
int pick_split(int nwords) {
  long long cover = (E.blen + CC_BLOCK - 1) / CC_BLOCK;
  long long split = 8192 / std::max(nwords, 1);
  split = std::max(1LL, std::min(split, cover));
  return (int)std::min(split, 65535LL);
}

// A v1 table is walked without per-entry bounds checks, so it is proven
// well-formed first: positions < n and distinct, 1 <= delta < ax_pos, zero
// byte terminates.  Any such pattern keeps every walked index inside X.
// field geometry of a table format: field bits, fields per pattern, and
// the shift / mask that split a field into (position, delta)
struct TblFmt {
  int fb, nf, ps;
  unsigned dm;
};
TblFmt tbl_fmt(int fmt) {
  if (fmt == CC_FMT_V1) return {8, CC_MAXPAT, 4, 15u};
  if (fmt == CC_FMT_WIDE) return {10, CC_MAXPAT_W, 4, 15u};
  return {16, CC_MAXPAT_X, 10, 1023u};
}

bool table_ok(const uint64_t *pat, long long len, int fmt) {
  const TblFmt tf = tbl_fmt(fmt);
  const int fb = tf.fb, nf = tf.nf;
  const uint64_t fm = fb == 64 ? ~0ull : (1ull << fb) - 1ull;
  for (long long t = 0; t < len; t++) {
    const uint64_t p = pat[t];
    if (fb * nf < 64 && (p >> (fb * nf))) return false;
    uint64_t seen = 0;
    bool ended = false;
    for (int k = 0; k < nf; k++) {
      const unsigned b = (unsigned)((p >> (fb * k)) & fm);
      if (!b) { ended = true; continue; }
      if (ended) return false;
      const int pos = (int)(b >> tf.ps), d = (int)(b & tf.dm);
      if (pos >= E.n || ((seen >> pos) & 1ull) || d < 1 || d >= E.ax[pos])
        return false;
      seen |= 1ull << pos;
    }
  }
  return true;
}

// Is a (well-formed) table exactly the metric ball of radius R?  Every
// pattern within the radius (Hamming: one per moved position; Lee: the
// circular distances sum; torus / grid: every circular distance <= R), no
// two patterns equal (fields compared in position order), and the count
// equals the ball's size (the coefficients 0..R of the per-axis weight
// enumerator product; a single coefficient for the L-inf families).  Only
// then does a set-side count equal the walk's.
bool table_is_ball(const uint64_t *pat, long long len, int fmt) {
  const TblFmt tf = tbl_fmt(fmt);
  const int fb = tf.fb, nf = tf.nf;
  const uint64_t fm = (1ull << fb) - 1ull;
  const bool linf = E.prob == CC_TORUS || E.prob == CC_GRID;
  auto wt = [&](int pos, int d) -> int {   // -1: outside the radius
    const int a = E.ax[pos], c = std::min(d, a - d);
    if (E.prob == CC_HAMMING) return 1;
    if (E.prob == CC_LEE) return c <= E.R ? c : -1;
    return c <= E.R ? 0 : -1;
  };
  std::vector<uint64_t> keys((size_t)len);
  for (long long t = 0; t < len; t++) {
    const uint64_t p = pat[t];
    uint64_t f[CC_MAXPAT] = {};
    int k = 0, w = 0;
    for (int i = 0; i < nf; i++) {
      const unsigned b = (unsigned)((p >> (fb * i)) & fm);
      if (!b) break;
      const int e = wt((int)(b >> tf.ps), (int)(b & tf.dm));
      if (e < 0) return false;
      w += e;
      f[k++] = b;
    }
    if (w > E.R) return false;
    std::sort(f, f + k);   // position-major: pos << shift | delta
    uint64_t key = 0;
    for (int i = 0; i < k; i++) key |= f[i] << (fb * i);
    keys[(size_t)t] = key;
  }
  std::sort(keys.begin(), keys.end());
  if (std::adjacent_find(keys.begin(), keys.end()) != keys.end()) return false;
  const int W = linf ? 0 : E.R;
  std::vector<unsigned long long> poly(W + 1, 0);
  poly[0] = 1;
  const unsigned long long cap = 1ull << 62;
  for (int i = 0; i < E.n; i++) {
    std::vector<unsigned long long> nx(W + 1, 0);
    for (int w = 0; w <= W; w++) {
      if (!poly[w]) continue;
      nx[w] = std::min(cap, nx[w] + poly[w]);
      for (int d = 1; d < E.ax[i]; d++) {
        const int e = wt(i, d);
        if (e >= 0 && w + e <= W) nx[w + e] = std::min(cap, nx[w + e] + poly[w]);
      }
    }
    poly = nx;
  }
  unsigned long long tot = 0;
  for (int w = 0; w <= W; w++) tot = std::min(cap, tot + poly[w]);
  return tot == (unsigned long long)len;
}

// ---- sweep schedule ------------------------------------------------------
// Greedy grouping of consecutive axes.  The first group starts at axis 0,
// so its tile is one contiguous run of P words; later groups stage TL = 32
// columns (a 128-byte line per row).  A group grows while its LDS tile
// (live layers x P x TL int32) stays within the budget.
// layers that can be nonzero after axes 0..c-1: the largest distance those
// axes can produce is one per axis (Hamming) or floor(ax/2) per axis (Lee)
int layers_after(int c) {
  if (E.L == 1) return 1;
  long long w = 0;
  for (int j = 0; j < c; j++) w += E.prob == CC_LEE ? E.ax[j] / 2 : 1;
  return (int)std::min<long long>(w, E.R) + 1;
}

// axes per register round of the round kernel for axis size q (q^RR <= 16)
int round_axes(int q) { return q == 2 ? 4 : (q <= 4 ? 2 : 1); }

int padded(int i) { return i + (i >> 4); }
// host copy of the packed kernels' LDS padding (pk_pad)
int pk_pad_h(int i) { return i + (i >> 6); }

void build_plan() {
  E.plan.clear();
  int j = 0;
  while (j < E.n) {
    Sweep sw;
    sw.j0 = j;
    sw.c0 = j;
    sw.s = E.pw[j];
    // strided groups read whole 128-byte rows: 32 int32 or 64 uint16 columns
    const int TL0 = (j == 0) ? 1
                             : (int)std::min<long long>(E.l16_field ? 64 : 32, sw.s);
    // a homogeneous Hamming group with q <= 5 runs the register-round
    // kernel; in 16-bit mode it packs two layers per LDS dword, and a group
    // of binary or quaternary axes is capped at one 16-word cube per thread
    // (4,096 words per tile) to keep its round count low
    const int q = E.ax[j];
    // (the round kernels of axis size <= 5 exchange layer planes only: with
    // compact words on, such a group takes the LDS tile, which has them)
    const bool round_ok = !E.generic_only && E.prob == CC_HAMMING && q <= 10 &&
                          (q > 5 || !E.compact);
    long long P = 1;
    int k = j;
    while (k < E.n) {
      const bool homog = E.ax[k] == q;
      const long long P2 = P * E.ax[k];
      const int lay = layers_after(k + 1);
      const bool packed = round_ok && homog && E.l16_field;
      const size_t need = packed ? (size_t)((lay + 1) / 2) * P2 * TL0 * 4
                                 : (size_t)lay * cc_gplane((int)std::min<long long>(
                                       P2 * TL0, 1 << 24)) * 4;
      const bool cap = packed && (q == 2 || q == 4) && P2 * TL0 > 4096;
      if (k > j && (need > (size_t)E.lds_budget || P2 > (1LL << 20) ||
                    k - j >= CC_MAXSW || cap))
        break;
      P = P2;
      k++;
    }
    sw.nax = k - j;
    sw.P = (int)P;
    sw.lmax = layers_after(k);
    // an L-inf group of one long axis (up to 1024 words) may not fit the
    // budget at full row width: narrow its tile (fewer columns) until it
    // does (the generic tile addresses every column; no contiguity assumed)
    int TL = TL0;
    while (j > 0 && TL > 1 && (E.prob == CC_TORUS || E.prob == CC_GRID) &&
           (size_t)sw.lmax * cc_gplane((int)std::min<long long>(P * TL, 1 << 24)) * 4 >
               (size_t)E.lds_budget)
      TL /= 2;
    // widen short strided tiles to ~4096 words so every thread has work
    while (j > 0 && TL * 2LL <= sw.s && P * TL < 4096 &&
           (size_t)sw.lmax * cc_gplane((int)(P * TL * 2)) * 4 <= (size_t)E.lds_budget)
      TL *= 2;
    sw.TL = TL;
    sw.ncolb = (sw.s + sw.TL - 1) / sw.TL;
    sw.ntiles = (E.X / (sw.s * P)) * sw.ncolb;
    sw.rq = 0;
    bool homog = true;
    for (int i = j; i < k; i++) homog = homog && E.ax[i] == q;
    // the round kernel's cubes hold the layers live after the group
    const int Lr = std::max(2, sw.lmax);
    const size_t lds_r = (size_t)Lr * (padded(P * sw.TL) + 1) * 4;
    const size_t lds_r16 = (size_t)((Lr + 1) / 2) * (padded(P * sw.TL) + 1) * 4;
    // (in 16-bit mode every transform of the cell runs 16-bit, so only the
    // packed footprint has to fit)
    if (round_ok && homog && sw.nax >= round_axes(q) &&
        (E.l16_field ? lds_r16 : lds_r) <= (size_t)E.lds_max)
      sw.rq = q;
    sw.threads = CC_BLOCK;
    if (sw.rq) {
      // one cube per thread: a tile with few cubes (a first sweep of large
      // axes: 64 cubes of 8 words) runs smaller workgroups, more per CU
      int C = 1;
      for (int i = 0; i < std::min(round_axes(q), sw.nax); i++) C *= q;
      const long long ncube = (long long)P * sw.TL / C;
      sw.threads = (int)std::min<long long>(CC_BLOCK, std::max<long long>(64, (ncube + 63) / 64 * 64));
      const bool one = sw.nax <= round_axes(q);
      sw.lds = one ? 0 : lds_r;
      sw.lds16 = one ? 0 : lds_r16;
    } else {
      sw.lds = sw.lds16 = (size_t)sw.lmax * cc_gplane((int)(P * sw.TL)) * 4;
    }
    E.plan.push_back(sw);
    j = k;
  }
}

// L: the layers the sweep's cubes hold -- those that can be nonzero after
// its axes (layers above them are zero and stay zero), so an early sweep
// (distances below R) carries fewer registers and does less arithmetic
hipError_t launch_round(int q, const SweepArgs &a, int L, int blocks, size_t lds,
                        bool l16, bool cmp, int threads) {
  return cc_sweep_round(a, L, q, blocks, lds, l16, cmp,
                        LaunchCtx{E.st, E.ncu, E.grid, threads});
}

hipError_t launch_sweep(const SweepArgs &a, int blocks, size_t lds, bool l16,
                        bool cmp) {
  const LaunchCtx lc{E.st, E.ncu, E.grid};
  switch (E.prob) {
    case CC_TORUS:
    case CC_GRID: return cc_sweep_linf(a, E.prob, blocks, lds, lc);
    case CC_HAMMING: return cc_sweep_hamming(a, E.L, blocks, lds, l16, cmp, lc);
    case CC_LEE: return cc_sweep_lee(a, E.L, blocks, lds, l16, cmp, lc);
  }
  return hipErrorInvalidValue;
}

// ---- compact words between unpacked sweeps -------------------------------
// After the first c axes, layer d at a word counts seed words that agree
// with it on the axes >= c and lie at partial distance exactly d over the
// axes < c.  There are N_d(c) such positions (Hamming: the elementary
// symmetric polynomial e_d of (q_i - 1), i < c; Lee: the coefficient of x^d
// in the product of the per-axis Lee weight enumerators), so the layer is at
// most s0 * N_d(c) for a seed whose words are <= s0, and never more than the
// seed's total cap.  Each layer gets bitlen of that bound; the fields are
// packed into at most CC_CMAXND dwords (largest first, each into the first
// dword with room), never straddling a dword.
int bitlen(double x);

struct CFmt {
  int nd = 0;
  unsigned char dw[CC_MAXL] = {}, off[CC_MAXL] = {}, bits[CC_MAXL] = {};
};

bool cmp_fmt(int c, double s0, double cap, CFmt &f) {
  std::vector<double> poly(E.L, 0.0);
  poly[0] = 1;
  for (int i = 0; i < c; i++) {
    std::vector<double> nx(E.L, 0.0);
    for (int w = 0; w < E.L; w++) {
      if (poly[w] == 0) continue;
      for (int a = 0; a < E.ax[i]; a++) {
        const int e = E.prob == CC_LEE ? std::min(a, E.ax[i] - a) : (a ? 1 : 0);
        if (w + e < E.L) nx[w + e] += poly[w];
      }
    }
    poly = nx;
  }
  f = CFmt();
  int order[CC_MAXL];
  for (int d = 0; d < E.L; d++) {
    const double b = std::min(s0 * poly[d], cap);
    const int nb = b >= 1 ? bitlen(b) : 0;
    if (nb > 31) return false;
    f.bits[d] = (unsigned char)nb;
    order[d] = d;
  }
  std::stable_sort(order, order + E.L,
                   [&](int x, int y) { return f.bits[x] > f.bits[y]; });
  int used[CC_CMAXND] = {};
  for (int i = 0; i < E.L; i++) {
    const int d = order[i];
    if (!f.bits[d]) continue;
    int k = 0;
    while (k < CC_CMAXND && used[k] + f.bits[d] > 32) k++;
    if (k == CC_CMAXND) return false;
    f.dw[d] = (unsigned char)k;
    f.off[d] = (unsigned char)used[k];
    used[k] += f.bits[d];
    f.nd = std::max(f.nd, k + 1);
  }
  // the compact planes live in the layer block (R planes)
  return f.nd >= 1 && f.nd <= E.R;
}

// formats after every sweep but the last of the unpacked plan, for a seed
// bounded by (s0, cap); false when the plan cannot exchange compact words
// (single-plane families, register-round sweeps, a bound past 31 bits or
// more dwords than the layer block holds)
bool compact_plan(double s0, double cap, std::vector<CFmt> &fm) {
  fm.clear();
  if (!E.compact || E.L < 2 || (E.prob != CC_HAMMING && E.prob != CC_LEE))
    return false;
  // (the register-round kernels of axis size <= 5 have no compact variant)
  for (const Sweep &sw : E.plan)
    if (sw.rq && sw.rq <= 5) return false;
  for (size_t i = 0; i + 1 < E.plan.size(); i++) {
    CFmt f;
    if (!cmp_fmt(E.plan[i].j0 + E.plan[i].nax, s0, cap, f)) return false;
    fm.push_back(f);
  }
  return true;
}

// seed bounds of the pending code (largest multiplicity, code size)
void code_bounds(double *s0, double *cap) {
  *s0 = std::max(E.uploaded.maxm, 1);
  *cap = std::max(E.codeM, 1);
}

// Run the whole transform: in_kind seeds layer 0, out_kind says where the
// result goes.  pmode is the predicate for CC_IN_PRED.
int run_transform(int in_kind, int out_kind, int pmode) {
  E.map_fold = false;
  E.cnt_fold_pending = false;
  const int ns = (int)E.plan.size();
  const bool l16 = out_kind == CC_OUT_CNT ? E.l16_count
                   : out_kind == CC_OUT_LAYERS && in_kind == CC_IN_CODE
                       ? E.l16_count
                       : E.l16_field;
  E.layers_live = out_kind == CC_OUT_LAYERS ? (l16 ? 2 : 1) : 0;
  E.nswept = 0;
  std::vector<CFmt> fm;
  bool cmp = false;
  if (out_kind != CC_OUT_LAYERS && ns > 1) {
    double s0 = 1, cap = 1e300;
    if (in_kind == CC_IN_CODE) code_bounds(&s0, &cap);
    cmp = compact_plan(s0, cap, fm);
  }
  for (int i = 0; i < ns; i++) {
    const Sweep &sw = E.plan[i];
    SweepArgs a;
    memset(&a, 0, sizeof(a));
    for (int d = 0; d < CC_MAXL; d++) {
      a.A[d] = E.A[d];
      a.H[d] = E.H[d];
    }
    a.cnt = E.cnt;
    a.s = sw.s;
    a.ncolb = sw.ncolb;
    a.ntiles = sw.ntiles;
    a.P = sw.P;
    a.TL = sw.TL;
    a.nax = sw.nax;
    a.c0 = sw.c0;
    a.R = E.R;
    a.mu = E.mu;
    a.pmode = pmode;
    a.in_kind = (i == 0) ? in_kind : CC_IN_LAYERS;
    a.out_kind = (i == ns - 1) ? out_kind : CC_OUT_LAYERS;
    // the layered families never change layer 0 after the first sweep
    // seeds it; the single-plane families rewrite it every sweep
    // 16-bit mode keeps layer 0 in H[0] too, so the first sweep writes it
    // whatever seeded it (A[0] holds the code indicator only as int32)
    a.write_l0 = (E.L == 1) || (i == 0 && (in_kind == CC_IN_PRED || l16));
    long long sk = 1;
    for (int k = 0; k < sw.nax; k++) {
      a.Q[k] = E.ax[sw.j0 + k];
      a.sk[k] = (int)sk;
      sk *= a.Q[k];
    }
    for (int k = 0; k <= sw.nax; k++) a.live[k] = layers_after(sw.j0 + k);
    if (cmp) {
      a.pk = (uint32_t *)E.lay;
      a.pitch = E.pitch;
      if (i > 0) {
        const CFmt &f = fm[i - 1];
        a.in_kind = CC_IN_COMPACT;
        a.in_nd = f.nd;
        for (int d = 0; d < E.L; d++) {
          a.idw[d] = f.dw[d]; a.ioff[d] = f.off[d]; a.ibits[d] = f.bits[d];
        }
      }
      if (i < ns - 1) {
        const CFmt &f = fm[i];
        a.out_kind = CC_OUT_COMPACT;
        a.out_nd = f.nd;
        for (int d = 0; d < E.L; d++) {
          a.odw[d] = f.dw[d]; a.ooff[d] = f.off[d]; a.obits[d] = f.bits[d];
        }
      }
    }
    const int blocks = (int)std::min<long long>(sw.ntiles, 1LL << 20);
    if (E.prof) CK(hipEventRecord(E.ev[2 * i], E.st));
    const bool cm = a.in_kind == CC_IN_COMPACT || a.out_kind == CC_OUT_COMPACT;
    if (sw.rq)
      CK(launch_round(sw.rq, a, std::max(2, sw.lmax), blocks,
                      l16 ? sw.lds16 : sw.lds, l16, cm, sw.threads));
    else CK(launch_sweep(a, blocks, sw.lds, l16, cm));
    if (E.prof) CK(hipEventRecord(E.ev[2 * i + 1], E.st));
  }
  E.nswept = ns;
  return 0;
}

// ---- packed-layer schedule (Hamming, homogeneous axes, cc_psweep.hpp) ----
// Bounds of the layer values for a seed whose words are <= s0 (1 for the
// gain / loss predicates, the largest multiplicity for a code) and whose
// total is <= cap (the code size; unbounded for a predicate): after c axes
// layer d counts seed words at partial distance exactly d, at most
// s0 * C(c,d) * (Q-1)^d of them, and a fiber sum along the next axis is at
// most Q times that.  Every count is also <= cap.
struct PBounds {
  int Q, R;
  double s0, cap;
  double lay(int c, int d) const {
    if (d > c) return 0;
    double b = s0;
    for (int i = 0; i < d; i++) b = b * (c - i) / (i + 1) * (Q - 1);
    return std::min(b, cap);
  }
  double tsum(int c, int d) const { return std::min(lay(c, d) * Q, cap); }
  // do W-bit fields hold every layer through axes [j, k) of the transform?
  bool fits(int W, int k) const {
    const double lim = std::ldexp(1.0, W) - 1;
    for (int d = 0; d <= R; d++)
      if (lay(k, d) > lim) return false;
    if (Q > 2)
      for (int d = 0; d < R; d++)
        if (tsum(k - 1, d) > lim) return false;
    return true;
  }
};

int bitlen(double x) {
  int b = 0;
  while (b < 53 && std::ldexp(1.0, b) <= x) b++;
  return b;
}

// compact HBM format of the layers after c axes: field d takes bitlen of its
// bound, packed greedily into at most two dword planes
bool compact_fmt(const PBounds &B, int c, int L, int *dw, int *off, int *bits,
                 int *nd, int *w8, int *nar) {
  *w8 = 0;
  *nar = 0;
  if (L >= 4 && E.pk_narrow) {
    // narrow word: every field packed into 16 or 24 bits, stored as a
    // uint16 plane (and a uint8 plane) in the last layer plane of `lay`,
    // which the dword formats (planes 0 and 1) never touch, so a sweep may
    // switch between the two without its output overwriting its input
    int tot = 0;
    for (int d = 0; d < L; d++) tot += bitlen(B.lay(c, d));
    if (tot <= 24) {
      int used = 0;
      for (int d = 0; d < L; d++) {
        dw[d] = 0; off[d] = used; bits[d] = bitlen(B.lay(c, d));
        used += bits[d];
      }
      *nd = 1;
      *nar = tot <= 16 ? 1 : 2;
      return true;
    }
  }
  if (L <= 4) {
    bool bytes = true;
    for (int d = 0; d < L; d++) bytes = bytes && B.lay(c, d) <= 255.0;
    if (bytes) {
      // one byte per layer: converts to the register formats with permutes
      for (int d = 0; d < L; d++) { dw[d] = 0; off[d] = 8 * d; bits[d] = 8; }
      *nd = 1;
      *w8 = 1;
      return true;
    }
  }
  int w = 0, used = 0;
  for (int d = 0; d < L; d++) {
    const int b = bitlen(B.lay(c, d));
    if (b > 32) return false;
    if (used + b > 32) { w++; used = 0; }
    if (w > 1) return false;
    dw[d] = w; off[d] = used; bits[d] = b;
    used += b;
  }
  *nd = w + 1;
  return true;
}

// largest register round of the kernel table for axis size q (cube of
// q^RR words of nd dwords in at most 128 registers)
int pk_rrmax(int q, int nd) {
  const int table = q == 2 ? 6 : (q == 3 ? 4 : 3);
  int rr = 0;
  long long c = 1;
  while (rr < table && c * q * nd <= 128) { c *= q; rr++; }
  return rr;
}

// Rounds of one sweep: round 0 is the cube of the group's lowest RR axes
// (contiguous in the first sweep: vector loads), middle rounds take the
// axes above in ascending chunks, the last round the top RR axes (whose
// cube elements are P / C words apart: coalesced stores in the first sweep).
void pk_rounds(PSweep &w, int q) {
  const int m = w.nax, RR = w.RR;
  int nr = 0;
  auto add = [&](int ca, int mask) {
    int qa = 1;
    for (int i = 0; i < ca; i++) qa *= q;
    w.ca[nr] = ca; w.amask[nr] = mask; w.qa[nr] = qa;
    // LDS addresses pad((pb + c qa) TL + tl) are linear in c when no
    // padding step (every 64 elements) falls inside a cube
    const long long K = (long long)qa * w.TL;
    int C = 1;
    for (int i = 0; i < w.RR; i++) C *= q;
    if (K % 64 == 0) { w.lin[nr] = 1; w.kp[nr] = (int)(K + K / 64); }
    else if (K == 1 && 64 % C == 0) { w.lin[nr] = 1; w.kp[nr] = 1; }
    else { w.lin[nr] = 0; w.kp[nr] = 0; }
    nr++;
  };
  add(0, (1 << RR) - 1);
  if (m > RR) {
    const int top = m - RR;
    int done = RR;
    for (int lo = RR; lo < top; lo += RR) {
      const int hi = std::min(lo + RR, top);
      add(lo, (1 << (hi - lo)) - 1);
      done = hi;
    }
    int mask = 0;
    for (int i = 0; i < RR; i++)
      if (top + i >= done) mask |= 1 << i;
    add(top, mask);
  }
  w.nrounds = nr;
}

// Launch a packed sweep kernel, or with blocks == 0 only check that its
// instantiation exists.
hipError_t pk_dispatch(const PSweep &w, const PSweepArgs &a, int blocks,
                       hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  switch (E.ax[0]) {
    case 2: return cc_psweep_q2(a, w.RR, w.W, w.ND, blocks, w.threads, w.lds, E.st, e0, e1);
    case 3: return cc_psweep_q3(a, w.RR, w.W, w.ND, blocks, w.threads, w.lds, E.st, e0, e1);
    case 4: return cc_psweep_q4(a, w.RR, w.W, w.ND, blocks, w.threads, w.lds, E.st, e0, e1);
    case 5: return cc_psweep_q5(a, w.RR, w.W, w.ND, blocks, w.threads, w.lds, E.st, e0, e1);
  }
  return hipErrorInvalidValue;
}

// One sweep over axes [j, k); prev is the sweep before it (its compact
// output is this sweep's input).
bool pk_sweep(const PBounds &B, int j, int k, bool pred_seed, bool cnt_out,
              const PSweep *prev, PSweep &w) {
  const int q = B.Q, L = E.L, m = k - j;
  memset(&w, 0, sizeof(w));
  w.j0 = j; w.nax = m;
  w.s = E.pw[j];
  const long long P = E.pw[k] / E.pw[j];
  if (P > (1LL << 20) || P * w.s > (1LL << 30)) return false;
  w.P = (int)P;
  if (B.fits(8, k)) { w.W = 8; w.ND = (L + 3) / 4; }
  else if (B.fits(16, k)) { w.W = 16; w.ND = (L + 1) / 2; }
  else return false;
  if (w.ND > 4 || (w.W == 8 && w.ND > 2)) return false;
  const int rrmax = pk_rrmax(q, w.ND);
  if (rrmax < 1) return false;
  w.RR = std::min(rrmax, m);
  auto qpow = [q](int e) { int c = 1; for (int i = 0; i < e; i++) c *= q; return c; };
  if (j == 0 && m > 1) {
    // first sweep: a cube small enough that the tile has whole waves of
    // cubes (64..256 threads, one cube each) when possible
    int best = 0;
    for (int r = std::min(rrmax, m); r >= 1 && !best; r--) {
      const long long nc = P / qpow(r);
      if (nc >= 64 && nc <= CC_BLOCK && nc % 64 == 0) best = r;
    }
    for (int r = std::min(rrmax, m); r >= 1 && !best; r--)
      if (P / qpow(r) >= 64) best = r;
    if (best) w.RR = best;
  } else if (j > 0 && E.pk_col >= 2 && q == 2 && m == rrmax + 1 && m >= 4 &&
             w.s % 128 == 0 && qpow(m - 1) * w.ND <= 128) {
    // a lane pair holds the cube: one more axis than a lane alone
    w.RR = m;
  }
  int C = qpow(w.RR);
  w.kern = 0;
  if (j == 0 && E.pk_first == 2 && q == 2 && m >= 10 && m <= 12 &&
      (1 << (m - 6)) * w.ND <= 64) {
    // LDS-free first sweep: m - 6 axes in-thread, 6 across the lanes
    w.kern = 3;
    w.RR = m - 6;
    w.TL = 1;
    w.lds = 0;
    w.threads = CC_BLOCK;
  } else if (j == 0) {
    w.TL = 1;
    w.lds = m > w.RR ? (size_t)w.ND * (pk_pad_h(w.P) + 1) * 4 : 0;
    if (w.lds > (size_t)E.lds_budget) return false;
    const int ncube = w.P / C;
    w.threads = std::min(256, (ncube + 63) / 64 * 64);
    // persistent grid with next-tile prefetch: one cube per thread (the
    // last wave's spare lanes idle), at least two rounds (a code seed takes
    // it too, seeded from the code's sorted word list)
    if (E.pk_first == 1 && (pred_seed || E.sparse) && m > w.RR && ncube <= CC_BLOCK)
      w.kern = 4;
  } else if (m == w.RR) {
    w.lds = 0;
    // column kernel: one cube per lane (or per lane pair, binary axes
    // with big cubes), whole workgroups of consecutive columns
    const int col = E.pk_col;
    const bool split = col >= 2 && q == 2 && w.RR >= 4 && w.s % 128 == 0;
    const bool whole = col >= 1 && w.RR >= 2 && w.s % 256 == 0 &&
                       (long long)C * w.ND <= 128;
    if (split || whole) {
      w.kern = split ? 2 : 1;
      w.TL = 1;
      w.threads = CC_BLOCK;
    } else {
      w.TL = (int)std::min<long long>(256, w.s);
      w.threads = std::min(256, (w.TL + 63) / 64 * 64);
    }
  } else {
    // two rounds (round 0 the low RR axes, round 1 the rest): the tile may
    // cross LDS as one compact dword per word, the format of the layers
    // after round 0, instead of ND register dwords
    w.lcmp = 0;
    if (E.pk_lcmp && w.ND > 1 && m <= 2 * w.RR) {
      int dw[CC_MAXL], nd = 0, nar = 0;
      if (compact_fmt(B, j + w.RR, L, dw, w.moff, w.mbits, &nd, &w.mid_w8, &nar) &&
          nd == 1 && !nar)
        w.lcmp = 1;
    }
    const int lnd = w.lcmp ? 1 : w.ND;
    w.TL = 0;
    for (int tl : {64, 32, 16}) {
      if (tl > w.s) continue;
      const size_t lds = (size_t)lnd * (pk_pad_h(w.P * tl) + 1) * 4;
      if (lds <= (size_t)E.lds_budget) { w.TL = tl; w.lds = lds; break; }
    }
    if (!w.TL) return false;
    w.threads = std::min(256, (w.P / C * w.TL + 63) / 64 * 64);
  }
  if (w.kern == 3) {
    w.ncolb = 1;
    w.ntiles = E.X / ((long long)64 << w.RR);   // wave tiles
  } else if (w.kern) {
    w.ncolb = 1;
    w.ntiles = (E.X / (w.s * P)) * w.s;   // cubes (one per column)
  } else {
    w.ncolb = (w.s + w.TL - 1) / w.TL;
    w.ntiles = (E.X / (w.s * P)) * w.ncolb;
  }
  if (prev) {
    w.in_nd = prev->out_nd;
    w.in_w8 = prev->out_w8;
    w.in_nar = prev->out_nar;
    for (int d = 0; d < L; d++) {
      w.idw[d] = prev->odw[d]; w.ioff[d] = prev->ooff[d]; w.ibits[d] = prev->obits[d];
    }
    w.in_bytes = prev->out_bytes;
  } else {
    w.in_nd = 0;
    // a code seed read from the A0 plane, or (persistent first sweep) from
    // the code's word list: no plane bytes
    w.in_bytes = pred_seed ? 2 : (w.kern == 4 ? 0 : 4);
  }
  if (k == E.n) {
    w.out_nd = 0;
    w.out_bytes = cnt_out ? 2 : 4;
  } else {
    if (!compact_fmt(B, k, L, w.odw, w.ooff, w.obits, &w.out_nd, &w.out_w8,
                     &w.out_nar))
      return false;
    if (w.out_nd > L - 1) return false;   // compact planes live in `lay`
    w.out_bytes = w.out_nar ? 1 + w.out_nar : 4 * w.out_nd;
  }
  pk_rounds(w, q);
  {
    // only schedule sweeps whose kernel instantiation exists
    PSweepArgs a;
    memset(&a, 0, sizeof(a));
    a.kern = w.kern;
    a.in_kind = prev ? CC_IN_PACKED : (pred_seed ? CC_IN_PRED : CC_IN_CODE);
    a.out_kind = k == E.n ? (cnt_out ? CC_OUT_CNT : CC_OUT_MAP) : CC_OUT_PACKED;
    a.in_nd = w.in_nd;
    a.out_nd = w.out_nd;
    if (pk_dispatch(w, a, 0) != hipSuccess) return false;
  }
  // relative time: algorithmic bytes over the bandwidth this kernel shape
  // reaches (MI355X measurements, K_2(24,3), K_2(30,4), K_5(10,3) sweeps,
  // normalised to the split column kernel at ~4.8 TB/s)
  double f;
  if (w.kern == 2) f = 1.0;
  else if (w.kern == 1) f = 1.35;
  else if (w.kern == 4) f = 1.25;
  else if (w.kern == 3) f = 1.8;
  else if (j == 0) {
    const long long nc = P / C;
    f = 2.0 * (nc < 64 ? 64.0 / (double)std::max(1LL, nc) : 1.0);
  } else if (w.nrounds > 1) {
    f = w.TL >= 64 ? 1.8 : (w.TL >= 32 ? 2.2 : 2.6);
  } else {
    f = w.TL >= 64 ? 1.8 : 2.6;
  }
  if (j == 0 && w.nrounds > 2) f *= 1.0 + 0.25 * (w.nrounds - 2);
  if (j == 0 && w.lds > 0) {
    // workgroups a CU holds at once, by LDS: few resident waves leave the
    // memory pipe idle between a tile's load and its stores
    const double waves = (double)(160 * 1024 / (long long)w.lds) * (w.threads / 64);
    if (waves < 8) f *= 8.0 / std::max(1.0, waves);
  }
  w.cost = (w.in_bytes + w.out_bytes) * f;
  return true;
}

// Schedule with the least estimated time (algorithmic bytes weighted by the
// kernel shape's measured bandwidth), then the fewest bytes and sweeps: a
// shortest path over sweep boundaries (the compact format at a boundary
// depends only on how many axes precede it, so a sweep's cost depends only
// on its two ends).
// Two sweep shapes that let a binary cell take fewer, larger sweeps
// (kern 5 and 6, CC_PK_WIDE): a first sweep of 13-14 axes whose byte-format
// LDS tile widens to 16-bit fields in its last round (k_pfirstw), and a
// strided sweep of 2 RA axes in two register rounds through a one-dword
// compact LDS tile (k_pstrip).  K_2(24,3): axes 0-13 and 14-23, 2 + 4 and
// 4 + 4 bytes per word (the three-sweep schedule moves 22).
bool pk_sweep_ext(const PBounds &B, int j, int k, bool pred_seed, bool cnt_out,
                  const PSweep *prev, PSweep &w) {
  const int q = B.Q, L = E.L, m = k - j;
  if (!E.pk_wide || q != 2 || L > 4 || L < 3) return false;
  memset(&w, 0, sizeof(w));
  w.j0 = j; w.nax = m;
  w.s = E.pw[j];
  const long long P = E.pw[k] / E.pw[j];
  if (P > (1LL << 20) || P * w.s > (1LL << 29)) return false;
  w.P = (int)P;
  if (!B.fits(16, k)) return false;
  w.W = 16; w.ND = 2;
  w.TL = 1;
  auto add = [&](int ca, int mask, int C) {
    const int r = w.nrounds;
    const int qa = 1 << ca;
    w.ca[r] = ca; w.amask[r] = mask; w.qa[r] = qa;
    const long long K = qa;
    if (K % 64 == 0) { w.lin[r] = 1; w.kp[r] = (int)(K + K / 64); }
    else if (K == 1 && 64 % C == 0) { w.lin[r] = 1; w.kp[r] = 1; }
    else { w.lin[r] = 0; w.kp[r] = 0; }
    w.nrounds++;
  };
  if (j == 0) {
    // k_pfirstw: 64-word cubes (RR = 6), byte rounds of 6 axes up to the
    // last axis count whose layers fit bytes, then the top NA axes widened
    if (m < 13 || m > 14 || !(pred_seed || E.sparse)) return false;
    int kb = 0;
    while (kb < m && B.fits(8, kb + 1)) kb++;
    const int a0 = std::min(kb, m - 1);
    w.na = m - a0;
    if (w.na < 1 || w.na > 2 || a0 < m - 6) return false;
    w.RR = 6;
    const int C = 64;
    for (int lo = 0; lo < a0; lo += 6) {
      const int hi = std::min(lo + 6, a0);
      const int ca = std::min(lo, m - 6);
      add(ca, ((1 << (hi - lo)) - 1) << (lo - ca), C);
    }
    add(m - 6, ((1 << w.na) - 1) << (6 - w.na), C);
    if (w.nrounds > CC_PK_MAXRD) return false;
    w.threads = (int)(P / C);
    w.lds = (size_t)(pk_pad_h(w.P) + 1) * 4;
    if (w.lds > (size_t)E.lds_max) return false;
    w.kern = 5;
    w.ncolb = 1;
    w.ntiles = E.X / P;
  } else {
    // k_pstrip: two rounds of RA axes, TL columns per tile, 512 threads
    if (m % 2) return false;
    const int RA = m / 2;
    if (RA < 4 || RA > 5) return false;
    const int C = 1 << RA;
    w.RR = RA;
    w.TL = 512 / C;
    if (w.s % w.TL) return false;
    int dw[CC_MAXL], nd = 0, nar = 0;
    if (!compact_fmt(B, j + RA, L, dw, w.moff, w.mbits, &nd, &w.mid_w8, &nar) ||
        nd != 1 || nar)
      return false;
    w.nrounds = 2;
    w.threads = w.TL * C;
    w.lds = (size_t)C * (C * w.TL + 16) * 4;
    if (w.lds > (size_t)E.lds_max) return false;
    w.kern = 6;
    w.ncolb = w.s / w.TL;
    w.ntiles = (E.X / (w.s * P)) * w.ncolb;
  }
  if (prev) {
    w.in_nd = prev->out_nd;
    w.in_w8 = prev->out_w8;
    w.in_nar = prev->out_nar;
    if (w.in_nar) return false;
    for (int d = 0; d < L; d++) {
      w.idw[d] = prev->odw[d]; w.ioff[d] = prev->ooff[d]; w.ibits[d] = prev->obits[d];
    }
    w.in_bytes = prev->out_bytes;
  } else {
    w.in_nd = 0;
    w.in_bytes = pred_seed ? 2 : 0;
  }
  if (k == E.n) {
    w.out_nd = 0;
    w.out_bytes = cnt_out ? 2 : 4;
  } else {
    if (!compact_fmt(B, k, L, w.odw, w.ooff, w.obits, &w.out_nd, &w.out_w8,
                     &w.out_nar))
      return false;
    if (w.out_nd > L - 1 || w.out_nar) return false;
    w.out_bytes = 4 * w.out_nd;
  }
  {
    PSweepArgs a;
    memset(&a, 0, sizeof(a));
    a.kern = w.kern;
    a.na = w.na;
    a.in_kind = prev ? CC_IN_PACKED : (pred_seed ? CC_IN_PRED : CC_IN_SPARSE);
    a.out_kind = k == E.n ? (cnt_out ? CC_OUT_CNT : CC_OUT_MAP) : CC_OUT_PACKED;
    a.in_nd = w.in_nd;
    a.out_nd = w.out_nd;
    if (pk_dispatch(w, a, 0) != hipSuccess) return false;
  }
  w.cost = (w.in_bytes + w.out_bytes) * (w.kern == 5 ? 1.25 : 1.1);
  return true;
}

PPlan pk_plan(double s0, double cap, bool pred_seed, bool cnt_out) {
  PPlan best;
  if (!E.packed || E.prob != CC_HAMMING || E.L < 2) return best;
  const int q = E.ax[0];
  for (int i = 1; i < E.n; i++)
    if (E.ax[i] != q) return best;
  if (q < 2 || q > 5) return best;
  const PBounds B{q, E.R, s0, cap};
  const int n = E.n;
  struct Cost { double t; long long bytes; int sweeps; bool ok; };
  auto better = [](const Cost &x, const Cost &y) {
    if (!y.ok) return x.ok;
    if (!x.ok) return false;
    if (x.t != y.t) return x.t < y.t;
    if (x.bytes != y.bytes) return x.bytes < y.bytes;
    return x.sweeps < y.sweeps;
  };
  std::vector<Cost> dp(n + 1, Cost{0, 0, 0, false});
  std::vector<PSweep> via(n + 1);
  std::vector<int> from(n + 1, -1);
  dp[0] = Cost{0, 0, 0, true};
  for (int j = 0; j < n; j++) {
    if (!dp[j].ok) continue;
    PSweep prev;
    memset(&prev, 0, sizeof(prev));
    if (j > 0 && !compact_fmt(B, j, E.L, prev.odw, prev.ooff, prev.obits,
                              &prev.out_nd, &prev.out_w8, &prev.out_nar))
      continue;
    prev.out_bytes = prev.out_nar ? 1 + prev.out_nar : 4 * prev.out_nd;
    // forced boundaries: the sweep from j ends at the next cut (or at n)
    int kf = n;
    for (int c = j + 1; g_cuts && c < n; c++)
      if ((g_cuts >> c) & 1) { kf = c; break; }
    for (int k = j + 1; k <= std::min(n, j + CC_MAXSW); k++) {
      if (g_cuts && k != kf) continue;
      PSweep w;
      const bool okw = pk_sweep(B, j, k, pred_seed, cnt_out, j ? &prev : nullptr, w);
      if (getenv("CC_PK_DEBUG"))
        fprintf(stderr, "pk_plan [%d,%d) ok=%d kern=%d RR=%d W=%d ND=%d TL=%d cost=%.2f\n",
                j, k, (int)okw, w.kern, w.RR, w.W, w.ND, w.TL, okw ? w.cost : -1.0);
      if (okw) {
        const Cost c{dp[j].t + w.cost, dp[j].bytes + w.in_bytes + w.out_bytes,
                     dp[j].sweeps + 1, true};
        if (better(c, dp[k])) { dp[k] = c; via[k] = w; from[k] = j; }
      }
      PSweep x;
      if (pk_sweep_ext(B, j, k, pred_seed, cnt_out, j ? &prev : nullptr, x)) {
        const Cost c{dp[j].t + x.cost, dp[j].bytes + x.in_bytes + x.out_bytes,
                     dp[j].sweeps + 1, true};
        if (better(c, dp[k])) { dp[k] = c; via[k] = x; from[k] = j; }
      }
    }
  }
  if (!dp[n].ok) return best;
  for (int k = n; k > 0; k = from[k]) best.sw.insert(best.sw.begin(), via[k]);
  best.ok = true;
  best.bytes = dp[n].bytes;
  return best;
}

// census regions: field scans, the set-side size census, the folded map
// maximum and the folded cnt deficit
enum { CEN_SCAN = 0, CEN_ZL = 1, CEN_FMAP = 2, CEN_FDEF = 3, CEN_N = 4 };
constexpr int CEN_MAXSTRIDE = 512;   // uint64 words (4 KB)

unsigned long long *cen(int r) {
  return E.d_cen + (size_t)r * CC_NSLOT * CEN_MAXSTRIDE;
}

int cen_clear(int r) {
  hipLaunchKernelGGL(k_clear_slots, dim3(1), dim3(64), 0, E.st, cen(r), E.cen_stride);
  CK(hipGetLastError());
  return 0;
}

// enqueue the copy of region r's slots to its host mirror h_cen[r]
int cen_copy(int r) {
  CK(hipMemcpy2DAsync(E.h_cen + (size_t)r * CC_NSLOT, 8, cen(r),
                      (size_t)E.cen_stride * 8, 8, CC_NSLOT, hipMemcpyDeviceToHost,
                      E.st));
  return 0;
}

// region r's host mirror combined (sum, or mx: maximum)
unsigned long long cen_combine(int r, bool mx) {
  const unsigned long long *h = E.h_cen + (size_t)r * CC_NSLOT;
  unsigned long long v = 0;
  for (int i = 0; i < CC_NSLOT; i++) v = mx ? std::max(v, h[i]) : v + h[i];
  return v;
}

// the slots of region r back on the host, combined
int cen_fetch(int r, bool mx, unsigned long long *out) {
  if (int rc = cen_copy(r)) return rc;
  CK(hipStreamSynchronize(E.st));
  *out = cen_combine(r, mx);
  return 0;
}

// a folded region's value: the transform that folded it also queued the copy
// of its slots, so the host mirror is final once the stream is idle
int cen_folded(int r, bool mx, unsigned long long *out) {
  CK(hipStreamSynchronize(E.st));
  *out = cen_combine(r, mx);
  return 0;
}

// the deficit that the recount which made this cnt state folded (the
// deficit at mu = 1 is count_eq(0)); false when none is held
bool folded_deficit(unsigned long long *r) {
  if (!E.fold_cnt_ver || E.fold_cnt_ver != E.cnt_ver) return false;
  return cen_folded(CEN_FDEF, false, r) == 0;
}

hipError_t pk_launch(const PSweep &w, const PSweepArgs &a, hipEvent_t e0 = nullptr,
                     hipEvent_t e1 = nullptr) {
  const long long units = w.kern == 3 ? (w.ntiles + CC_BLOCK / 64 - 1) / (CC_BLOCK / 64)
                         : (w.kern == 1 || w.kern == 2) ? w.ntiles / (CC_BLOCK / w.kern)
                                                        : w.ntiles;
  const int blocks = (int)std::max(1LL, std::min<long long>(units, 1LL << 30));
  return pk_dispatch(w, a, blocks, e0, e1);
}

// Run a packed transform: in_kind CC_IN_PRED (predicate [plo <= cnt <= phi])
// or CC_IN_CODE (A0), out_kind CC_OUT_MAP or CC_OUT_CNT.
int run_ptransform(const PPlan &pl, int in_kind, int out_kind, int plo,
                   int phi) {
  const int ns = (int)pl.sw.size();
  E.layers_live = 0;
  E.nswept = 0;
  E.map_fold = false;
  E.cnt_fold_pending = false;
  // the last sweep folds the map maximum or the cnt deficit into census
  // region CEN_FMAP or CEN_FDEF (32-bit wave deficit sums: X mu < 2^32),
  // which the first sweep clears.  A one-sweep transform (a small cell)
  // does not fold: clearing would cost a launch of its own, as much as the
  // scan it saves.
  unsigned long long *fslot = nullptr;
  if (E.fold && ns > 1 &&
      (out_kind == CC_OUT_MAP ||
       (out_kind == CC_OUT_CNT && (double)E.X * E.mu < 4294967296.0)))
    fslot = cen(out_kind == CC_OUT_MAP ? CEN_FMAP : CEN_FDEF);
  for (int i = 0; i < ns; i++) {
    const PSweep &w = pl.sw[i];
    PSweepArgs a;
    memset(&a, 0, sizeof(a));
    a.fold = i == ns - 1 ? fslot : nullptr;
    a.fold_zero = i == 0 ? fslot : nullptr;
    a.fold_mu = E.mu;
    a.fold_stride = E.cen_stride;
    a.cnt = E.cnt;
    a.code = E.A[0];
    a.cw = E.d_cw;
    a.cm = E.d_cm;
    a.toff = E.toff_cur;
    a.pk = (uint32_t *)E.lay;
    a.map = E.A[0];
    a.cnt_out = E.cnt;
    a.pitch = E.pitch;
    a.X = E.X;
    a.s = w.s; a.ncolb = w.ncolb; a.ntiles = w.ntiles;
    a.P = w.P; a.TL = w.TL; a.nax = w.nax; a.L = E.L;
    a.kern = w.kern;
    a.in_kind = i == 0 ? (in_kind == CC_IN_CODE && (w.kern == 4 || w.kern == 5)
                              ? CC_IN_SPARSE : in_kind)
                       : CC_IN_PACKED;
    a.out_kind = i == ns - 1 ? out_kind : CC_OUT_PACKED;
    a.plo = plo; a.phi = phi;
    a.in_nd = w.in_nd; a.out_nd = w.out_nd;
    a.nrounds = w.nrounds;
    for (int r = 0; r < w.nrounds; r++) {
      a.ca[r] = w.ca[r]; a.amask[r] = w.amask[r]; a.qa[r] = w.qa[r];
      a.lin[r] = w.lin[r]; a.kp[r] = w.kp[r];
    }
    a.in_w8 = w.in_w8;
    a.out_w8 = w.out_w8;
    a.in_nar = w.in_nar;
    a.out_nar = w.out_nar;
    a.nar_off = (long long)(E.L - 2) * E.pitch;
    a.out_bmask = E.L >= 4 ? 0xFFFFFFFFu : (uint32_t)((1u << (8 * E.L)) - 1u);
    a.st_nt = (E.pk_nt >> i) & 1;
    a.na = w.na;
    a.mid_w8 = w.mid_w8;
    a.lcmp = w.lcmp;
    for (int d = 0; d < E.L; d++) { a.moff[d] = w.moff[d]; a.mbits[d] = w.mbits[d]; }
    // real-layer slots: layer d = slot * ND + dword (interleaved layout)
    const int S = 32 / w.W;
    for (int k = 0; k < w.ND; k++) {
      uint32_t fm = 0;
      for (int sl = 0; sl < S; sl++)
        if (sl * w.ND + k < E.L) fm |= (uint32_t)((1ull << w.W) - 1) << (w.W * sl);
      a.fmask[k] = fm;
    }
    for (int d = 0; d < E.L; d++) {
      a.idw[d] = w.idw[d]; a.ioff[d] = w.ioff[d]; a.ibits[d] = w.ibits[d];
      a.odw[d] = w.odw[d]; a.ooff[d] = w.ooff[d]; a.obits[d] = w.obits[d];
    }
    // timing: the launch itself stamps the kernel's start and end
    if (E.prof) CK(pk_launch(w, a, E.ev[2 * i], E.ev[2 * i + 1]));
    else CK(pk_launch(w, a));
  }
  E.nswept = ns;
  if (fslot)   // the census comes back with the transform (no second round trip)
    if (int rc = cen_copy(out_kind == CC_OUT_MAP ? CEN_FMAP : CEN_FDEF)) return rc;
  E.map_fold = fslot && out_kind == CC_OUT_MAP;
  E.cnt_fold_pending = fslot && out_kind == CC_OUT_CNT;
  return 0;
}

// toff[t] = first entry of the ascending word list cw[0..n) at or above
// t * P (a lower bound by binary search), t = 0..ntiles
__global__ void k_tile_offsets(const long long *__restrict__ cw, int n, long long P,
                               long long ntiles, int32_t *toff) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  const long long key = t * P;
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cw[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  toff[t] = lo;
}

template <typename T>
int dalloc(T **p, size_t bytes);

// the pending code list on the device (for the indicator-plane scatter)
int ensure_code_dev() {
  if (E.code_dev_ok) return 0;
  const int M = (int)E.code_raw.size();
  for (int off = 0; off < M; off += CODECAP) {
    const int nb = std::min(M - off, CODECAP);
    memcpy(E.h_words, E.code_raw.data() + off, (size_t)nb * 8);
    CK(hipMemcpyAsync(E.d_code + off, E.h_words, (size_t)nb * 8,
                      hipMemcpyHostToDevice, E.st));
    CK(hipStreamSynchronize(E.st));
  }
  E.code_dev_ok = true;
  return 0;
}

// Stage the pending code for a sparse seed of first sweep w: its distinct
// words ascending and their multiplicities (through the pinned swap
// staging, free outside swap_eval), and the entry offsets of w's tiles.
// Everything is queued on the engine's stream; the transform's closing
// synchronisation releases the staging.
int stage_sparse_code(const PSweep &w) {
  // one upload: the n words, then their n multiplicities right after them
  const int n = (int)E.uploaded.w.size();
  long long *hw = (long long *)E.h_swap;
  memcpy(hw, E.uploaded.w.data(), (size_t)n * 8);
  memcpy(hw + n, E.uploaded.m.data(), (size_t)n * 4);
  E.d_cm = (int32_t *)(E.d_cw + n);
  const long long ntiles = w.ntiles;
  const size_t with_offs = (size_t)n * 12 + (size_t)(ntiles + 1) * 4;
  if (with_offs <= (size_t)CODECAP * 12 && with_offs <= SWAP_IN_BYTES) {
    // few tiles: their entry offsets by one merge of the ascending words
    // against the tile bounds here, uploaded behind the multiplicities in
    // the same copy (no offset kernel)
    int32_t *ht = (int32_t *)hw + 3LL * n;
    const long long *cw = E.uploaded.w.data();
    long long i = 0;
    for (long long t = 0; t <= ntiles; t++) {
      const long long key = t * (long long)w.P;
      while (i < n && cw[i] < key) i++;
      ht[t] = (int32_t)i;
    }
    E.toff_cur = (int32_t *)E.d_cw + 3LL * n;
    CK(hipMemcpyAsync(E.d_cw, hw, with_offs, hipMemcpyHostToDevice, E.st));
    return 0;
  }
  if (n)
    CK(hipMemcpyAsync(E.d_cw, hw, (size_t)n * 12, hipMemcpyHostToDevice, E.st));
  E.toff_cur = nullptr;
  if (ntiles + 1 > E.toff_cap) {
    CK(hipStreamSynchronize(E.st));
    if (E.d_toff) (void)hipFree(E.d_toff);
    E.d_toff = nullptr;
    E.toff_cap = 0;
    if (dalloc(&E.d_toff, (size_t)(ntiles + 1) * 4)) return -3;
    E.toff_cap = ntiles + 1;
  }
  hipLaunchKernelGGL(k_tile_offsets, dim3((unsigned)((ntiles + 256) / 256)), dim3(256), 0,
                     E.st, E.d_cw, n, (long long)w.P, ntiles, E.d_toff);
  CK(hipGetLastError());
  E.toff_cur = E.d_toff;
  return 0;
}

// packed plan of a code seed: multiplicities of the uploaded multiset
PPlan pk_code_plan(bool cnt_out) {
  const double s0 = std::max(E.uploaded.maxm, 1), cap = std::max(E.codeM, 1);
  // the planner is a shortest path over every sweep boundary pair: keep its
  // result while the seed bounds and the schedule options stay the same
  const int opts[8] = {(int)cnt_out, E.lds_budget, E.pk_col, E.pk_first,
                       E.pk_narrow | (E.pk_wide << 4) | (E.pk_lcmp << 5),
                       E.packed, E.sparse, (int)(g_cuts & 0x7FFFFFFF) ^ (int)(g_cuts >> 31)};
  if (E.cp_ok && E.cp_s0 == s0 && E.cp_cap == cap && !memcmp(opts, E.cp_opts, sizeof(opts)))
    return E.cp_plan;
  PPlan p = pk_plan(s0, cap, false, cnt_out);
  if (!p.ok && g_cuts) {
    // forced boundaries a code seed cannot use: the planner's own schedule
    const uint64_t keep = g_cuts;
    g_cuts = 0;
    p = pk_plan(s0, cap, false, cnt_out);
    g_cuts = keep;
  }
  E.cp_ok = true; E.cp_s0 = s0; E.cp_cap = cap;
  memcpy(E.cp_opts, opts, sizeof(opts));
  E.cp_plan = p;
  return p;
}

int sync_stream() {
  CK(hipStreamSynchronize(E.st));
  return 0;
}

// zero the reduction scalar; n = 8 also clears the running transition
// counter (init only: update_delta reads it as a difference)
int zero_scalars(int n = 1) {
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, E.st, E.d_scal, n);
  CK(hipGetLastError());
  return 0;
}

int fetch_scalar(int i, unsigned long long *r) {
  CK(hipMemcpyAsync(E.h_scal, E.d_scal + i, 8, hipMemcpyDeviceToHost, E.st));
  CK(hipStreamSynchronize(E.st));
  *r = E.h_scal[0];
  return 0;
}


// copy nb words to d_words through pinned staging (caller synchronises
// before the staging buffer is reused)
int stage_words(const long long *w, int nb) {
  memcpy(E.h_words, w, (size_t)nb * 8);
  CK(hipMemcpyAsync(E.d_words, E.h_words, (size_t)nb * 8,
                    hipMemcpyHostToDevice, E.st));
  return 0;
}

const uint64_t *walk_table() { return E.fmt == 1 ? E.pat2 : E.pat; }

template <int KIND>
hipError_t launch_walk(const WordsArg &wa, const long long *wptr, int nb,
                       int split, int delta, int pmode, int target,
                       int32_t *out) {
  const Geom g = make_geom();
  dim3 grid((unsigned)nb * split);
  const bool wide = E.fmt == 0 && E.tfmt == CC_FMT_WIDE;
  const bool x16 = E.fmt == 0 && E.tfmt == CC_FMT_X16;
#define CC_WALK_LAUNCH(GR, FM)                                              \
  hipLaunchKernelGGL((k_walk<KIND, GR, FM>), grid, dim3(CC_BLOCK), 0, E.st, \
                     g, wa, wptr, split, delta, E.cnt, walk_table(), E.blen,\
                     pmode, target, out, E.d_scal + 1)
  if (x16) {
    if (E.prob == CC_GRID) CC_WALK_LAUNCH(true, CC_FMT_X16);
    else CC_WALK_LAUNCH(false, CC_FMT_X16);
  } else if (E.prob == CC_GRID) {
    CC_WALK_LAUNCH(true, CC_FMT_V1);   // grid cells have n <= 8
  } else if (wide) {
    CC_WALK_LAUNCH(false, CC_FMT_WIDE);
  } else {
    CC_WALK_LAUNCH(false, CC_FMT_V1);
  }
#undef CC_WALK_LAUNCH
  return hipGetLastError();
}

// cnt += delta over the balls of nwords validated words; no wait for
// batches of at most ASYNC_UPDATE_MAX words
template <int KIND>
int update_inner(const long long *words, int nwords, int delta) {
  if (nwords <= 0) return 0;
  E.cnt_ver++;
  if (delta > 0) E.plus_words += nwords;
  for (int off = 0; off < nwords; off += CODECAP) {
    const int nb = std::min(nwords - off, CODECAP);
    if (nb <= ASYNC_UPDATE_MAX) {
      for (int o2 = 0; o2 < nb; o2 += NV_WORDS) {
        const int k = std::min(nb - o2, NV_WORDS);
        WordsArg wa;
        memset(&wa, 0, sizeof(wa));
        memcpy(wa.w, words + off + o2, (size_t)k * 8);
        CK((launch_walk<KIND>(wa, nullptr, k, pick_split(k), delta, 0, 0,
                              nullptr)));
      }
    } else {
      WordsArg wa;
      memset(&wa, 0, sizeof(wa));
      if (int rc = stage_words(words + off, nb)) return rc;
      CK((launch_walk<KIND>(wa, E.d_words, nb, pick_split(nb), delta, 0, 0,
                            nullptr)));
      CK(hipStreamSynchronize(E.st));
    }
  }
  return 0;
}

// Z = {x : pred(cnt[x])} listed for the current cnt (reference dct_hc.cu:
// 607-646 for the caching contract).  Returns 1 when the list is valid, 0
// when Z has more than lim entries (the caller walks), < 0 on error.  The
// predicates listed shrink under placements, so a listing that overflowed
// by F words still overflows while F - (words placed since) x |B| > lim:
// that case skips the scan.
int ensure_zlist(int pmode, int target, long long lim) {
  if (E.z_mode == pmode && E.z_target == target && E.z_mu == E.mu &&
      E.z_ver == E.cnt_ver)
    return 1;
  if (E.o_mode == pmode && E.o_target == target && E.o_mu == E.mu &&
      E.o_F - (E.plus_words - E.o_plus) * E.blen > lim)
    return 0;
  lim = std::min(lim, E.zcap);
  // size first: one vector-load counting pass (the count of cnt == target,
  // or for the deficient predicate the deficit, an upper bound on the
  // count) decides whether a listing can fit, so a state far from a cover
  // never pays for a listing of millions of words
  unsigned long long bound = 0;
  if (!((pmode == CC_DEF || (target == 0 && E.mu == 1)) && folded_deficit(&bound))) {
    if (int rc = cen_clear(CEN_ZL)) return rc;
    const Slabs sl = scan_slabs(E.X);
    if (pmode == CC_DEF)
      hipLaunchKernelGGL(k_cnt_census<CC_DEF>, dim3(sl.blocks), dim3(CC_BLOCK), 0,
                         E.st, E.cnt, E.X, sl.slab, 0, E.mu, cen(CEN_ZL), E.cen_stride);
    else
      hipLaunchKernelGGL(k_cnt_census<CC_EQ>, dim3(sl.blocks), dim3(CC_BLOCK), 0,
                         E.st, E.cnt, E.X, sl.slab, target, E.mu, cen(CEN_ZL),
                         E.cen_stride);
    CK(hipGetLastError());
    if (int rc = cen_fetch(CEN_ZL, false, &bound)) return rc;
  }
  if ((long long)bound > lim) {
    E.stats[3]++;
    E.o_mode = pmode; E.o_target = target; E.o_mu = E.mu;
    E.o_F = (long long)bound; E.o_plus = E.plus_words;
    E.z_mode = -1;
    return 0;
  }
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, E.st, E.d_scal + 2, 1);
  CK(hipGetLastError());
  const Dig dg = make_dig();
  if (pmode == CC_DEF)
    hipLaunchKernelGGL(k_zlist<CC_DEF>, dim3(pass_blocks(E.X)), dim3(CC_BLOCK), 0,
                       E.st, E.cnt, E.X, target, dg, E.zbuf, lim, E.d_scal + 2);
  else
    hipLaunchKernelGGL(k_zlist<CC_EQ>, dim3(pass_blocks(E.X)), dim3(CC_BLOCK), 0,
                       E.st, E.cnt, E.X, target, dg, E.zbuf, lim, E.d_scal + 2);
  CK(hipGetLastError());
  unsigned long long found = 0;
  if (int rc = fetch_scalar(2, &found)) return rc;
  E.stats[2]++;
  if ((long long)found <= lim) {
    E.z_mode = pmode; E.z_target = target; E.z_mu = E.mu;
    E.z_ver = E.cnt_ver; E.zn = (long long)found;
    return 1;
  }
  E.stats[3]++;
  E.o_mode = pmode; E.o_target = target; E.o_mu = E.mu;
  E.o_F = (long long)found; E.o_plus = E.plus_words;
  E.z_mode = -1;
  return 0;
}

// set side for this chunk?  Only shrinking predicates (deficient, cnt == 0),
// which are the small sets near a cover; by default when the walk would
// read at least max(|X| / 8, 2^21) cnt entries (reference dct_hc.cu:648-659)
bool want_set(int pmode, int target, int nb) {
  if (!set_ok() || !E.zbuf || E.set_policy == 0) return false;
  if (!(pmode == CC_DEF || (pmode == CC_EQ && target == 0))) return false;
  if (E.set_policy == 2) return true;
  const long long lim = E.set_min >= 0 ? E.set_min
                                       : std::max(E.X / 8, 1LL << 21);
  return (long long)nb * E.blen >= lim;
}

template <int PROB>
void launch_gather_set(const long long *dw, int nb, int32_t *dout) {
  const int gy = (nb + CCS_CB - 1) / CCS_CB;
  long long gx = std::max(1LL, (E.zn + 2047) / 2048);
  gx = std::min(gx, std::max(1LL, 4096LL / gy));
  const long long zper = (E.zn + gx - 1) / gx;
  hipLaunchKernelGGL(k_gather_set<PROB>, dim3((unsigned)gx, (unsigned)gy),
                     dim3(CC_BLOCK), 0, E.st, dw, nb, E.zbuf, E.zn, zper,
                     make_dig(), dout);
}

// the gather of nb device words into dout, set side or support walk, no
// final synchronisation; zeroed: dout already holds zeros
int gather_launch(int pmode, int target, int nb, const long long *dw,
                  int32_t *dout, bool zeroed) {
  const long long lim = std::min(E.blen * E.set_ratio, E.zcap);
  const int zs = want_set(pmode, target, nb) ? ensure_zlist(pmode, target, lim)
                                             : 0;
  if (zs < 0) return zs;
  if (zs == 1) {
    E.stats[0]++;
    if (E.zn > 32768) E.stats[7]++;
    if (!zeroed) {
      hipLaunchKernelGGL(k_fill32, dim3((nb + 255) / 256), dim3(256), 0, E.st,
                         dout, 0, (long long)nb);
      CK(hipGetLastError());
    }
    if (E.zn > 0) {
      switch (E.prob) {
        case CC_HAMMING: launch_gather_set<CC_HAMMING>(dw, nb, dout); break;
        case CC_LEE: launch_gather_set<CC_LEE>(dw, nb, dout); break;
        case CC_TORUS: launch_gather_set<CC_TORUS>(dw, nb, dout); break;
        default: launch_gather_set<CC_GRID>(dw, nb, dout); break;
      }
      CK(hipGetLastError());
    }
    return 0;
  }
  E.stats[1]++;
  const int split = pick_split(nb);
  if (split > 1 && !zeroed) {
    hipLaunchKernelGGL(k_fill32, dim3((nb + 255) / 256), dim3(256), 0, E.st,
                       dout, 0, (long long)nb);
    CK(hipGetLastError());
  }
  WordsArg wa;
  memset(&wa, 0, sizeof(wa));
  CK((launch_walk<CC_WALK_GATHER>(wa, dw, nb, split, 0, pmode, target, dout)));
  return 0;
}

int gather_impl(const long long *words, int nwords, int pmode, int target,
                int32_t *out) {
  if (nwords <= 0) return 0;
  if (!out) return -1;
  if (nwords <= NV_WORDS && E.h_out_dev && pick_split(nwords) == 1 &&
      !want_set(pmode, target, nwords)) {
    // a small walked gather with one workgroup per word: the words travel
    // in the kernel arguments and each sum is stored straight into the
    // pinned host buffer -- one launch and the wait, no copies
    E.stats[1]++;
    WordsArg wa;
    memset(&wa, 0, sizeof(wa));
    memcpy(wa.w, words, (size_t)nwords * 8);
    CK((launch_walk<CC_WALK_GATHER>(wa, nullptr, nwords, 1, 0, pmode, target,
                                    E.h_out_dev)));
    CK(hipStreamSynchronize(E.st));
    memcpy(out, E.h_out, (size_t)nwords * 4);
    return 0;
  }
  for (int off = 0; off < nwords; off += CODECAP) {
    const int nb = std::min(nwords - off, CODECAP);
    if (int rc = stage_words(words + off, nb)) return rc;
    if (int rc = gather_launch(pmode, target, nb, E.d_words, E.d_out, false))
      return rc;
    CK(hipMemcpyAsync(E.h_out, E.d_out, (size_t)nb * 4, hipMemcpyDeviceToHost,
                      E.st));
    CK(hipStreamSynchronize(E.st));
    memcpy(out + off, E.h_out, (size_t)nb * 4);
  }
  return 0;
}

// host-side validation of an update (reference dct_hc.cu:567-577)
bool update_ok(const long long *words, int nwords, int delta,
               std::unordered_map<long long, int> &batch) {
  if (!valid_words(words, nwords) || (delta != 1 && delta != -1)) return false;
  if (delta == 1 && nwords > MAX_CODE - E.active_size) return false;
  for (int i = 0; i < nwords; i++) ++batch[words[i]];
  if (delta == -1)
    for (const auto &kv : batch) {
      const auto &act = active_map();
      auto it = act.find(kv.first);
      if (it == act.end() || it->second < kv.second) return false;
    }
  return true;
}

void apply_batch(const std::unordered_map<long long, int> &batch, int delta,
                 int nwords) {
  auto &act = active_mut();
  for (const auto &kv : batch) {
    int &c = act[kv.first];
    c += delta * kv.second;
    if (c == 0) act.erase(kv.first);
  }
  E.active_size += delta * nwords;
}

template <typename T>
int dalloc(T **p, size_t bytes) {
  if (hipMalloc((void **)p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return -1;
  }
  E.bytes_dev += (long long)bytes;
  return 0;
}

void release_all() {
  if (E.st) (void)hipStreamSynchronize(E.st);
  if (E.A[0]) (void)hipFree(E.A[0]);
  if (E.lay) (void)hipFree(E.lay);
  E.lay = nullptr;
  for (int d = 0; d < CC_MAXL; d++) { E.A[d] = nullptr; E.H[d] = nullptr; }
  void *bufs[] = {E.cnt, E.pat, E.pat2, E.d_code, E.d_cw, E.d_toff,
                  E.d_words, E.d_out,
                  E.d_scal, E.d_cen, E.d_hist, E.d_eidx64, E.d_eidx32, E.d_eval,
                  E.d_bcnt, E.d_boff, E.zbuf, E.lbuf, E.ldrop, E.d_swap,
                  E.d_sres};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  E.cnt = nullptr; E.pat = E.pat2 = nullptr; E.d_code = E.d_words = nullptr;
  E.d_cw = nullptr; E.d_cm = nullptr; E.d_toff = nullptr; E.toff_cap = 0;
  E.toff_cur = nullptr;
  E.d_out = nullptr; E.d_scal = nullptr; E.d_hist = nullptr; E.d_cen = nullptr;
  E.d_eidx64 = nullptr; E.d_eidx32 = nullptr; E.d_eval = nullptr;
  E.d_bcnt = E.d_boff = nullptr;
  E.zbuf = nullptr; E.lbuf = nullptr; E.ldrop = nullptr;
  E.d_swap = nullptr; E.d_sres = nullptr;
  E.zcap = E.lcap = 0;
  E.set_ok1 = E.set_ok2 = 0;
  E.z_mode = E.o_mode = -1;
  if (E.h_swap) { (void)hipHostFree(E.h_swap); E.h_swap = nullptr; }
  if (E.h_sres) { (void)hipHostFree(E.h_sres); E.h_sres = nullptr; }
  if (E.h_words) { (void)hipHostFree(E.h_words); E.h_words = nullptr; }
  if (E.h_out) { (void)hipHostFree(E.h_out); E.h_out = nullptr; }
  E.h_out_dev = nullptr;
  if (E.h_scal) { (void)hipHostFree(E.h_scal); E.h_scal = nullptr; }
  if (E.h_cen) { (void)hipHostFree(E.h_cen); E.h_cen = nullptr; }
  for (auto &e : E.ev)
    if (e) { (void)hipEventDestroy(e); e = nullptr; }
  if (E.st) { (void)hipStreamDestroy(E.st); E.st = nullptr; }
  E.uploaded.clear();
  E.active.clear();
  E.code_raw.clear();
  E.code_dev_ok = false;
  E.active_shared = false;
  E.cp_ok = false;
  E.active_size = 0;
  E.codeM = 0;
  E.live = false;
  E.mu = 1;
  E.fmt = 0;
  E.nswept = 0;
  E.pfield = PPlan();
  g_cuts = 0;
}

}  // namespace

namespace {
template <typename IDX>
long long extract_impl(int32_t thr, IDX *dev_idx, IDX *out_idx, int32_t *out_val,
                       int cap) {
  if (!E.live || cap < 0) return -1;
  if (cap > E.ecap) cap = E.ecap;
  const unsigned nb = (unsigned)E.nchunks;
  hipLaunchKernelGGL(k_extract_count, dim3(nb), dim3(CC_BLOCK), 0, E.st,
                     E.A[0], E.X, thr, EXTRACT_CH, E.d_bcnt);
  CK(hipGetLastError());
  hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(1024), 0, E.st, E.d_bcnt,
                     E.nchunks, E.d_boff);
  CK(hipGetLastError());
  hipLaunchKernelGGL((k_extract_write<IDX>), dim3(nb), dim3(CC_BLOCK), 0, E.st,
                     E.A[0], E.X, thr, EXTRACT_CH, E.d_boff, (long long)cap,
                     dev_idx, E.d_eval);
  CK(hipGetLastError());
  long long found = 0;
  CK(hipMemcpyAsync(E.h_scal, E.d_boff + E.nchunks, 8, hipMemcpyDeviceToHost,
                    E.st));
  CK(hipStreamSynchronize(E.st));
  found = (long long)E.h_scal[0];
  const long long m = std::min<long long>(found, cap);
  if (m > 0) {
    if (!out_idx || !out_val) return -1;
    CK(hipMemcpy(out_idx, dev_idx, (size_t)m * sizeof(IDX), hipMemcpyDeviceToHost));
    CK(hipMemcpy(out_val, E.d_eval, (size_t)m * 4, hipMemcpyDeviceToHost));
  }
  return found;
}
}  // namespace

#pragma GCC visibility push(default)
extern "C" {

// ---- lifecycle (reference dct_hc.cu:224-428) -------------------------------
static int init_cell(int problem, int n, const int *ax, int R,
                     const uint64_t *ballpat, long long balllen, int tfmt,
                     int extract_cap, long long *bytes, int use_owner) {
  const int nmax = tfmt == CC_FMT_V1 ? CC_MAXN_V1 : CC_MAXN;
  if (n < 2 || n > nmax || R < 1 || !ax || !ballpat || balllen < 1 ||
      balllen > INT_MAX || extract_cap < 1)
    return -1;
  // the long-axis table reaches L-inf axes up to CC_MAXAX_X (the runtime
  // window of the sweeps holds R <= 3); the nibble formats stop at 16
  const int axmax = tfmt == CC_FMT_X16 ? CC_MAXAX_X : CC_MAXAX;
  bool longax = false;
  for (int i = 0; i < n; i++) longax = longax || ax[i] > CC_MAXAX;
  int L;
  if (problem == CC_HAMMING) {
    if (R > CC_MAXR) return -1;
    for (int i = 0; i < n; i++)
      if (ax[i] < 2 || ax[i] > 10) return -1;
    L = R + 1;
  } else if (problem == CC_TORUS) {
    for (int i = 0; i < n; i++)
      if (ax[i] < 3 || ax[i] > axmax || R > (ax[i] - 1) / 2) return -1;
    if (n > CC_MAXPAT || (longax && R > 3)) return -1;
    L = 1;
  } else if (problem == CC_LEE) {
    if (R > CC_MAXR) return -1;
    for (int i = 0; i < n; i++)
      if (ax[i] < 2 || ax[i] > CC_MAXAX) return -1;
    L = R + 1;
  } else if (problem == CC_GRID) {
    for (int i = 0; i < n; i++)
      if (ax[i] < 2 || ax[i] > axmax) return -1;
    if (n > CC_MAXPAT || (longax && R > 3)) return -1;
    L = 1;
  } else {
    return -1;
  }
  long long X = 1;
  for (int i = 0; i < n; i++) {
    if (X > (1LL << 40) / ax[i]) return -2;
    X *= ax[i];
  }
  if (E.live) release_all();
  E.prob = problem; E.n = n; E.R = R; E.L = L; E.X = X; E.mu = 1;
  E.pw[0] = 1;
  for (int i = 0; i < n; i++) {
    E.ax[i] = ax[i];
    E.pw[i + 1] = E.pw[i] * ax[i];
  }
  if (!table_ok(ballpat, balllen, tfmt)) return -1;
  E.tfmt = tfmt;
  E.use_owner = use_owner;
  E.bytes_dev = 0;
  memset(E.stats, 0, sizeof(E.stats));
  {
    int dev = 0, mx = 0;
    CK(hipGetDevice(&dev));
    CK(hipDeviceGetAttribute(&mx, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    // (less the packed kernels' static census LDS)
    E.lds_max = (mx > 0 ? mx : 64 * 1024) - 256;
    int ncu = 0;
    CK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    E.ncu = ncu > 0 ? ncu : 256;
  }
  const char *lb = getenv("CC_SWEEP_LDS");
  E.lds_budget = lb ? std::max(4096, atoi(lb)) : 64 * 1024;
  E.lds_budget = std::min(E.lds_budget, E.lds_max);
  const char *sg = getenv("CC_SWEEP_GRID");
  E.grid = sg ? atoi(sg) : 0;   // measured: one workgroup per tile is best
  const char *sk = getenv("CC_SWEEP_KERNEL");
  E.generic_only = sk && !strcmp(sk, "generic");
  {
    // the plan built below depends on it (a cell never inherits the last
    // cell's tuned value)
    const char *pcm = getenv("CC_COMPACT");
    E.compact = pcm ? atoi(pcm) : 1;
  }
  CK(hipStreamCreateWithFlags(&E.st, hipStreamNonBlocking));
  // A[0] (map / code indicator) on its own; layers 1..R as one block whose
  // planes are 64-word aligned, viewed as int32 planes A[1..R] or, in 16-bit
  // mode, as L uint16 planes H[0..R] (2 L pitch <= 4 R pitch for R >= 1)
  E.pitch = (X + 63) & ~63LL;
  if (dalloc(&E.A[0], (size_t)X * 4)) { release_all(); return -3; }
  if (L > 1) {
    if (dalloc(&E.lay, (size_t)(L - 1) * E.pitch * 4)) { release_all(); return -3; }
    for (int d = 1; d < L; d++) E.A[d] = E.lay + (size_t)(d - 1) * E.pitch;
    for (int d = 0; d < L; d++)
      E.H[d] = (uint16_t *)E.lay + (size_t)d * E.pitch;
  }
  E.cnt_len = (X + 1) & ~1LL;
  if (dalloc(&E.cnt, (size_t)E.cnt_len * 2)) { release_all(); return -4; }
  hipLaunchKernelGGL(k_fill16, dim3(pass_blocks(E.cnt_len)), dim3(CC_BLOCK), 0,
                     E.st, E.cnt, (uint16_t)0, E.cnt_len);
  CK(hipGetLastError());
  E.blen = balllen;
  if (dalloc(&E.pat, (size_t)balllen * 8)) { release_all(); return -3; }
  CK(hipMemcpy(E.pat, ballpat, (size_t)balllen * 8, hipMemcpyHostToDevice));
  if (dalloc(&E.d_code, (size_t)CODECAP * 8) ||
      dalloc(&E.d_cw, (size_t)CODECAP * 12) ||
      dalloc(&E.d_words, (size_t)CODECAP * 8) ||
      dalloc(&E.d_out, (size_t)CODECAP * 4) || dalloc(&E.d_scal, 64) ||
      dalloc(&E.d_cen, (size_t)CEN_N * CC_NSLOT * CEN_MAXSTRIDE * 8) ||
      dalloc(&E.d_hist, 4096 * 4)) {
    release_all();
    return -3;
  }
  E.ecap = extract_cap;
  E.nchunks = (X + EXTRACT_CH - 1) / EXTRACT_CH;
  if (dalloc(&E.d_eidx64, (size_t)extract_cap * 8) ||
      dalloc(&E.d_eidx32, (size_t)extract_cap * 4) ||
      dalloc(&E.d_eval, (size_t)extract_cap * 4) ||
      dalloc(&E.d_bcnt, (size_t)E.nchunks * 8) ||
      dalloc(&E.d_boff, (size_t)(E.nchunks + 1) * 8)) {
    release_all();
    return -3;
  }
  CK(hipHostMalloc((void **)&E.h_words, (size_t)CODECAP * 8, hipHostMallocDefault));
  CK(hipHostMalloc((void **)&E.h_swap, SWAP_IN_BYTES, hipHostMallocDefault));
  CK(hipHostMalloc((void **)&E.h_sres, SWAP_RES_BYTES, hipHostMallocDefault));
  CK(hipHostMalloc((void **)&E.h_out, (size_t)std::max(CODECAP, 4096) * 4,
                   hipHostMallocDefault));
  if (hipHostGetDevicePointer((void **)&E.h_out_dev, E.h_out, 0) != hipSuccess) {
    (void)hipGetLastError();
    E.h_out_dev = nullptr;     // small gathers then copy back like large ones
  }
  CK(hipHostMalloc((void **)&E.h_scal, 64, hipHostMallocDefault));
  CK(hipHostMalloc((void **)&E.h_cen, (size_t)CEN_N * CC_NSLOT * 8,
                   hipHostMallocDefault));
  {
    const char *cs = getenv("CC_SLOT_STRIDE");
    const int st = cs ? atoi(cs) : 32;
    E.cen_stride = st >= 1 && st <= CEN_MAXSTRIDE ? st : 32;
  }
  for (auto &e : E.ev) CK(hipEventCreate(&e));
  const char *pe = getenv("DCT_TIMING");
  E.prof = pe ? atoi(pe) : 0;
  if (int rc = zero_scalars(8)) return rc;
  E.trans_last = 0;
  {
    // set-side gathers and the swap fast path (exact only on a proven ball)
    const char *sp = getenv("CC_SET_POLICY");
    E.set_policy = sp ? atoi(sp) : 1;
    const char *sr = getenv("CC_SET_RATIO");
    E.set_ratio = sr ? std::max(1, atoi(sr)) : 2;
    const char *sm = getenv("CC_SET_MIN");
    E.set_min = sm ? atoll(sm) : -1;
    int amax = 2;
    for (int i = 0; i < n; i++) amax = std::max(amax, ax[i]);
    E.dbits = 1;
    while ((1 << E.dbits) < amax) E.dbits++;
    const bool ball = n * E.dbits <= 64 && table_is_ball(ballpat, balllen, tfmt);
    E.set_ok1 = E.set_policy != 0 && ball;
    E.set_ok2 = 0;
    E.region_walk = ball && E.prob == CC_HAMMING &&
                    (tfmt == CC_FMT_V1 || tfmt == CC_FMT_WIDE);
    const char *rw = getenv("CC_REGION_WALK");
    E.region_walk_mode = rw ? atoi(rw) : 1;
    E.cnt_ver = 1;
    E.map_fold = false;
    E.cnt_fold_pending = false;
    E.fold_cnt_ver = 0;
    const char *sb = getenv("CC_SCAN_BLOCKS");
    E.scan_blocks = sb ? std::max(64, std::min(65536, atoi(sb))) : 2048;
    const char *fo = getenv("CC_FOLD");
    E.fold = fo ? (atoi(fo) != 0) : 0;
    E.plus_words = 0;
    E.z_mode = E.o_mode = -1;
    if (E.set_ok1) {
      E.zcap = std::min(X, std::min(1LL << 27,
                                    std::max(4096LL, balllen * E.set_ratio)));
      E.lcap = std::min(1LL << 22, std::max(4096LL, 64 * balllen));
      if (dalloc(&E.zbuf, (size_t)E.zcap * 8) ||
          dalloc(&E.lbuf, (size_t)E.lcap * 8) ||
          dalloc(&E.ldrop, (size_t)E.lcap * 4)) {
        release_all();
        return -3;
      }
    }
    if (dalloc(&E.d_swap, SWAP_IN_BYTES) || dalloc(&E.d_sres, SWAP_RES_BYTES)) {
      release_all();
      return -3;
    }
  }
  {
    // exact ball size: coefficients 0..R of prod_j (per-axis enumerator)
    std::vector<double> poly(R + 1, 0.0);
    poly[0] = 1;
    for (int i = 0; i < n; i++) {
      std::vector<double> nx(R + 1, 0.0);
      for (int w = 0; w <= R; w++) {
        if (poly[w] == 0) continue;
        nx[w] += poly[w];
        for (int dl = 1; dl < ax[i]; dl++) {
          const int c = std::min(dl, ax[i] - dl);
          const int e = problem == CC_LEE ? c : 1;
          if (w + e <= R) nx[w + e] += poly[w];
        }
      }
      poly = nx;
    }
    E.ball = 0;
    for (double v : poly) E.ball += v;
    const char *l16 = getenv("CC_LAYER16");
    const bool on = L > 1 && !(l16 && !strcmp(l16, "0"));
    E.l16_field = on && E.ball <= 65535.0;
    E.l16_count = on;          // layer values never exceed M <= 65,535
  }
  build_plan();
  {
    const char *pe = getenv("CC_PACKED");
    E.packed = !(pe && !strcmp(pe, "0")) && !E.generic_only;
    const char *pc = getenv("CC_PK_COL");
    E.pk_col = pc ? atoi(pc) : 2;
    const char *pf = getenv("CC_PK_FIRST");
    E.pk_first = pf ? atoi(pf) : 1;
    const char *pnt = getenv("CC_PK_NT");
    // default: every sweep after the first (its words are read right away
    // by the next sweep, which non-temporal stores measured slower for)
    E.pk_nt = pnt ? (int)strtol(pnt, nullptr, 0) : ~1;
    const char *pwd = getenv("CC_PK_WIDE");
    E.pk_wide = pwd ? atoi(pwd) : 0;
    const char *plc = getenv("CC_PK_LCMP");
    E.pk_lcmp = plc ? atoi(plc) : 0;
    const char *pn = getenv("CC_PK_NARROW");
    E.pk_narrow = pn ? atoi(pn) : 0;
    const char *pcm = getenv("CC_COMPACT");
    E.compact = pcm ? atoi(pcm) : 1;
    const char *psc = getenv("CC_SPARSE_CODE");
    E.sparse = psc ? atoi(psc) : 1;
    // an explicit schedule setting turns the automatic timing off
    const char *pa = getenv("CC_AUTOTUNE");
    E.autotune = pa ? atoi(pa) : 1;
    for (const char *k : {"CC_PACKED", "CC_SWEEP_KERNEL", "CC_SWEEP_LDS",
                          "CC_PK_COL", "CC_PK_FIRST", "CC_PK_NARROW", "CC_LAYER16",
                          "CC_COMPACT"})
      if (getenv(k) && !pa) E.autotune = 0;
    E.tuned = 0;
    g_cuts = 0;
    E.pfield = pk_plan(1.0, 1e300, true, false);
  }
  E.uploaded.clear();
  E.active.clear();
  E.code_raw.clear();
  E.code_dev_ok = false;
  E.active_shared = false;
  E.cp_ok = false;
  E.active_size = 0;
  E.codeM = 0;
  E.fmt = 0;
  E.live = true;
  CK(hipStreamSynchronize(E.st));
  if (bytes) *bytes = E.bytes_dev;
  return 0;
}

int dct_init3(int problem, int n, const int *ax, int R,
              const uint64_t *ballpat, long long balllen, int extract_cap,
              long long *bytes, int layers_mode, int cnt_mode, int use_owner) {
  (void)layers_mode;
  (void)cnt_mode;
  return init_cell(problem, n, ax, R, ballpat, balllen, CC_FMT_V1,
                   extract_cap, bytes, use_owner);
}

// additive: the same cell with a wide-format table (10-bit pos<<4|delta
// fields, at most 6 per pattern), for Hamming / Lee spaces of up to 40 axes
int dct_init3w(int problem, int n, const int *ax, int R,
               const uint64_t *widepat, long long balllen, int extract_cap,
               long long *bytes, int layers_mode, int cnt_mode,
               int use_owner) {
  (void)layers_mode;
  (void)cnt_mode;
  return init_cell(problem, n, ax, R, widepat, balllen, CC_FMT_WIDE,
                   extract_cap, bytes, use_owner);
}

// additive: the same cell with a long-axis table (16-bit pos<<10|delta
// fields, at most 4 per pattern), for king tori and grid boxes with axes
// longer than 16 (up to 1024, R <= 3)
int dct_init3x(int problem, int n, const int *ax, int R,
               const uint64_t *xpat, long long balllen, int extract_cap,
               long long *bytes, int layers_mode, int cnt_mode,
               int use_owner) {
  (void)layers_mode;
  (void)cnt_mode;
  return init_cell(problem, n, ax, R, xpat, balllen, CC_FMT_X16,
                   extract_cap, bytes, use_owner);
}

int dct_init2(int q, int n, int R, const uint64_t *ballpat, long long balllen,
              int extract_cap, long long *bytes, int layers_mode, int cnt_mode,
              int use_owner) {
  int ax[CC_MAXN_V1];
  if (n < 2 || n > CC_MAXN_V1) return -1;
  for (int i = 0; i < n; i++) ax[i] = q;
  return dct_init3(CC_HAMMING, n, ax, R, ballpat, balllen, extract_cap, bytes,
                   layers_mode, cnt_mode, use_owner);
}

int dct_init(int q, int n, int R, const uint64_t *ballpat, long long balllen,
             int extract_cap, long long *bytes) {
  return dct_init2(q, n, R, ballpat, balllen, extract_cap, bytes, 0, 0, 0);
}

int dct_free(void) {
  if (E.live) release_all();
  E.mu = 1;
  return 0;
}

long long dct_host_bytes(void) { return 0; }

// ---- mu-fold covering (reference dct_hc.cu:210-222) -----------------------
int dct_set_mu(int mu) {
  if (mu < 1 || mu > 65535) return -1;
  E.mu = mu;
  return 0;
}

int dct_get_mu(void) { return E.mu; }

// ---- solution and transforms (reference dct_hc.cu:452-521) ----------------
int dct_set_code(const long long *idx, int M) {
  if (!E.live || !valid_words(idx, M)) return -1;
  E.codeM = M;
  // the represented multiset may still stand for the old pending one: hand
  // it over before the pending one is rebuilt
  if (E.active_shared) {
    std::swap(E.active, E.uploaded);
    E.active_shared = false;
  }
  E.uploaded.assign(idx, M);
  // the device copy of the list is made only for a transform that reads
  // the code through the indicator plane (ensure_code_dev)
  E.code_raw.assign(idx, idx + M);
  E.code_dev_ok = false;
  return 0;
}

// init_mode: 0 = code multiset, 1 = [cnt < mu] (deficient), 2 = [1 <= cnt <=
// mu] (critical); out_mode: 1 = layer sum -> A0, 2 = clamped sum -> cnt
// (init_mode 0 only).  At mu = 1 the indicators are [cnt==0] / [cnt==1].
int dct_transform(int init_mode, int out_mode) {
  if (!E.live) return -1;
  if (init_mode < 0 || init_mode > 2 || out_mode < 1 || out_mode > 2 ||
      (out_mode == 2 && init_mode != 0))
    return -1;
  int in_kind = CC_IN_PRED, pmode = CC_DEF;
  const int out_kind = out_mode == 1 ? CC_OUT_MAP : CC_OUT_CNT;
  PPlan code_plan;
  if (init_mode == 0) {
    in_kind = CC_IN_CODE;
    code_plan = pk_code_plan(out_kind == CC_OUT_CNT);
    if (code_plan.ok && (code_plan.sw[0].kern == 4 || code_plan.sw[0].kern == 5)) {
      // the persistent first sweep seeds from the code's word list: no
      // plane fill, no indicator plane read
      if (int rc = stage_sparse_code(code_plan.sw[0])) return rc;
    } else {
      hipLaunchKernelGGL(k_fill32, dim3(pass_blocks(E.X)), dim3(CC_BLOCK), 0,
                         E.st, E.A[0], 0, E.X);
      CK(hipGetLastError());
      if (E.codeM > 0) {
        if (int rc = ensure_code_dev()) return rc;
        hipLaunchKernelGGL(k_scatter_code, dim3((E.codeM + 255) / 256), dim3(256),
                           0, E.st, E.A[0], E.d_code, E.codeM);
        CK(hipGetLastError());
      }
    }
  } else if (init_mode == 2) {
    pmode = CC_CRIT;
  }
  E.layers_live = 0;
  if (in_kind == CC_IN_PRED && E.autotune && !E.tuned && E.X >= (1LL << 22)) {
    // big cell, first gain / loss transform: time the candidate schedules
    // once (the transform below overwrites the map they leave)
    E.tuned = 1;
    if (int rc = dct_autotune(2, nullptr); rc < 0) return rc;
  }
  const PPlan &pp = in_kind == CC_IN_PRED ? E.pfield : code_plan;
  if (pp.ok) {
    const int plo = pmode == CC_CRIT ? 1 : 0;
    const int phi = pmode == CC_CRIT ? E.mu : E.mu - 1;
    if (int rc = run_ptransform(pp, in_kind, out_kind, plo, phi)) return rc;
  } else if (int rc = run_transform(in_kind, out_kind, pmode)) {
    return rc;
  }
  CK(hipStreamSynchronize(E.st));
  if (E.prof)
    for (int i = 0; i < E.nswept; i++)
      CK(hipEventElapsedTime(&E.sweep_ms[i], E.ev[2 * i], E.ev[2 * i + 1]));
  if (out_mode == 2) {
    E.active_shared = true;          // the represented multiset is the pending one
    E.active_size = E.codeM;
    E.cnt_ver++;
    E.o_mode = -1;
    E.fold_cnt_ver = E.cnt_fold_pending ? E.cnt_ver : 0;
    E.cnt_fold_pending = false;
  }
  return 0;
}

// ---- incremental updates and exact gathers (reference dct_hc.cu:533-799) --
int dct_ball_update(const long long *words, int nwords, int delta) {
  if (!E.live) return -1;
  std::unordered_map<long long, int> batch;
  if (!update_ok(words, nwords, delta, batch)) return -1;
  if (int rc = update_inner<CC_WALK_UPDATE>(words, nwords, delta)) return rc;
  apply_batch(batch, delta, nwords);
  return 0;
}

int dct_ball_gather(const long long *words, int nwords, int target,
                    int32_t *out) {
  if (!E.live) return -1;
  if (nwords <= 0) return 0;
  if (!words) return -1;
  for (int i = 0; i < nwords; i++)
    if (words[i] < 0 || words[i] >= E.X) return -1;
  return gather_impl(words, nwords, CC_EQ, target, out);
}

int dct_ball_gather_mu(const long long *words, int nwords, int role,
                       int32_t *out) {
  if (!E.live) return -1;
  if (nwords <= 0) return 0;
  if (role < 0 || role > 1 || !words) return -1;
  for (int i = 0; i < nwords; i++)
    if (words[i] < 0 || words[i] >= E.X) return -1;
  return gather_impl(words, nwords, role == 0 ? CC_DEF : CC_CRIT, 0, out);
}

int dct_ball_update_delta(const long long *words, int nwords, int delta,
                          long long *dd) {
  if (!dd) return -1;
  *dd = 0;
  if (!E.live) return -1;
  std::unordered_map<long long, int> batch;
  if (!update_ok(words, nwords, delta, batch)) return -1;
  if (nwords == 0) return 0;
  if (int rc = update_inner<CC_WALK_TRANS>(words, nwords, delta)) return rc;
  unsigned long long tr;
  if (int rc = fetch_scalar(1, &tr)) return rc;
  const long long diff = (long long)(tr - E.trans_last);
  E.trans_last = tr;
  *dd = delta > 0 ? -diff : diff;
  apply_batch(batch, delta, nwords);
  return 0;
}

// read-only batched swap evaluation (reference dct_hc.cu:868-977): the exact
// drop / gather / undo sequence, cnt restored before returning
int dct_swap_eval(const long long *drops, int k, const long long *cand,
                  const long long *off, int32_t *gain, int32_t *loss) {
  if (!E.live) return -1;
  if (k <= 0) return 0;
  if (k > CODECAP || !off || off[0] != 0 || !drops || !loss) return -1;
  for (int i = 0; i < k; i++)
    if (off[i + 1] < off[i]) return -1;
  const long long nc = off[k];
  if (nc > CODECAP || (nc && (!cand || !gain))) return -1;
  if (!valid_words(drops, k) || !valid_words(cand, (int)nc)) return -1;
  for (int i = 0; i < k; i++)
    if (active_map().find(drops[i]) == active_map().end()) return -1;
  const int lmode = E.mu == 1 ? CC_EQ : CC_CRIT, lt = E.mu == 1 ? 1 : 0;
  const int gmode = E.mu == 1 ? CC_EQ : CC_DEF;
  if (set_ok() && E.lbuf) {
    // fast path: loss and drop lists by one walk per drop, the candidates'
    // gains on the current cnt, then the per-drop corrections; cnt is never
    // modified.  One upload, one cleared result region, one download.
    const size_t in_drop = (size_t)nc * 8, in_cdrop = (size_t)(nc + k) * 8;
    const size_t in_bytes = in_cdrop + (size_t)nc * 4;
    const size_t r_loss = (size_t)nc * 4;
    const size_t r_cnt = ((size_t)(nc + k) * 4 + 7) & ~(size_t)7;
    const size_t r_bytes = r_cnt + 8;
    if (nc) memcpy(E.h_swap, cand, (size_t)nc * 8);
    memcpy(E.h_swap + in_drop, drops, (size_t)k * 8);
    int32_t *hc = (int32_t *)(E.h_swap + in_cdrop);
    for (int i = 0; i < k; i++)
      for (long long j = off[i]; j < off[i + 1]; j++) hc[j] = i;
    const long long *dcand = (const long long *)E.d_swap;
    const long long *ddrop = (const long long *)(E.d_swap + in_drop);
    const int32_t *dcdrop = (const int32_t *)(E.d_swap + in_cdrop);
    int32_t *dgain = (int32_t *)E.d_sres;
    int32_t *dloss = (int32_t *)(E.d_sres + r_loss);
    unsigned long long *dlcnt = (unsigned long long *)(E.d_sres + r_cnt);
    CK(hipMemcpyAsync(E.d_swap, E.h_swap, in_bytes, hipMemcpyHostToDevice, E.st));
    hipLaunchKernelGGL(k_fill32, dim3((unsigned)((r_bytes / 4 + 255) / 256)),
                       dim3(256), 0, E.st, (int32_t *)E.d_sres, 0,
                       (long long)(r_bytes / 4));
    CK(hipGetLastError());
    const Geom g = make_geom();
    const Dig dg = make_dig();
    const int split = pick_split(k);
    const dim3 grid((unsigned)k * split);
    const bool wide = E.fmt == 0 && E.tfmt == CC_FMT_WIDE;
#define CC_DROP(GR, FM)                                                       \
    hipLaunchKernelGGL((k_drop_walk<GR, FM>), grid, dim3(CC_BLOCK), 0, E.st, g, \
                       dg, ddrop, split, E.cnt, walk_table(), E.blen, dloss,    \
                       E.lbuf, E.ldrop, E.lcap, dlcnt)
    const bool x16 = E.fmt == 0 && E.tfmt == CC_FMT_X16;
    if (x16 && E.prob == CC_GRID) CC_DROP(true, CC_FMT_X16);
    else if (x16) CC_DROP(false, CC_FMT_X16);
    else if (E.prob == CC_GRID) CC_DROP(true, CC_FMT_V1);
    else if (wide) CC_DROP(false, CC_FMT_WIDE);
    else CC_DROP(false, CC_FMT_V1);
#undef CC_DROP
    CK(hipGetLastError());
    if (nc) {
      if (int rc = gather_launch(gmode, 0, (int)nc, dcand, dgain, true))
        return rc;
      const int cs = (int)std::max(1LL, std::min(64LL, 8192 / nc));
      const dim3 g2((unsigned)(nc * cs));
#define CC_CORR(P)                                                             \
      hipLaunchKernelGGL(k_swap_corr<P>, g2, dim3(CC_BLOCK), 0, E.st, dcand,    \
                         dcdrop, cs, E.lbuf, E.ldrop, E.lcap, dlcnt, dg, dgain)
      switch (E.prob) {
        case CC_HAMMING: CC_CORR(CC_HAMMING); break;
        case CC_LEE: CC_CORR(CC_LEE); break;
        case CC_TORUS: CC_CORR(CC_TORUS); break;
        default: CC_CORR(CC_GRID); break;
      }
#undef CC_CORR
      CK(hipGetLastError());
    }
    CK(hipMemcpyAsync(E.h_sres, E.d_sres, r_bytes, hipMemcpyDeviceToHost, E.st));
    CK(hipStreamSynchronize(E.st));
    const unsigned long long lfound = *(const unsigned long long *)(E.h_sres + r_cnt);
    if ((long long)lfound <= E.lcap) {
      memcpy(loss, E.h_sres + r_loss, (size_t)k * 4);
      if (nc) memcpy(gain, E.h_sres, (size_t)nc * 4);
      E.stats[4]++;
      return 0;
    }
  }
  for (int i = 0; i < k; i++) {
    if (int rc = gather_impl(drops + i, 1, lmode, lt, loss + i)) return rc;
    const long long ni = off[i + 1] - off[i];
    if (!ni) continue;
    if (int rc = update_inner<CC_WALK_UPDATE>(drops + i, 1, -1)) return rc;
    int rc = gather_impl(cand + off[i], (int)ni, gmode, 0, gain + off[i]);
    int rc2 = update_inner<CC_WALK_UPDATE>(drops + i, 1, +1);
    if (rc) return rc;
    if (rc2) return rc2;
  }
  CK(hipStreamSynchronize(E.st));
  E.stats[5]++;
  return 0;
}

int dct_stats(long long *out, int n) {
  if (!out) return -1;
  if (n > 8) n = 8;
  for (int i = 0; i < n; i++) out[i] = E.stats[i];
  return n;
}

// owner-trick loss (reference dct_hc.cu:994-1022): rebuild cnt from the full
// solution, then per-codeword critical counts.  A word is critical to
// exactly the codewords covering it, so the critical gather reproduces the
// owner scan's integers (and is what the reference itself runs for mu > 1).
int dct_loss_owner(const long long *words, int M, int32_t *out) {
  if (!E.live || !E.use_owner) return -1;
  if (!valid_words(words, M) || (M && !out)) return -2;
  hipLaunchKernelGGL(k_fill16, dim3(pass_blocks(E.cnt_len)), dim3(CC_BLOCK), 0,
                     E.st, E.cnt, (uint16_t)0, E.cnt_len);
  CK(hipGetLastError());
  E.cnt_ver++;
  E.o_mode = -1;
  if (int rc = update_inner<CC_WALK_UPDATE>(words, M, +1)) return rc;
  E.active_shared = false;
  E.active.assign(words, M);
  E.active_size = M;
  if (M == 0) return sync_stream();
  return gather_impl(words, M, E.mu == 1 ? CC_EQ : CC_CRIT, E.mu == 1 ? 1 : 0,
                     out);
}

// ---- reductions and queries (reference dct_hc.cu:979-1157) ---------------
long long dct_count_eq(int target) {
  if (!E.live) return -1;
  unsigned long long fr;
  if (target == 0 && E.mu == 1 && folded_deficit(&fr)) return (long long)fr;
  if (int rc = cen_clear(CEN_SCAN)) return rc;
  const Slabs sl = scan_slabs(E.X);
  hipLaunchKernelGGL(k_cnt_census<CC_EQ>, dim3(sl.blocks), dim3(CC_BLOCK), 0,
                     E.st, E.cnt, E.X, sl.slab, target, E.mu, cen(CEN_SCAN), E.cen_stride);
  CK(hipGetLastError());
  unsigned long long r;
  if (int rc = cen_fetch(CEN_SCAN, false, &r)) return rc;
  return (long long)r;
}

long long dct_count_deficit(void) {
  if (!E.live) return -1;
  unsigned long long fr;
  if (folded_deficit(&fr)) return (long long)fr;
  if (int rc = cen_clear(CEN_SCAN)) return rc;
  const Slabs sl = scan_slabs(E.X);
  hipLaunchKernelGGL(k_cnt_census<CC_DEF>, dim3(sl.blocks), dim3(CC_BLOCK), 0,
                     E.st, E.cnt, E.X, sl.slab, 0, E.mu, cen(CEN_SCAN), E.cen_stride);
  CK(hipGetLastError());
  unsigned long long r;
  if (int rc = cen_fetch(CEN_SCAN, false, &r)) return rc;
  return (long long)r;
}

int dct_map_max(int32_t *out_max) {
  if (!E.live || !out_max) return -1;
  unsigned long long r;
  if (E.map_fold) {
    // folded by the transform's last sweep
    if (int rc = cen_folded(CEN_FMAP, true, &r)) return rc;
    *out_max = (int32_t)(r & 0xFFFFFFFFull);
    return 0;
  }
  if (int rc = cen_clear(CEN_SCAN)) return rc;
  const Slabs sl = scan_slabs(E.X);
  hipLaunchKernelGGL(k_map_peak, dim3(sl.blocks), dim3(CC_BLOCK), 0, E.st,
                     E.A[0], E.X, sl.slab, cen(CEN_SCAN), E.cen_stride);
  CK(hipGetLastError());
  if (int rc = cen_fetch(CEN_SCAN, true, &r)) return rc;
  *out_max = (int32_t)(r & 0xFFFFFFFFull);
  return 0;
}

int dct_map_hist(int nbins, int32_t vmax, int32_t *out_hist) {
  if (!E.live || nbins > 4096 || nbins < 1 || !out_hist) return -1;
  hipLaunchKernelGGL(k_fill32, dim3((nbins + 255) / 256), dim3(256), 0, E.st,
                     E.d_hist, 0, (long long)nbins);
  CK(hipGetLastError());
  if (vmax > 0) {
    // at most 1,024 slabs (no 32-bit partials here): each adds its nonzero
    // bins to the global ones
    const Slabs sl = scan_slabs(E.X, 1024, 1LL << 40);
    hipLaunchKernelGGL(k_map_hist, dim3(sl.blocks), dim3(CC_BLOCK), (size_t)nbins * 4,
                       E.st, E.A[0], E.X, sl.slab, nbins, vmax, E.d_hist);
    CK(hipGetLastError());
  }
  CK(hipMemcpyAsync(E.h_out, E.d_hist, (size_t)nbins * 4,
                    hipMemcpyDeviceToHost, E.st));
  CK(hipStreamSynchronize(E.st));
  memcpy(out_hist, E.h_out, (size_t)nbins * 4);
  return 0;
}


long long dct_map_extract(int32_t thr, uint32_t *out_idx, int32_t *out_val,
                          int cap) {
  return extract_impl<uint32_t>(thr, E.d_eidx32, out_idx, out_val, cap);
}

long long dct_map_extract64(int32_t thr, int64_t *out_idx, int32_t *out_val,
                            int cap) {
  return extract_impl<int64_t>(thr, E.d_eidx64, out_idx, out_val, cap);
}

int dct_map_read_at(const long long *idx, int k, int32_t *out) {
  if (!E.live) return -1;
  if (k <= 0) return 0;
  if (!idx || !out) return -1;
  for (int i = 0; i < k; i++)
    if (idx[i] < 0 || idx[i] >= E.X) return -1;
  for (int off = 0; off < k; off += CODECAP) {
    const int nb = std::min(k - off, CODECAP);
    if (int rc = stage_words(idx + off, nb)) return rc;
    hipLaunchKernelGGL(k_read_at, dim3((nb + 255) / 256), dim3(256), 0, E.st,
                       E.A[0], E.d_words, nb, E.d_out);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(E.h_out, E.d_out, (size_t)nb * 4, hipMemcpyDeviceToHost,
                      E.st));
    CK(hipStreamSynchronize(E.st));
    memcpy(out + off, E.h_out, (size_t)nb * 4);
  }
  return 0;
}

int dct_read_cnt(long long off, long long len, uint16_t *out) {
  if (!E.live || off < 0 || len < 0 || off + len > E.X || (len && !out))
    return -1;
  CK(hipStreamSynchronize(E.st));
  if (len) CK(hipMemcpy(out, E.cnt + off, (size_t)len * 2, hipMemcpyDeviceToHost));
  return 0;
}

// additive: a contiguous range of the map plane to the host (the map-plane
// counterpart of dct_read_cnt; one copy, no index list)
int dct_map_read(long long off, long long len, int32_t *out) {
  if (!E.live || off < 0 || len < 0 || off + len > E.X || (len && !out))
    return -1;
  CK(hipStreamSynchronize(E.st));
  if (len) CK(hipMemcpy(out, E.A[0] + off, (size_t)len * 4, hipMemcpyDeviceToHost));
  return 0;
}

// ---- optional ox table (reference dct_hc.cu:430-450) ----------------------
// The ox table must enumerate the same support set; walks are order-free, so
// it is converted to the v1 byte form once and walked by the same kernels.
int dct_set_tbl2(const uint16_t *mask, const uint64_t *offs, long long len) {
  if (!E.live || len != E.blen || !mask || !offs || E.n > CC_MAXN_V1) return -1;
  std::vector<uint64_t> v1((size_t)len);
  for (long long t = 0; t < len; t++) {
    uint64_t p = 0;
    int k = 0;
    for (int pos = 0; pos < E.n; pos++)
      if ((mask[t] >> pos) & 1u) {
        if (k >= CC_MAXPAT) return -1;
        const uint64_t d = (offs[t] >> (4 * pos)) & 0xFull;
        p |= (((uint64_t)pos << 4) | d) << (8 * k);
        k++;
      }
    if (mask[t] >> E.n) return -1;
    v1[(size_t)t] = p;
  }
  if (!table_ok(v1.data(), len, CC_FMT_V1)) return -1;
  E.set_ok2 = E.set_ok1 && table_is_ball(v1.data(), len, CC_FMT_V1);
  if (!E.pat2 && dalloc(&E.pat2, (size_t)len * 8)) return -3;
  CK(hipMemcpy(E.pat2, v1.data(), (size_t)len * 8, hipMemcpyHostToDevice));
  return 0;
}

int dct_use_fmt(int fmt) {
  if (!E.live || fmt < 0 || fmt > 1) return -1;
  if (fmt == 1 && !E.pat2) return -1;
  if (fmt == 1 && E.prob == CC_GRID) return -1;  // the grid walk needs v1
  E.fmt = fmt;
  return 0;
}

// ---- additive (not in the reference ABI): exact distance layers ---------
// dct_transform_layers(init_mode): the transform of dct_transform(init_mode,
// 1) with the last sweep keeping the layers instead of summing them, so
// layer d at x is the exact number of seed words at distance d from x:
// init 0 codewords (multiset of set_code; the represented solution and cnt
// are unchanged), 1 deficient words, 2 critical words.  The map plane is
// overwritten.  Layered families only (Hamming, Lee).
int dct_transform_layers(int init_mode) {
  if (!E.live || E.L < 2 || init_mode < 0 || init_mode > 2) return -1;
  int in_kind = CC_IN_PRED, pmode = CC_DEF;
  if (init_mode == 0) {
    in_kind = CC_IN_CODE;
    hipLaunchKernelGGL(k_fill32, dim3(pass_blocks(E.X)), dim3(CC_BLOCK), 0,
                       E.st, E.A[0], 0, E.X);
    CK(hipGetLastError());
    if (E.codeM > 0) {
      if (int rc = ensure_code_dev()) return rc;
      hipLaunchKernelGGL(k_scatter_code, dim3((E.codeM + 255) / 256), dim3(256),
                         0, E.st, E.A[0], E.d_code, E.codeM);
      CK(hipGetLastError());
    }
  } else if (init_mode == 2) {
    pmode = CC_CRIT;
  }
  if (int rc = run_transform(in_kind, CC_OUT_LAYERS, pmode)) return rc;
  CK(hipStreamSynchronize(E.st));
  return 0;
}

// ---- additive: ruin-region features of the learned proposal ---------------
// Five windows of R+1 exact layer fractions per region of k words (layout
// [region][window][d]): 0 mean over the region of the critical layers, 1
// mean and 2 max of the deficient layers, 3 mean of the codeword layers, 4
// min of the critical layers; a layer count at distance d is divided by the
// distance-d shell size (double precision, rounded to float at the end).
}  // extern "C"

struct FeatShell { double s[CC_MAXL]; };

template <typename LT>
__global__ void k_region_feat(const SweepArgs a, int L, int Lv,
                              const long long *__restrict__ words, int nreg, int k,
                              FeatShell sh, int kind, float *out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nreg) return;
  float *o = out + (long long)r * 5 * L;
  for (int d = 0; d < L; d++) {
    double sum = 0, mx = 0, mn = 0;
    for (int j = 0; j < k; j++) {
      const long long w = words[(long long)r * k + j];
      double v = 0;
      if (d < Lv) {
        if constexpr (sizeof(LT) == 2) v = (double)a.H[d][w];
        else v = (double)a.A[d][w];
      }
      v = v / sh.s[d];
      sum += v;
      mx = j == 0 ? v : (v > mx ? v : mx);
      mn = j == 0 ? v : (v < mn ? v : mn);
    }
    const double mean = sum / k;
    if (kind == 1) { o[1 * L + d] = (float)mean; o[2 * L + d] = (float)mx; }
    else if (kind == 2) { o[0 * L + d] = (float)mean; o[4 * L + d] = (float)mn; }
    else o[3 * L + d] = (float)mean;
  }
}

// Exact layers at listed words without whole-space transforms, for cells
// where walking each word's ball is far cheaper than three layer transforms
// (the table must be exactly the Hamming ball, so a pattern's distance is
// its number of fields): block i walks words[i]'s ball and counts the
// deficient (cnt < mu) and critical (1 <= cnt <= mu) words at each
// distance; out[i][0][d], out[i][1][d].
template <int FMT>
__global__ void __launch_bounds__(256)
    k_region_walk(Geom g, const long long *__restrict__ words,
                  const uint16_t *__restrict__ cnt, const uint64_t *__restrict__ pat,
                  long long blen, int L, int32_t *out) {
  __shared__ long long disp[CC_MAXN * CC_CSTRIDE];
  __shared__ unsigned char okf[CC_MAXN * CC_CSTRIDE];
  __shared__ int digs[CC_MAXN];
  __shared__ int red[2 * CC_MAXL];
  const long long w = words[blockIdx.x];
  cc_walk_disp(g, w, disp, okf, digs);
  if (threadIdx.x < 2 * CC_MAXL) red[threadIdx.x] = 0;
  __syncthreads();
  constexpr int FB = FMT == CC_FMT_V1 ? 8 : 10;
  constexpr int NF = FMT == CC_FMT_V1 ? CC_MAXPAT : CC_MAXPAT_W;
  constexpr uint64_t FM = (1u << FB) - 1u;
  int def[CC_MAXL], crit[CC_MAXL];
#pragma unroll
  for (int d = 0; d < CC_MAXL; d++) def[d] = crit[d] = 0;
  for (long long t = threadIdx.x; t < blen; t += blockDim.x) {
    const uint64_t pt = pat[t];
    int nf = 0;
#pragma unroll
    for (int f = 0; f < NF; f++) nf += ((pt >> (FB * f)) & FM) != 0;
    bool ok;
    const long long x = cc_walk_word<false, FMT>(pt, w, disp, okf, &ok);
    const int c = (int)cnt[x];
#pragma unroll
    for (int d = 0; d < CC_MAXL; d++)
      if (d == nf) {
        def[d] += c < g.mu;
        crit[d] += c >= 1 && c <= g.mu;
      }
  }
  for (int d = 0; d < L; d++) {
    int a = def[d], b = crit[d];
    for (int o = 32; o > 0; o >>= 1) {
      a += __shfl_xor(a, o, 64);
      b += __shfl_xor(b, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&red[d], a);
      atomicAdd(&red[CC_MAXL + d], b);
    }
  }
  __syncthreads();
  if (threadIdx.x < L) {
    out[(long long)blockIdx.x * 2 * L + threadIdx.x] = red[threadIdx.x];
    out[(long long)blockIdx.x * 2 * L + L + threadIdx.x] = red[CC_MAXL + threadIdx.x];
  }
}

// the code's words as packed digits
__global__ void k_pack_words(Dig g, const long long *__restrict__ w, int n,
                             uint64_t *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = cc_pack(g, w[i]);
}

// codeword layers at listed words: out[i][d] = code words at Hamming
// distance exactly d from words[i] (packed digits: xor, fold each digit's
// bits onto its low bit, popcount)
__global__ void __launch_bounds__(256)
    k_region_code(Dig g, const long long *__restrict__ words,
                  const uint64_t *__restrict__ cpk, int M, int L, int32_t *out) {
  __shared__ int red[CC_MAXL];
  if (threadIdx.x < CC_MAXL) red[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t pw = cc_pack(g, words[blockIdx.x]);
  int acc[CC_MAXL];
#pragma unroll
  for (int d = 0; d < CC_MAXL; d++) acc[d] = 0;
  for (int j = threadIdx.x; j < M; j += blockDim.x) {
    const uint64_t x = pw ^ cpk[j];
    uint64_t y = x;
    for (int sft = 1; sft < g.b; sft++) y |= x >> sft;
    const int d = __popcll(y & g.lsb);
#pragma unroll
    for (int e = 0; e < CC_MAXL; e++) acc[e] += d == e;
  }
  for (int d = 0; d < L; d++) {
    int a = acc[d];
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(&red[d], a);
  }
  __syncthreads();
  if (threadIdx.x < L) out[(long long)blockIdx.x * L + threadIdx.x] = red[threadIdx.x];
}

// the five windows of nreg regions of k words from per-word layers (walk:
// [word][deficient, critical][d], code: [word][d]), as k_region_feat
__global__ void k_region_feat_words(const int32_t *__restrict__ walk,
                                    const int32_t *__restrict__ codel, int L, int Lv,
                                    int nreg, int k, FeatShell sh, float *out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nreg) return;
  float *o = out + (long long)r * 5 * L;
  for (int kind = 0; kind < 3; kind++)
    for (int d = 0; d < L; d++) {
      double sum = 0, mx = 0, mn = 0;
      for (int j = 0; j < k; j++) {
        const long long i = (long long)r * k + j;
        double v = 0;
        if (d < Lv)
          v = kind == 1 ? (double)walk[i * 2 * L + d]
            : kind == 2 ? (double)walk[i * 2 * L + L + d]
                        : (double)codel[i * L + d];
        v = v / sh.s[d];
        sum += v;
        mx = j == 0 ? v : (v > mx ? v : mx);
        mn = j == 0 ? v : (v < mn ? v : mn);
      }
      const double mean = sum / k;
      if (kind == 1) { o[1 * L + d] = (float)mean; o[2 * L + d] = (float)mx; }
      else if (kind == 2) { o[0 * L + d] = (float)mean; o[4 * L + d] = (float)mn; }
      else o[3 * L + d] = (float)mean;
    }
}

// Ruin regions on the device: block b takes anchor code position anc[b]
// and picks the k code words nearest to it in digit distance, ties broken
// by a 24-bit hash of (seed, anchor position, word position) and then by
// the position; the key (distance << 40 | hash << 16 | position) is unique,
// so round r takes the smallest key above round r-1's.  Writes the
// positions (ascending key) to pos[b * k ..] and the words to words[..].
__device__ __forceinline__ uint64_t cc_mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ void __launch_bounds__(256)
    k_region_pick(Geom g, const long long *__restrict__ code, int M,
                  const int32_t *__restrict__ anc, int k, uint64_t seed,
                  int32_t *pos, long long *words) {
  __shared__ unsigned long long red[4];
  __shared__ int adig[CC_MAXN];
  const int b = blockIdx.x;
  const int a = anc[b];
  const long long wa = code[a];
  for (int i = threadIdx.x; i < g.n; i += blockDim.x)
    adig[i] = (int)((wa / g.pw[i]) % g.ax[i]);
  __syncthreads();
  unsigned long long last = 0;
  for (int r = 0; r < k; r++) {
    unsigned long long best = ~0ull;
    for (int j = threadIdx.x; j < M; j += blockDim.x) {
      const long long w = code[j];
      int d = 0;
      for (int i = 0; i < g.n; i++) d += (int)((w / g.pw[i]) % g.ax[i]) != adig[i];
      const uint64_t h = cc_mix64(seed ^ ((uint64_t)a << 32) ^ (uint64_t)j) & 0xFFFFFFull;
      const unsigned long long key = ((unsigned long long)d << 40) | (h << 16) | (unsigned)j;
      if ((r == 0 || key > last) && key < best) best = key;
    }
    // block minimum: wave minima through shuffles, then the four waves
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long x = __shfl_xor(best, o, 64);
      best = x < best ? x : best;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
    __syncthreads();
    unsigned long long m = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); i++) m = red[i] < m ? red[i] : m;
    __syncthreads();
    last = m;
    if (threadIdx.x == 0) {
      const int j = (int)(m & 0xFFFFull);
      pos[b * k + r] = j;
      words[b * k + r] = code[j];
    }
  }
}

// the five windows of nreg regions whose words are in d_words (region r:
// d_words[r * k ..]) from the three layer transforms of the pending code,
// into the feature block at d_out (no synchronisation)
extern "C" int dct_transform_layers(int init_mode);
int region_windows(int nreg, int k) {
  FeatShell sh;
  // distance-d shell sizes: the elementary symmetric polynomials of the
  // (axis size - 1), at least 1
  std::vector<double> poly(E.L, 0.0);
  poly[0] = 1;
  for (int i = 0; i < E.n; i++)
    for (int d = E.L - 1; d >= 1; d--) poly[d] += (E.ax[i] - 1) * poly[d - 1];
  for (int d = 0; d < CC_MAXL; d++) sh.s[d] = d < E.L ? std::max(poly[d], 1.0) : 1.0;
  float *dfeat = (float *)E.d_out;
  const int Lv = layers_after(E.n);
  const long long nw = (long long)nreg * k;
  // scratch behind the feature block and the positions region_propose
  // keeps at d_out + 4096: per-word walk and code layers
  const bool fits = 8192 + nw * 3 * E.L <= CODECAP &&
                    (long long)E.codeM * 8 <= (long long)CODECAP * 12;
  // walks when they read less than an eighth of one whole-space pass, or
  // at most 2^22 entries (three layer transforms then cost more in launches
  // and waits than the walks in reads)
  const bool cheap = nw * E.blen * 8 < E.X || nw * E.blen <= (1LL << 22);
  const bool walk = E.region_walk && E.fmt == 0 && E.codeM > 0 && fits &&
                    (E.region_walk_mode == 2 || (E.region_walk_mode == 1 && cheap));
  if (walk) {
    // per-word layers by ball walks and code comparisons (the words are in
    // d_words): no layer transform of the whole space
    if (int rc = ensure_code_dev()) return rc;
    int32_t *d_walk = E.d_out + 8192;
    int32_t *d_codel = d_walk + nw * 2 * E.L;
    uint64_t *d_cpk = (uint64_t *)E.d_cw;   // the code packed (sparse-seed staging)
    const Dig dg = make_dig();
    hipLaunchKernelGGL(k_pack_words, dim3((E.codeM + 255) / 256), dim3(256), 0, E.st, dg,
                       E.d_code, E.codeM, d_cpk);
    CK(hipGetLastError());
    const Geom g = make_geom();
    if (E.tfmt == CC_FMT_WIDE)
      hipLaunchKernelGGL(k_region_walk<CC_FMT_WIDE>, dim3((unsigned)nw), dim3(256), 0,
                         E.st, g, E.d_words, E.cnt, walk_table(), E.blen, E.L, d_walk);
    else
      hipLaunchKernelGGL(k_region_walk<CC_FMT_V1>, dim3((unsigned)nw), dim3(256), 0,
                         E.st, g, E.d_words, E.cnt, walk_table(), E.blen, E.L, d_walk);
    CK(hipGetLastError());
    hipLaunchKernelGGL(k_region_code, dim3((unsigned)nw), dim3(256), 0, E.st, dg,
                       E.d_words, d_cpk, E.codeM, E.L, d_codel);
    CK(hipGetLastError());
    hipLaunchKernelGGL(k_region_feat_words, dim3((nreg + 63) / 64), dim3(64), 0, E.st,
                       d_walk, d_codel, E.L, Lv, nreg, k, sh, dfeat);
    CK(hipGetLastError());
    return 0;
  }
  for (int kind : {1, 2, 0}) {
    if (int rc = dct_transform_layers(kind)) return rc;
    SweepArgs a;
    memset(&a, 0, sizeof(a));
    for (int d = 0; d < CC_MAXL; d++) { a.A[d] = E.A[d]; a.H[d] = E.H[d]; }
    if (E.layers_live == 2)
      hipLaunchKernelGGL(k_region_feat<uint16_t>, dim3((nreg + 63) / 64), dim3(64), 0,
                         E.st, a, E.L, Lv, E.d_words, nreg, k, sh, kind, dfeat);
    else
      hipLaunchKernelGGL(k_region_feat<int32_t>, dim3((nreg + 63) / 64), dim3(64), 0,
                         E.st, a, E.L, Lv, E.d_words, nreg, k, sh, kind, dfeat);
    CK(hipGetLastError());
  }
  return 0;
}

extern "C" {

// dct_region_propose(code, M, anchors, na, k, seed, regions, feats): for
// each of the na anchor positions the region of the k nearest code words
// (k_region_pick, positions into code, written to regions[b * k ..]) and
// its five feature windows (as dct_region_features, written to
// feats[b * 5 (R+1) ..]): one upload, one download, no host work per
// candidate.  Hamming only; M <= 65,535, na <= 64, k <= min(M, 64).
int dct_region_propose(const long long *code, int M, const int32_t *anchors,
                       int na, int k, unsigned long long seed, int32_t *regions,
                       float *feats) {
  if (!E.live || E.prob != CC_HAMMING || E.L < 2 || M < 1 || na < 1 || na > 64 ||
      k < 1 || k > 64 || k > M || !anchors || !regions || !feats)
    return -1;
  for (int i = 0; i < na; i++)
    if (anchors[i] < 0 || anchors[i] >= M) return -1;
  if (int rc = dct_set_code(code, M)) return rc;
  if (int rc = ensure_code_dev()) return rc;
  // anchors through the word staging (int32 in the int64 buffer)
  int32_t *h_anc = (int32_t *)E.h_words;
  memcpy(h_anc, anchors, (size_t)na * 4);
  int32_t *d_anc = (int32_t *)E.d_words + 2 * CODECAP - 64;  // the buffer's tail
  CK(hipMemcpyAsync(d_anc, h_anc, (size_t)na * 4, hipMemcpyHostToDevice, E.st));
  int32_t *d_pos = E.d_out + 4096;          // after the feature block
  hipLaunchKernelGGL(k_region_pick, dim3(na), dim3(256), 0, E.st, make_geom(),
                     E.d_code, M, d_anc, k, (uint64_t)seed, d_pos, E.d_words);
  CK(hipGetLastError());
  if (int rc = region_windows(na, k)) return rc;
  // one download: the features, then the positions
  float *dfeat = (float *)E.d_out;
  const size_t fb = (size_t)na * 5 * E.L * 4;
  CK(hipMemcpyAsync(E.h_out, dfeat, fb, hipMemcpyDeviceToHost, E.st));
  CK(hipMemcpyAsync(E.h_out + 4096, d_pos, (size_t)na * k * 4, hipMemcpyDeviceToHost,
                    E.st));
  CK(hipStreamSynchronize(E.st));
  memcpy(feats, E.h_out, fb);
  memcpy(regions, E.h_out + 4096, (size_t)na * k * 4);
  return 0;
}

// dct_region_features(code, M, words, nreg, k, out): sets the pending code
// to code[0..M) and, from three layer transforms (deficient, critical and
// codeword seeds) on the device, writes the five windows of each of the
// nreg regions words[r*k .. r*k+k) to out[r * 5 (R+1) ..]: one call, one
// download, no layer planes to the host.  Hamming only, nreg * k <= 65,536;
// cnt and the represented code are unchanged, the map plane is overwritten.
int dct_region_features(const long long *code, int M, const long long *words,
                        int nreg, int k, float *out) {
  if (!E.live || E.prob != CC_HAMMING || E.L < 2 || nreg < 1 || k < 1 ||
      (long long)nreg * k > CODECAP || (long long)nreg * 5 * E.L > CODECAP || !out)
    return -1;
  if (!valid_words(words, nreg * k)) return -1;
  if (int rc = dct_set_code(code, M)) return rc;
  // the code goes up first: its upload reuses the pinned word staging, so it
  // must be done before the region words are staged there
  if (int rc = ensure_code_dev()) return rc;
  if (int rc = stage_words(words, nreg * k)) return rc;
  if (int rc = region_windows(nreg, k)) return rc;
  float *dfeat = (float *)E.d_out;
  const size_t nb = (size_t)nreg * 5 * E.L * 4;
  CK(hipMemcpyAsync(E.h_out, dfeat, nb, hipMemcpyDeviceToHost, E.st));
  CK(hipStreamSynchronize(E.st));
  memcpy(out, E.h_out, nb);
  return 0;
}

// out[i * (R+1) + d] = layer d at idx[i] after dct_transform_layers
int dct_layers_read_at(const long long *idx, int k, int32_t *out) {
  if (!E.live || !E.layers_live) return -1;
  if (k <= 0) return 0;
  if (!idx || !out) return -1;
  for (int i = 0; i < k; i++)
    if (idx[i] < 0 || idx[i] >= E.X) return -1;
  SweepArgs a;
  memset(&a, 0, sizeof(a));
  for (int d = 0; d < CC_MAXL; d++) { a.A[d] = E.A[d]; a.H[d] = E.H[d]; }
  const int L = E.L;
  const int chunk = CODECAP / L;
  for (int off = 0; off < k; off += chunk) {
    const int nb = std::min(k - off, chunk);
    if (int rc = stage_words(idx + off, nb)) return rc;
    if (E.layers_live == 2)
      hipLaunchKernelGGL(k_layers_read_at<uint16_t>, dim3((nb + 255) / 256),
                         dim3(256), 0, E.st, a, L, E.d_words, nb, E.d_out);
    else
      hipLaunchKernelGGL(k_layers_read_at<int32_t>, dim3((nb + 255) / 256),
                         dim3(256), 0, E.st, a, L, E.d_words, nb, E.d_out);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(E.h_out, E.d_out, (size_t)nb * L * 4,
                      hipMemcpyDeviceToHost, E.st));
    CK(hipStreamSynchronize(E.st));
    memcpy(out + (size_t)off * L, E.h_out, (size_t)nb * L * 4);
  }
  // distances beyond the space's diameter are never written: they are 0
  const int Lv = layers_after(E.n);
  for (int i = 0; i < k && Lv < L; i++)
    for (int d = Lv; d < L; d++) out[(size_t)i * L + d] = 0;
  return 0;
}

// ---- additive (not in the reference ABI): planner probe --------------------
// dct_plan_probe: the packed schedule the engine would plan for a Hamming
// cell (homogeneous axes), without a device and without a live cell: kind
// 0 gain / loss, kind 1 recount of a code with multiplicities <= s0 and M
// words.  Rows as dct_sweep_plan2.  Returns the sweep count (0: the cell
// runs the unpacked kernels), -1 on bad arguments or while a cell is live.
int dct_plan_probe(int n, const int *ax, int R, int kind, int s0, int M,
                   int *out, int cap) {
  if (E.live || n < 1 || n > CC_MAXN || !ax || R < 1 || R > CC_MAXR || !out ||
      kind < 0 || kind > 1 || s0 < 1 || M < 1)
    return -1;
  Engine saved = E;
  E.prob = CC_HAMMING; E.n = n; E.R = R; E.L = R + 1;
  E.pw[0] = 1;
  for (int i = 0; i < n; i++) {
    if (ax[i] < 2 || ax[i] > 10) { E = saved; return -1; }
    E.ax[i] = ax[i];
    E.pw[i + 1] = E.pw[i] * ax[i];
    if (E.pw[i + 1] > (1LL << 40)) { E = saved; return -1; }
  }
  E.X = E.pw[n];
  E.lds_max = 160 * 1024 - 256;
  const char *lb = getenv("CC_SWEEP_LDS");
  E.lds_budget = std::min(lb ? std::max(4096, atoi(lb)) : 64 * 1024, E.lds_max);
  const char *pe = getenv("CC_PACKED");
  const char *sk = getenv("CC_SWEEP_KERNEL");
  E.packed = !(pe && !strcmp(pe, "0")) && !(sk && !strcmp(sk, "generic"));
  const char *pc = getenv("CC_PK_COL");
  E.pk_col = pc ? atoi(pc) : 2;
  const char *pf = getenv("CC_PK_FIRST");
  E.pk_first = pf ? atoi(pf) : 1;
  const char *pn = getenv("CC_PK_NARROW");
  E.pk_narrow = pn ? atoi(pn) : 0;
  const char *psc = getenv("CC_SPARSE_CODE");
  E.sparse = psc ? atoi(psc) : 1;
  const char *pwd = getenv("CC_PK_WIDE");
  E.pk_wide = pwd ? atoi(pwd) : 0;
  const char *plc = getenv("CC_PK_LCMP");
  E.pk_lcmp = plc ? atoi(plc) : 0;
  const PPlan pl = kind == 0 ? pk_plan(1.0, 1e300, true, false)
                             : pk_plan((double)s0, (double)M, false, true);
  int ns = 0;
  if (pl.ok) {
    ns = (int)pl.sw.size();
    for (int i = 0; i < ns && i < cap; i++) {
      const PSweep &w = pl.sw[i];
      int *o = out + 13 * i;
      o[0] = w.j0; o[1] = w.nax; o[2] = w.P; o[3] = w.TL; o[4] = w.RR;
      o[5] = w.W; o[6] = w.ND; o[7] = w.in_bytes; o[8] = w.out_bytes;
      o[9] = (int)w.lds; o[10] = w.threads; o[11] = w.nrounds;
      o[12] = 1 + w.kern;
    }
  }
  E = saved;
  return ns;
}

// ---- additive (not in the reference ABI): schedule autotuner -------------
// dct_autotune(reps, best_us): time the candidate schedules of the gain /
// loss transform on the live cell and keep the fastest.  Candidates are the
// planner's schedule under every combination of LDS budget (16-128 KB),
// column-kernel mode (off / whole cube / lane pairs) and first-sweep kernel
// (LDS-tiled / persistent with prefetch / LDS-free across the lanes), plus
// the unpacked kernels under every LDS budget with compact words or layer
// planes and with or without register rounds; duplicates are timed once.  Each is run once untimed and then
// reps times between HIP events on the engine's stream.  The winning
// options also steer the recount schedule.  Like a transform, it
// overwrites the map plane; cnt and the represented solution are unchanged.
// Returns the number of schedules timed (best_us: the winner's
// microseconds per transform), or < 0.
int dct_autotune(int reps, float *best_us) {
  if (!E.live || reps < 1) return -1;
  E.tuned = 1;
  struct Opt { int lds, col, first, packed, narrow, cmp, gen, wide; };
  std::vector<Opt> opts;
  for (int lds : {16384, 32768, 65536, 98304, 131072}) {
    if (lds > E.lds_max) continue;
    for (int col : {2, 1, 0})
      for (int first : {1, 0, 2})
        for (int wide : {0, 1})
          opts.push_back({lds, col, first, 1, E.pk_narrow, E.compact, E.generic_only,
                          wide});
  }
  // the unpacked schedule under each LDS budget (its plan is rebuilt per
  // budget: tile sizes, sweep grouping and the workgroups a CU holds move
  // together), with compact words between sweeps and with layer planes,
  // with register-round sweeps for homogeneous groups and with LDS tiles only
  for (int lds : {16384, 32768, 65536, 98304, 131072})
    if (lds <= E.lds_max)
      for (int gen : {0, 1})
        for (int cmp : {1, 0})
          opts.push_back({lds, E.pk_col, E.pk_first, 0, E.pk_narrow, cmp, gen,
                          E.pk_wide});
  const Opt saved{E.lds_budget, E.pk_col, E.pk_first, E.packed, E.pk_narrow,
                  E.compact, E.generic_only, E.pk_wide};
  std::vector<std::string> seen;
  float best = 1e30f;
  Opt win = saved;
  PPlan wplan = E.pfield;
  int timed = 0;
  hipEvent_t e0 = E.ev[0], e1 = E.ev[1];
  for (const Opt &o : opts) {
    E.lds_budget = o.lds; E.pk_col = o.col; E.pk_first = o.first; E.packed = o.packed;
    E.pk_narrow = o.narrow;
    E.pk_wide = o.wide;
    E.compact = o.cmp;
    E.generic_only = o.gen;
    PPlan pl = o.packed ? pk_plan(1.0, 1e300, true, false) : PPlan();
    std::string sig = "unpacked|";
    if (!o.packed) {
      build_plan();   // (a plan equal to an earlier one has its signature)
      std::vector<CFmt> fm;
      if (!compact_plan(1.0, 1e300, fm)) {
        if (o.cmp) continue;         // the same plan as without compact words
      } else {
        sig += o.cmp ? "compact|" : "planes|";
      }
      for (const Sweep &w : E.plan) {
        char b[96];
        snprintf(b, sizeof(b), "%d.%d.%d.%d.%d|", w.j0, w.nax, w.P, w.TL, w.rq);
        sig += b;
      }
    } else if (pl.ok) {
      sig.clear();
      for (const PSweep &w : pl.sw) {
        char b[96];
        snprintf(b, sizeof(b), "%d.%d.%d.%d.%d.%d.%d.%zu.%d|", w.j0, w.nax, w.RR,
                 w.W, w.ND, w.TL, w.kern, w.lds, w.out_nar);
        sig += b;
      }
    } else if (o.packed) {
      continue;                      // no packed schedule for this cell
    }
    if (std::find(seen.begin(), seen.end(), sig) != seen.end()) continue;
    seen.push_back(sig);
    const int prof = E.prof;
    E.prof = 0;
    float ms = 0;
    for (int r = 0; r <= reps; r++) {
      if (r == 1) CK(hipEventRecord(e0, E.st));
      int rc = pl.ok ? run_ptransform(pl, CC_IN_PRED, CC_OUT_MAP, 0, E.mu - 1)
                     : run_transform(CC_IN_PRED, CC_OUT_MAP, CC_DEF);
      if (rc) { E.prof = prof; return rc; }
    }
    CK(hipEventRecord(e1, E.st));
    CK(hipEventSynchronize(e1));
    CK(hipEventElapsedTime(&ms, e0, e1));
    E.prof = prof;
    timed++;
    if (ms < best) { best = ms; win = o; wplan = pl; }
  }
  E.lds_budget = win.lds; E.pk_col = win.col; E.pk_first = win.first;
  E.packed = win.packed;
  E.pk_narrow = win.narrow;
  E.pk_wide = win.wide;
  E.compact = win.cmp;
  E.generic_only = win.gen;
  E.pfield = wplan;
  E.cp_ok = false;
  if (wplan.ok && !getenv("CC_PK_NT")) {
    // the winner's stores: non-temporal after the first sweep, or plain
    float tnt[2] = {0, 0};
    const int masks[2] = {~1, 0};
    const int prof = E.prof;
    E.prof = 0;
    for (int v = 0; v < 2; v++) {
      E.pk_nt = masks[v];
      for (int r = 0; r <= reps; r++) {
        if (r == 1) CK(hipEventRecord(e0, E.st));
        if (int rc = run_ptransform(wplan, CC_IN_PRED, CC_OUT_MAP, 0, E.mu - 1)) {
          E.prof = prof;
          return rc;
        }
      }
      CK(hipEventRecord(e1, E.st));
      CK(hipEventSynchronize(e1));
      CK(hipEventElapsedTime(&tnt[v], e0, e1));
      timed++;
    }
    E.prof = prof;
    E.pk_nt = tnt[1] < tnt[0] ? masks[1] : masks[0];
    best = std::min(best, std::min(tnt[0], tnt[1]));
  }
  build_plan();   // the unpacked plan of the winning budget
  if (best_us) *best_us = best * 1000.0f / reps;
  return timed;
}

// ---- additive (not in the reference ABI): schedule and sweep timing ------
// dct_sweep_plan: per sweep 8 ints (first axis, axes, P, TL, layers in,
// layers out, round-kernel axis size or 0, LDS bytes) and, after the last
// sweep, 2 ints: inter-sweep layer bytes of gain/loss and of recount (2 or
// 4).  Returns the sweep count.
int dct_sweep_plan(int *out, int cap) {
  if (!E.live || !out) return -1;
  const int ns = (int)E.plan.size();
  for (int i = 0; i < ns && i < cap; i++) {
    const Sweep &s = E.plan[i];
    out[8 * i + 0] = s.j0;
    out[8 * i + 1] = s.nax;
    out[8 * i + 2] = s.P;
    out[8 * i + 3] = s.TL;
    out[8 * i + 4] = layers_after(s.c0);
    out[8 * i + 5] = s.lmax;
    out[8 * i + 6] = s.rq;
    out[8 * i + 7] = (int)(E.l16_field ? s.lds16 : s.lds);
  }
  if (ns < cap) {
    out[8 * ns + 0] = E.l16_field ? 2 : 4;
    out[8 * ns + 1] = E.l16_count ? 2 : 4;
  }
  return ns;
}

// dct_sweep_plan2(kind, out, cap): the schedule a transform of this kind
// runs now (kind 0: gain / loss, 1: recount of the pending code), 13 ints
// per sweep: first axis, axes, P, TL, register round RR (0: unpacked
// kernels), field width W (32: unpacked layer planes), dwords per word in
// registers, algorithmic bytes per word read and written, LDS bytes per
// workgroup, threads per workgroup, rounds, kernel (0 unpacked, 1 LDS-tiled
// packed, 2 / 3 column kernel with 1 / 2 lanes per cube, 4 LDS-free first
// sweep, 5 persistent first sweep with prefetch).  Returns the sweep count.
int dct_sweep_plan2(int kind, int *out, int cap) {
  if (!E.live || !out || kind < 0 || kind > 1) return -1;
  PPlan cp;
  if (kind == 1) cp = pk_code_plan(true);
  const PPlan &pp = kind == 0 ? E.pfield : cp;
  if (pp.ok) {
    const int ns = (int)pp.sw.size();
    for (int i = 0; i < ns && i < cap; i++) {
      const PSweep &w = pp.sw[i];
      int *o = out + 13 * i;
      o[0] = w.j0; o[1] = w.nax; o[2] = w.P; o[3] = w.TL; o[4] = w.RR;
      o[5] = w.W; o[6] = w.ND; o[7] = w.in_bytes; o[8] = w.out_bytes;
      o[9] = (int)w.lds; o[10] = w.threads; o[11] = w.nrounds;
      o[12] = 1 + w.kern;
    }
    return ns;
  }
  const int ns = (int)E.plan.size();
  const bool field = kind == 0;
  const int lb = (field ? E.l16_field : E.l16_count) ? 2 : 4;
  std::vector<CFmt> fm;
  bool cmp = false;
  if (ns > 1) {
    double s0 = 1, cp = 1e300;
    if (!field) code_bounds(&s0, &cp);
    cmp = compact_plan(s0, cp, fm);
  }
  for (int i = 0; i < ns && i < cap; i++) {
    const Sweep &sw = E.plan[i];
    const int lin = layers_after(sw.c0), lout = sw.lmax;
    int rd = i == 0 ? (field ? 2 : 4) : lb * lin;
    int wr;
    if (i == ns - 1) wr = field ? 4 : 2;
    else if (i == 0) wr = (field || lb == 2) ? lb * lout : lb * (lout - 1);
    else wr = lb * (lout - 1);
    if (E.L == 1) {           // single plane: int32 in and out
      rd = i == 0 ? (field ? 2 : 4) : 4;
      wr = i == ns - 1 ? (field ? 4 : 2) : 4;
    }
    if (cmp) {                // compact words of 1..4 dwords between sweeps
      rd = i == 0 ? (field ? 2 : 4) : 4 * fm[i - 1].nd;
      wr = i == ns - 1 ? (field ? 4 : 2) : 4 * fm[i].nd;
    }
    int *o = out + 13 * i;
    o[0] = sw.j0; o[1] = sw.nax; o[2] = sw.P; o[3] = sw.TL; o[4] = 0;
    o[5] = 32; o[6] = E.L; o[7] = rd; o[8] = wr;
    o[9] = (int)(E.l16_field ? sw.lds16 : sw.lds); o[10] = sw.threads; o[11] = 0;
    o[12] = 0;
  }
  return ns;
}

// dct_sweep_force(cuts, ncuts): pin the sweep boundaries of the packed
// schedules (ascending axis counts 1..n-1; each sweep ends at the next
// cut).  On a live cell the gain / loss schedule is replanned at once and
// the automatic timing is switched off; a recount whose code seed cannot
// use the cuts keeps the planner's own schedule.  ncuts = 0 clears them.
// With no live cell the cuts steer dct_plan_probe.  Returns the gain / loss
// sweep count (0 with no live cell), -1 on bad cuts or when no packed
// schedule has these boundaries (the cuts are then cleared).
int dct_sweep_force(const int *cuts, int ncuts) {
  if (ncuts < 0 || ncuts >= CC_MAXN || (ncuts > 0 && !cuts)) return -1;
  uint64_t m = 0;
  int last = 0;
  for (int i = 0; i < ncuts; i++) {
    if (cuts[i] <= last || cuts[i] >= CC_MAXN || (E.live && cuts[i] >= E.n))
      return -1;
    last = cuts[i];
    m |= 1ull << cuts[i];
  }
  g_cuts = m;
  E.cp_ok = false;
  if (!E.live) return 0;
  E.autotune = 0;
  E.tuned = 1;
  E.pfield = pk_plan(1.0, 1e300, true, false);
  if (!E.pfield.ok && m) {
    g_cuts = 0;
    E.pfield = pk_plan(1.0, 1e300, true, false);
    return -1;
  }
  return E.pfield.ok ? (int)E.pfield.sw.size() : 0;
}

// dct_sweep_profile(1): record HIP events around every sweep of later
// transforms (on the engine's stream); dct_sweep_times copies the last
// transform's per-sweep milliseconds and returns the sweep count.
int dct_sweep_profile(int on) {
  if (!E.live) return -1;
  E.prof = on ? 1 : 0;
  return 0;
}

int dct_sweep_times(float *ms, int cap) {
  if (!E.live || !ms) return -1;
  for (int i = 0; i < E.nswept && i < cap; i++) ms[i] = E.sweep_ms[i];
  return E.nswept;
}

// dct_set_fold(1): later packed transforms fold their census into the last
// sweep (gain / loss: the map maximum; recount: the cnt deficit), which
// dct_map_max, dct_count_eq(0) and dct_count_deficit then read instead of
// scanning the field.  The transform pays a workgroup epilogue and one
// atomic per workgroup.  Returns the previous setting.
int dct_set_fold(int on) {
  if (!E.live || on < 0 || on > 1) return -1;
  const int was = E.fold;
  E.fold = on;
  return was;
}

}  // extern "C"
#pragma GCC visibility pop

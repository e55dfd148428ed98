// cc_sweep.hpp -- the separable per-axis DP transform as LDS-tiled sweeps.
//
// Per-axis fiber recurrences (one fiber = the Q words that differ only in
// one axis digit a = 0..Q-1):
//   hamming  newA_d[a] = A_d[a] + T_{d-1} - A_{d-1}[a],  T_d = sum_a A_d[a]
//            (reference src/plugins/covering/dctcore_hamming.cuh:4-8, 111-154)
//   lee      newA_d[a] = A_d[a] + sum_{w=1..min(d,Q/2)} (A_{d-w}[a-w] +
//            A_{d-w}[a+w]), the two terms counted once when they coincide
//            (reference dctcore_lee.cuh:25-40)
//   torus    newA[a] = sum_{|e|<=R} A[(a+e) mod Q]     (dctcore_torus.cuh:25-48)
//   grid     newA[a] = sum_{b=max(0,a-R)}^{min(Q-1,a+R)} A[b]
//            (dctcore_grid.cuh:44-66)
// Only the layers up to the largest distance the axes done so far can
// produce are nonzero (Hamming: one per axis, reference dctcore_hamming.cuh:
// 98-110; Lee: floor(Q/2) per axis), so a pass reads only the live layers
// and writes only the ones it can make nonzero.  Layer 0 never changes for the two layered
// families.  Arithmetic is modulo 2^32, so int32 wrap-around (if a cell ever
// reached it) is reproduced exactly.
//
// MI355X design: a sweep stages a whole tile (several axes' worth of words)
// in LDS and runs every axis of its group there; global traffic is one read
// of the live layers and one write of the produced ones per group, instead
// of one round trip per axis.  Strided groups stage TL = 32 consecutive
// columns so each global row access is a full 128-byte line.
#pragma once
#include "cc_common.hpp"

// RW > 0: the L-inf radius as a compile-time constant (the window's terms
// are then fixed: 2R+1 adds per word instead of Q predicated ones)
template <int PROB, int R1, int Q, bool PAD, int RW = 0>
__device__ __forceinline__ void cc_fiber(int32_t *__restrict__ S, int PT,
                                         int e0, int es, int lv, int lw,
                                         int R) {
  uint32_t v[R1][Q];
#pragma unroll
  for (int d = 0; d < R1; d++)
#pragma unroll
    for (int a = 0; a < Q; a++)
      v[d][a] = (d < lv) ? (uint32_t)S[d * PT + cc_gpad<PAD>(e0 + a * es)] : 0u;
  if (PROB == CC_HAMMING) {
    uint32_t T[R1];
#pragma unroll
    for (int d = 0; d < R1; d++) {
      T[d] = 0;
#pragma unroll
      for (int a = 0; a < Q; a++) T[d] += v[d][a];
    }
#pragma unroll
    for (int d = 1; d < R1; d++)
      if (d < lw)
#pragma unroll
        for (int a = 0; a < Q; a++)
          S[d * PT + cc_gpad<PAD>(e0 + a * es)] = (int32_t)(v[d][a] + T[d - 1] - v[d - 1][a]);
  } else if (PROB == CC_LEE) {
    constexpr int W = Q / 2;
#pragma unroll
    for (int d = 1; d < R1; d++)
      if (d < lw)
#pragma unroll
        for (int a = 0; a < Q; a++) {
          uint32_t s = v[d][a];
#pragma unroll
          for (int w = 1; w <= W; w++) {
            if (w <= d) {
              const int am = (a - w + Q) % Q, ap = (a + w) % Q;
              s += v[d - w][am];
              if (ap != am) s += v[d - w][ap];
            }
          }
          S[d * PT + cc_gpad<PAD>(e0 + a * es)] = (int32_t)s;
        }
  } else {
    // single-plane L-inf window sums (torus: circular, grid: clamped)
#pragma unroll
    for (int a = 0; a < Q; a++) {
      uint32_t s = 0;
#pragma unroll
      for (int b = 0; b < Q; b++) {
        int dd = a > b ? a - b : b - a;
        if (PROB == CC_TORUS && Q - dd < dd) dd = Q - dd;
        if constexpr (RW > 0) {
          if (dd <= RW) s += v[0][b];
        } else {
          s += (dd <= R) ? v[0][b] : 0u;
        }
      }
      S[cc_gpad<PAD>(e0 + a * es)] = (int32_t)s;
    }
  }
}

template <int PROB, int R1, int Q, bool PAD, int RW>
__device__ __forceinline__ void cc_step_r(int32_t *__restrict__ S, int PT,
                                          int TL, int nf, int sk, int lv,
                                          int lw, int R) {
  const uint32_t mT = cc_magic((uint32_t)TL), mS = cc_magic((uint32_t)sk);
  for (int fi = threadIdx.x; fi < nf; fi += blockDim.x) {
    const int f = cc_div(fi, mT), tl = fi - f * TL;
    const int hi = cc_div(f, mS), lo = f - hi * sk;
    const int e0 = (hi * sk * Q + lo) * TL + tl;
    cc_fiber<PROB, R1, Q, PAD, RW>(S, PT, e0, sk * TL, lv, lw, R);
  }
}

// one axis pass over the tile's fibers; the L-inf families take their
// radius (1..3, the sizes the workloads use) as a compile-time constant
template <int PROB, int R1, int Q, bool PAD>
__device__ __forceinline__ void cc_step(int32_t *__restrict__ S, int PT,
                                        int TL, int nf, int sk, int lv,
                                        int lw, int R) {
  if constexpr (PROB == CC_TORUS || PROB == CC_GRID) {
    if (R == 1) return cc_step_r<PROB, R1, Q, PAD, 1>(S, PT, TL, nf, sk, lv, lw, R);
    if (R == 2) return cc_step_r<PROB, R1, Q, PAD, 2>(S, PT, TL, nf, sk, lv, lw, R);
    if (R == 3) return cc_step_r<PROB, R1, Q, PAD, 3>(S, PT, TL, nf, sk, lv, lw, R);
  }
  cc_step_r<PROB, R1, Q, PAD, 0>(S, PT, TL, nf, sk, lv, lw, R);
}

#define CC_STEP_CASE(QQ)                                                    \
  case QQ:                                                                  \
    if (QQ >= QMIN && QQ <= QMAX)                                           \
      cc_step<PROB, R1, (QQ >= QMIN && QQ <= QMAX) ? QQ : QMIN, PAD>(       \
          S, PTP, a.TL, nf, sk, lv, lw, a.R);                               \
    break;

template <int PROB>
struct CcQRange {
  static constexpr int lo = PROB == CC_TORUS ? 3 : 2;
  static constexpr int hi = PROB == CC_HAMMING ? 10 : 16;
};

// runtime-Q window sums for the L-inf families with axes longer than the
// unrolled table (axes above 16, up to 1024, R <= 3).  A fiber's words are
// cut into segments of CC_LSEG outputs, one (fiber, segment) per thread, so
// a long axis keeps the whole workgroup busy (one thread per fiber left a
// 1024-long torus axis to a single lane).  A pass covers whole fibers: every
// thread first reads its segment's window (CC_LSEG + 2R words, wrapped on
// the torus, clamped on the grid) into registers, the workgroup
// synchronises, then the outputs are written in place; fibers of different
// passes share no words, so no pass reads what an earlier one wrote.
#define CC_LSEG 8
template <int R, bool TORUS, bool PAD>
__device__ __forceinline__ void cc_step_linf_rt(int32_t *__restrict__ S, int TL,
                                                 int nf, int sk, int Q) {
  const int segs = (Q + CC_LSEG - 1) / CC_LSEG;           // <= 128 for Q <= 1024
  const int fpp = max(1, (int)blockDim.x / segs);          // fibers per pass
  const uint32_t mT = cc_magic((uint32_t)TL), mS = cc_magic((uint32_t)sk);
  const uint32_t mG = cc_magic((uint32_t)segs);
  const int es = sk * TL;
  const int fl = cc_div((int)threadIdx.x, mG), sg = (int)threadIdx.x - fl * segs;
  const int a0 = sg * CC_LSEG;
  for (int f0 = 0; f0 < nf; f0 += fpp) {
    const int fi = f0 + fl;
    const bool act = fl < fpp && fi < nf;
    uint32_t out[CC_LSEG];
    int e0 = 0;
    if (act) {
      const int f = cc_div(fi, mT), tl = fi - f * TL;
      const int hi = cc_div(f, mS), lo = f - hi * sk;
      e0 = (hi * sk * Q + lo) * TL + tl;
      uint32_t w[CC_LSEG + 2 * R];
#pragma unroll
      for (int i = 0; i < CC_LSEG + 2 * R; i++) {
        int b = a0 + i - R;              // -R .. Q + CC_LSEG + R - 2
        uint32_t v = 0u;
        if (TORUS) {
          b += b < 0 ? Q : (b >= Q ? -Q : 0);
          v = (uint32_t)S[cc_gpad<PAD>(e0 + b * es)];
        } else if (b >= 0 && b < Q) {
          v = (uint32_t)S[cc_gpad<PAD>(e0 + b * es)];
        }
        w[i] = v;
      }
#pragma unroll
      for (int i = 0; i < CC_LSEG; i++) {
        uint32_t sum = 0u;
#pragma unroll
        for (int j = 0; j <= 2 * R; j++) sum += w[i + j];
        out[i] = sum;
      }
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int i = 0; i < CC_LSEG; i++)
        if (a0 + i < Q) S[cc_gpad<PAD>(e0 + (a0 + i) * es)] = (int32_t)out[i];
    }
    __syncthreads();
  }
}

// the same window sums with one fiber per thread, in place, through a
// shift register of the 2R+1 original values under the window and the
// first R originals kept for the circular wrap (no indexed arrays, so
// nothing spills to scratch): fewer registers and no extra barriers, which
// the batched level kernel (axes up to 32 in its workloads) measured faster
template <int R, bool TORUS, bool PAD>
__device__ __forceinline__ void cc_step_linf_seq(int32_t *__restrict__ S, int TL,
                                                 int nf, int sk, int Q) {
  const uint32_t mT = cc_magic((uint32_t)TL), mS = cc_magic((uint32_t)sk);
  for (int fi = threadIdx.x; fi < nf; fi += blockDim.x) {
    const int f = cc_div(fi, mT), tl = fi - f * TL;
    const int hi = cc_div(f, mS), lo = f - hi * sk;
    const int e0 = (hi * sk * Q + lo) * TL + tl, es = sk * TL;
    uint32_t head[R], win[2 * R + 1];
#pragma unroll
    for (int i = 0; i < R; i++) head[i] = (uint32_t)S[cc_gpad<PAD>(e0 + i * es)];
#pragma unroll
    for (int i = 0; i <= 2 * R; i++) {
      const int b = i - R;   // -R .. R  (Q >= 2R+1 for the torus)
      if (b < 0) win[i] = TORUS ? (uint32_t)S[cc_gpad<PAD>(e0 + (b + Q) * es)] : 0u;
      else win[i] = b < Q ? (uint32_t)S[cc_gpad<PAD>(e0 + b * es)] : 0u;
    }
    for (int a = 0; a < Q; a++) {
      uint32_t s = 0;
#pragma unroll
      for (int i = 0; i <= 2 * R; i++) s += win[i];
      S[cc_gpad<PAD>(e0 + a * es)] = (int32_t)s;
      // slide: the next window ends at a + 1 + R
      const int b = a + 1 + R;
      uint32_t nx = 0;
      if (b < Q) {
        nx = (uint32_t)S[cc_gpad<PAD>(e0 + b * es)];
      } else if (TORUS) {
#pragma unroll
        for (int i = 0; i < R; i++) nx = (b - Q == i) ? head[i] : nx;
      }
#pragma unroll
      for (int i = 0; i < 2 * R; i++) win[i] = win[i + 1];
      win[2 * R] = nx;
    }
  }
}


// Inter-sweep layer planes are int32 (LT = int32_t, planes A[d]) or, when
// every layer value is provably below 2^16, uint16 (LT = uint16_t, planes
// H[d]).  A layer value is a count of solution elements at one partial
// distance, so it is bounded by the ball size for the gain / loss
// indicators and by the code size M <= 65,535 for a recount.
template <typename LT>
__device__ __forceinline__ uint32_t cc_lay_ld(const SweepArgs &a, int d,
                                              long long w) {
  if constexpr (sizeof(LT) == 2) return (uint32_t)a.H[d][w];
  else return (uint32_t)a.A[d][w];
}

template <typename LT>
__device__ __forceinline__ void cc_lay_st(const SweepArgs &a, int d,
                                          long long w, uint32_t v) {
  if constexpr (sizeof(LT) == 2) a.H[d][w] = (uint16_t)v;
  else a.A[d][w] = (int32_t)v;
}

// One tile of a sweep: stage, run every axis of the group in LDS, write
// back (the batched transform calls it for tiles of many cells).
#define CC_STAGE 8   // tile elements per thread whose loads are issued together

// SEG: long L-inf axes by segments (true) or one fiber per thread (false)
// CMP: the sweep reads or writes compact words (a separate instantiation,
// so the layer-plane kernels carry none of its code)
template <int PROB, int R1, typename LT, bool PAD = true, bool SEG = true,
          bool CMP = false>
__device__ __forceinline__ void cc_sweep_tile(const SweepArgs &a, long long tile,
                                              int32_t *__restrict__ S) {
  constexpr int QMIN = CcQRange<PROB>::lo, QMAX = CcQRange<PROB>::hi;
  const int PT = a.P * a.TL;
  const int PTP = cc_gplane<PAD>(PT);
  const int Lmax = a.live[a.nax];
  const int Lin = a.live[0];
  const long long h = tile / a.ncolb;
  const long long l0 = (tile - h * a.ncolb) * a.TL;
  const int ncol = (int)min((long long)a.TL, a.s - l0);
  const long long base = h * a.s * a.P + l0;
  const uint32_t mTL = cc_magic((uint32_t)a.TL);
  // ---- stage the tile: every live layer of every word, coalesced along tl;
  // CC_STAGE elements per thread are loaded before any is stored to LDS, so
  // their global loads are in flight together
  constexpr int U = CC_STAGE;
  for (int e0 = threadIdx.x; e0 < PT; e0 += U * blockDim.x) {
    uint32_t val[U][R1];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int e = e0 + u * (int)blockDim.x;
#pragma unroll
      for (int d = 0; d < R1; d++) val[u][d] = 0u;
      if (e >= PT) continue;
      const int p = cc_div(e, mTL), tl = e - p * a.TL;
      if (tl >= ncol) continue;
      const long long w = base + (long long)p * a.s + tl;
      if (CMP && a.in_kind == CC_IN_COMPACT) {
        // dword planes in turn, each field extracted from its own plane
        // (uniform layout: scalar branches, one bit-field extract each)
        for (int k = 0; k < CC_CMAXND; k++) {
          if (k >= a.in_nd) break;
          const uint32_t q = a.pk[(long long)k * a.pitch + w];
#pragma unroll
          for (int d = 0; d < R1; d++)
            if (a.idw[d] == k && a.ibits[d])
              val[u][d] = __builtin_amdgcn_ubfe(q, a.ioff[d], a.ibits[d]);
        }
        continue;
      }
      if (a.in_kind == CC_IN_PRED)
        val[u][0] = cc_pred(a.pmode, (int)a.cnt[w], 0, a.mu) ? 1u : 0u;
      else if (a.in_kind == CC_IN_CODE)
        val[u][0] = (uint32_t)a.A[0][w];
      else
        val[u][0] = cc_lay_ld<LT>(a, 0, w);
      if (a.in_kind == CC_IN_LAYERS)
#pragma unroll
        for (int d = 1; d < R1; d++)
          if (d < Lin) val[u][d] = cc_lay_ld<LT>(a, d, w);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int e = e0 + u * (int)blockDim.x;
      if (e >= PT) continue;
      const int ep = cc_gpad<PAD>(e);
#pragma unroll
      for (int d = 0; d < R1; d++)
        if (d < Lmax) S[d * PTP + ep] = (int32_t)val[u][d];
    }
  }
  __syncthreads();
  // ---- every axis of the group, in LDS
  for (int k = 0; k < a.nax; k++) {
    const int lv = a.live[k], lw = a.live[k + 1];
    const int Q = a.Q[k], sk = a.sk[k];
    const int nf = (a.P / Q) * a.TL;
    switch (Q) {
      CC_STEP_CASE(2) CC_STEP_CASE(3) CC_STEP_CASE(4) CC_STEP_CASE(5)
      CC_STEP_CASE(6) CC_STEP_CASE(7) CC_STEP_CASE(8) CC_STEP_CASE(9)
      CC_STEP_CASE(10) CC_STEP_CASE(11) CC_STEP_CASE(12) CC_STEP_CASE(13)
      CC_STEP_CASE(14) CC_STEP_CASE(15) CC_STEP_CASE(16)
      default:
        // L-inf axes above 16 (R <= 3): the runtime-length window
        if constexpr (PROB == CC_TORUS || PROB == CC_GRID) {
          constexpr bool T = PROB == CC_TORUS;
          if constexpr (SEG) {
            if (a.R == 1) cc_step_linf_rt<1, T, PAD>(S, a.TL, nf, sk, Q);
            else if (a.R == 2) cc_step_linf_rt<2, T, PAD>(S, a.TL, nf, sk, Q);
            else cc_step_linf_rt<3, T, PAD>(S, a.TL, nf, sk, Q);
          } else {
            if (a.R == 1) cc_step_linf_seq<1, T, PAD>(S, a.TL, nf, sk, Q);
            else if (a.R == 2) cc_step_linf_seq<2, T, PAD>(S, a.TL, nf, sk, Q);
            else cc_step_linf_seq<3, T, PAD>(S, a.TL, nf, sk, Q);
          }
        }
        break;
    }
    __syncthreads();
  }
  // ---- write back: layers, or the layer sum into the map / cnt plane
  unsigned long long best = 0;
  for (int e = threadIdx.x; e < PT; e += blockDim.x) {
    const int p = cc_div(e, mTL), tl = e - p * a.TL;
    if (tl >= ncol) continue;
    const long long w = base + (long long)p * a.s + tl;
    const int ep = cc_gpad<PAD>(e);
    if (CMP && a.out_kind == CC_OUT_COMPACT) {
      for (int k = 0; k < CC_CMAXND; k++) {
        if (k >= a.out_nd) break;
        uint32_t q = 0u;
#pragma unroll
        for (int d = 0; d < R1; d++)
          if (d < Lmax && a.obits[d] && a.odw[d] == k)
            q |= (uint32_t)S[d * PTP + ep] << a.ooff[d];
        a.pk[(long long)k * a.pitch + w] = q;
      }
    } else if (a.out_kind == CC_OUT_LAYERS) {
      for (int d = a.write_l0 ? 0 : 1; d < Lmax; d++)
        cc_lay_st<LT>(a, d, w, (uint32_t)S[d * PTP + ep]);
    } else {
      uint32_t sum = 0;
      for (int d = 0; d < Lmax; d++) sum += (uint32_t)S[d * PTP + ep];
      if (a.out_kind == CC_OUT_MAP) {
        a.A[0][w] = (int32_t)sum;
        if (a.amax) cc_key_max(best, sum, w);
      } else {
        const int32_t si = (int32_t)sum;
        a.cnt[w] = si > 65535 ? (uint16_t)65535 : (uint16_t)si;
      }
    }
  }
  if (a.amax) cc_key_flush(best, a.amax);
  __syncthreads();
}

template <int PROB, int R1, typename LT, bool CMP = false>
__global__ void __launch_bounds__(CC_BLOCK) k_sweep(SweepArgs a) {
  extern __shared__ int32_t S[];
  for (long long tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x)
    cc_sweep_tile<PROB, R1, LT, true, true, CMP>(a, tile, S);
}
#undef CC_STEP_CASE

// ---------------------------------------------------------------------------
// Round-based sweep for a homogeneous group (every axis of the group has size
// Q): each thread holds a sub-cube of RR consecutive group axes (C = Q^RR
// words x R1 layers) in registers and applies those RR axis passes there, so
// one LDS round trip covers RR axes instead of one.  The first round loads
// straight from HBM, the last stores straight to HBM; only the rounds in
// between exchange through LDS (a padded layout, idx + idx/16, keeps the
// stride-C accesses of the first round off a single bank; in 16-bit mode
// two layers share one LDS dword, halving the tile's LDS and doubling the
// workgroups a CU holds).  The grid is persistent (occupancy x CUs) and each
// thread loads the next tile's first-round inputs while the current tile
// runs its LDS rounds.  A group whose
// axis count m is not a multiple of RR ends with a cube placed at the top
// RR axes that applies only the axes not yet done.  Layers above the live
// count are zero in registers and stay zero under every recurrence, so the
// arithmetic is the per-fiber arithmetic of cc_fiber, unconditioned.
__device__ __forceinline__ int cc_pad(int i) { return i + (i >> 4); }

template <int PROB, int R1, int Q, int C, int ST>
__device__ __forceinline__ void cc_cube_axis(uint32_t (&v)[R1][C], int R) {
#pragma unroll
  for (int b = 0; b < C; b++) {
    if ((b / ST) % Q != 0) continue;
    uint32_t f[R1][Q];
#pragma unroll
    for (int d = 0; d < R1; d++)
#pragma unroll
      for (int a = 0; a < Q; a++) f[d][a] = v[d][b + a * ST];
    if (PROB == CC_HAMMING && Q == 2) {
      // binary fiber: T_{d-1} - A_{d-1}[a] is the other word's A_{d-1}
#pragma unroll
      for (int d = 1; d < R1; d++) {
        v[d][b] = f[d][0] + f[d - 1][1];
        v[d][b + ST] = f[d][1] + f[d - 1][0];
      }
    } else if (PROB == CC_HAMMING) {
#pragma unroll
      for (int d = 1; d < R1; d++) {
        uint32_t t = 0;
#pragma unroll
        for (int a = 0; a < Q; a++) t += f[d - 1][a];
#pragma unroll
        for (int a = 0; a < Q; a++) v[d][b + a * ST] = f[d][a] + t - f[d - 1][a];
      }
    } else if (PROB == CC_LEE) {
      constexpr int W = Q / 2;
#pragma unroll
      for (int d = 1; d < R1; d++)
#pragma unroll
        for (int a = 0; a < Q; a++) {
          uint32_t s = f[d][a];
#pragma unroll
          for (int w = 1; w <= W; w++)
            if (w <= d) {
              const int am = (a - w + Q) % Q, ap = (a + w) % Q;
              s += f[d - w][am];
              if (ap != am) s += f[d - w][ap];
            }
          v[d][b + a * ST] = s;
        }
    } else {
#pragma unroll
      for (int a = 0; a < Q; a++) {
        uint32_t s = 0;
#pragma unroll
        for (int bb = 0; bb < Q; bb++) {
          int dd = a > bb ? a - bb : bb - a;
          if (PROB == CC_TORUS && Q - dd < dd) dd = Q - dd;
          s += (dd <= R) ? f[0][bb] : 0u;
        }
        v[0][b + a * ST] = s;
      }
    }
  }
}

template <int PROB, int R1, int Q, int C, int I, int RR>
__device__ __forceinline__ void cc_cube_round(uint32_t (&v)[R1][C], int ilo,
                                              int R) {
  if constexpr (I < RR) {
    constexpr int ST = I == 0 ? 1 : (I == 1 ? Q : (I == 2 ? Q * Q : Q * Q * Q));
    if (I >= ilo) cc_cube_axis<PROB, R1, Q, C, ST>(v, R);
    cc_cube_round<PROB, R1, Q, C, I + 1, RR>(v, ilo, R);
  }
}

// First-round inputs of one cube: predicate(cnt), the code indicator, or the
// live layers (the seed kind is block-uniform and branched on once per
// cube).  With TL = 1 the first round's cube is C consecutive words, read
// with 16-byte vector loads.
template <int R1, int C, typename LT, bool CMP = false>
__device__ __forceinline__ void cc_cube_load(const SweepArgs &a, long long base,
                                             int pb, int tl, bool live,
                                             int Lin, uint32_t (&v)[R1][C]) {
#pragma unroll
  for (int d = 0; d < R1; d++)
#pragma unroll
    for (int c = 0; c < C; c++) v[d][c] = 0u;
  if (CMP && a.in_kind == CC_IN_COMPACT) {
    // dword plane k of the cube's words, then every field that lives in it
    // (the field layout is uniform: the branches are scalar and outside the
    // per-word loops, one bit-field extract per layer and word)
    if (!live) return;
    const uint32_t *w0 = a.pk + base + (long long)pb * a.s + tl;
    for (int k = 0; k < CC_CMAXND; k++) {
      if (k >= a.in_nd) break;
      uint32_t q[C];
#pragma unroll
      for (int c = 0; c < C; c++) q[c] = w0[(long long)k * a.pitch + c * a.s];
#pragma unroll
      for (int d = 0; d < R1; d++)
        if (a.idw[d] == k && a.ibits[d])
#pragma unroll
          for (int c = 0; c < C; c++)
            v[d][c] = __builtin_amdgcn_ubfe(q[c], a.ioff[d], a.ibits[d]);
    }
    return;
  }
  if (a.TL == 1 && C % 8 == 0) {
    const long long w0 = base + pb;              // C-aligned, 16-byte aligned
    if (a.in_kind == CC_IN_PRED) {
#pragma unroll
      for (int c8 = 0; c8 < C; c8 += 8) {
        const uint4 q = *(const uint4 *)(a.cnt + w0 + c8);
        const uint32_t wv[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int h = 0; h < 8; h++)
          v[0][c8 + h] = cc_pred(a.pmode, (int)((wv[h >> 1] >> (16 * (h & 1))) & 0xFFFFu),
                                 0, a.mu) ? 1u : 0u;
      }
    } else if (a.in_kind == CC_IN_CODE) {
#pragma unroll
      for (int c4 = 0; c4 < C; c4 += 4) {
        const int4 q = *(const int4 *)(a.A[0] + w0 + c4);
        v[0][c4] = (uint32_t)q.x; v[0][c4 + 1] = (uint32_t)q.y;
        v[0][c4 + 2] = (uint32_t)q.z; v[0][c4 + 3] = (uint32_t)q.w;
      }
    } else {
#pragma unroll
      for (int d = 0; d < R1; d++) {
        if (d >= Lin) break;
        if constexpr (sizeof(LT) == 2) {
#pragma unroll
          for (int c8 = 0; c8 < C; c8 += 8) {
            const uint4 q = *(const uint4 *)(a.H[d] + w0 + c8);
            const uint32_t wv[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int h = 0; h < 8; h++)
              v[d][c8 + h] = (wv[h >> 1] >> (16 * (h & 1))) & 0xFFFFu;
          }
        } else {
#pragma unroll
          for (int c4 = 0; c4 < C; c4 += 4) {
            const int4 q = *(const int4 *)(a.A[d] + w0 + c4);
            v[d][c4] = (uint32_t)q.x; v[d][c4 + 1] = (uint32_t)q.y;
            v[d][c4 + 2] = (uint32_t)q.z; v[d][c4 + 3] = (uint32_t)q.w;
          }
        }
      }
    }
    return;
  }
  if (!live) return;
  const long long w0 = base + (long long)pb * a.s + tl;
  if (a.in_kind == CC_IN_PRED) {
#pragma unroll
    for (int c = 0; c < C; c++)
      v[0][c] = cc_pred(a.pmode, (int)a.cnt[w0 + c * a.s], 0, a.mu) ? 1u : 0u;
  } else if (a.in_kind == CC_IN_CODE) {
#pragma unroll
    for (int c = 0; c < C; c++) v[0][c] = (uint32_t)a.A[0][w0 + c * a.s];
  } else {
#pragma unroll
    for (int d = 0; d < R1; d++) {
      if (d >= Lin) break;
#pragma unroll
      for (int c = 0; c < C; c++) v[d][c] = cc_lay_ld<LT>(a, d, w0 + c * a.s);
    }
  }
}

// Last-round output of one cube (pb, stride qa in p): the produced layers,
// or the layer sum into the map / clamped into cnt.
template <int R1, int C, typename LT, bool CMP = false>
__device__ __forceinline__ void cc_cube_store(const SweepArgs &a, long long base,
                                              int pb, int qa, int tl,
                                              int Lmax, const uint32_t (&v)[R1][C],
                                              unsigned long long &best) {
  const long long sq = (long long)qa * a.s;
  const long long w0 = base + (long long)pb * a.s + tl;
  if (CMP && a.out_kind == CC_OUT_COMPACT) {
    // dword plane k of the cube's words from the fields that live in it
    uint32_t *o0 = a.pk + w0;
    for (int k = 0; k < CC_CMAXND; k++) {
      if (k >= a.out_nd) break;
      uint32_t q[C];
#pragma unroll
      for (int c = 0; c < C; c++) q[c] = 0u;
#pragma unroll
      for (int d = 0; d < R1; d++)
        if (d < Lmax && a.obits[d] && a.odw[d] == k)
#pragma unroll
          for (int c = 0; c < C; c++) q[c] |= v[d][c] << a.ooff[d];
#pragma unroll
      for (int c = 0; c < C; c++) o0[(long long)k * a.pitch + c * sq] = q[c];
    }
    return;
  }
  if (a.out_kind == CC_OUT_LAYERS) {
#pragma unroll
    for (int d = 0; d < R1; d++) {
      if (d >= Lmax) break;
      if (d == 0 && !a.write_l0) continue;
#pragma unroll
      for (int c = 0; c < C; c++) cc_lay_st<LT>(a, d, w0 + c * sq, v[d][c]);
    }
  } else if (a.out_kind == CC_OUT_MAP) {
#pragma unroll
    for (int c = 0; c < C; c++) {
      uint32_t sum = 0;
#pragma unroll
      for (int d = 0; d < R1; d++) sum += v[d][c];
      a.A[0][w0 + c * sq] = (int32_t)sum;
      if (a.amax) cc_key_max(best, sum, w0 + c * sq);
    }
  } else {
#pragma unroll
    for (int c = 0; c < C; c++) {
      uint32_t sum = 0;
#pragma unroll
      for (int d = 0; d < R1; d++) sum += v[d][c];
      const int32_t si = (int32_t)sum;
      a.cnt[w0 + c * sq] = si > 65535 ? (uint16_t)65535 : (uint16_t)si;
    }
  }
}

// LDS exchange of one cube: all R1 layers (two 16-bit layers per dword in
// 16-bit mode), so the unrolled code has no per-layer conditions.
template <int R1, int C, typename LT>
__device__ __forceinline__ void cc_cube_lds_st(int32_t *__restrict__ S, int PTP,
                                               int pb, int qa, int TL, int tl,
                                               const uint32_t (&v)[R1][C]) {
#pragma unroll
  for (int c = 0; c < C; c++) {
    const int e = cc_pad((pb + c * qa) * TL + tl);
    if constexpr (sizeof(LT) == 2) {
#pragma unroll
      for (int d = 0; d < R1; d += 2)
        S[(d >> 1) * PTP + e] = (int32_t)(v[d][c] | (d + 1 < R1 ? v[d + 1][c] << 16 : 0u));
    } else {
#pragma unroll
      for (int d = 0; d < R1; d++) S[d * PTP + e] = (int32_t)v[d][c];
    }
  }
}

template <int R1, int C, typename LT>
__device__ __forceinline__ void cc_cube_lds_ld(const int32_t *__restrict__ S,
                                               int PTP, int pb, int qa, int TL,
                                               int tl, uint32_t (&v)[R1][C]) {
#pragma unroll
  for (int c = 0; c < C; c++) {
    const int e = cc_pad((pb + c * qa) * TL + tl);
    if constexpr (sizeof(LT) == 2) {
#pragma unroll
      for (int d = 0; d < R1; d += 2) {
        const uint32_t wv = (uint32_t)S[(d >> 1) * PTP + e];
        v[d][c] = wv & 0xFFFFu;
        if (d + 1 < R1) v[d + 1][c] = wv >> 16;
      }
    } else {
#pragma unroll
      for (int d = 0; d < R1; d++) v[d][c] = (uint32_t)S[d * PTP + e];
    }
  }
}

// the axis passes of one round; ilo (block-uniform) skips cube axes an
// earlier round already applied
template <int PROB, int R1, int Q, int C, int RR>
__device__ __forceinline__ void cc_cube_apply(uint32_t (&v)[R1][C], int ilo,
                                              int R) {
  switch (ilo) {
    case 0: cc_cube_round<PROB, R1, Q, C, 0, RR>(v, 0, R); break;
    case 1: cc_cube_round<PROB, R1, Q, C, 1, RR>(v, 0, R); break;
    case 2: cc_cube_round<PROB, R1, Q, C, 2, RR>(v, 0, R); break;
    default: cc_cube_round<PROB, R1, Q, C, 3, RR>(v, 0, R); break;
  }
}

// one tile of the round sweep (the batched transform calls it for the
// homogeneous Hamming groups of many cells)
template <int PROB, int R1, int Q, int RR, typename LT, bool CMP = false>
__device__ __forceinline__ void cc_round_tile(const SweepArgs &a, long long tile,
                                              int32_t *__restrict__ S) {
  constexpr int C = RR == 1 ? Q : (RR == 2 ? Q * Q : (RR == 3 ? Q * Q * Q : Q * Q * Q * Q));
  const int m = a.nax;
  const int PT = a.P * a.TL;
  const int PTP = cc_pad(PT) + 1;               // padded layer pitch
  const int Lmax = a.live[m];
  const int Lin = a.live[0];
  const int nrounds = (m + RR - 1) / RR;
  const int ncube = PT / C;
  const long long h = tile / a.ncolb;
  const long long l0 = (tile - h * a.ncolb) * a.TL;
  const int ncol = (int)min((long long)a.TL, a.s - l0);
  const long long base = h * a.s * a.P + l0;
  unsigned long long best = 0;   // argmax key of the map words this thread wrote
  // ---- round 0: HBM -> registers (cube = group axes 0..RR-1)
  {
    const int ca = min(0, m - RR);
    for (int cu = threadIdx.x; cu < ncube; cu += blockDim.x) {
      const int tl = cu % a.TL, pb = (cu / a.TL) * C;
      uint32_t v[R1][C];
      cc_cube_load<R1, C, LT, CMP>(a, base, pb, tl, tl < ncol, Lin, v);
      cc_cube_apply<PROB, R1, Q, C, RR>(v, -ca, a.R);
      if (nrounds == 1) {
        if (tl < ncol) cc_cube_store<R1, C, LT, CMP>(a, base, pb, 1, tl, Lmax, v, best);
      } else {
        cc_cube_lds_st<R1, C, LT>(S, PTP, pb, 1, a.TL, tl, v);
      }
    }
    __syncthreads();
  }
  // ---- rounds 1..: LDS -> registers -> LDS, the last one -> HBM
  for (int rd = 1; rd < nrounds; rd++) {
    const bool last = rd == nrounds - 1;
    const int lo = rd * RR;                     // first axis of this round
    const int ca = min(lo, m - RR);             // cube's lowest group axis
    int qa = 1;                                 // Q^ca: in-tile stride
    for (int i = 0; i < ca; i++) qa *= Q;
    for (int cu = threadIdx.x; cu < ncube; cu += blockDim.x) {
      const int tl = cu % a.TL, u = cu / a.TL;
      const int pb = (u / qa) * qa * C + (u % qa);
      uint32_t v[R1][C];
      cc_cube_lds_ld<R1, C, LT>(S, PTP, pb, qa, a.TL, tl, v);
      cc_cube_apply<PROB, R1, Q, C, RR>(v, lo - ca, a.R);
      if (last) {
        if (tl < ncol) cc_cube_store<R1, C, LT, CMP>(a, base, pb, qa, tl, Lmax, v, best);
      } else {
        cc_cube_lds_st<R1, C, LT>(S, PTP, pb, qa, a.TL, tl, v);
      }
    }
    __syncthreads();
  }
  if (a.amax) cc_key_flush(best, a.amax);
}

template <int PROB, int R1, int Q, int RR, typename LT, bool CMP = false>
__global__ void __launch_bounds__(CC_BLOCK) k_sweep_r(SweepArgs a) {
  extern __shared__ int32_t S[];
  for (long long tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x)
    cc_round_tile<PROB, R1, Q, RR, LT, CMP>(a, tile, S);
}

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
// After axis c only layers 0..min(c,R) can be nonzero, so a pass reads
// lv = min(c,R)+1 layers and writes lw = min(c+1,R)+1 (reference
// dctcore_hamming.cuh:98-110).  Layer 0 never changes for the two layered
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

template <int PROB, int R1, int Q>
__device__ __forceinline__ void cc_fiber(int32_t *__restrict__ S, int PT,
                                         int e0, int es, int lv, int lw,
                                         int R) {
  uint32_t v[R1][Q];
#pragma unroll
  for (int d = 0; d < R1; d++)
#pragma unroll
    for (int a = 0; a < Q; a++)
      v[d][a] = (d < lv) ? (uint32_t)S[d * PT + e0 + a * es] : 0u;
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
          S[d * PT + e0 + a * es] = (int32_t)(v[d][a] + T[d - 1] - v[d - 1][a]);
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
          S[d * PT + e0 + a * es] = (int32_t)s;
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
        s += (dd <= R) ? v[0][b] : 0u;
      }
      S[e0 + a * es] = (int32_t)s;
    }
  }
}

template <int PROB, int R1, int Q>
__device__ __forceinline__ void cc_step(int32_t *__restrict__ S, int PT,
                                        int TL, int nf, int sk, int lv,
                                        int lw, int R) {
  for (int fi = threadIdx.x; fi < nf; fi += CC_BLOCK) {
    const int f = fi / TL, tl = fi - f * TL;
    const int hi = f / sk, lo = f - hi * sk;
    const int e0 = (hi * sk * Q + lo) * TL + tl;
    cc_fiber<PROB, R1, Q>(S, PT, e0, sk * TL, lv, lw, R);
  }
}

#define CC_STEP_CASE(QQ)                                                    \
  case QQ:                                                                  \
    if (QQ >= QMIN && QQ <= QMAX)                                           \
      cc_step<PROB, R1, (QQ >= QMIN && QQ <= QMAX) ? QQ : QMIN>(            \
          S, PT, a.TL, nf, sk, lv, lw, a.R);                                \
    break;

template <int PROB>
struct CcQRange {
  static constexpr int lo = PROB == CC_TORUS ? 3 : 2;
  static constexpr int hi = PROB == CC_HAMMING ? 10 : 16;
};

template <int PROB, int R1>
__global__ void __launch_bounds__(CC_BLOCK) k_sweep(SweepArgs a) {
  extern __shared__ int32_t S[];
  constexpr int QMIN = CcQRange<PROB>::lo, QMAX = CcQRange<PROB>::hi;
  const int PT = a.P * a.TL;
  const int Lmax = (R1 == 1) ? 1 : min(a.c0 + a.nax, a.R) + 1;
  const int Lin = (R1 == 1) ? 1 : min(a.c0, a.R) + 1;
  for (long long tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    const long long h = tile / a.ncolb;
    const long long l0 = (tile - h * a.ncolb) * a.TL;
    const int ncol = (int)min((long long)a.TL, a.s - l0);
    const long long base = h * a.s * a.P + l0;
    // ---- stage the tile: every live layer of every word, coalesced along tl
    for (int e = threadIdx.x; e < PT; e += CC_BLOCK) {
      const int p = e / a.TL, tl = e - p * a.TL;
      const bool live = tl < ncol;
      const long long w = base + (long long)p * a.s + tl;
      if (a.in_kind == CC_IN_PRED)
        S[e] = (live && cc_pred(a.pmode, (int)a.cnt[w], 0, a.mu)) ? 1 : 0;
      else
        S[e] = live ? a.A[0][w] : 0;
      for (int d = 1; d < Lmax; d++)
        S[d * PT + e] =
            (live && d < Lin && a.in_kind == CC_IN_LAYERS) ? a.A[d][w] : 0;
    }
    __syncthreads();
    // ---- every axis of the group, in LDS
    for (int k = 0; k < a.nax; k++) {
      const int c = a.c0 + k;
      const int lv = (R1 == 1) ? 1 : min(c, a.R) + 1;
      const int lw = (R1 == 1) ? 1 : min(c + 1, a.R) + 1;
      const int Q = a.Q[k], sk = a.sk[k];
      const int nf = (a.P / Q) * a.TL;
      switch (Q) {
        CC_STEP_CASE(2) CC_STEP_CASE(3) CC_STEP_CASE(4) CC_STEP_CASE(5)
        CC_STEP_CASE(6) CC_STEP_CASE(7) CC_STEP_CASE(8) CC_STEP_CASE(9)
        CC_STEP_CASE(10) CC_STEP_CASE(11) CC_STEP_CASE(12) CC_STEP_CASE(13)
        CC_STEP_CASE(14) CC_STEP_CASE(15) CC_STEP_CASE(16)
        default: break;
      }
      __syncthreads();
    }
    // ---- write back: layers, or the layer sum into the map / cnt plane
    for (int e = threadIdx.x; e < PT; e += CC_BLOCK) {
      const int p = e / a.TL, tl = e - p * a.TL;
      if (tl >= ncol) continue;
      const long long w = base + (long long)p * a.s + tl;
      if (a.out_kind == CC_OUT_LAYERS) {
        for (int d = a.write_l0 ? 0 : 1; d < Lmax; d++) a.A[d][w] = S[d * PT + e];
      } else {
        uint32_t sum = 0;
        for (int d = 0; d < Lmax; d++) sum += (uint32_t)S[d * PT + e];
        if (a.out_kind == CC_OUT_MAP) {
          a.A[0][w] = (int32_t)sum;
        } else {
          const int32_t si = (int32_t)sum;
          a.cnt[w] = si > 65535 ? (uint16_t)65535 : (uint16_t)si;
        }
      }
    }
    __syncthreads();
  }
}
#undef CC_STEP_CASE

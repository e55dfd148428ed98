// cc_psweep.hpp -- packed-layer sweeps of the Hamming transform.
//
// The R+1 distance layers of one word travel together in ND dwords of W-bit
// fields (W = 8 or 16): layer d sits in dword d / (32/W), slot d % (32/W);
// slots above R are don't-care.  The host proves from closed-form bounds
// that every layer, and every fiber sum T_d with d < R, stays below 2^W for
// the whole sweep, so whole-dword integer adds never carry between layers.
// The Hamming fiber recurrence of one axis
//     newA_d[a] = A_d[a] + T_{d-1} - A_{d-1}[a],   T_d = sum_a A_d[a]
// (reference src/plugins/covering/dctcore_hamming.cuh:4-8) becomes, per word
// of the fiber,
//     v[a] += up(T - v[a])        (T = sum_a v[a], fieldwise)
// where up() moves every field one slot up (layer d -> d + 1): a shift of the
// top dword and a funnel shift (alignbit) of the lower ones.  For Q = 2,
// T - v[a] is the other word of the pair, so a binary axis is one add of
// the partner shifted up.  Field R of T may overflow: its carry and borrow
// only reach don't-care slots or leave the top dword, and up() drops it.
//
// Between sweeps the layers go to HBM in a compact format: field d takes
// exactly the bits its bound needs (layer d after c axes of a seed <= s0 is
// at most s0 * C(c,d) * (Q-1)^d), packed into one or two dword planes
// (plane-major, so every access is a coalesced dword).  For K_2(24,3) the
// layers after 12 and 18 axes need 20 and 24 bits: 4 bytes per word instead
// of 8 or 16.
//
// Tile and rounds: a sweep applies the passes of nax consecutive axes.  As
// in the unpacked round kernel, a thread holds a cube of RR group axes
// (C = Q^RR words) in registers and applies those passes there; rounds
// exchange through LDS.  Round 0 takes the top RR axes, whose cube elements
// are P / C words apart, so consecutive threads read consecutive words even
// in the contiguous first sweep; the later rounds take the lower axes in
// ascending chunks.  A contiguous first sweep with several rounds writes its
// result back to LDS and stores it in the round-0 shape (coalesced).  A
// strided sweep stages TL consecutive columns per row, so a wave's accesses
// are coalesced in every round; when the whole group fits one cube the
// sweep needs no LDS at all (one fiber cube per thread).
#pragma once
#include "cc_common.hpp"

#define CC_IN_PACKED 3   // layers from the compact planes
#define CC_OUT_PACKED 3  // layers to the compact planes
#define CC_PK_MAXRD 12   // rounds per sweep

struct PSweepArgs {
  const uint16_t *cnt;   // CC_IN_PRED: predicate source
  const int32_t *code;   // CC_IN_CODE: code multiplicity plane
  uint32_t *pk;          // compact layer planes (in and out, in place)
  int32_t *map;          // CC_OUT_MAP
  uint16_t *cnt_out;     // CC_OUT_CNT
  long long pitch;       // dwords between compact planes
  long long s;           // stride (words) of the group's first axis
  long long ncolb;       // column blocks per outer index = ceil(s / TL)
  long long ntiles;
  int P, TL, nax, L;
  int in_kind, out_kind;
  int plo, phi;          // CC_IN_PRED: [plo <= cnt <= phi]
  int in_nd, out_nd;     // compact planes read / written
  int nrounds;
  int ca[CC_PK_MAXRD];     // lowest group axis of each round's cube
  int amask[CC_PK_MAXRD];  // cube axes the round applies (bit i: axis ca + i)
  int qa[CC_PK_MAXRD];     // Q^ca: in-tile stride of the round's cube
  uint32_t fmask[4];       // real-layer slots of each register dword
  int idw[CC_MAXL], ioff[CC_MAXL], ibits[CC_MAXL];
  int odw[CC_MAXL], ooff[CC_MAXL], obits[CC_MAXL];
};

constexpr int cc_ipow(int q, int e) { return e == 0 ? 1 : q * cc_ipow(q, e - 1); }

__device__ __forceinline__ int pk_pad(int i) { return i + (i >> 6); }

// Scheduling fence every 16 cube elements: the compiler may not hoist the
// address arithmetic of a whole 64-word cube (one VGPR per element) ahead of
// its memory operations.  Loads already issued stay in flight across it.
#define PK_FENCE(c)                                                     \
  do {                                                                  \
    if (((c) & 15) == 15) __builtin_amdgcn_sched_barrier(0);            \
  } while (0)

// x += up(y): every field of y moved one slot up, added fieldwise
template <int W, int ND>
__device__ __forceinline__ void pk_add_up(uint32_t (&x)[ND],
                                          const uint32_t (&y)[ND]) {
#pragma unroll
  for (int k = ND - 1; k > 0; k--)
    x[k] += __builtin_amdgcn_alignbit(y[k], y[k - 1], 32 - W);
  x[0] += y[0] << W;
}

// one axis of the cube: fibers of Q words ST apart
template <int Q, int C, int ST, int W, int ND>
__device__ __forceinline__ void pk_axis(uint32_t (&v)[C][ND]) {
#pragma unroll
  for (int b = 0; b < C; b++) {
    if ((b / ST) % Q != 0) continue;
    if constexpr (Q == 2) {
      uint32_t x0[ND], x1[ND];
#pragma unroll
      for (int k = 0; k < ND; k++) {
        x0[k] = v[b][k];
        x1[k] = v[b + ST][k];
      }
      pk_add_up<W, ND>(v[b], x1);
      pk_add_up<W, ND>(v[b + ST], x0);
    } else {
      uint32_t T[ND];
#pragma unroll
      for (int k = 0; k < ND; k++) {
        T[k] = 0;
#pragma unroll
        for (int a = 0; a < Q; a++) T[k] += v[b + a * ST][k];
      }
#pragma unroll
      for (int a = 0; a < Q; a++) {
        uint32_t D[ND];
#pragma unroll
        for (int k = 0; k < ND; k++) D[k] = T[k] - v[b + a * ST][k];
        pk_add_up<W, ND>(v[b + a * ST], D);
      }
    }
  }
}

template <int Q, int C, int I, int RR, int W, int ND>
__device__ __forceinline__ void pk_apply(uint32_t (&v)[C][ND], int mask) {
  if constexpr (I < RR) {
    if (mask & (1 << I)) pk_axis<Q, C, cc_ipow(Q, I), W, ND>(v);
    pk_apply<Q, C, I + 1, RR, W, ND>(v, mask);
  }
}

// Global addressing: the word of cube element c is ub + lane + c * usq,
// where ub (the tile's first word) is workgroup-uniform and lane + c * usq
// is a 32-bit element offset (the host checks that every offset inside a
// tile, in bytes, stays below 2^32).  The tile base is read once into
// scalar registers (readfirstlane makes it opaque, so the compiler keeps
// it there) and every access is the scalar-base + 32-bit VGPR-offset form.
#define CC_GLOBAL __attribute__((address_space(1)))
typedef uint32_t pk_u32x4 __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __forceinline__ CC_GLOBAL T *pk_tile(const T *plane, long long ub) {
  const uintptr_t x = (uintptr_t)(plane + ub);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
  return (CC_GLOBAL T *)(((uintptr_t)hi << 32) | lo);
}
template <typename T>
__device__ __forceinline__ T pk_ld(const CC_GLOBAL T *t, uint32_t e) {
  return *(const CC_GLOBAL T *)((const CC_GLOBAL char *)t + e * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void pk_st(CC_GLOBAL T *t, uint32_t e, T v) {
  *(CC_GLOBAL T *)((CC_GLOBAL char *)t + e * (uint32_t)sizeof(T)) = v;
}

// seed or compact layers of C words into register format.  Compact words
// are loaded straight into the register array (at most two dwords per
// word, never more than ND) and converted in place.  The predicate
// [plo <= cnt <= phi] covers EQ (plo = phi = t), DEF (0 .. mu-1) and CRIT
// (1 .. mu) without branches.
template <int C, int W, int ND>
__device__ __forceinline__ void pk_load(const PSweepArgs &a, long long ub,
                                        uint32_t usq, uint32_t lane,
                                        bool live, uint32_t (&v)[C][ND]) {
  constexpr int S = 32 / W;
#pragma unroll
  for (int c = 0; c < C; c++)
#pragma unroll
    for (int k = 0; k < ND; k++) v[c][k] = 0u;
  if (!live) return;
  if (a.in_kind == CC_IN_PRED) {
    const CC_GLOBAL uint16_t *t = pk_tile(a.cnt, ub);
    const uint32_t lo = (uint32_t)a.plo, span = (uint32_t)(a.phi - a.plo);
    if (C % 8 == 0 && usq == 1) {
      // contiguous cube (16-byte aligned): 8 counters per load
#pragma unroll
      for (int c = 0; c < C; c += 8) {
        const pk_u32x4 q = *(const CC_GLOBAL pk_u32x4 *)(t + lane + c);
        const uint32_t wv[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int h = 0; h < 8; h++)
          v[c + h][0] = (wv[h >> 1] >> (16 * (h & 1))) & 0xFFFFu;
      }
    } else {
#pragma unroll
      for (int c = 0; c < C; c++) {
        v[c][0] = (uint32_t)pk_ld(t, lane + c * usq);
        PK_FENCE(c);
      }
    }
#pragma unroll
    for (int c = 0; c < C; c++) v[c][0] = (v[c][0] - lo) <= span ? 1u : 0u;
  } else if (a.in_kind == CC_IN_CODE) {
    const CC_GLOBAL int32_t *t = pk_tile(a.code, ub);
    if (C % 4 == 0 && usq == 1) {
#pragma unroll
      for (int c = 0; c < C; c += 4) {
        const pk_u32x4 q = *(const CC_GLOBAL pk_u32x4 *)(t + lane + c);
        v[c][0] = q.x; v[c + 1][0] = q.y; v[c + 2][0] = q.z; v[c + 3][0] = q.w;
      }
    } else {
#pragma unroll
      for (int c = 0; c < C; c++) {
        v[c][0] = (uint32_t)pk_ld(t, lane + c * usq);
        PK_FENCE(c);
      }
    }
  } else {
    const CC_GLOBAL uint32_t *t = pk_tile((const uint32_t *)a.pk, ub);
#pragma unroll
    for (int c = 0; c < C; c++) {
      v[c][0] = pk_ld(t, lane + c * usq);
      PK_FENCE(c);
    }
    if constexpr (ND > 1) {
      if (a.in_nd > 1) {
        const CC_GLOBAL uint32_t *t1 = pk_tile((const uint32_t *)a.pk, ub + a.pitch);
#pragma unroll
        for (int c = 0; c < C; c++) {
          v[c][1] = pk_ld(t1, lane + c * usq);
          PK_FENCE(c);
        }
      }
    }
    // per-layer extraction, straight-line: layers >= L have width 0
    int off[ND * S], bits[ND * S];
    bool hi[ND * S];
#pragma unroll
    for (int d = 0; d < ND * S; d++) {
      const bool on = d < a.L;
      off[d] = on ? a.ioff[d] : 0;
      bits[d] = on ? a.ibits[d] : 0;
      hi[d] = on && a.idw[d] != 0;
    }
#pragma unroll
    for (int c = 0; c < C; c++) {
      const uint32_t x0 = v[c][0];
      const uint32_t x1 = ND > 1 ? v[c][ND > 1 ? 1 : 0] : 0u;
      uint32_t r[ND];
#pragma unroll
      for (int k = 0; k < ND; k++) r[k] = 0u;
#pragma unroll
      for (int d = 0; d < ND * S; d++)
        r[d / S] |= __builtin_amdgcn_ubfe(hi[d] ? x1 : x0, off[d], bits[d])
                    << (W * (d % S));
#pragma unroll
      for (int k = 0; k < ND; k++) v[c][k] = r[k];
    }
  }
}

// sum of the real layer fields of one word
template <int W, int ND>
__device__ __forceinline__ uint32_t pk_sum(const PSweepArgs &a,
                                           const uint32_t (&r)[ND]) {
  constexpr int S = 32 / W;
  constexpr uint32_t M = (1u << W) - 1u;
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < ND; k++) {
    const uint32_t x = r[k] & a.fmask[k];
#pragma unroll
    for (int i = 0; i < S; i++) sum += (x >> (W * i)) & M;
  }
  return sum;
}

// C words out (addressed as in pk_load): compact layers, the layer sum into
// the map, or the clamped sum into cnt
template <int C, int W, int ND, bool SUM>
__device__ __forceinline__ void pk_store(const PSweepArgs &a, long long ub,
                                         uint32_t usq, uint32_t lane,
                                         const uint32_t (&v)[C][ND]) {
  constexpr int S = 32 / W;
  constexpr uint32_t M = (1u << W) - 1u;
  if constexpr (!SUM) {
    // compact fields, straight-line: layer d goes to bit ooff[d] of plane
    // odw[d]; layers >= L are masked off
    uint32_t en[ND * S], lo[ND * S];
    int sh[ND * S];
#pragma unroll
    for (int d = 0; d < ND * S; d++) {
      const bool on = d < a.L;
      en[d] = on ? M : 0u;
      lo[d] = on && a.odw[d] == 0 ? 0xFFFFFFFFu : 0u;
      sh[d] = on ? a.ooff[d] : 0;
    }
    CC_GLOBAL uint32_t *t = pk_tile(a.pk, ub);
    if (ND > 1 && a.out_nd > 1) {
      CC_GLOBAL uint32_t *t1 = pk_tile(a.pk, ub + a.pitch);
#pragma unroll
      for (int c = 0; c < C; c++) {
        uint32_t o0 = 0u, o1 = 0u;
#pragma unroll
        for (int d = 0; d < ND * S; d++) {
          const uint32_t f = ((v[c][d / S] >> (W * (d % S))) & en[d]) << sh[d];
          o0 |= f & lo[d];
          o1 |= f & ~lo[d];
        }
        pk_st(t, lane + c * usq, o0);
        pk_st(t1, lane + c * usq, o1);
        PK_FENCE(c);
      }
    } else {
#pragma unroll
      for (int c = 0; c < C; c++) {
        uint32_t o0 = 0u;
#pragma unroll
        for (int d = 0; d < ND * S; d++)
          o0 |= ((v[c][d / S] >> (W * (d % S))) & en[d]) << sh[d];
        pk_st(t, lane + c * usq, o0);
        PK_FENCE(c);
      }
    }
  } else {
    // one loop for both sums, so the per-word sums are formed next to
    // their stores (not hoisted ahead of the branch for all C words)
    const bool to_map = a.out_kind == CC_OUT_MAP;
    CC_GLOBAL int32_t *tm = pk_tile(a.map, ub);
    CC_GLOBAL uint16_t *tc = pk_tile(a.cnt_out, ub);
#pragma unroll
    for (int c = 0; c < C; c++) {
      const uint32_t s = pk_sum<W, ND>(a, v[c]);
      if (to_map) pk_st(tm, lane + c * usq, (int32_t)s);
      else pk_st(tc, lane + c * usq, s > 65535u ? (uint16_t)65535 : (uint16_t)s);
      PK_FENCE(c);
    }
  }
}

template <int C, int ND>
__device__ __forceinline__ void pk_lds_st(uint32_t *__restrict__ S, int PTP,
                                          int pb, int qa, int TL, int tl,
                                          const uint32_t (&v)[C][ND]) {
#pragma unroll
  for (int c = 0; c < C; c++) {
    const int e = pk_pad((pb + c * qa) * TL + tl);
#pragma unroll
    for (int k = 0; k < ND; k++) S[k * PTP + e] = v[c][k];
    PK_FENCE(c);
  }
}

template <int C, int ND>
__device__ __forceinline__ void pk_lds_ld(const uint32_t *__restrict__ S,
                                          int PTP, int pb, int qa, int TL,
                                          int tl, uint32_t (&v)[C][ND]) {
#pragma unroll
  for (int c = 0; c < C; c++) {
    const int e = pk_pad((pb + c * qa) * TL + tl);
#pragma unroll
    for (int k = 0; k < ND; k++) v[c][k] = S[k * PTP + e];
    PK_FENCE(c);
  }
}

// SUM: the sweep ends the transform (layer sum into the map or cnt);
// otherwise it writes compact layers
template <int Q, int RR, int W, int ND, bool SUM>
__global__ void __launch_bounds__(CC_BLOCK) k_psweep(PSweepArgs a) {
  extern __shared__ uint32_t Sp[];
  constexpr int C = cc_ipow(Q, RR);
  const int PT = a.P * a.TL;
  const int PTP = pk_pad(PT) + 1;
  const int ncube = PT / C;
  const int nt = blockDim.x;
  for (long long tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    const long long h = tile / a.ncolb;
    const long long l0 = (tile - h * a.ncolb) * a.TL;
    const int ncol = (int)min((long long)a.TL, a.s - l0);
    const long long base = h * a.s * a.P + l0;
    for (int rd = 0; rd < a.nrounds; rd++) {
      const int qa = a.qa[rd];
      const bool first = rd == 0, last = rd == a.nrounds - 1;
      for (int cu = threadIdx.x; cu < ncube; cu += nt) {
        const int tl = cu % a.TL, u = cu / a.TL;
        const int pb = (u / qa) * qa * C + (u % qa);
        const uint32_t lane = (uint32_t)((long long)pb * a.s + tl);
        const uint32_t usq = (uint32_t)((long long)qa * a.s);
        uint32_t v[C][ND];
        if (first)
          pk_load<C, W, ND>(a, base, usq, lane, tl < ncol, v);
        else
          pk_lds_ld<C, ND>(Sp, PTP, pb, qa, a.TL, tl, v);
        pk_apply<Q, C, 0, RR, W, ND>(v, a.amask[rd]);
        if (last) {
          if (tl < ncol) pk_store<C, W, ND, SUM>(a, base, usq, lane, v);
        } else {
          pk_lds_st<C, ND>(Sp, PTP, pb, qa, a.TL, tl, v);
        }
      }
      __syncthreads();
    }
  }
}

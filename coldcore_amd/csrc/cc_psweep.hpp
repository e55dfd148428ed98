// cc_psweep.hpp -- packed-layer sweeps of the Hamming transform.
//
// The R+1 distance layers of one word travel together in ND dwords of W-bit
// fields (W = 8 or 16), interleaved: layer d sits in dword d % ND, slot
// d / ND.  Slots whose layer index is above R are don't-care.  The host
// proves from closed-form bounds that every layer, and every fiber sum T_d
// with d < R, stays below 2^W for the whole sweep, so whole-dword integer
// adds never carry between layers.  The Hamming fiber recurrence of one axis
//     newA_d[a] = A_d[a] + T_{d-1} - A_{d-1}[a],   T_d = sum_a A_d[a]
// (reference src/plugins/covering/dctcore_hamming.cuh:4-8) becomes, per word
// of the fiber,
//     v[a] += up(T - v[a])        (T = sum_a v[a], fieldwise)
// where up() moves every layer d to d + 1.  With the interleaved layout that
// is one add per dword: dword k (k > 0) receives dword k-1 at the same
// slots, dword 0 receives dword ND-1 one slot up.  For Q = 2, T - v[a] is the
// other word of the pair, so a binary axis costs ND adds per word.  Field R
// of T may overflow: its carry and borrow only reach don't-care slots (the
// slots above a real layer in the same dword hold higher layers) or leave
// the dword, and up() drops layer R.
//
// Between sweeps the layers go to HBM in a compact format: either one byte
// per layer in one dword (when R <= 3 and every bound is <= 255), or field d
// taking exactly the bits its bound needs (layer d after c axes of a seed
// <= s0 is at most s0 * C(c,d) * (Q-1)^d), packed into one or two dword
// planes (plane-major, so every access is a coalesced dword).  For K_2(24,3)
// the layers after 12 and 18 axes need 20 and 24 bits: 4 bytes per word.
// The byte format converts to and from the register formats with byte
// permutes (v_perm_b32).
//
// Tile and rounds: a sweep applies the passes of nax consecutive axes.  A
// thread holds a cube of RR group axes (C = Q^RR words) in registers and
// applies those passes there; rounds exchange through LDS.  Round 0 takes
// the group's lowest RR axes (contiguous words in the first sweep: vector
// loads), the middle rounds the axes above in ascending chunks, the last
// round the top RR axes, whose cube elements are P / C words apart, so
// consecutive threads store consecutive words even in the contiguous first
// sweep.  A strided sweep stages TL consecutive columns per row, so a
// wave's accesses are coalesced in every round; when the whole group fits
// one cube the sweep needs no LDS at all (one fiber cube per thread).
#pragma once
#include "cc_common.hpp"

#define CC_IN_PACKED 3   // layers from the compact planes
#define CC_IN_SPARSE 4   // layer 0 = the code's multiplicities, from its sorted
                         // word list (k_pfirstp only: no plane is read)
#define CC_OUT_PACKED 3  // layers to the compact planes
#define CC_PK_MAXRD 12   // rounds per sweep
// cache-policy bits of a store: 0 default, CC_ST_NT non-temporal (gfx950:
// the line is kept in L2 for a short time only)
#define CC_ST_NT 2

struct PSweepArgs {
  const uint16_t *cnt;   // CC_IN_PRED: predicate source
  const int32_t *code;   // CC_IN_CODE: code multiplicity plane
  // CC_IN_SPARSE: the code's distinct words ascending (cw) with their
  // multiplicities (cm); the words of first-sweep tile t are entries
  // [toff[t], toff[t + 1])
  const long long *cw;
  const int32_t *cm;
  const int32_t *toff;
  uint32_t *pk;          // compact layer planes (in and out, in place)
  int32_t *map;          // CC_OUT_MAP
  uint16_t *cnt_out;     // CC_OUT_CNT
  // the last sweep folds a census of what it writes into the census slots
  // at fold (null: none; cc_census_flush): the map maximum or the cnt
  // deficit sum max(0, fold_mu - cnt); the first sweep of the transform
  // zeroes the slots at fold_zero first
  unsigned long long *fold;
  unsigned long long *fold_zero;
  int fold_mu;
  int fold_stride;
  long long pitch;       // dwords between compact planes
  long long X;           // words of every plane (buffer bounds)
  long long s;           // stride (words) of the group's first axis
  long long ncolb;       // column blocks per outer index = ceil(s / TL)
  long long ntiles;
  int P, TL, nax, L;
  int kern;              // 0 k_psweep, 1 / 2 k_pcol (lanes per cube), 3 k_pfirst,
                         // 4 k_pfirstp
  int in_kind, out_kind;
  int plo, phi;          // CC_IN_PRED: [plo <= cnt <= phi]
  int in_nd, out_nd;     // compact planes read / written
  int in_w8, out_w8;     // compact format is one byte per layer
  int in_nar, out_nar;   // narrow compact word: 1 = 16 bits (a uint16
                         // plane), 2 = 24 bits (a uint16 and a uint8 plane)
  long long nar_off;     // dwords from pk to the narrow planes (the uint8
                         // plane starts 2 * pitch bytes after the uint16 one)
  uint32_t out_bmask;    // byte format: bytes of real layers
  int st_nt;             // non-temporal stores
  // k_pstrip: the one-dword format of the LDS exchange between its rounds
  // (layer d in bits [moff, moff + mbits)); mid_w8: one byte per layer
  int mid_w8;
  int moff[CC_MAXL], mbits[CC_MAXL];
  int na;                // k_pfirstw: axes of its last (widening) round
  int lcmp;              // k_psweep with two rounds: the LDS tile holds one
                         // compact dword per word (format moff / mbits)
  int nrounds;
  int ca[CC_PK_MAXRD];     // lowest group axis of each round's cube
  int amask[CC_PK_MAXRD];  // cube axes the round applies (bit i: axis ca + i)
  int qa[CC_PK_MAXRD];     // Q^ca: in-tile stride of the round's cube
  int lin[CC_PK_MAXRD];    // round's LDS addresses are e0 + c * kp
  int kp[CC_PK_MAXRD];
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

// x += up(y): every layer of y moved to the next layer index, added
template <int W, int ND>
__device__ __forceinline__ void pk_add_up(uint32_t (&x)[ND],
                                          const uint32_t (&y)[ND]) {
#pragma unroll
  for (int k = ND - 1; k > 0; k--) x[k] += y[k - 1];
  x[0] += y[ND - 1] << W;
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

// Global addressing: raw buffer operations.  The word of cube element c is
// ub + lane + c * usq: the tile base ub is workgroup-uniform and becomes a
// buffer descriptor (scalar registers), the per-thread part `lane` is one
// 32-bit byte offset shared by every element, and c * usq is a scalar
// offset -- no per-element vector address arithmetic.  The descriptor's
// record count is the plane's remaining length, so the hardware drops any
// access outside the plane (loads return 0).  The host checks that every
// offset inside a tile, in bytes, stays below 2^32.
typedef uint32_t pk_u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pk_rsrc(const void *plane,
                                                          long long ub,
                                                          long long len,
                                                          int esize) {
  const uintptr_t x = (uintptr_t)plane + (uintptr_t)(ub * esize);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
  const long long rem = (len - ub) * esize;
  const uint32_t n = rem <= 0 ? 0u : (rem >= 0xFFFFFFFFLL ? 0xFFFFFFFFu : (uint32_t)rem);
  return __builtin_amdgcn_make_buffer_rsrc(
      (void *)(((uintptr_t)hi << 32) | lo), (short)0,
      (int)__builtin_amdgcn_readfirstlane(n), 0x00020000);
}
#define PK_LD32(r, c) \
  __builtin_amdgcn_raw_buffer_load_b32((r), (int)(lane * 4u), (int)((c) * usq * 4u), 0)
#define PK_LD16(r, c) \
  ((uint32_t)__builtin_amdgcn_raw_buffer_load_b16((r), (int)(lane * 2u), (int)((c) * usq * 2u), 0))
#define PK_ST32(r, c, val) \
  __builtin_amdgcn_raw_buffer_store_b32((uint32_t)(val), (r), (int)(lane * 4u), (int)((c) * usq * 4u), ST_AUX)
#define PK_ST16(r, c, val) \
  __builtin_amdgcn_raw_buffer_store_b16((unsigned short)(val), (r), (int)(lane * 2u), (int)((c) * usq * 2u), ST_AUX)
#define PK_LD8(r, c) \
  ((uint32_t)__builtin_amdgcn_raw_buffer_load_b8((r), (int)lane, (int)((c) * usq), 0))
#define PK_ST8(r, c, val) \
  __builtin_amdgcn_raw_buffer_store_b8((unsigned char)(val), (r), (int)lane, (int)((c) * usq), ST_AUX)

// the narrow compact planes of a sweep
__device__ __forceinline__ const uint16_t *pk_nar16(const PSweepArgs &a) {
  return (const uint16_t *)(a.pk + a.nar_off);
}
__device__ __forceinline__ const uint8_t *pk_nar8(const PSweepArgs &a) {
  return (const uint8_t *)(a.pk + a.nar_off) + 2 * a.pitch;
}

// byte-format dword (byte d = layer d, R <= 3) -> register format
template <int W, int ND>
__device__ __forceinline__ void pk_from_bytes(uint32_t x, uint32_t (&r)[ND]) {
#pragma unroll
  for (int k = 0; k < ND; k++) r[k] = 0u;
  if constexpr (W == 8) {
    r[0] = x;                                           // same layout
  } else if constexpr (ND == 1) {
    r[0] = __builtin_amdgcn_perm(0u, x, 0x0C010C00u);   // [A0, A1]
  } else {
    r[0] = __builtin_amdgcn_perm(0u, x, 0x0C020C00u);   // [A0, A2]
    r[1] = __builtin_amdgcn_perm(0u, x, 0x0C030C01u);   // [A1, A3]
  }
}

// register format -> byte-format dword (real layers only)
template <int W, int ND>
__device__ __forceinline__ uint32_t pk_to_bytes(const uint32_t (&r)[ND],
                                                uint32_t bmask) {
  if constexpr (W == 8) {
    return r[0] & bmask;
  } else if constexpr (ND == 1) {
    return __builtin_amdgcn_perm(0u, r[0], 0x0C0C0200u) & bmask;
  } else {
    return __builtin_amdgcn_perm(r[1], r[0], 0x06020400u) & bmask;
  }
}

// compact words (in v[c][0], and v[c][1] for a second plane) -> register
// format, in place
template <int C, int W, int ND, bool IN2>
__device__ __forceinline__ void pk_unpack(const PSweepArgs &a, uint32_t (&v)[C][ND]) {
  if (!IN2 && a.in_w8) {
#pragma unroll
    for (int c = 0; c < C; c++) {
      uint32_t r[ND];
      pk_from_bytes<W, ND>(v[c][0], r);
#pragma unroll
      for (int k = 0; k < ND; k++) v[c][k] = r[k];
    }
  } else {
    // general compact fields: layer d is bits [off, off + bits) of
    // dword plane idw[d]; layers >= L have width 0
    constexpr int NS = ND * (32 / W);
    int off[NS], bits[NS];
    bool hi[NS];
#pragma unroll
    for (int d = 0; d < NS; d++) {
      const bool on = d < a.L;
      off[d] = on ? a.ioff[d] : 0;
      bits[d] = on ? a.ibits[d] : 0;
      hi[d] = on && a.idw[d] != 0;
    }
#pragma unroll
    for (int c = 0; c < C; c++) {
      const uint32_t x0 = v[c][0];
      const uint32_t x1 = (ND > 1 && IN2) ? v[c][ND > 1 ? 1 : 0] : 0u;
      uint32_t r[ND];
#pragma unroll
      for (int k = 0; k < ND; k++) r[k] = 0u;
#pragma unroll
      for (int d = 0; d < NS; d++) {
        const uint32_t src = (ND > 1 && IN2) ? (hi[d] ? x1 : x0) : x0;
        r[d % ND] |= __builtin_amdgcn_ubfe(src, off[d], bits[d]) << (W * (d / ND));
      }
#pragma unroll
      for (int k = 0; k < ND; k++) v[c][k] = r[k];
    }
  }
}

// seed or compact layers of C words into register format.  Compact words
// are loaded straight into the register array (at most two dwords per
// word, never more than ND) and converted in place.  The predicate
// [plo <= cnt <= phi] covers EQ (plo = phi = t), DEF (0 .. mu-1) and CRIT
// (1 .. mu) without branches.
// SEED (compile-time): the first sweep reads a seed (predicate or code
// plane, branched on at run time); later sweeps read compact layers.
template <int C, int W, int ND, bool SEED, bool IN2>
__device__ __forceinline__ void pk_load(const PSweepArgs &a, long long ub,
                                        uint32_t usq, uint32_t lane,
                                        bool live, uint32_t (&v)[C][ND]) {
#pragma unroll
  for (int c = 0; c < C; c++)
#pragma unroll
    for (int k = 0; k < ND; k++) v[c][k] = 0u;
  if (!live) return;
  if (SEED && a.in_kind == CC_IN_PRED) {
    const __amdgpu_buffer_rsrc_t r = pk_rsrc(a.cnt, ub, a.X, 2);
    const uint32_t lo = (uint32_t)a.plo, span = (uint32_t)(a.phi - a.plo);
    if (C % 8 == 0 && usq == 1) {
      // contiguous cube (16-byte aligned): 8 counters per load
#pragma unroll
      for (int c = 0; c < C; c += 8) {
        const pk_u32x4 q = __builtin_bit_cast(pk_u32x4,
            __builtin_amdgcn_raw_buffer_load_b128(r, (int)(lane * 2u), c * 2, 0));
        const uint32_t wv[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int h = 0; h < 8; h++)
          v[c + h][0] = (wv[h >> 1] >> (16 * (h & 1))) & 0xFFFFu;
      }
    } else {
#pragma unroll
      for (int c = 0; c < C; c++) {
        v[c][0] = PK_LD16(r, c);
        PK_FENCE(c);
      }
    }
#pragma unroll
    for (int c = 0; c < C; c++) v[c][0] = (v[c][0] - lo) <= span ? 1u : 0u;
  } else if (SEED) {
    const __amdgpu_buffer_rsrc_t r = pk_rsrc(a.code, ub, a.X, 4);
    if (C % 4 == 0 && usq == 1) {
#pragma unroll
      for (int c = 0; c < C; c += 4) {
        const pk_u32x4 q = __builtin_bit_cast(pk_u32x4,
            __builtin_amdgcn_raw_buffer_load_b128(r, (int)(lane * 4u), c * 4, 0));
        v[c][0] = q.x; v[c + 1][0] = q.y; v[c + 2][0] = q.z; v[c + 3][0] = q.w;
      }
    } else {
#pragma unroll
      for (int c = 0; c < C; c++) {
        v[c][0] = PK_LD32(r, c);
        PK_FENCE(c);
      }
    }
  } else if (a.in_nar) {
    // narrow word: 16 bits from the uint16 plane, bits 16..23 from the
    // uint8 plane; its fields are then extracted as from one dword
    const __amdgpu_buffer_rsrc_t r16 = pk_rsrc(pk_nar16(a), ub, a.X, 2);
#pragma unroll
    for (int c = 0; c < C; c++) {
      v[c][0] = PK_LD16(r16, c);
      PK_FENCE(c);
    }
    if (a.in_nar == 2) {
      const __amdgpu_buffer_rsrc_t r8 = pk_rsrc(pk_nar8(a), ub, a.X, 1);
#pragma unroll
      for (int c = 0; c < C; c++) {
        v[c][0] |= PK_LD8(r8, c) << 16;
        PK_FENCE(c);
      }
    }
    pk_unpack<C, W, ND, false>(a, v);
  } else {
    const __amdgpu_buffer_rsrc_t r = pk_rsrc(a.pk, ub, a.X, 4);
#pragma unroll
    for (int c = 0; c < C; c++) {
      v[c][0] = PK_LD32(r, c);
      PK_FENCE(c);
    }
    if constexpr (ND > 1 && IN2) {
      const __amdgpu_buffer_rsrc_t r1 = pk_rsrc(a.pk + a.pitch, ub, a.X, 4);
#pragma unroll
      for (int c = 0; c < C; c++) {
        v[c][1] = PK_LD32(r1, c);
        PK_FENCE(c);
      }
    }
    pk_unpack<C, W, ND, IN2>(a, v);
  }
}

// sum of the real layer fields of one word (byte sums as a dot product)
template <int W, int ND>
__device__ __forceinline__ uint32_t pk_sum(const PSweepArgs &a,
                                           const uint32_t (&r)[ND]) {
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < ND; k++) {
    const uint32_t x = r[k] & a.fmask[k];
    if constexpr (W == 8) sum = __builtin_amdgcn_udot4(x, 0x01010101u, sum, false);
    else {
      typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
      const u16x2 ones = {1, 1};
      sum = __builtin_amdgcn_udot2(__builtin_bit_cast(u16x2, x), ones, sum, false);
    }
  }
  return sum;
}

// C words out (addressed as in pk_load): compact layers, the layer sum into
// the map, or the clamped sum into cnt
// compact output word(s) of one word in register format
template <int W, int ND, bool OUT2>
__device__ __forceinline__ uint64_t pk_compact(const PSweepArgs &a,
                                               const uint32_t (&v)[ND],
                                               const int (&sh)[ND * (32 / W)]) {
  constexpr int NS = ND * (32 / W);
  constexpr uint32_t M = (1u << W) - 1u;
  uint32_t x[ND];
#pragma unroll
  for (int k = 0; k < ND; k++) x[k] = v[k] & a.fmask[k];
  if constexpr (OUT2) {
    uint64_t o = 0;
#pragma unroll
    for (int d = 0; d < NS; d++)
      o |= (uint64_t)((x[d % ND] >> (W * (d / ND))) & M) << sh[d];
    return o;
  } else {
    uint32_t o = 0;
#pragma unroll
    for (int d = 0; d < NS; d++) o |= ((x[d % ND] >> (W * (d / ND))) & M) << sh[d];
    return o;
  }
}

// C words out (addressed as in pk_load): compact layers, the layer sum into
// the map, or the clamped sum into cnt.  VEC: the cube is C consecutive
// words (usq == 1, 16-byte aligned): 16-byte vector stores.
template <int C, int W, int ND, bool SUM, bool OUT2, bool VEC, int ST_AUX>
__device__ __forceinline__ uint32_t pk_store_aux(const PSweepArgs &a, long long ub,
                                             uint32_t usq, uint32_t lane,
                                             const uint32_t (&v)[C][ND]) {
  constexpr int NS = ND * (32 / W);
  if constexpr (!SUM) {
    if (a.out_nar) {
      // narrow word: low 16 bits to the uint16 plane, bits 16..23 to the
      // uint8 plane
      int sh[NS];
#pragma unroll
      for (int d = 0; d < NS; d++) sh[d] = d < a.L ? a.ooff[d] : 0;
      const __amdgpu_buffer_rsrc_t r16 = pk_rsrc(pk_nar16(a), ub, a.X, 2);
      const __amdgpu_buffer_rsrc_t r8 = pk_rsrc(pk_nar8(a), ub, a.X, 1);
      const bool two = a.out_nar == 2;
#pragma unroll
      for (int c = 0; c < C; c++) {
        const uint32_t o = (uint32_t)pk_compact<W, ND, false>(a, v[c], sh);
        PK_ST16(r16, c, o);
        if (two) PK_ST8(r8, c, o >> 16);
        PK_FENCE(c);
      }
      return 0u;
    }
    const __amdgpu_buffer_rsrc_t r = pk_rsrc(a.pk, ub, a.X, 4);
    // layer d goes to bit ooff[d] of plane odw[d] (OUT2: two planes);
    // don't-care slots are masked to zero first, so every slot is packed
    // unconditionally
    int sh[NS];
#pragma unroll
    for (int d = 0; d < NS; d++)
      sh[d] = d < a.L ? a.ooff[d] + (OUT2 && a.odw[d] ? 32 : 0) : 0;
    const __amdgpu_buffer_rsrc_t r1 = pk_rsrc(a.pk + (OUT2 ? a.pitch : 0), ub, a.X, 4);
    const bool w8 = a.out_w8 != 0;
    if constexpr (VEC && C % 4 == 0) {
      if (w8) {
#pragma unroll
        for (int c = 0; c < C; c += 4) {
          pk_u32x4 o;
#pragma unroll
          for (int i = 0; i < 4; i++) o[i] = pk_to_bytes<W, ND>(v[c + i], a.out_bmask);
          __builtin_amdgcn_raw_buffer_store_b128(o, r, (int)(lane * 4u), c * 4, ST_AUX);
        }
      } else {
#pragma unroll
        for (int c = 0; c < C; c += 4) {
          pk_u32x4 o, o1;
#pragma unroll
          for (int i = 0; i < 4; i++) {
            const uint64_t x = pk_compact<W, ND, OUT2>(a, v[c + i], sh);
            o[i] = (uint32_t)x;
            o1[i] = (uint32_t)(x >> 32);
          }
          __builtin_amdgcn_raw_buffer_store_b128(o, r, (int)(lane * 4u), c * 4, ST_AUX);
          if (OUT2)
            __builtin_amdgcn_raw_buffer_store_b128(o1, r1, (int)(lane * 4u), c * 4, ST_AUX);
        }
      }
    } else if (w8) {
#pragma unroll
      for (int c = 0; c < C; c++) {
        PK_ST32(r, c, (pk_to_bytes<W, ND>(v[c], a.out_bmask)));
        PK_FENCE(c);
      }
    } else {
#pragma unroll
      for (int c = 0; c < C; c++) {
        const uint64_t o = pk_compact<W, ND, OUT2>(a, v[c], sh);
        PK_ST32(r, c, (uint32_t)o);
        if (OUT2) PK_ST32(r1, c, (uint32_t)(o >> 32));
        PK_FENCE(c);
      }
    }
    return 0u;
  } else {
    // one loop for both sums, so the per-word sums are formed next to
    // their stores (not hoisted ahead of the branch for all C words)
    // f: the census of the stored words (their maximum, or their deficits'
    // sum below fold_mu)
    const bool to_map = a.out_kind == CC_OUT_MAP;
    const __amdgpu_buffer_rsrc_t rm = pk_rsrc(a.map, ub, a.X, 4);
    const __amdgpu_buffer_rsrc_t rc = pk_rsrc(a.cnt_out, ub, a.X, 2);
    const uint32_t mu = (uint32_t)a.fold_mu;
    uint32_t f = 0u;
    if constexpr (VEC && C % 8 == 0) {
#pragma unroll
      for (int c = 0; c < C; c += 8) {
        uint32_t sm[8];
#pragma unroll
        for (int i = 0; i < 8; i++) sm[i] = pk_sum<W, ND>(a, v[c + i]);
#pragma unroll
        for (int i = 0; i < 8; i++)
          f = to_map ? max(f, sm[i]) : f + (mu > min(sm[i], 65535u) ? mu - min(sm[i], 65535u) : 0u);
        if (to_map) {
          const pk_u32x4 o0 = {sm[0], sm[1], sm[2], sm[3]};
          const pk_u32x4 o1 = {sm[4], sm[5], sm[6], sm[7]};
          __builtin_amdgcn_raw_buffer_store_b128(o0, rm, (int)(lane * 4u), c * 4, ST_AUX);
          __builtin_amdgcn_raw_buffer_store_b128(o1, rm, (int)(lane * 4u), c * 4 + 16, ST_AUX);
        } else {
          pk_u32x4 o;
#pragma unroll
          for (int i = 0; i < 4; i++)
            o[i] = min(sm[2 * i], 65535u) | (min(sm[2 * i + 1], 65535u) << 16);
          __builtin_amdgcn_raw_buffer_store_b128(o, rc, (int)(lane * 2u), c * 2, ST_AUX);
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < C; c++) {
        const uint32_t s = pk_sum<W, ND>(a, v[c]);
        const uint32_t sc = s > 65535u ? 65535u : s;
        if (to_map) PK_ST32(rm, c, s);
        else PK_ST16(rc, c, sc);
        f = to_map ? max(f, s) : f + (mu > sc ? mu - sc : 0u);
        PK_FENCE(c);
      }
    }
    return f;
  }
}

// the store cache policy is chosen per sweep (st_nt: non-temporal stores
// for words that the next kernel does not read soon, such as the map)
template <int C, int W, int ND, bool SUM, bool OUT2, bool VEC = false>
__device__ __forceinline__ uint32_t pk_store(const PSweepArgs &a, long long ub,
                                             uint32_t usq, uint32_t lane,
                                             const uint32_t (&v)[C][ND]) {
  if (a.st_nt) return pk_store_aux<C, W, ND, SUM, OUT2, VEC, CC_ST_NT>(a, ub, usq, lane, v);
  return pk_store_aux<C, W, ND, SUM, OUT2, VEC, 0>(a, ub, usq, lane, v);
}

// a thread's running census over the cubes it stored
__device__ __forceinline__ uint32_t pk_fold(const PSweepArgs &a, uint32_t acc, uint32_t f) {
  return a.out_kind == CC_OUT_MAP ? max(acc, f) : acc + f;
}

// kernel prologue: the first sweep clears the census slots
__device__ __forceinline__ void pk_fold_zero(const PSweepArgs &a) {
  if (a.fold_zero && blockIdx.x == 0 && threadIdx.x < CC_NSLOT)
    a.fold_zero[(size_t)threadIdx.x * a.fold_stride] = 0ull;
}

// kernel epilogue, reached by every thread of the workgroup (all launches
// are whole waves): the workgroup's census, one atomic per workgroup
template <bool SUM>
__device__ __forceinline__ void pk_fold_flush(const PSweepArgs &a, uint32_t acc) {
  if constexpr (SUM) {
    if (a.fold) {
      if (a.out_kind == CC_OUT_MAP) cc_census_flush<true>(acc, a.fold, a.fold_stride);
      else cc_census_flush<false>(acc, a.fold, a.fold_stride);
    }
  }
}

// LDS exchange: element (row p, column tl) of the tile lives at
// pad(p * TL + tl) of each dword plane; the cube's element c is row
// pb + c * qa.  When the round's addresses are linear in c (lin: the host
// checked that no padding step falls inside a cube) they are e0 + c * kp.
template <int C, int ND>
__device__ __forceinline__ void pk_lds_st(uint32_t *__restrict__ S, int PTP,
                                          int pb, int qa, int TL, int tl,
                                          bool lin, int kp,
                                          const uint32_t (&v)[C][ND]) {
  if (lin) {
    const int e0 = pk_pad(pb * TL + tl);
#pragma unroll
    for (int c = 0; c < C; c++) {
#pragma unroll
      for (int k = 0; k < ND; k++) S[k * PTP + e0 + c * kp] = v[c][k];
      PK_FENCE(c);
    }
  } else {
#pragma unroll
    for (int c = 0; c < C; c++) {
      const int e = pk_pad((pb + c * qa) * TL + tl);
#pragma unroll
      for (int k = 0; k < ND; k++) S[k * PTP + e] = v[c][k];
      PK_FENCE(c);
    }
  }
}

template <int C, int ND>
__device__ __forceinline__ void pk_lds_ld(const uint32_t *__restrict__ S,
                                          int PTP, int pb, int qa, int TL,
                                          int tl, bool lin, int kp,
                                          uint32_t (&v)[C][ND]) {
  if (lin) {
    const int e0 = pk_pad(pb * TL + tl);
#pragma unroll
    for (int c = 0; c < C; c++) {
#pragma unroll
      for (int k = 0; k < ND; k++) v[c][k] = S[k * PTP + e0 + c * kp];
      PK_FENCE(c);
    }
  } else {
#pragma unroll
    for (int c = 0; c < C; c++) {
      const int e = pk_pad((pb + c * qa) * TL + tl);
#pragma unroll
      for (int k = 0; k < ND; k++) v[c][k] = S[k * PTP + e];
      PK_FENCE(c);
    }
  }
}

template <int W, int ND>
__device__ __forceinline__ uint32_t pk_mid_pack(const PSweepArgs &a,
                                                const uint32_t (&r)[ND]) {
  if (a.mid_w8) return pk_to_bytes<W, ND>(r, a.out_bmask);
  constexpr int NS = ND * (32 / W);
  constexpr uint32_t M = (1u << W) - 1u;
  uint32_t o = 0;
#pragma unroll
  for (int d = 0; d < NS; d++) {
    if (d >= a.L) continue;
    o |= ((r[d % ND] >> (W * (d / ND))) & M) << a.moff[d];
  }
  return o;
}

template <int W, int ND>
__device__ __forceinline__ void pk_mid_unpack(const PSweepArgs &a, uint32_t x,
                                              uint32_t (&r)[ND]) {
  if (a.mid_w8) {
    pk_from_bytes<W, ND>(x, r);
    return;
  }
  constexpr int NS = ND * (32 / W);
#pragma unroll
  for (int k = 0; k < ND; k++) r[k] = 0u;
#pragma unroll
  for (int d = 0; d < NS; d++) {
    if (d >= a.L) continue;
    r[d % ND] |= __builtin_amdgcn_ubfe(x, a.moff[d], a.mbits[d]) << (W * (d / ND));
  }
}

// the LDS exchange of k_psweep in one compact dword per word (addresses as
// pk_lds_st / pk_lds_ld, plane 0 only)
template <int C, int W, int ND>
__device__ __forceinline__ void pk_lds_st_c(const PSweepArgs &a, uint32_t *__restrict__ S,
                                            int pb, int qa, int TL, int tl, bool lin,
                                            int kp, const uint32_t (&v)[C][ND]) {
#pragma unroll
  for (int c = 0; c < C; c++) {
    const int e = lin ? pk_pad(pb * TL + tl) + c * kp : pk_pad((pb + c * qa) * TL + tl);
    S[e] = pk_mid_pack<W, ND>(a, v[c]);
    PK_FENCE(c);
  }
}

template <int C, int W, int ND>
__device__ __forceinline__ void pk_lds_ld_c(const PSweepArgs &a,
                                            const uint32_t *__restrict__ S, int pb,
                                            int qa, int TL, int tl, bool lin, int kp,
                                            uint32_t (&v)[C][ND]) {
#pragma unroll
  for (int c = 0; c < C; c++) {
    const int e = lin ? pk_pad(pb * TL + tl) + c * kp : pk_pad((pb + c * qa) * TL + tl);
    pk_mid_unpack<W, ND>(a, S[e], v[c]);
    PK_FENCE(c);
  }
}

// SEED: the sweep starts the transform (reads the predicate or code plane).
// SUM: the sweep ends the transform (layer sum into the map or cnt);
// otherwise it writes compact layers.  IN2 / OUT2: its compact input /
// output has two dword planes (compile-time: a runtime plane count made the
// compiler keep both register images of the cube)
template <int Q, int RR, int W, int ND, bool SEED, bool SUM, bool IN2, bool OUT2>
__global__ void __launch_bounds__(CC_BLOCK) k_psweep(PSweepArgs a) {
  extern __shared__ uint32_t Sp[];
  pk_fold_zero(a);
  uint32_t facc = 0u;
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
      const bool lin = a.lin[rd] != 0;
      const int kp = a.kp[rd];
      for (int cu = threadIdx.x; cu < ncube; cu += nt) {
        const int tl = cu % a.TL, u = cu / a.TL;
        const int pb = (u / qa) * qa * C + (u % qa);
        const uint32_t lane = (uint32_t)((long long)pb * a.s + tl);
        const uint32_t usq = (uint32_t)((long long)qa * a.s);
        uint32_t v[C][ND];
        if (first)
          pk_load<C, W, ND, SEED, IN2>(a, base, usq, lane, tl < ncol, v);
        else if (ND > 1 && a.lcmp)
          pk_lds_ld_c<C, W, ND>(a, Sp, pb, qa, a.TL, tl, lin, kp, v);
        else
          pk_lds_ld<C, ND>(Sp, PTP, pb, qa, a.TL, tl, lin, kp, v);
        pk_apply<Q, C, 0, RR, W, ND>(v, a.amask[rd]);
        if (last) {
          if (tl < ncol) facc = pk_fold(a, facc, pk_store<C, W, ND, SUM, OUT2>(a, base, usq, lane, v));
        } else if (ND > 1 && a.lcmp) {
          pk_lds_st_c<C, W, ND>(a, Sp, pb, qa, a.TL, tl, lin, kp, v);
        } else {
          pk_lds_st<C, ND>(Sp, PTP, pb, qa, a.TL, tl, lin, kp, v);
        }
      }
      __syncthreads();
    }
  }
  pk_fold_flush<SUM>(a, facc);
}

// ---------------------------------------------------------------------------
// Strided sweep whose whole group is one register cube: every thread (or,
// with SPLIT = 2 on binary axes, every pair of adjacent lanes) owns the Q^RR
// words of one fiber cube of one column, rows s apart.  Consecutive lanes
// take consecutive columns, so each row access of a wave is coalesced; a
// workgroup covers 256 / SPLIT consecutive columns of one outer block
// (the host requires s to be a multiple of that), so the row address is a
// workgroup-uniform scalar base plus one per-lane 32-bit offset shared by
// every element: no per-element address arithmetic.  With SPLIT = 2 a lane
// holds half the cube (the rows whose top cube digit equals its lane
// parity) and the top axis is exchanged with the partner lane by a DPP
// quad permute, which halves the registers per lane.
template <int W, int ND>
__device__ __forceinline__ void pk_add_up_xlane(uint32_t (&x)[ND]) {
  uint32_t p[ND];
#pragma unroll
  for (int k = 0; k < ND; k++)
    p[k] = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x[k], 0xB1, 0xF, 0xF, false);
  pk_add_up<W, ND>(x, p);
}

template <int Q, int RR, int W, int ND, bool SUM, bool IN2, bool OUT2, int SPLIT>
__global__ void __launch_bounds__(CC_BLOCK) k_pcol(PSweepArgs a) {
  pk_fold_zero(a);
  uint32_t facc = 0u;
  constexpr int C = cc_ipow(Q, RR);
  constexpr int CH = C / SPLIT;              // words per lane
  constexpr int RI = SPLIT == 2 ? RR - 1 : RR;  // axes inside the lane
  constexpr int CPB = CC_BLOCK / SPLIT;      // columns per workgroup
  const long long ncubes = a.ntiles;         // one cube per column
  const long long nblk = ncubes / CPB;
  const int half = SPLIT == 2 ? (int)(threadIdx.x & 1) : 0;
  const int colw = (int)threadIdx.x / SPLIT;
  const uint32_t lane = (uint32_t)(colw + (long long)half * CH * a.s);
  const uint32_t usq = (uint32_t)a.s;
  for (long long b = blockIdx.x; b < nblk; b += gridDim.x) {
    const long long cube0 = b * CPB;         // the block's first column
    const long long h = cube0 / a.s;
    const long long base = h * a.s * a.P + (cube0 - h * a.s);
    uint32_t v[CH][ND];
    pk_load<CH, W, ND, false, IN2>(a, base, usq, lane, true, v);
    pk_apply<Q, CH, 0, RI, W, ND>(v, (1 << RI) - 1);
    if constexpr (SPLIT == 2) {
#pragma unroll
      for (int c = 0; c < CH; c++) pk_add_up_xlane<W, ND>(v[c]);
    }
    facc = pk_fold(a, facc, pk_store<CH, W, ND, SUM, OUT2>(a, base, usq, lane, v));
  }
  pk_fold_flush<SUM>(a, facc);
}

// ---------------------------------------------------------------------------
// First sweep of a binary cell without LDS: a wave owns a tile of 2^(RR+6)
// consecutive words, lane l the C = 2^RR contiguous words l*C .. l*C+C-1
// (vector loads and stores).  The RR low axes are in-thread; the six axes
// above are the lane index bits and are applied across lanes: xor 1, 2, 4, 8
// with DPP row permutes (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror
// then quad_perm [3,2,1,0], row_ror:8), xor 16 and 32 with permlane16 /
// permlane32 swaps of register pairs, which put each lane's cross-row
// partner into its own registers (the swap is its own inverse, so a second
// swap restores the layout).  No LDS, so occupancy is bounded by registers
// alone.
template <int CTRL>
__device__ __forceinline__ uint32_t pk_dpp(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
}

template <int W, int ND, int XOR>
__device__ __forceinline__ void pk_xlane_row(uint32_t (&x)[ND]) {
  uint32_t p[ND];
#pragma unroll
  for (int k = 0; k < ND; k++) {
    if constexpr (XOR == 1) p[k] = pk_dpp<0xB1>(x[k]);
    else if constexpr (XOR == 2) p[k] = pk_dpp<0x4E>(x[k]);
    else if constexpr (XOR == 4) p[k] = pk_dpp<0x1B>(pk_dpp<0x141>(x[k]));
    else p[k] = pk_dpp<0x128>(x[k]);
  }
  pk_add_up<W, ND>(x, p);
}

// registers a and b (one dword of two words): swap halves so that each lane
// holds a cross-row pair, apply the axis to the pair, swap back
template <int W, int ND, int XOR>
__device__ __forceinline__ void pk_xlane_swap(uint32_t (&a)[ND], uint32_t (&b)[ND]) {
  uint32_t s0[ND], s1[ND];
#pragma unroll
  for (int k = 0; k < ND; k++) {
    auto r = XOR == 16 ? __builtin_amdgcn_permlane16_swap(a[k], b[k], false, false)
                       : __builtin_amdgcn_permlane32_swap(a[k], b[k], false, false);
    s0[k] = r[0];
    s1[k] = r[1];
  }
  uint32_t t0[ND], t1[ND];
#pragma unroll
  for (int k = 0; k < ND; k++) { t0[k] = s0[k]; t1[k] = s1[k]; }
  pk_add_up<W, ND>(s0, t1);
  pk_add_up<W, ND>(s1, t0);
#pragma unroll
  for (int k = 0; k < ND; k++) {
    auto r = XOR == 16 ? __builtin_amdgcn_permlane16_swap(s0[k], s1[k], false, false)
                       : __builtin_amdgcn_permlane32_swap(s0[k], s1[k], false, false);
    a[k] = r[0];
    b[k] = r[1];
  }
}

template <int RR, int W, int ND, bool SUM, bool OUT2>
__global__ void __launch_bounds__(CC_BLOCK) k_pfirst(PSweepArgs a) {
  pk_fold_zero(a);
  uint32_t facc = 0u;
  constexpr int C = 1 << RR;
  constexpr int TW = C * 64;                  // words per wave tile
  const int wave = (int)threadIdx.x >> 6;
  const uint32_t ln = threadIdx.x & 63;
  const uint32_t lane = ln * C;               // first word of this lane
  const long long ntiles = a.X / TW;
  for (long long t0 = (long long)blockIdx.x * (CC_BLOCK / 64); t0 < ntiles;
       t0 += (long long)gridDim.x * (CC_BLOCK / 64)) {
    const long long tile = t0 + wave;
    if (tile >= ntiles) break;                // wave-uniform
    const long long base = tile * TW;
    uint32_t v[C][ND];
    pk_load<C, W, ND, true, false>(a, base, 1u, lane, true, v);
    pk_apply<2, C, 0, RR, W, ND>(v, (1 << RR) - 1);
    // (scheduling fences every 8 words keep the compiler from interleaving
    // the cross-lane chains of the whole cube, which costs a temporary per
    // word)
#pragma unroll
    for (int c = 0; c < C; c++) {
      pk_xlane_row<W, ND, 1>(v[c]);
      pk_xlane_row<W, ND, 2>(v[c]);
      pk_xlane_row<W, ND, 4>(v[c]);
      pk_xlane_row<W, ND, 8>(v[c]);
      if ((c & 7) == 7) __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int c = 0; c < C / 2; c++) {
      pk_xlane_swap<W, ND, 16>(v[c], v[c + C / 2]);
      pk_xlane_swap<W, ND, 32>(v[c], v[c + C / 2]);
      if ((c & 7) == 7) __builtin_amdgcn_sched_barrier(0);
    }
    facc = pk_fold(a, facc, pk_store<C, W, ND, SUM, OUT2, true>(a, base, 1u, lane, v));
  }
  pk_fold_flush<SUM>(a, facc);
}

// ---------------------------------------------------------------------------
// First sweep with an LDS exchange on a persistent grid, predicate seed (or
// a code seed from the code's sorted word list):
// one cube per thread, and while a tile runs its rounds each thread already
// has the next tile's counters in flight (raw 16-byte loads into registers),
// so the memory pipe never waits for a tile's compute.  Round 0 is the
// contiguous cube (vector loads), the last round the top cube (coalesced
// stores), as in k_psweep.
// a cube's counters held between tiles: packed pairs from 16-byte loads
// when the cube is whole groups of 8 counters, else one per register
template <int C>
struct PkPf {
  static constexpr int N = C % 8 == 0 ? C / 2 : C;
};

template <int C>
__device__ __forceinline__ void pk_cnt_raw(const PSweepArgs &a, long long ub,
                                           uint32_t lane, uint32_t (&pf)[PkPf<C>::N]) {
  const __amdgpu_buffer_rsrc_t r = pk_rsrc(a.cnt, ub, a.X, 2);
  if constexpr (C % 8 == 0) {
#pragma unroll
    for (int c = 0; c < C; c += 8) {
      const pk_u32x4 q = __builtin_bit_cast(pk_u32x4,
          __builtin_amdgcn_raw_buffer_load_b128(r, (int)(lane * 2u), c * 2, 0));
      pf[c / 2] = q.x; pf[c / 2 + 1] = q.y; pf[c / 2 + 2] = q.z; pf[c / 2 + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int c = 0; c < C; c++)
      pf[c] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, (int)(lane * 2u), c * 2, 0);
  }
}

template <int C>
__device__ __forceinline__ uint32_t pk_cnt_at(const uint32_t (&pf)[PkPf<C>::N], int c) {
  if constexpr (C % 8 == 0) return (pf[c >> 1] >> (16 * (c & 1))) & 0xFFFFu;
  else return pf[c];
}

template <int Q, int RR, int W, int ND, bool SUM, bool OUT2>
__global__ void __launch_bounds__(CC_BLOCK) k_pfirstp(PSweepArgs a) {
  extern __shared__ uint32_t Sp[];
  pk_fold_zero(a);
  uint32_t facc = 0u;
  constexpr int C = cc_ipow(Q, RR);
  const int PTP = pk_pad(a.P) + 1;
  const int u = threadIdx.x;                 // this thread's round-0 cube
  // the tile's P / C cubes fill whole waves but the last: its idle lanes
  // only join the barriers
  const bool act = u < a.P / C;
  const uint32_t lane0 = (uint32_t)u * C;
  const uint32_t lo = (uint32_t)a.plo, span = (uint32_t)(a.phi - a.plo);
  const bool sparse = a.in_kind == CC_IN_SPARSE;
  uint32_t pf[PkPf<C>::N];
#pragma unroll
  for (int i = 0; i < PkPf<C>::N; i++) pf[i] = 0u;
  long long tile = blockIdx.x;
  if (tile < a.ntiles && !sparse && act) pk_cnt_raw<C>(a, tile * a.P, lane0, pf);
  for (; tile < a.ntiles; tile += gridDim.x) {
    const long long base = tile * a.P;
    uint32_t v[C][ND];
    if (sparse) {
      // code seed from the sorted word list: the tile's entries (a
      // workgroup-uniform range, usually empty or one word) are matched
      // against this thread's cube; no plane is read
#pragma unroll
      for (int c = 0; c < C; c++)
#pragma unroll
        for (int k = 0; k < ND; k++) v[c][k] = 0u;
      const int e1 = a.toff[tile + 1];
      const long long cube0 = base + lane0;
      for (int e = a.toff[tile]; e < e1 && act; e++) {
        const long long d = a.cw[e] - cube0;
        if (d >= 0 && d < C) {
          const uint32_t m = (uint32_t)a.cm[e];
#pragma unroll
          for (int c = 0; c < C; c++) v[c][0] = d == c ? m : v[c][0];
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < C; c++) {
#pragma unroll
        for (int k = 0; k < ND; k++) v[c][k] = 0u;
        const uint32_t cv = pk_cnt_at<C>(pf, c);
        v[c][0] = (cv - lo) <= span ? 1u : 0u;
      }
      const long long next = tile + gridDim.x;
      if (next < a.ntiles && act) pk_cnt_raw<C>(a, next * a.P, lane0, pf);
    }
    for (int rd = 0; rd < a.nrounds; rd++) {
      const int qa = a.qa[rd];
      const int pb = (u / qa) * qa * C + (u % qa);
      const bool lin = a.lin[rd] != 0;
      const int kp = a.kp[rd];
      if (act) {
        if (rd > 0) pk_lds_ld<C, ND>(Sp, PTP, pb, qa, 1, 0, lin, kp, v);
        pk_apply<Q, C, 0, RR, W, ND>(v, a.amask[rd]);
        if (rd == a.nrounds - 1)
          facc = pk_fold(a, facc, pk_store<C, W, ND, SUM, OUT2>(a, base, (uint32_t)qa, (uint32_t)pb, v));
        else
          pk_lds_st<C, ND>(Sp, PTP, pb, qa, 1, 0, lin, kp, v);
      }
      __syncthreads();
    }
  }
  pk_fold_flush<SUM>(a, facc);
}

// ---------------------------------------------------------------------------
// First sweep whose layers outgrow bytes near its end (k_pfirstw): the
// rounds that stay within byte bounds run as in k_pfirstp in the W = 8
// single-dword format (the LDS tile holds one byte per layer), and the last
// round, which applies only the top NA axes of its cube, widens each chunk
// of Q^NA words to the (W2, ND2) format on the way out.  A 14-axis binary
// first sweep (16,384-word tiles) keeps a 66 KB byte tile in LDS instead of
// 133 KB of 16-bit fields.
template <int Q, int RR, int W2, int ND2, int NA, bool SUM, bool OUT2>
__global__ void __launch_bounds__(CC_BLOCK) k_pfirstw(PSweepArgs a) {
  extern __shared__ uint32_t Sp[];
  pk_fold_zero(a);
  uint32_t facc = 0u;
  constexpr int C = cc_ipow(Q, RR);
  constexpr int CA = cc_ipow(Q, NA);         // words of one chunk
  constexpr int F = C / CA;                  // chunks of the last round's cube
  static_assert(C % 8 == 0 && NA >= 1 && NA <= RR, "cube shape");
  const int PTP = pk_pad(a.P) + 1;
  const int u = threadIdx.x;
  const uint32_t lane0 = (uint32_t)u * C;
  const uint32_t lo = (uint32_t)a.plo, span = (uint32_t)(a.phi - a.plo);
  const bool sparse = a.in_kind == CC_IN_SPARSE;
  const int last = a.nrounds - 1;
  uint32_t pf[PkPf<C>::N];
  long long tile = blockIdx.x;
  if (tile < a.ntiles && !sparse) pk_cnt_raw<C>(a, tile * a.P, lane0, pf);
  for (; tile < a.ntiles; tile += gridDim.x) {
    const long long base = tile * a.P;
    uint32_t v[C][1];
#pragma unroll
    for (int c = 0; c < C; c++) v[c][0] = 0u;
    if (sparse) {
      const int e1 = a.toff[tile + 1];
      const long long cube0 = base + lane0;
      for (int e = a.toff[tile]; e < e1; e++) {
        const long long d = a.cw[e] - cube0;
        if (d >= 0 && d < C) {
          const uint32_t m = (uint32_t)a.cm[e];
#pragma unroll
          for (int c = 0; c < C; c++) v[c][0] = d == c ? m : v[c][0];
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < C; c++) {
        const uint32_t cv = pk_cnt_at<C>(pf, c);
        v[c][0] = (cv - lo) <= span ? 1u : 0u;
      }
      const long long next = tile + gridDim.x;
      if (next < a.ntiles) pk_cnt_raw<C>(a, next * a.P, lane0, pf);
    }
    // byte rounds
    for (int rd = 0; rd < last; rd++) {
      const int qa = a.qa[rd];
      const int pb = (u / qa) * qa * C + (u % qa);
      if (rd > 0) pk_lds_ld<C, 1>(Sp, PTP, pb, qa, 1, 0, a.lin[rd] != 0, a.kp[rd], v);
      pk_apply<Q, C, 0, RR, 8, 1>(v, a.amask[rd]);
      pk_lds_st<C, 1>(Sp, PTP, pb, qa, 1, 0, a.lin[rd] != 0, a.kp[rd], v);
      __syncthreads();
    }
    // last round: the top NA cube axes, chunk by chunk in the wide format;
    // chunk f holds the cube elements f + F g (g < Q^NA), F * qa words apart
    {
      const int qa = a.qa[last];
      const int pb = (u / qa) * qa * C + (u % qa);
      pk_lds_ld<C, 1>(Sp, PTP, pb, qa, 1, 0, a.lin[last] != 0, a.kp[last], v);
#pragma unroll
      for (int f = 0; f < F; f++) {
        uint32_t w[CA][ND2];
#pragma unroll
        for (int g = 0; g < CA; g++) pk_from_bytes<W2, ND2>(v[f + F * g][0], w[g]);
        pk_apply<Q, CA, 0, NA, W2, ND2>(w, (1 << NA) - 1);
        facc = pk_fold(a, facc, pk_store<CA, W2, ND2, SUM, OUT2>(a, base, (uint32_t)(F * qa),
                                         (uint32_t)(pb + f * qa), w));
      }
      __syncthreads();
    }
  }
  pk_fold_flush<SUM>(a, facc);
}

// ---------------------------------------------------------------------------
// Strided sweep of 2 RA axes in two register rounds (k_pstrip): a tile is TL
// consecutive columns of one outer block, each a fiber of P = C^2 rows
// (C = Q^RA) s words apart.  Round A: thread (tl, u) holds the cube of
// rows e + C u, e < C (the low RA axes, coalesced loads: consecutive
// threads take consecutive columns), applies them and writes the rows to
// LDS in a one-dword compact format; round B: the same thread holds rows
// u + C k, k < C (the top RA axes), applies them and stores.  LDS row r
// lives at r TL + (r / C) PADR so that the C-row groups of a round-A wave
// fall on different banks; round-B waves read consecutive dwords.
#define CC_STRIP_PADR 16

template <int Q, int RA, int W, int ND, bool SUM, bool IN2, bool OUT2>
__global__ void __launch_bounds__(512) k_pstrip(PSweepArgs a) {
  extern __shared__ uint32_t Sp[];
  pk_fold_zero(a);
  uint32_t facc = 0u;
  constexpr int C = cc_ipow(Q, RA);
  const int TL = a.TL;
  const int tl = (int)threadIdx.x % TL;
  const int u = (int)threadIdx.x / TL;            // 0 .. C-1
  const uint32_t s = (uint32_t)a.s;
  const int rstride = C * TL + CC_STRIP_PADR;      // LDS dwords per C rows
  for (long long tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    const long long h = tile / a.ncolb;
    const long long base = h * a.s * a.P + (tile - h * a.ncolb) * TL;
    uint32_t v[C][ND];
    // round A: rows e + C u
    pk_load<C, W, ND, false, IN2>(a, base, s, (uint32_t)tl + (uint32_t)(C * u) * s,
                                  true, v);
    pk_apply<Q, C, 0, RA, W, ND>(v, (1 << RA) - 1);
#pragma unroll
    for (int e = 0; e < C; e++) {
      Sp[u * rstride + e * TL + tl] = pk_mid_pack<W, ND>(a, v[e]);
      PK_FENCE(e);
    }
    __syncthreads();
    // round B: rows u + C k
#pragma unroll
    for (int k = 0; k < C; k++) {
      pk_mid_unpack<W, ND>(a, Sp[k * rstride + u * TL + tl], v[k]);
      PK_FENCE(k);
    }
    pk_apply<Q, C, 0, RA, W, ND>(v, (1 << RA) - 1);
    facc = pk_fold(a, facc, pk_store<C, W, ND, SUM, OUT2>(a, base, (uint32_t)C * s, (uint32_t)tl + (uint32_t)u * s,
                                  v));
    __syncthreads();                 // the next tile rewrites the LDS rows
  }
  pk_fold_flush<SUM>(a, facc);
}

// cc_common.hpp -- shared constants, device helpers and launch arguments of
// the MI355X exact-transform engine (gfx950, wave64).
//
// The engine holds one problem cell: a mixed-radix product space
// X = Z_{a_0} x ... x Z_{a_{n-1}} (little-endian index, axis 0 contiguous),
// L = R+1 int32 distance-layer planes (1 plane for the L-inf families), a
// uint16 multiplicity plane `cnt` padded to an even length so packed 32-bit
// atomics never leave it, and the v1 support table.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define CC_MAXN 40      // axes (|X| <= 2^40); the v1 table format caps n at 16
#define CC_MAXN_V1 16   // axes addressable by the v1 byte format
#define CC_MAXSW 24     // axes inside one sweep (its tile has <= 2^20 words)
#define CC_MAXAX 16     // axis size (nibble-packed support deltas)
#define CC_MAXR 6       // Hamming / Lee radius (layer planes R+1 <= 7)
#define CC_MAXL (CC_MAXR + 1)
#define CC_MAXPAT 8     // bytes per packed v1 pattern
#define CC_MAXPAT_W 6   // 10-bit fields per packed wide pattern

// support-table formats the walks decode
#define CC_FMT_V1 0     // uint64, byte k = pos<<4 | delta (pos < 16), <= 8
#define CC_FMT_WIDE 1   // uint64, bits [10k, 10k+10) = pos<<4 | delta, <= 6
#define CC_FMT_X16 2    // uint64, bits [16k, 16k+16) = pos<<10 | delta, <= 4
#define CC_MAXPAT_X 4   // 16-bit fields per long-axis pattern
#define CC_MAXAX_X 1024 // L-inf axis size reachable with the long-axis format
#define CC_CSTRIDE 15   // per-position slots of the walk displacement table
#define CC_BLOCK 256    // threads per workgroup (4 waves of 64)

// problem ids (ABI values, reference src/include/coldcore/plugin.h:45-53)
#define CC_HAMMING 0
#define CC_TORUS 1
#define CC_LEE 2
#define CC_GRID 3

// cnt predicates (reference src/plugins/covering/dctcore_mu.cuh:29-40)
#define CC_EQ 0     // cnt == target
#define CC_DEF 1    // cnt < mu          (deficient: gain role)
#define CC_CRIT 2   // 1 <= cnt <= mu    (critical: loss role)

__device__ __forceinline__ bool cc_pred(int mode, int c, int target, int mu) {
  if (mode == CC_DEF) return c < mu;
  if (mode == CC_CRIT) return c >= 1 && c <= mu;
  return c == target;
}

// ---------------------------------------------------------------- wave64
__device__ __forceinline__ long long cc_wave_sum(long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ int cc_wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// wave64 total by DPP: three row shifts give each lane the sum of itself
// and the three lanes before it in its row of 16, shifts by 4 and 8 (upper
// banks only) complete each row's sum in its lane 15, and two row
// broadcasts (lane 15 into the next row, lane 31 into the upper two rows)
// carry the rows into lane 63, which is read back as a wave-uniform value
__device__ __forceinline__ uint32_t cc_dpp_wave_add(uint32_t v) {
  const int x = (int)v;
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, true);   // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, true);   // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, x, 0x113, 0xF, 0xF, true);   // row_shr:3
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xE, false);  // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xC, false);  // row_shr:8
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// the same shape for the maximum of non-negative values (0 is the identity
// of the lanes a shift leaves empty; the result is max(0, wave maximum))
__device__ __forceinline__ int cc_dpp_wave_max0(int v) {
  const int x = v;
  v = max(v, __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, true));
  v = max(v, __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, true));
  v = max(v, __builtin_amdgcn_update_dpp(0, x, 0x113, 0xF, 0xF, true));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xE, false));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xC, false));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false));
  return max(0, __builtin_amdgcn_readlane(v, 63));
}

// Census slots: CC_NSLOT uint64 accumulators `stride` words apart.  A
// workgroup combines its waves' totals in LDS and adds (or maxes) the
// workgroup total into slot blockIdx % CC_NSLOT, so the same-address
// atomics of a launch spread over CC_NSLOT addresses (and memory channels);
// the host reads the slots back and combines them.  Every thread of the
// workgroup calls it (whole waves).
#define CC_NSLOT 32
template <bool MAX>
__device__ __forceinline__ void cc_census_flush(uint32_t v, unsigned long long *slots,
                                                int stride) {
  __shared__ uint32_t cw[16];
  const uint32_t t = MAX ? (uint32_t)cc_dpp_wave_max0((int)v) : cc_dpp_wave_add(v);
  if ((threadIdx.x & 63) == 0) cw[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s = 0;
    const int nw = (int)(blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; i++)
      s = MAX ? (s > cw[i] ? s : (unsigned long long)cw[i]) : s + cw[i];
    if (s) {
      unsigned long long *p = slots + (size_t)(blockIdx.x % CC_NSLOT) * stride;
      if (MAX) atomicMax(p, s);
      else atomicAdd(p, s);
    }
  }
}

// inclusive prefix sum over the 64 lanes of a wave
__device__ __forceinline__ int cc_wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

// block-wide sum for CC_BLOCK threads; the result is valid in every thread
__device__ __forceinline__ long long cc_block_sum(long long v) {
  __shared__ long long red[CC_BLOCK / 64];
  v = cc_wave_sum(v);
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[wid] = v;
  __syncthreads();
  long long t = 0;
#pragma unroll
  for (int i = 0; i < CC_BLOCK / 64; i++) t += red[i];
  __syncthreads();
  return t;
}

// ---------------------------------------------------------------- sweeps
// One sweep applies the per-axis passes of a group of consecutive axes
// [j0, j0 + nax) to every tile of the space.  A tile is the set of words
// base + p * s + tl, p < P (the group's sub-space, in its own mixed radix)
// and tl < TL consecutive columns (TL = 1 when the group starts at axis 0,
// where the tile is one contiguous run).  The tile's live layers are staged
// in LDS as S[d][p * TL + tl]; every axis of the group runs in LDS, so one
// HBM round trip covers the whole group.
#define CC_IN_LAYERS 0   // layers 0..Lin-1 come from the planes
#define CC_IN_PRED 1     // layer 0 = predicate(cnt), the rest zero (first sweep)
#define CC_IN_CODE 2     // layer 0 = code indicator already in A0 (first sweep)
#define CC_OUT_LAYERS 0  // write back the layers this sweep produced
#define CC_OUT_MAP 1     // write sum_d A_d into A0 (gain / loss map)
#define CC_OUT_CNT 2     // write min(sum_d A_d, 65535) into cnt (recount)
// compact inter-sweep words: every layer of a word (layer 0 included) is a
// bit field of the word's 1..4 dwords, sized by a proven bound of its value
#define CC_IN_COMPACT 3
#define CC_OUT_COMPACT 3
#define CC_CMAXND 4      // dwords per compact word

struct SweepArgs {
  int32_t *A[CC_MAXL];   // int32 planes: A[0] is the map / code indicator
  uint16_t *H[CC_MAXL];  // uint16 inter-sweep layer planes (16-bit mode)
  uint16_t *cnt;
  long long s;       // stride (words) of the group's first axis
  long long ncolb;   // column blocks per outer index = ceil(s / TL)
  long long ntiles;  // tiles in the space = (|X| / (s * P)) * ncolb
  int P;             // words per column inside the group
  int TL;            // columns per tile
  int nax;           // axes in the group
  int c0;            // axes transformed before this sweep
  int R;
  int in_kind, out_kind, pmode, mu;
  int write_l0;      // CC_OUT_LAYERS: also store layer 0 (first sweep only)
  int Q[CC_MAXSW];   // axis sizes of the group
  int sk[CC_MAXSW];  // in-tile stride (units of p) of each group axis
  // live[k] = layers that can be nonzero before group axis k (live[nax]
  // after the group): min(sum of the per-axis maximal distances so far, R)+1,
  // one per axis for Hamming and floor(ax/2) for Lee
  int live[CC_MAXSW + 1];
  int rr;            // batched sweeps: axes per register round (0: LDS tile)
  // compact words: dword k of word w is pk[k * pitch + w]; layer d is the
  // field of bits[d] bits at bit off[d] of dword dw[d] (0 bits: the layer
  // is provably zero); in_* describe the input words, o* the output
  uint32_t *pk;
  long long pitch;
  int in_nd, out_nd;
  unsigned char idw[CC_MAXL], ioff[CC_MAXL], ibits[CC_MAXL];
  unsigned char odw[CC_MAXL], ooff[CC_MAXL], obits[CC_MAXL];
  // batched gain sweeps that write the map: the largest (gain << 32 |
  // 0xFFFFFFFF - index) key of the written words is folded in here with
  // atomicMax (the cell's argmax without a separate reduction); nullptr: off
  unsigned long long *amax;
};

// fold (gain, index) into a running argmax key: larger gain first, then the
// smaller index; zero gains never enter
__device__ __forceinline__ void cc_key_max(unsigned long long &best, uint32_t g,
                                           long long w) {
  if ((int32_t)g > 0) {
    const unsigned long long key = ((unsigned long long)g << 32) |
                                   (0xFFFFFFFFull - (unsigned long long)w);
    best = key > best ? key : best;
  }
}

// workgroup maximum of the keys, one atomicMax per workgroup: the waves'
// maxima meet in LDS (every thread of the workgroup calls it, after a
// barrier that follows the previous call's read of the LDS slots).  One
// atomic per wave made a large cell's tiles queue ~10 ns each on the one
// key address.
__device__ __forceinline__ void cc_key_flush(unsigned long long best,
                                             unsigned long long *amax) {
  __shared__ unsigned long long kb[16];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long b2 = __shfl_xor(best, o, 64);
    best = b2 > best ? b2 : best;
  }
  if ((threadIdx.x & 63) == 0) kb[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (int)(blockDim.x + 63) >> 6;
    unsigned long long m = 0;
    for (int i = 0; i < nw; i++) m = kb[i] > m ? kb[i] : m;
    if (m) atomicMax(amax, m);
  }
}

// LDS layout of a generic tile: layer plane d starts at d * cc_gplane(PT),
// tile element e sits at cc_gpad(e).  With PAD, one pad word every 32
// spreads the stride-Q fiber accesses (Q = 2, 4, 8, ...) over the LDS banks
// (measured faster for the singleton's mixed-radix sweeps, slower for the
// batched tiles, which use PAD = false).
template <bool PAD = true>
__device__ __forceinline__ int cc_gpad(int e) { return PAD ? e + (e >> 5) : e; }
template <bool PAD = true>
__host__ __device__ __forceinline__ int cc_gplane(int pt) {
  return PAD ? pt + (pt >> 5) + 1 : pt;
}

// x / d for x, d < 2^16 by one high multiply: m = floor((2^32 - 1) / d) + 1
// (m = 0 stands for d = 1); exact because x * d < 2^32
__device__ __forceinline__ uint32_t cc_magic(uint32_t d) {
  return d == 1 ? 0u : 0xFFFFFFFFu / d + 1u;
}
__device__ __forceinline__ int cc_div(int x, uint32_t m) {
  return m ? (int)__umulhi((uint32_t)x, m) : x;
}

// ---------------------------------------------------------------- walks
struct WordsArg {
  long long w[128];  // small batches travel by value (no staging, no sync)
};

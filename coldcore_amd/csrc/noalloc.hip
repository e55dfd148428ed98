// noalloc.hip -- the allocation-free variant of dct_map_extract64.
//
// dct_map_extract64 (reference src/plugins/covering/dct_hc.cu:1107-1150)
// compacts the live cell's map plane through chunk counts and offsets that
// the engine allocates at init.  dct_map_extract64_noalloc runs the same
// three kernels (chunk counts, one-workgroup exclusive scan, ordered write)
// on a caller's device map into caller device buffers, asynchronously on
// the caller's stream, with its chunk counts and offsets in a file-scope
// device array: no hipMalloc, no hipFree, no allocating library call, no
// live cell and no host synchronisation.  It is for callers that cannot
// allocate (code under HIP graph capture, tight serving loops).
//
// The array `scratch` holds NA_CHUNKS counts and NA_CHUNKS + 1 offsets
// (1 MiB): maps of at most NA_CHUNKS x EXTRACT_CH = 2^30 words.  Each GPU
// holds its own copy until the process ends; host code reaches it only
// through the per-device address that hipGetSymbolAddress returns (the
// symbol's name is a host-side stand-in, not a device address).  Every call
// on one GPU shares that GPU's copy: calls that overlap on one GPU (two
// streams, two threads) race and must be serialised by the caller.  Nothing
// needs to start at zero: the count kernel writes every chunk's count and
// the scan every offset before they are read.
#include <cstdint>
#include <cstdio>

#include <hip/hip_runtime.h>

#include "cc_common.hpp"
#include "cc_extract.hpp"

namespace {

constexpr long long NA_CH = 1 << 14;          // words per chunk
constexpr long long NA_CHUNKS = 1 << 16;      // chunks the scratch holds
constexpr long long NA_MAX_WORDS = NA_CH * NA_CHUNKS;   // 2^30
constexpr int NA_MAX_DEV = 64;

// [0, NA_CHUNKS) chunk counts, [NA_CHUNKS, 2 NA_CHUNKS] offsets (+ total)
__device__ __align__(256) long long scratch[2 * NA_CHUNKS + 1];

// per-device address of `scratch`, looked up once per device
long long *na_ptr[NA_MAX_DEV] = {};

__global__ void k_na_found(const long long *total, long long *found) {
  if (threadIdx.x == 0) *found = *total;
}

// device address of this GPU's scratch (null on failure, rc in *err)
long long *na_scratch(hipError_t *err) {
  int dev = 0;
  *err = hipGetDevice(&dev);
  if (*err != hipSuccess) return nullptr;
  if (dev < 0 || dev >= NA_MAX_DEV) { *err = hipErrorInvalidDevice; return nullptr; }
  if (!na_ptr[dev]) {
    void *p = nullptr;
    *err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (*err != hipSuccess || !p) return nullptr;
    na_ptr[dev] = (long long *)p;
  }
  return na_ptr[dev];
}

}  // namespace

#pragma GCC visibility push(default)
extern "C" {

// (idx, val) of map[i] >= thr, i < n, in ascending index order: the first
// min(found, cap) hits go to idx[0..) / val[0..) and *found (device int64)
// receives the exact hit count.  map, idx, val and found are device
// pointers; stream is a hipStream_t (null: the null stream).  Returns 0
// once the work is queued, -1 on bad arguments or n above 2^30 words (the
// scratch's limit), -(100 + hipError_t) on a device error.
int dct_map_extract64_noalloc(const int32_t *map, long long n, int32_t thr,
                              int64_t *idx, int32_t *val, long long cap,
                              long long *found, void *stream) {
  if (n < 0 || cap < 0 || !found || (n > 0 && !map) || (cap > 0 && (!idx || !val)))
    return -1;
  // the count and write kernels read the map with 16-byte loads
  if (((uintptr_t)map & 15) || ((uintptr_t)idx & 7) || ((uintptr_t)val & 3) ||
      ((uintptr_t)found & 7))
    return -1;
  if (n > NA_MAX_WORDS) {
    fprintf(stderr,
            "dct_map_extract64_noalloc: %lld words exceed the static scratch's "
            "limit of %lld (2^30)\n", n, NA_MAX_WORDS);
    return -1;
  }
  hipError_t e;
  long long *s = na_scratch(&e);
  if (!s) {
    fprintf(stderr, "dct_map_extract64_noalloc: scratch address: %s\n",
            hipGetErrorString(e));
    return -(100 + (int)e);
  }
  hipStream_t st = (hipStream_t)stream;
  const long long nb = (n + NA_CH - 1) / NA_CH;
  long long *bcnt = s, *boff = s + NA_CHUNKS;
  if (nb > 0) {
    hipLaunchKernelGGL(k_extract_count, dim3((unsigned)nb), dim3(CC_BLOCK), 0, st,
                       map, n, thr, NA_CH, bcnt);
  }
  // with n = 0 the scan writes the total 0 (nb = 0 reads no counts)
  hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(1024), 0, st, bcnt, nb, boff);
  if (nb > 0) {
    hipLaunchKernelGGL((k_extract_write<int64_t>), dim3((unsigned)nb), dim3(CC_BLOCK),
                       0, st, map, n, thr, NA_CH, boff, cap, idx, val);
  }
  hipLaunchKernelGGL(k_na_found, dim3(1), dim3(64), 0, st, boff + nb, found);
  e = hipGetLastError();
  if (e != hipSuccess) {
    fprintf(stderr, "dct_map_extract64_noalloc: %s\n", hipGetErrorString(e));
    return -(100 + (int)e);
  }
  return 0;
}

}  // extern "C"
#pragma GCC visibility pop

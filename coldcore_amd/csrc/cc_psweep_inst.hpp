// cc_psweep_inst.hpp -- launch table of the packed sweep kernels for one
// axis size Q.  A kernel exists for every round size RR (cube C = Q^RR
// words) and register format (W-bit fields in ND dwords) whose cube stays
// within 128 registers (C * ND <= 128).
#pragma once
#include <hip/hip_ext.h>
#include "cc_launch.hpp"
#include "cc_psweep.hpp"

constexpr bool pk_fits(int q, int rr, int nd) { return cc_ipow(q, rr) * nd <= 128; }

// blocks == 0: only report whether the kernel instantiation exists (the
// planner checks every sweep it schedules).
// The kernel instantiation for one (Q, RR, W, ND): SUM (last sweep), IN2 /
// OUT2 (two-plane compact input / output, only with ND > 1), and for the
// column kernel the lanes per cube.
template <int Q, int RR, int W, int ND, bool COL, int SPLIT, bool SEED, bool SUM,
          bool IN2, bool OUT2>
constexpr auto pk_kernel() {
  if constexpr (COL) return &k_pcol<Q, RR, W, ND, SUM, IN2, OUT2, SPLIT>;
  else return &k_psweep<Q, RR, W, ND, SEED, SUM, IN2, OUT2>;
}

template <int Q, int RR, int W, int ND, bool COL, int SPLIT, bool SEED>
void (*pk_select(bool sum, bool in2, bool out2))(PSweepArgs) {
  if constexpr (ND == 1 || SEED) {
    // one compact input plane (a seed has none)
    if (sum) return pk_kernel<Q, RR, W, ND, COL, SPLIT, SEED, true, false, false>();
    if constexpr (ND > 1)
      if (out2) return pk_kernel<Q, RR, W, ND, COL, SPLIT, SEED, false, false, true>();
    return pk_kernel<Q, RR, W, ND, COL, SPLIT, SEED, false, false, false>();
  } else {
    if (sum)
      return in2 ? pk_kernel<Q, RR, W, ND, COL, SPLIT, SEED, true, true, false>()
                 : pk_kernel<Q, RR, W, ND, COL, SPLIT, SEED, true, false, false>();
    if (in2)
      return out2 ? pk_kernel<Q, RR, W, ND, COL, SPLIT, SEED, false, true, true>()
                  : pk_kernel<Q, RR, W, ND, COL, SPLIT, SEED, false, true, false>();
    return out2 ? pk_kernel<Q, RR, W, ND, COL, SPLIT, SEED, false, false, true>()
                : pk_kernel<Q, RR, W, ND, COL, SPLIT, SEED, false, false, false>();
  }
}

// launch kf; with events, the runtime stamps the kernel's own start and end
// (the kernel duration, not the gaps around it)
inline void pk_run(void (*kf)(PSweepArgs), int grid, int threads, size_t lds,
                   hipStream_t st, hipEvent_t e0, hipEvent_t e1,
                   const PSweepArgs &a) {
  if (e0 || e1)
    hipExtLaunchKernelGGL(kf, dim3(grid), dim3(threads), (uint32_t)lds, st, e0, e1,
                          0u, a);
  else
    hipLaunchKernelGGL(kf, dim3(grid), dim3(threads), lds, st, a);
}

// resident workgroups of a persistent launch, queried once per (kernel,
// threads, LDS) -- the attribute and occupancy calls cost microseconds of
// host time -- after allowing the kernel more than 64 KB of LDS
inline long long pk_resident(void (*kf)(PSweepArgs), int threads, size_t lds) {
  struct Ent { void (*kf)(PSweepArgs); int threads; size_t lds; long long res; };
  static Ent cache[32];
  static int n = 0;
  for (int i = 0; i < n; i++)
    if (cache[i].kf == kf && cache[i].threads == threads && cache[i].lds == lds)
      return cache[i].res;
  if (lds > 64 * 1024 &&
      hipFuncSetAttribute((const void *)kf, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)lds) != hipSuccess)
    return -1;
  int per_cu = 0, dev = 0, ncu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)kf, threads,
                                                   lds) != hipSuccess || per_cu < 1)
    per_cu = 1;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) !=
          hipSuccess || ncu < 1)
    ncu = 256;
  const long long res = (long long)per_cu * ncu;
  cache[n < 32 ? n++ : 31] = Ent{kf, threads, lds, res};
  return res;
}

// persistent launch of kf: at most as many workgroups as the GPU holds
inline hipError_t pk_persist(void (*kf)(PSweepArgs), int blocks, int threads,
                             size_t lds, hipStream_t st, hipEvent_t e0, hipEvent_t e1,
                             const PSweepArgs &a) {
  const long long resident = pk_resident(kf, threads, lds);
  if (resident < 0) return hipErrorInvalidValue;
  const int g = (int)(blocks < resident ? blocks : resident);
  pk_run(kf, g, threads, lds, st, e0, e1, a);
  return hipGetLastError();
}

template <int Q, int RR, int W2, int ND2, int NA>
void (*pk_wide_select(bool sum, bool out2))(PSweepArgs) {
  if (sum) return &k_pfirstw<Q, RR, W2, ND2, NA, true, false>;
  if constexpr (ND2 > 1)
    if (out2) return &k_pfirstw<Q, RR, W2, ND2, NA, false, true>;
  return &k_pfirstw<Q, RR, W2, ND2, NA, false, false>;
}

template <int Q, int RA, int W, int ND>
void (*pk_strip_select(bool sum, bool in2, bool out2))(PSweepArgs) {
  if constexpr (ND > 1) {
    if (sum) return in2 ? &k_pstrip<Q, RA, W, ND, true, true, false>
                        : &k_pstrip<Q, RA, W, ND, true, false, false>;
    if (in2) return out2 ? &k_pstrip<Q, RA, W, ND, false, true, true>
                         : &k_pstrip<Q, RA, W, ND, false, true, false>;
    return out2 ? &k_pstrip<Q, RA, W, ND, false, false, true>
                : &k_pstrip<Q, RA, W, ND, false, false, false>;
  } else {
    if (sum) return &k_pstrip<Q, RA, W, ND, true, false, false>;
    return &k_pstrip<Q, RA, W, ND, false, false, false>;
  }
}

template <int Q, int RR, int W, int ND>
hipError_t pk_go(const PSweepArgs &a, int blocks, int threads, size_t lds,
                 hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
  const bool sum = a.out_kind != CC_OUT_PACKED;
  const bool in2 = ND > 1 && a.in_kind == CC_IN_PACKED && a.in_nd > 1;
  const bool out2 = ND > 1 && !sum && a.out_nd > 1;
  void (*kf)(PSweepArgs) = nullptr;
  if (a.kern == 1 || a.kern == 2) {
    // column kernel (k_pcol): kern = lanes per cube, whole workgroups
    if constexpr (Q == 2 && RR >= 4 && pk_fits(Q, RR - 1, ND)) {
      if (a.kern == 2) kf = pk_select<Q, RR, W, ND, true, 2, false>(sum, in2, out2);
    }
    if constexpr (RR >= 2 && pk_fits(Q, RR, ND)) {
      if (a.kern == 1) kf = pk_select<Q, RR, W, ND, true, 1, false>(sum, in2, out2);
    }
    if (!kf) return hipErrorInvalidValue;
    if (blocks == 0) return hipSuccess;
    pk_run(kf, blocks, CC_BLOCK, 0, st, e0, e1, a);
    return hipGetLastError();
  }
  if (a.kern == 4) {
    // first sweep, LDS exchange, persistent grid with next-tile prefetch
    if constexpr (pk_fits(Q, RR, ND)) {
      if (sum) kf = &k_pfirstp<Q, RR, W, ND, true, false>;
      else if constexpr (ND > 1) kf = out2 ? &k_pfirstp<Q, RR, W, ND, false, true>
                                           : &k_pfirstp<Q, RR, W, ND, false, false>;
      else kf = &k_pfirstp<Q, RR, W, ND, false, false>;
    }
    if (!kf) return hipErrorInvalidValue;
    if (blocks == 0) return hipSuccess;
    if (lds > 64 * 1024) {
      hipError_t e = hipFuncSetAttribute((const void *)kf,
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)lds);
      if (e != hipSuccess) return e;
    }
    // resident workgroups: queried once per (kernel, threads, LDS); the
    // occupancy and attribute calls cost microseconds of host time per call
    static void (*c_kf)(PSweepArgs) = nullptr;
    static int c_threads = -1;
    static size_t c_lds = 0;
    static long long c_resident = 0;
    if (kf != c_kf || threads != c_threads || lds != c_lds) {
      int per_cu = 0, dev = 0, ncu = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)kf,
                                                       threads, lds) != hipSuccess ||
          per_cu < 1)
        per_cu = 1;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) !=
              hipSuccess || ncu < 1)
        ncu = 256;
      c_kf = kf; c_threads = threads; c_lds = lds;
      c_resident = (long long)per_cu * ncu;
    }
    const long long resident = c_resident;
    const int g = (int)(blocks < resident ? blocks : resident);
    pk_run(kf, g, threads, lds, st, e0, e1, a);
    return hipGetLastError();
  }
  if (a.kern == 5) {
    // first sweep whose last round widens the byte tile (binary cells)
    if constexpr (Q == 2 && RR == 6 && W == 16 && ND == 2) {
      if (a.na == 1) kf = pk_wide_select<Q, RR, W, ND, 1>(sum, out2);
      else if (a.na == 2) kf = pk_wide_select<Q, RR, W, ND, 2>(sum, out2);
    }
    if (!kf) return hipErrorInvalidValue;
    if (blocks == 0) return hipSuccess;
    return pk_persist(kf, blocks, threads, lds, st, e0, e1, a);
  }
  if (a.kern == 6) {
    // strided sweep of two register rounds through a compact LDS tile
    if constexpr (Q == 2 && (RR == 4 || RR == 5) && W == 16 && ND == 2)
      kf = pk_strip_select<Q, RR, W, ND>(sum, in2, out2);
    if (!kf) return hipErrorInvalidValue;
    if (blocks == 0) return hipSuccess;
    return pk_persist(kf, blocks, threads, lds, st, e0, e1, a);
  }
  if (a.kern == 3) {
    // LDS-free first sweep (k_pfirst): RR in-thread axes + 6 lane axes
    if constexpr (Q == 2 && RR >= 4 && (1 << RR) * ND <= 64) {
      if (sum) kf = &k_pfirst<RR, W, ND, true, false>;
      else if constexpr (ND > 1) kf = out2 ? &k_pfirst<RR, W, ND, false, true>
                                           : &k_pfirst<RR, W, ND, false, false>;
      else kf = &k_pfirst<RR, W, ND, false, false>;
    }
    if (!kf) return hipErrorInvalidValue;
    if (blocks == 0) return hipSuccess;
    pk_run(kf, blocks, CC_BLOCK, 0, st, e0, e1, a);
    return hipGetLastError();
  }
  if constexpr (pk_fits(Q, RR, ND)) {
    kf = a.in_kind == CC_IN_PACKED
             ? pk_select<Q, RR, W, ND, false, 1, false>(sum, in2, out2)
             : pk_select<Q, RR, W, ND, false, 1, true>(sum, false, out2);
    if (blocks == 0) return hipSuccess;
    if (lds > 64 * 1024) {
      hipError_t e = hipFuncSetAttribute((const void *)kf,
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)lds);
      if (e != hipSuccess) return e;
    }
    pk_run(kf, blocks, threads, lds, st, e0, e1, a);
    return hipGetLastError();
  } else {
    return hipErrorInvalidValue;
  }
}

template <int Q, int RR>
hipError_t pk_by_fmt(const PSweepArgs &a, int W, int ND, int blocks,
                     int threads, size_t lds, hipStream_t st, hipEvent_t e0,
                     hipEvent_t e1) {
  if (W == 8 && ND == 1) return pk_go<Q, RR, 8, 1>(a, blocks, threads, lds, st, e0, e1);
  if (W == 8 && ND == 2) return pk_go<Q, RR, 8, 2>(a, blocks, threads, lds, st, e0, e1);
  if (W == 16 && ND == 1) return pk_go<Q, RR, 16, 1>(a, blocks, threads, lds, st, e0, e1);
  if (W == 16 && ND == 2) return pk_go<Q, RR, 16, 2>(a, blocks, threads, lds, st, e0, e1);
  if (W == 16 && ND == 3) return pk_go<Q, RR, 16, 3>(a, blocks, threads, lds, st, e0, e1);
  if (W == 16 && ND == 4) return pk_go<Q, RR, 16, 4>(a, blocks, threads, lds, st, e0, e1);
  return hipErrorInvalidValue;
}

template <int Q, int MAXRR>
hipError_t pk_by_rr(const PSweepArgs &a, int RR, int W, int ND, int blocks,
                    int threads, size_t lds, hipStream_t st, hipEvent_t e0,
                    hipEvent_t e1) {
  switch (RR) {
    case 1: return pk_by_fmt<Q, 1>(a, W, ND, blocks, threads, lds, st, e0, e1);
    case 2: return pk_by_fmt<Q, 2>(a, W, ND, blocks, threads, lds, st, e0, e1);
    case 3: if constexpr (MAXRR >= 3) return pk_by_fmt<Q, 3>(a, W, ND, blocks, threads, lds, st, e0, e1); break;
    case 4: if constexpr (MAXRR >= 4) return pk_by_fmt<Q, 4>(a, W, ND, blocks, threads, lds, st, e0, e1); break;
    case 5: if constexpr (MAXRR >= 5) return pk_by_fmt<Q, 5>(a, W, ND, blocks, threads, lds, st, e0, e1); break;
    case 6: if constexpr (MAXRR >= 6) return pk_by_fmt<Q, 6>(a, W, ND, blocks, threads, lds, st, e0, e1); break;
  }
  return hipErrorInvalidValue;
}

// cc_psweep_inst.hpp -- launch table of the packed sweep kernels for one
// axis size Q.  A kernel exists for every round size RR (cube C = Q^RR
// words) and register format (W-bit fields in ND dwords) whose cube stays
// within 128 registers (C * ND <= 128).
#pragma once
#include "cc_launch.hpp"
#include "cc_psweep.hpp"

constexpr bool pk_fits(int q, int rr, int nd) { return cc_ipow(q, rr) * nd <= 128; }

template <int Q, int RR, int W, int ND>
hipError_t pk_go(const PSweepArgs &a, int blocks, int threads, size_t lds,
                 hipStream_t st) {
  if constexpr (pk_fits(Q, RR, ND)) {
    auto kf = a.out_kind == CC_OUT_PACKED ? k_psweep<Q, RR, W, ND, false>
                                          : k_psweep<Q, RR, W, ND, true>;
    if (lds > 64 * 1024) {
      hipError_t e = hipFuncSetAttribute((const void *)kf,
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)lds);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kf, dim3(blocks), dim3(threads), lds, st, a);
    return hipGetLastError();
  } else {
    return hipErrorInvalidValue;
  }
}

template <int Q, int RR>
hipError_t pk_by_fmt(const PSweepArgs &a, int W, int ND, int blocks,
                     int threads, size_t lds, hipStream_t st) {
  if (W == 8 && ND == 1) return pk_go<Q, RR, 8, 1>(a, blocks, threads, lds, st);
  if (W == 8 && ND == 2) return pk_go<Q, RR, 8, 2>(a, blocks, threads, lds, st);
  if (W == 16 && ND == 1) return pk_go<Q, RR, 16, 1>(a, blocks, threads, lds, st);
  if (W == 16 && ND == 2) return pk_go<Q, RR, 16, 2>(a, blocks, threads, lds, st);
  if (W == 16 && ND == 3) return pk_go<Q, RR, 16, 3>(a, blocks, threads, lds, st);
  if (W == 16 && ND == 4) return pk_go<Q, RR, 16, 4>(a, blocks, threads, lds, st);
  return hipErrorInvalidValue;
}

template <int Q, int MAXRR>
hipError_t pk_by_rr(const PSweepArgs &a, int RR, int W, int ND, int blocks,
                    int threads, size_t lds, hipStream_t st) {
  switch (RR) {
    case 1: return pk_by_fmt<Q, 1>(a, W, ND, blocks, threads, lds, st);
    case 2: return pk_by_fmt<Q, 2>(a, W, ND, blocks, threads, lds, st);
    case 3: if constexpr (MAXRR >= 3) return pk_by_fmt<Q, 3>(a, W, ND, blocks, threads, lds, st); break;
    case 4: if constexpr (MAXRR >= 4) return pk_by_fmt<Q, 4>(a, W, ND, blocks, threads, lds, st); break;
    case 5: if constexpr (MAXRR >= 5) return pk_by_fmt<Q, 5>(a, W, ND, blocks, threads, lds, st); break;
    case 6: if constexpr (MAXRR >= 6) return pk_by_fmt<Q, 6>(a, W, ND, blocks, threads, lds, st); break;
  }
  return hipErrorInvalidValue;
}

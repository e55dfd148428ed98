// cc_round_inst.hpp -- launch table of the register-round sweep kernels
// (k_sweep_r) of homogeneous Hamming groups, one (R1, Q, RR) per entry; with
// WCMP the table also holds the compact-word variants.
#pragma once
#include "cc_launch.hpp"
#include "cc_sweep.hpp"

template <int R1, int Q, int RR, bool WCMP>
static hipError_t cc_round_go(const SweepArgs &a, int blocks, size_t lds, bool l16,
                              bool cmp, const LaunchCtx &lc) {
  if constexpr (WCMP) {
    if (cmp && l16)
      return cc_launch_k(k_sweep_r<CC_HAMMING, R1, Q, RR, uint16_t, true>, a, blocks, lds, lc, lc.threads);
    if (cmp)
      return cc_launch_k(k_sweep_r<CC_HAMMING, R1, Q, RR, int32_t, true>, a, blocks, lds, lc, lc.threads);
  } else {
    if (cmp) return hipErrorInvalidValue;
  }
  if (l16)
    return cc_launch_k(k_sweep_r<CC_HAMMING, R1, Q, RR, uint16_t>, a, blocks, lds, lc, lc.threads);
  return cc_launch_k(k_sweep_r<CC_HAMMING, R1, Q, RR, int32_t>, a, blocks, lds, lc, lc.threads);
}

template <int Q, int RR, bool WCMP>
static hipError_t cc_round_by_l(const SweepArgs &a, int L, int blocks, size_t lds,
                                bool l16, bool cmp, const LaunchCtx &lc) {
  switch (L) {
    case 2: return cc_round_go<2, Q, RR, WCMP>(a, blocks, lds, l16, cmp, lc);
    case 3: return cc_round_go<3, Q, RR, WCMP>(a, blocks, lds, l16, cmp, lc);
    case 4: return cc_round_go<4, Q, RR, WCMP>(a, blocks, lds, l16, cmp, lc);
    case 5: return cc_round_go<5, Q, RR, WCMP>(a, blocks, lds, l16, cmp, lc);
    case 6: return cc_round_go<6, Q, RR, WCMP>(a, blocks, lds, l16, cmp, lc);
    case 7: return cc_round_go<7, Q, RR, WCMP>(a, blocks, lds, l16, cmp, lc);
  }
  return hipErrorInvalidValue;
}

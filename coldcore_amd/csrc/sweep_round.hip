// sweep_round.hip -- register-round sweeps of homogeneous Hamming groups.
#include "cc_launch.hpp"
#include "cc_sweep.hpp"

template <int R1, int Q, int RR>
static hipError_t go(const SweepArgs &a, int blocks, size_t lds, bool l16,
                     const LaunchCtx &lc) {
  if (l16)
    return cc_launch_k(k_sweep_r<CC_HAMMING, R1, Q, RR, uint16_t>, a, blocks, lds, lc);
  return cc_launch_k(k_sweep_r<CC_HAMMING, R1, Q, RR, int32_t>, a, blocks, lds, lc);
}

template <int R1>
static hipError_t by_q(const SweepArgs &a, int q, int blocks, size_t lds,
                       bool l16, const LaunchCtx &lc) {
  switch (q) {
    case 2: return go<R1, 2, 4>(a, blocks, lds, l16, lc);
    case 3: return go<R1, 3, 2>(a, blocks, lds, l16, lc);
    case 4: return go<R1, 4, 2>(a, blocks, lds, l16, lc);
    case 5: return go<R1, 5, 1>(a, blocks, lds, l16, lc);
  }
  return hipErrorInvalidValue;
}

hipError_t cc_sweep_round(const SweepArgs &a, int L, int q, int blocks,
                          size_t lds, bool l16, const LaunchCtx &lc) {
  switch (L) {
    case 2: return by_q<2>(a, q, blocks, lds, l16, lc);
    case 3: return by_q<3>(a, q, blocks, lds, l16, lc);
    case 4: return by_q<4>(a, q, blocks, lds, l16, lc);
    case 5: return by_q<5>(a, q, blocks, lds, l16, lc);
    case 6: return by_q<6>(a, q, blocks, lds, l16, lc);
    case 7: return by_q<7>(a, q, blocks, lds, l16, lc);
  }
  return hipErrorInvalidValue;
}

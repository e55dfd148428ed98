// sweep_round_q67.hip -- register-round sweeps, axis sizes 6 and 7 (one
// axis per round), layer planes or compact words between sweeps.
#include "cc_round_inst.hpp"

hipError_t cc_sweep_round_q67(const SweepArgs &a, int L, int q, int blocks,
                              size_t lds, bool l16, bool cmp, const LaunchCtx &lc) {
  if (q == 6) return cc_round_by_l<6, 1, true>(a, L, blocks, lds, l16, cmp, lc);
  if (q == 7) return cc_round_by_l<7, 1, true>(a, L, blocks, lds, l16, cmp, lc);
  return hipErrorInvalidValue;
}

// sweep_round_q810.hip -- register-round sweeps, axis sizes 8, 9 and 10
// (one axis per round), layer planes or compact words between sweeps.
#include "cc_round_inst.hpp"

hipError_t cc_sweep_round_q810(const SweepArgs &a, int L, int q, int blocks,
                               size_t lds, bool l16, bool cmp, const LaunchCtx &lc) {
  if (q == 8) return cc_round_by_l<8, 1, true>(a, L, blocks, lds, l16, cmp, lc);
  if (q == 9) return cc_round_by_l<9, 1, true>(a, L, blocks, lds, l16, cmp, lc);
  if (q == 10) return cc_round_by_l<10, 1, true>(a, L, blocks, lds, l16, cmp, lc);
  return hipErrorInvalidValue;
}

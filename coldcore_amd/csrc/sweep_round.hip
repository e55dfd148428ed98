// sweep_round.hip -- register-round sweeps of homogeneous Hamming groups,
// axis sizes 2..5 (layer planes between sweeps); larger axes go to their
// own units.
#include "cc_round_inst.hpp"

hipError_t cc_sweep_round(const SweepArgs &a, int L, int q, int blocks,
                          size_t lds, bool l16, bool cmp, const LaunchCtx &lc) {
  switch (q) {
    case 2: return cc_round_by_l<2, 4, false>(a, L, blocks, lds, l16, cmp, lc);
    case 3: return cc_round_by_l<3, 2, false>(a, L, blocks, lds, l16, cmp, lc);
    case 4: return cc_round_by_l<4, 2, false>(a, L, blocks, lds, l16, cmp, lc);
    case 5: return cc_round_by_l<5, 1, false>(a, L, blocks, lds, l16, cmp, lc);
    case 6: case 7: return cc_sweep_round_q67(a, L, q, blocks, lds, l16, cmp, lc);
    case 8: case 9: case 10:
      return cc_sweep_round_q810(a, L, q, blocks, lds, l16, cmp, lc);
  }
  return hipErrorInvalidValue;
}

// sweep_lee.hip -- per-axis sweep kernels of the Lee family.
#include "cc_launch.hpp"
#include "cc_sweep.hpp"

template <int R1>
static hipError_t go(const SweepArgs &a, int blocks, size_t lds, bool l16,
                     bool cmp, const LaunchCtx &lc) {
  if (cmp) return cc_launch_k(k_sweep<CC_LEE, R1, int32_t, true>, a, blocks, lds, lc);
  if (l16) return cc_launch_k(k_sweep<CC_LEE, R1, uint16_t>, a, blocks, lds, lc);
  return cc_launch_k(k_sweep<CC_LEE, R1, int32_t>, a, blocks, lds, lc);
}

hipError_t cc_sweep_lee(const SweepArgs &a, int L, int blocks, size_t lds,
                            bool l16, bool cmp, const LaunchCtx &lc) {
  switch (L) {
    case 2: return go<2>(a, blocks, lds, l16, cmp, lc);
    case 3: return go<3>(a, blocks, lds, l16, cmp, lc);
    case 4: return go<4>(a, blocks, lds, l16, cmp, lc);
    case 5: return go<5>(a, blocks, lds, l16, cmp, lc);
    case 6: return go<6>(a, blocks, lds, l16, cmp, lc);
    case 7: return go<7>(a, blocks, lds, l16, cmp, lc);
  }
  return hipErrorInvalidValue;
}

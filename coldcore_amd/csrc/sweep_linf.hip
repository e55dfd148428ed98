// sweep_linf.hip -- single-plane window-sum sweeps (torus and open grid).
#include "cc_launch.hpp"
#include "cc_sweep.hpp"

hipError_t cc_sweep_linf(const SweepArgs &a, int prob, int blocks, size_t lds,
                         const LaunchCtx &lc) {
  if (prob == CC_TORUS)
    return cc_launch_k(k_sweep<CC_TORUS, 1, int32_t>, a, blocks, lds, lc);
  return cc_launch_k(k_sweep<CC_GRID, 1, int32_t>, a, blocks, lds, lc);
}

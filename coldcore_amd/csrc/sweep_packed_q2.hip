// sweep_packed_q2.hip -- packed-layer Hamming sweeps for axis size 2.
#include "cc_psweep_inst.hpp"

hipError_t cc_psweep_q2(const PSweepArgs &a, int RR, int W, int ND,
                        int blocks, int threads, size_t lds, hipStream_t st,
                        hipEvent_t e0, hipEvent_t e1) {
  return pk_by_rr<2, 6>(a, RR, W, ND, blocks, threads, lds, st, e0, e1);
}

// sweep_packed_q4.hip -- packed-layer Hamming sweeps for axis size 4.
#include "cc_psweep_inst.hpp"

hipError_t cc_psweep_q4(const PSweepArgs &a, int RR, int W, int ND,
                        int blocks, int threads, size_t lds, hipStream_t st,
                        hipEvent_t e0, hipEvent_t e1) {
  return pk_by_rr<4, 3>(a, RR, W, ND, blocks, threads, lds, st, e0, e1);
}

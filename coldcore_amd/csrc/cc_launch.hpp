// cc_launch.hpp -- sweep launchers.  Each family's kernel instantiations
// live in their own translation unit (sweep_*.hip) so the library builds in
// parallel; the engine calls them through these declarations.
#pragma once
#include "cc_common.hpp"

struct LaunchCtx {
  hipStream_t st;
  int ncu;
  int grid;   // workgroups per CU for the persistent grid; 0 = one per tile,
              // < 0 = the occupancy the runtime reports
  int threads = CC_BLOCK;   // workgroup size of the round kernels
};

// Launch a sweep kernel on a persistent grid: as many workgroups as the
// CUs hold at once (occupancy from the kernel's registers and LDS), each
// looping over tiles, capped by the tile count.
// threads: workgroup size (a multiple of 64, at most CC_BLOCK); the round
// kernels of a first sweep with few cubes per tile run smaller workgroups
template <typename K>
hipError_t cc_launch_k(K kf, const SweepArgs &a, int blocks, size_t lds,
                       const LaunchCtx &lc, int threads = CC_BLOCK) {
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void *)kf,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
  }
  int per_cu = lc.grid;
  if (per_cu < 0 &&
      (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kf, threads, lds) !=
           hipSuccess || per_cu < 1))
    per_cu = 1;
  const long long resident = per_cu == 0 ? blocks : (long long)per_cu * lc.ncu;
  const int g = (int)(blocks < resident ? blocks : resident);
  hipLaunchKernelGGL(kf, dim3(g), dim3(threads), lds, lc.st, a);
  return hipGetLastError();
}

// per-axis sweep kernels: L = layer planes (1 for the L-inf families); cmp:
// the sweep reads or writes compact words
hipError_t cc_sweep_hamming(const SweepArgs &a, int L, int blocks, size_t lds,
                            bool l16, bool cmp, const LaunchCtx &lc);
hipError_t cc_sweep_lee(const SweepArgs &a, int L, int blocks, size_t lds,
                        bool l16, bool cmp, const LaunchCtx &lc);
hipError_t cc_sweep_linf(const SweepArgs &a, int prob, int blocks, size_t lds,
                         const LaunchCtx &lc);
// register-round kernel for homogeneous Hamming groups, axis size q <= 10
// (compact words for q >= 6); the q >= 6 kernels live in their own units
hipError_t cc_sweep_round(const SweepArgs &a, int L, int q, int blocks,
                          size_t lds, bool l16, bool cmp, const LaunchCtx &lc);
hipError_t cc_sweep_round_q67(const SweepArgs &a, int L, int q, int blocks,
                              size_t lds, bool l16, bool cmp, const LaunchCtx &lc);
hipError_t cc_sweep_round_q810(const SweepArgs &a, int L, int q, int blocks,
                               size_t lds, bool l16, bool cmp, const LaunchCtx &lc);

// packed-layer Hamming sweeps (cc_psweep.hpp), one unit per axis size Q:
// RR axes per register round, W-bit fields in ND dwords per word; with
// events e0 / e1 the launch records the kernel's own start and end
struct PSweepArgs;
hipError_t cc_psweep_q2(const PSweepArgs &a, int RR, int W, int ND, int blocks,
                        int threads, size_t lds, hipStream_t st,
                        hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
hipError_t cc_psweep_q3(const PSweepArgs &a, int RR, int W, int ND, int blocks,
                        int threads, size_t lds, hipStream_t st,
                        hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
hipError_t cc_psweep_q4(const PSweepArgs &a, int RR, int W, int ND, int blocks,
                        int threads, size_t lds, hipStream_t st,
                        hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
hipError_t cc_psweep_q5(const PSweepArgs &a, int RR, int W, int ND, int blocks,
                        int threads, size_t lds, hipStream_t st,
                        hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);

/* coldcore_amd/batch.h -- the batched multi-cell engine (additive C ABI).
 *
 * No counterpart in the reference: its plugin holds one problem cell per
 * process (reference docs/PLUGIN.md:73-76, src/include/coldcore/plugin.h:
 * 66-80 -- dct_init3 replaces the one cell).  These entry points live in the
 * same library as the dct_* surface (libcoldcore_covering.so) and hold a
 * separate arena, so a host can keep one singleton cell and one batch.
 *
 * A batch is many uneven cells at once: Hamming, Lee (R <= 3, the same v1 /
 * wide support tables as dct_init3 / dct_init3w), king torus and open grid
 * (n <= 8, axes <= 1024, R <= 3; torus axes >= 2R + 1), each with at most
 * 2^26 words.  The exact transform of every cell runs as one persistent
 * launch per sweep level (the l-th sweep of every cell together; no
 * workgroup waits on another); a batched large-neighbourhood-search round
 * (ruin k random code
 * words per cell, greedy recreate by exact gain, keep or revert) is a fixed
 * launch sequence captured once into a HIP graph.
 *
 * Return codes: 0 ok; -1 invalid argument or no batch; -3 allocation;
 * -(100 + e) HIP error e.
 * Calls are not thread-safe; the batch owns one HIP stream.
 */
#ifndef COLDCORE_AMD_BATCH_H
#define COLDCORE_AMD_BATCH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ncells cells: problem[i] (0 Hamming, 1 torus L-inf, 2 Lee, 3 open grid,
 * the dct_init3 ids), n[i] axes (all cells' axes concatenated in axes),
 * radius R[i] in 1..3; Hamming / Lee cells bring npats[i] support patterns
 * (concatenated in pats) of format fmt[i] (0 v1: n <= 16, 1 wide); the
 * L-inf cells take npats[i] = 0.  cap = code words per cell (<= 65535).
 * Replaces any previous batch. */
int ccb_init(int ncells, const int *problem, const int *n, const int *axes,
             const int *R, const uint64_t *pats, const long long *npats,
             const int *fmt, int cap);
int ccb_free(void);

/* the code of one cell: M word indices (a set or a multiset) */
int ccb_set_code(int cell, const long long *idx, int M);

/* the transform of every cell, one persistent launch per sweep level; (init_mode,
 * out_mode) as dct_transform: (0, 2) recount cnt from the code, (1, 1) gain
 * map into the map plane, (2, 1) loss map (mu = 1) */
int ccb_transform(int init_mode, int out_mode);

/* a cell's cnt (X uint16) or map (X int32) plane */
int ccb_read_cnt(int cell, uint16_t *out);
int ccb_read_map(int cell, int32_t *out);

/* steps greedy steps on every cell (gain, argmax with the smallest index on
 * ties, add; cells already covered are unchanged); cnt must be current.
 * Returns the number of cells still uncovered afterwards (>= 0). */
int ccb_greedy(int steps);

/* LNS state from the current cnt (call after a recount or greedy): per-cell
 * uncovered counts, random streams from seed, best = current if feasible */
int ccb_lns_init(unsigned long long seed);

/* rounds LNS rounds with k (1..8) ruined words per cell and round: a covered
 * cell re-adds k - 1 words (one fewer), an uncovered one k; a round that
 * leaves more uncovered words than it started with is reverted with
 * probability 0.9.  graph != 0 replays one captured round (captured on the
 * first call for this k).  *ms = wall milliseconds of the rounds. */
int ccb_lns_run(int rounds, int k, int graph, float *ms);

/* which 0: the current code, 1: the best covering code found; returns its
 * size (or < 0) and copies at most cap words */
int ccb_get_code(int cell, int which, long long *out, int cap);

/* out[0..5] = code size, uncovered words, best covering size (-1 none),
 * rounds, accepted rounds, reverted rounds */
int ccb_stats(int cell, long long *out);

/* out[0..6] = cells, sweeps, work items, sweep levels (persistent launches
 * per transform), workgroups per launch at most, LDS budget of the plan,
 * threads per workgroup of the level launches */
int ccb_info(long long *out);

/* the batch's schedule autotuner: times reps batched gain transforms under
 * each LDS budget (8 to 128 KiB per tile; axis grouping, tile widths and
 * the persistent grid follow from it) and each level-launch workgroup size
 * (256, 512, 1,024 threads per tile) and keeps the fastest plan.
 * Overwrites the map planes (cnt unchanged) and drops a captured graph.
 * Returns the chosen budget in bytes; *best_us = its time per transform. */
int ccb_autotune(int reps, float *best_us);

#ifdef __cplusplus
}
#endif
#endif

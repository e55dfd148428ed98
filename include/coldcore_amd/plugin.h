/* coldcore_amd/plugin.h -- C ABI of the MI355X exact-transform engine.
 *
 * The engine library libcoldcore_covering.so (built from
 * coldcore_amd/csrc/dct_engine.hip for gfx950) is a drop-in for the
 * reference's CUDA plugin of the same name.  Every entry point keeps the
 * reference's name, signature, argument meaning and return codes; the
 * comment on each cites the reference declaration it replaces
 * (reference src/include/coldcore/plugin.h, implementation
 * src/plugins/covering/dct_hc.cu).
 *
 * Conventions (reference plugin.h:7-19, docs/PLUGIN.md:71-140):
 *  - singleton: one cell per process; not thread-safe;
 *  - return codes: 0 ok, -1 invalid argument or rejected input (state left
 *    untouched), -2 space above 2^40, -3/-4/-5 allocation failure,
 *    -(100 + hipError_t) device error (message on stderr);
 *  - caller-allocated outputs; inputs are read only during the call;
 *  - all device work runs in call order on one stream owned by the engine;
 *    calls that return data synchronise once, ball_update of <= 1024 words
 *    returns without waiting (a device fault surfaces at the next
 *    synchronising call);
 *  - every plane lives in HBM: layers_mode / cnt_mode are accepted and
 *    ignored, dct_host_bytes() returns 0.
 * Strict refinements of the reference contract: map_extract(64) returns the
 * hits in ascending index order; read_cnt and the gathers validate their
 * ranges and return -1 instead of reading out of bounds; init rejects a
 * support table whose patterns are malformed (position >= n, repeated
 * position, delta outside 1..ax-1).
 */
#ifndef COLDCORE_AMD_PLUGIN_H
#define COLDCORE_AMD_PLUGIN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define COLDCORE_PLUGIN_ABI_VERSION 1

/* problem ids (reference plugin.h:45-53) */
#define CCP_PROB_HAMMING    0
#define CCP_PROB_TORUS_LINF 1
#define CCP_PROB_LEE        2
#define CCP_PROB_GRID_LINF  3

/* ---- lifecycle ---------------------------------------------------------- */
/* reference plugin.h:62-65, dct_hc.cu:230-374.  ballpat: v1 support table,
 * one uint64 per pattern, byte k = pos<<4 | delta; entry 0 = the centre. */
int  dct_init3(int problem, int n, const int *ax, int R,
               const uint64_t *ballpat, long long balllen, int extract_cap,
               long long *bytes, int layers_mode, int cnt_mode,
               int use_owner);
/* reference plugin.h:66, dct_hc.cu:393-428 (also resets mu to 1) */
int  dct_free(void);
/* reference plugin.h:67, dct_hc.cu:208 */
long long dct_host_bytes(void);
/* reference plugin.h:70-73, dct_hc.cu:376-391 (homogeneous Hamming) */
int  dct_init2(int q, int n, int R, const uint64_t *ballpat,
               long long balllen, int extract_cap, long long *bytes,
               int layers_mode, int cnt_mode, int use_owner);
int  dct_init (int q, int n, int R, const uint64_t *ballpat,
               long long balllen, int extract_cap, long long *bytes);

/* ---- solution / transforms ---------------------------------------------- */
/* reference plugin.h:79, dct_hc.cu:452-459: pending code multiset (<= 65535
 * occurrences); it becomes the represented solution at transform(0, 2). */
int  dct_set_code(const long long *idx, int M);
/* reference plugin.h:80, dct_hc.cu:465-521.  init_mode 0 = code multiset,
 * 1 = [cnt < mu], 2 = [1 <= cnt <= mu]; out_mode 1 = sum -> map (A0),
 * 2 = clamped sum -> cnt (init_mode 0 only). */
int  dct_transform(int init_mode, int out_mode);

/* ---- mu-fold covering ---------------------------------------------------- */
/* reference plugin.h:89-90, dct_hc.cu:214-222 */
int  dct_set_mu(int mu);
int  dct_get_mu(void);

/* ---- incremental / exact per-candidate ---------------------------------- */
/* reference plugin.h:93, dct_hc.cu:567-587 */
int  dct_ball_update(const long long *words, int nwords, int delta);
/* reference plugin.h:94-95, dct_hc.cu:785-788: literal cnt == target */
int  dct_ball_gather(const long long *words, int nwords, int target,
                     int32_t *out);
/* reference plugin.h:98-99, dct_hc.cu:794-799: role 0 deficient, 1 critical */
int  dct_ball_gather_mu(const long long *words, int nwords, int role,
                        int32_t *out);
/* reference plugin.h:100, dct_hc.cu:997-1022 (needs use_owner = 1) */
int  dct_loss_owner(const long long *words, int M, int32_t *out);

/* ---- v0.3 optional ------------------------------------------------------- */
/* reference plugin.h:110-111, dct_hc.cu:826-866 */
int  dct_ball_update_delta(const long long *words, int nwords, int delta,
                           long long *dd);
/* reference plugin.h:112-114, dct_hc.cu:882-977 */
int  dct_swap_eval(const long long *drops, int k, const long long *cand,
                   const long long *off, int32_t *gain, int32_t *loss);
/* reference plugin.h:115, dct_hc.cu:779-783.  out[0..7]: set-side gathers,
 * walk gathers, predicate listings, listing overflows, fast swap_eval
 * calls, fallback swap_eval calls, (unused), set-side gathers over more
 * than 32,768 listed words.  The set side is taken only when the walked
 * table is proven to be exactly the metric ball; CC_SET_POLICY (0 never,
 * 1 auto, 2 always), CC_SET_RATIO and CC_SET_MIN are read at init. */
int  dct_stats(long long *out, int n);

/* ---- reductions / map queries ------------------------------------------- */
/* reference plugin.h:121, dct_hc.cu:1024-1054 */
long long dct_count_eq(int target);
/* reference plugin.h:124, dct_hc.cu:981-992 */
long long dct_count_deficit(void);
/* reference plugin.h:125, dct_hc.cu:1056-1069 */
int  dct_map_max(int32_t *out);
/* reference plugin.h:126, dct_hc.cu:1071-1089 */
int  dct_map_hist(int nbins, int32_t vmax, int32_t *out);
/* reference plugin.h:127-130, dct_hc.cu:1091-1134 (ascending order here) */
long long dct_map_extract  (int32_t thr, uint32_t *idx, int32_t *val,
                            int cap);
long long dct_map_extract64(int32_t thr, int64_t *idx, int32_t *val,
                            int cap);
/* reference plugin.h:131, dct_hc.cu:1136-1150 */
int  dct_map_read_at(const long long *idx, int k, int32_t *out);
/* reference plugin.h:132, dct_hc.cu:1152-1157 */
int  dct_read_cnt(long long off, long long len, uint16_t *out);

/* ---- optional ox support-table format ----------------------------------- */
/* reference plugin.h:135-137, dct_hc.cu:430-450 */
int  dct_set_tbl2(const uint16_t *mask, const uint64_t *offs,
                  long long len);
int  dct_use_fmt(int fmt);

/* ---- additive, no counterpart in the reference -------------------------- */
/* dct_init3 with a wide-format table: each pattern a uint64 of up to six
 * 10-bit fields (pos<<4 | delta) at bits [10k, 10k+10), zero field ends.
 * v1 bytes hold a position in a nibble, so the reference ABI stops at 16
 * axes (reference plugin.h:22, dct_hc.cu:233); this entry accepts up to 40
 * (|X| <= 2^40), which the binary benchmark cells K_2(24,3) and K_2(30,4)
 * need.  Same cell, same semantics, same return codes otherwise. */
int  dct_init3w(int problem, int n, const int *ax, int R,
                const uint64_t *widepat, long long balllen, int extract_cap,
                long long *bytes, int layers_mode, int cnt_mode,
                int use_owner);
/* dct_init3 with a long-axis table: each pattern a uint64 of up to four
 * 16-bit fields (pos<<10 | delta) at bits [16k, 16k+16), zero field ends.
 * v1 and ox hold a delta in a nibble, so the reference ABI stops at axes of
 * 16 (reference docs/PLUGIN.md:142-152); this entry takes torus_linf and
 * grid_linf cells with axes up to 1024 (R <= 3 once an axis exceeds 16),
 * e.g. the k x k king tori for k = 17..31.  Same cell, same semantics,
 * same return codes otherwise. */
int  dct_init3x(int problem, int n, const int *ax, int R,
                const uint64_t *xpat, long long balllen, int extract_cap,
                long long *bytes, int layers_mode, int cnt_mode,
                int use_owner);
/* The transform of dct_transform(init_mode, 1) with its last sweep keeping
 * the exact distance layers: after it, layer d at x = number of seed words
 * at distance exactly d from x (init 0 codewords of set_code, 1 deficient,
 * 2 critical).  cnt and the represented solution are unchanged, the map
 * plane is overwritten.  Hamming and Lee only (-1 otherwise).  The layers
 * feed the learned ruin proposal. */
int  dct_transform_layers(int init_mode);
/* out[i*(R+1) + d] = layer d at idx[i] after dct_transform_layers. */
int  dct_layers_read_at(const long long *idx, int k, int32_t *out);
/* Ruin-region features of the learned proposal, computed on the device:
 * sets the pending code to code[0..M), runs the deficient, critical and
 * codeword layer transforms and writes, for each of the nreg regions of k
 * words (words[r*k .. r*k+k)), five windows of R+1 floats to out[r*5*(R+1)]:
 * mean of the critical layers, mean and max of the deficient layers, mean
 * of the codeword layers, min of the critical layers, each layer count
 * divided by its distance shell size.  Hamming only, nreg*k <= 65,536.
 * cnt and the represented code are unchanged; the map plane is overwritten.
 * No reference counterpart (the reference has no learned operator). */
int  dct_region_features(const long long *code, int M, const long long *words,
                         int nreg, int k, float *out);
/* Candidate ruin regions and their features in one device call: for each of
 * the na anchor positions into code[0..M), the k code words nearest to the
 * anchor in digit distance (ties by a 24-bit splitmix64 hash of (seed,
 * anchor position, word position), then by position), written as positions
 * to regions[b*k ..] in that order, and their five windows (as
 * dct_region_features) to feats[b*5*(R+1) ..].  Hamming only; M <= 65,535,
 * na <= 64, k <= min(M, 64).  No reference counterpart. */
int  dct_region_propose(const long long *code, int M, const int32_t *anchors,
                        int na, int k, unsigned long long seed, int32_t *regions,
                        float *feats);
/* Schedule of the LDS-tiled sweeps: 8 ints per sweep (first axis, axes in
 * the group, P, TL, layers in, layers out, axis size of the register-round
 * kernel or 0 for the per-axis kernel, LDS bytes); returns the sweep count.
 * For benchmarks and the roofline accounting of each sweep's traffic. */
int  dct_sweep_plan(int *out, int cap);
/* The schedule a transform runs now (kind 0: gain / loss; 1: recount of the
 * pending code): 13 ints per sweep -- first axis, axes, P, TL, register
 * round RR (0: unpacked per-layer kernels), field width W of the packed
 * layers (32: unpacked planes), dwords per word in registers, algorithmic
 * bytes per word read and written, LDS bytes, threads per workgroup,
 * rounds, kernel (0 unpacked, 1 LDS-tiled packed, 2 / 3 column kernel with
 * 1 / 2 lanes per cube, 4 LDS-free first sweep, 5 persistent first sweep
 * with prefetch).  Returns the sweep count.  Homogeneous Hamming cells with axis
 * size <= 5 run packed-layer sweeps: all R+1 layers of a word in bit
 * fields of one or two dwords, with field widths proven from closed-form
 * layer bounds, so inter-sweep traffic is 4-8 bytes per word. */
int  dct_sweep_plan2(int kind, int *out, int cap);
/* The packed schedule the engine would plan for a homogeneous Hamming cell
 * (axes ax[0..n-1], radius R), computed on the host with no device and no
 * live cell: kind 0 gain / loss, kind 1 recount of a code of M words with
 * multiplicities <= s0.  Rows as dct_sweep_plan2; returns the sweep count
 * (0: unpacked kernels), -1 on bad arguments or while a cell is live.  Every
 * scheduled sweep's kernel instantiation exists (the planner checks). */
int  dct_plan_probe(int n, const int *ax, int R, int kind, int s0, int M,
                    int *out, int cap);
/* Schedule autotuner: time the candidate schedules of the gain / loss
 * transform on the live cell (every combination of LDS budget, column-kernel
 * mode and first-sweep kernel the planner offers, plus the unpacked
 * per-layer kernels; reps timed runs each) and keep the fastest; the
 * winning options also steer the recount schedule.  Overwrites the map
 * plane like a transform; cnt and the represented solution are unchanged.
 * Returns the number of schedules timed (*best_us: microseconds per
 * transform of the winner), or < 0. */
int  dct_autotune(int reps, float *best_us);
/* Copy map-plane entries [off, off+len) (int32) to out.  The map-plane
 * counterpart of dct_read_cnt (no reference counterpart: the reference
 * reads the map only at listed indices, dct_map_read_at); 0 or -1. */
int  dct_map_read(long long off, long long len, int32_t *out);
/* Allocation-free ordered extraction: the dct_map_extract64 compaction of
 * a caller's device map (int32 map[0..n), 16-byte aligned) into caller
 * device buffers idx[0..cap) / val[0..cap), ascending index order, with
 * the exact hit count written to the device int64 *found; queued on
 * `stream` (a hipStream_t; NULL: the null stream) with no host
 * synchronisation.  No reference counterpart: it is for callers that cannot
 * allocate (code under HIP graph capture, serving loops).  It calls no
 * hipMalloc / hipFree and needs no live cell: its chunk counts and offsets
 * live in a static 1 MiB device array per GPU, so n is limited to 2^30
 * words (larger n returns -1 with a message on stderr), and calls that
 * overlap on one GPU race.  Returns 0 once queued, -1 on bad arguments,
 * -(100 + hipError_t) on a device error. */
int  dct_map_extract64_noalloc(const int32_t *map, long long n, int32_t thr,
                               int64_t *idx, int32_t *val, long long cap,
                               long long *found, void *stream);
/* Pin the sweep boundaries of the packed schedules: cuts[0..ncuts-1]
 * ascending axis counts in 1..n-1, each sweep ends at the next cut (the
 * planner still picks each sweep's kernel, tile and formats).  On a live
 * cell the gain / loss schedule is replanned at once and automatic timing
 * is switched off; a recount whose code seed cannot use the cuts keeps the
 * planner's schedule.  ncuts = 0 clears.  With no live cell the cuts steer
 * dct_plan_probe; init and free clear them.  Returns the gain / loss sweep
 * count (0 with no live cell), -1 on bad cuts or when no packed schedule
 * has these boundaries.  No reference counterpart: lets tests run every
 * kernel instantiation a large cell's schedules use on a small cell. */
int  dct_sweep_force(const int *cuts, int ncuts);
/* Record HIP events around every sweep of later transforms (on the
 * engine's stream; also enabled by DCT_TIMING=1 at init). */
int  dct_sweep_profile(int on);
/* Per-sweep milliseconds of the last transform; returns the sweep count. */
int  dct_sweep_times(float *ms, int cap);
/* on = 1: later packed transforms fold a census of what their last sweep
 * writes (gain / loss: max(0, max A0); recount: sum max(0, mu - cnt)) into
 * device slots, and dct_map_max / dct_count_eq(0) / dct_count_deficit read
 * it instead of scanning the field while the field is the one that
 * transform wrote.  Off at init (CC_FOLD=1 turns it on); costs the
 * transform a workgroup epilogue and one atomic per workgroup (~1.5 us on
 * 2^24 words), saves a 4 / 2 B/word scan per query.  Returns the previous
 * setting, -1 with no live cell or a bad argument.  No reference
 * counterpart (its reductions are separate passes). */
int  dct_set_fold(int on);

/* ---- additive: the learned ruin proposal's network -------------------
 * No counterpart in the reference (its search has no learned operator).
 * A 2-layer MLP, 16 fp32 inputs -> 256 relu -> 1 logit, trained with
 * binary cross-entropy; W1 and w2 are used as bf16 (MFMA), fp32 masters
 * and Adam state stay on the device.  One network per process; host
 * buffers in and out.  Parameter layout: W1 (256 x 16 row-major), b1
 * (256), w2 (256), b2 -- cc_mlp_param_count() floats. */
int  cc_mlp_init(int max_batch);
int  cc_mlp_free(void);
int  cc_mlp_param_count(void);
/* sets the fp32 masters and their bf16 copies; resets Adam */
int  cc_mlp_set_params(const float *p);
int  cc_mlp_get_params(float *p);
/* y[i] = logit of row i of X (B x 16), B <= max_batch */
int  cc_mlp_forward(const float *X, int B, float *y);
/* gradient of the mean BCE of B rows with targets t in [0, 1]; g and loss
 * are optional outputs (NULL skips the copy) */
int  cc_mlp_grad(const float *X, const float *t, int B, float *g, float *loss);
/* one Adam step on the gradient of the last cc_mlp_grad */
int  cc_mlp_adam(float lr, float beta1, float beta2, float eps);

#ifdef __cplusplus
}
#endif
#endif /* COLDCORE_AMD_PLUGIN_H */

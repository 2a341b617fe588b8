/*
 * poyhip.h -- C ABI of the MI355X (gfx950) implementation of POY's
 * direct-optimization pairwise alignment path.
 *
 * Every per-pair entry point replaces one OCaml `external` of POY 5.1's
 * Sequence.Align / Sequence module; the comment on each names the reference
 * stub (path:line in the POY source tree) whose arguments and results it
 * mirrors.  The batch and serving entry points have no reference
 * counterpart: they are the additions that run many pairs per launch.
 *
 * Conventions
 *  - Sequences are byte arrays of 5-bit nucleotide bitsets (A=1 C=2 G=4 T=8
 *    gap=16, IUPAC codes are ORs) and start with POY's leading gap (16).
 *    The per-pair entry points return POYHIP_EINVAL for a byte of 32 or
 *    more; batches and the server take device-resident sequences unchecked,
 *    and such a byte there gives undefined results (the kernels index
 *    32 x 32 tables with the raw codes).
 *  - Output sequences are written left to right into caller buffers.
 *  - Every function returns POYHIP_OK or a negative POYHIP_E* code and never
 *    aborts; poyhip_last_error() describes the last failure of a context.
 *  - A context owns one device stream and its scratch; calls on one context
 *    are not reentrant (POY's global Matrix.default had the same property).
 *  - No PyTorch or HIP types appear in signatures.  Pointers named d_* are
 *    device pointers (memory of the context's device); all others are host.
 */
#ifndef POYHIP_H
#define POYHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define POYHIP_OK 0
#define POYHIP_EINVAL -1   /* bad argument (e.g. longer sequence first)     */
#define POYHIP_ENOMEM -2   /* device or host allocation failed               */
#define POYHIP_EDEVICE -3  /* HIP runtime error                              */
#define POYHIP_ESTATE -4   /* call order violated (no cost matrix, no fill)  */
#define POYHIP_ECAP -5     /* output capacity too small                      */
#define POYHIP_EBUSY -6    /* serving: admission refused (memory budget)     */

typedef struct poyhip_ctx poyhip_ctx;

/* The fields of POY's struct cm (src/cm.h:33-76) used on the DNA path,
 * tables indexed (a << lcm) + b with lcm = 5.  Filled by the caller, as POY's
 * Cost_matrix.Two_D pushes its tables through cm_CAML_set_* (src/cm.c:
 * 2345-2407). */
typedef struct {
  int lcm;          /* 5 */
  int gap;          /* 16 */
  int a_sz;         /* 31 */
  int combinations; /* 1 */
  int cost_model;   /* 0 linear, 1 affine (cost_model_type) */
  int gap_open;     /* affine gap opening (cm_get_gap_opening_parameter) */
  int cost[2048];
  int worst[2048];
  int prepend[64];
  int tail[64];
  uint8_t median[2048];
} poyhip_cm;

/* ---- context ------------------------------------------------------------ */
/* Replaces the module initialisers (cm_CAML_initialize src/cm.c:2529,
 * mat_CAML_initialize src/matrices.c:345): opens `device`, creates a stream
 * and the context's scratch (the counterpart of Matrix.default). */
int poyhip_create(int device, poyhip_ctx **out);
int poyhip_destroy(poyhip_ctx *ctx);
const char *poyhip_last_error(const poyhip_ctx *ctx);
/* cm_CAML_create + cm_CAML_set_* (src/cm.c:2452, 2345-2407): installs the
 * 2D cost model used by every later call on the context. */
int poyhip_set_cm(poyhip_ctx *ctx, const poyhip_cm *cm);
/* mat_CAML_flush_memory (src/matrices.c:352): frees the context scratch,
 * and the device blocks destroyed batches left to the context. */
int poyhip_flush(poyhip_ctx *ctx);

/* ---- per-pair drop-ins (one launch per pair, synchronous) --------------- */

/* algn_CAML_simple_2 (src/algn.c:3134) = Align.c_cost_2
 * (src/sequence.ml:617): non-affine cost of s1 (not longer than s2) with
 * the band hint deltawh.  Like the reference it leaves the direction store
 * of this fill in the context for a following poyhip_backtrace_2d. */
int poyhip_cost_2(poyhip_ctx *ctx, const uint8_t *s1, int l1, const uint8_t *s2,
                  int l2, int deltawh, int *cost);

/* algn_CAML_backtrace_2d (src/algn.c:3424) = Align.extract_edited_2
 * (src/sequence.ml:944): aligned s1/s2 from the last poyhip_cost_2 on the
 * same sequences.  cap >= l1 + l2. */
int poyhip_backtrace_2d(poyhip_ctx *ctx, const uint8_t *s1, int l1,
                        const uint8_t *s2, int l2, uint8_t *r1, uint8_t *r2,
                        int cap, int *len, int swaped);

/* algn_CAML_limit_2 (src/algn.c:3161) = Align.c_cost_2_limit
 * (src/sequence.ml:620), called by Align.cost_2_limit (764).  The reference
 * fills the whole sequences (s1 not longer than s2, as poyhip_cost_2): the
 * limits enter only the full-plane-vs-band decision, whose height becomes
 * 50 + (s1_len - s2_len) + w (algn_nw_limit, src/algn.c:2954); h, s1_st and
 * s2_st are read nowhere.  Leaves the direction store for
 * poyhip_backtrace_2d_limit / poyhip_backtrace_2d like poyhip_cost_2. */
int poyhip_limit_2(poyhip_ctx *ctx, const uint8_t *s1, int l1, const uint8_t *s2, int l2,
                   int w, int h, int s1_st, int s2_st, int s1_len, int s2_len, int *cost);

/* algn_CAML_backtrace_2d_limit (src/algn.c:3449) = Align.extract_edited_2_limit
 * (src/sequence.ml:948): backtrace_2d (3278) of the last fill over a
 * window.  The walk starts at row algn_s1 - 1, column st_s2 + algn_s2 - 1,
 * reads s1 from index st_s1 + algn_s1 - 1 down (st_s1 shifts the symbols,
 * not the row, as in the reference) and s2 from st_s2 + algn_s2 - 1 down,
 * and ends when it reaches row 0 left of column st_s2.  Limits outside the
 * sequences: POYHIP_EINVAL.  Defined while the walk stays inside the band
 * the last fill kept (always after a full-plane fill); outside it the
 * reference reads a direction matrix that fill never wrote. */
int poyhip_backtrace_2d_limit(poyhip_ctx *ctx, const uint8_t *s1, int l1, const uint8_t *s2,
                              int l2, uint8_t *r1, uint8_t *r2, int cap, int *len, int st_s1,
                              int st_s2, int algn_s1, int algn_s2, int swaped);

/* algn_CAML_align_2d (src/algn.c:3503) = Align.c_align_2: fill + backtrace. */
int poyhip_align_2d(poyhip_ctx *ctx, const uint8_t *s1, int l1, const uint8_t *s2,
                    int l2, int deltawh, int swaped, uint8_t *r1, uint8_t *r2,
                    int cap, int *len, int *cost);

/* algn_CAML_cost_affine_3 (src/algn.c:2457) = Align.cost_2_affine
 * (src/sequence.ml:630): full affine DP without directions, either order. */
int poyhip_cost_affine_3(poyhip_ctx *ctx, const uint8_t *si, int li,
                         const uint8_t *sj, int lj, int *cost);

/* algn_CAML_align_affine_3 (src/algn.c:2359) = Align.c_align_affine_3
 * (src/sequence.ml:626): banded affine fill with T-doubling and
 * backtrace_aff.  li <= lj (else POYHIP_EINVAL, the reference's
 * "pass the shorter one as first").  Each output buffer holds cap >=
 * li + lj + 2 bytes. */
int poyhip_align_affine_3(poyhip_ctx *ctx, const uint8_t *si, int li,
                          const uint8_t *sj, int lj, int swaped, uint8_t *median,
                          int *median_len, uint8_t *medianwg, int *medianwg_len,
                          uint8_t *resi, uint8_t *resj, int *res_len, int cap,
                          int *cost);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): poyhip_align_affine_3 with the same arguments and
 * results, whose staging, direction store, workspace and outputs live in a
 * fixed 64 MiB device array per device, held until the process ends (about
 * li * lj + 16 * lj + 6 * (li + lj) bytes per pair: pairs up to ~8,000 x
 * 8,000).  A pair that does not fit returns POYHIP_ECAP without touching the
 * device.  Every call on one device shares that array: calls that overlap on
 * one device (several contexts or threads) race and must be serialised by the
 * caller. */
int poyhip_align_affine_3_noalloc(poyhip_ctx *ctx, const uint8_t *si, int li,
                                  const uint8_t *sj, int lj, int swaped, uint8_t *median,
                                  int *median_len, uint8_t *medianwg, int *medianwg_len,
                                  uint8_t *resi, uint8_t *resj, int *res_len, int cap,
                                  int *cost);

/* ---- diagonal-store Ukkonen aligner (src/newkkonen.c), affine model ----- */
/* newkkonen_CAML_algn_affine (src/newkkonen.c:1473) = NewkkAlign's
 * newkk_cost2_affine (src/sequence.ml:1840): cost of the affine band
 * aligner with T doubling of s1 against s2, l1 <= l2 (else POYHIP_EINVAL, the
 * reference's assertion).  swaped only matters to the backtrace, as in the
 * reference.  Non-affine tables return POYHIP_EINVAL: newkkonen_CAML_algn's
 * non-affine cells are defective (interior cells take cost 0) and, with no
 * reset between calls, its refill compares against the previous call's
 * matrix, so its result is not a function of its inputs.  Tables whose sums
 * could reach the aligner's infinity (INT_MAX / 2) return POYHIP_EINVAL. */
int poyhip_newkk_cost_2_affine(poyhip_ctx *ctx, const uint8_t *s1, int l1,
                               const uint8_t *s2, int l2, int swaped, int *cost);
/* newkkonen_CAML_algn_affine followed by newkkonen_CAML_backtrace_affine
 * (src/newkkonen.c:1847) on the same matrix, as NewkkAlign.get_alignment
 * pairs them (src/sequence.ml:1864-1881): r1, r2 (cap >= l1 + l2 bytes, the
 * reference's sz1 + sz2) receive the aligned pair, *len its length. */
int poyhip_newkk_align_2_affine(poyhip_ctx *ctx, const uint8_t *s1, int l1,
                                const uint8_t *s2, int l2, int swaped, int *cost,
                                uint8_t *r1, uint8_t *r2, int cap, int *len);

/* ---- unions of the implied alignment (src/union.c) ---------------------- */
/* union_CAML_make (src/union.c:318) = Sequence.Unions.make_union
 * (src/sequence.ml:2845): union_merge of the unions ua and ub of two
 * children, along their aligned sequences a, b and aligned median (all of
 * length lena), into a new union c.  A union is a sequence right-aligned in
 * a buffer of `cap` bytes plus three int32 arrays of `cap` entries (offset,
 * union_c1, union_c2: Unions.u, src/sequence.ml:2782), live parts at the
 * right end; ua and ub pass their sequence buffer and offsets.  c's buffer
 * and arrays (c_cap entries, Unions.union uses ua_len + ub_len + 2) are
 * written right-aligned, *c_len receives its length; entries the merge does
 * not write are 0 (the reference leaves them uninitialised).  The offsets
 * are 32-bit, the reference's USE_LONG_SEQUENCES layout.  POYHIP_EINVAL
 * where the reference fails ("infinit loop") or c_cap is too small. */
int poyhip_union_merge(poyhip_ctx *ctx, const uint8_t *a, const uint8_t *b,
                       const uint8_t *median, int lena, const uint8_t *ua_s, int ua_len,
                       int ua_cap, const int32_t *ua_off, const uint8_t *ub_s, int ub_len,
                       int ub_cap, const int32_t *ub_off, uint8_t *c_s, int c_cap,
                       int32_t *c_off, int32_t *c_c1, int32_t *c_c2, int *c_len);

/* algn_CAML_ancestor_2 (src/algn.c:3742): median of an aligned pair with
 * POY's gap rules; cap >= len + 1. */
int poyhip_ancestor_2(poyhip_ctx *ctx, const uint8_t *a, const uint8_t *b, int len,
                      uint8_t *out, int cap, int *out_len);
/* seq_CAML_median_2_no_gaps / _with_gaps (src/seq.c:275, 288). */
int poyhip_median_2(poyhip_ctx *ctx, const uint8_t *a, const uint8_t *b, int len,
                    uint8_t *out, int cap, int *out_len);
int poyhip_median_2_with_gaps(poyhip_ctx *ctx, const uint8_t *a, const uint8_t *b,
                              int len, uint8_t *out, int cap, int *out_len);
/* algn_CAML_worst_2 (src/algn.c:3108) behind Align.max_cost_2. */
int poyhip_worst_2(poyhip_ctx *ctx, const uint8_t *a, const uint8_t *b, int len,
                   int *cost);
/* The column map of Align.closest (src/sequence.ml:1180-1236; OCaml, no C
 * counterpart) for a table with combinations: over the aligned pair
 * (s1', s2') each column becomes Cost_matrix.Two_D.get_closest(s1'[i],
 * s2'[i]) (src/cost_matrix.ml:1387), then remove_gaps2 (src/sequence.ml:209)
 * drops the gaps and prepends one.  `equal` selects the branch for equal
 * s1 = s2, whose columns after the first lose their gap bit first.  cap >=
 * len + 1; POYHIP_EINVAL where the reference fails (a 0 symbol). */
int poyhip_closest_2(poyhip_ctx *ctx, const uint8_t *a, const uint8_t *b, int len,
                     int equal, uint8_t *out, int cap, int *out_len);

/* Many column maps in one launch (no reference counterpart): for job k,
 * median_2, median_2_with_gaps or the closest map of the aligned pair at
 * d_seq + a and d_seq + b (len bytes each) into d_seq + out (cap >= len + 1
 * bytes, the result right-aligned); d_out_len[k] and d_status[k] (0, or
 * POYHIP_EINVAL for a failing map or a job outside the blob).  Inputs and
 * outputs share the blob; jobs must not write what another reads. */
#define POYHIP_POST_MEDIAN_2 1           /* seq_CAML_median_2_no_gaps   */
#define POYHIP_POST_MEDIAN_2_WITH_GAPS 2 /* seq_CAML_median_2_with_gaps */
#define POYHIP_POST_CLOSEST 4            /* poyhip_closest_2, equal = 0  */
#define POYHIP_POST_CLOSEST_EQ 5         /* poyhip_closest_2, equal = 1  */
typedef struct {
  int64_t a, b, out;  /* d_seq offsets */
  int32_t len, cap;
} poyhip_post_job;
int poyhip_pairs_post_batch(poyhip_ctx *ctx, int op, int n, const poyhip_post_job *d_jobs,
                            uint8_t *d_seq, int64_t seq_bytes, int *d_out_len, int *d_status,
                            int sync);

/* ---- three sequences ----------------------------------------------------- */
/* Cost and median entries of POY's struct cm_3d (src/cm.h:253-280) for 5-bit
 * codes, index ((a << 5) + b) << 5 + c, as Cost_matrix.Three_D pushes them
 * through cm_CAML_create_3d / cm_CAML_set_cost_3d / cm_CAML_set_median_3d
 * (src/cm.c:2433, 2333, 2369). */
typedef struct {
  int cost[32768];
  uint8_t median[32768];
} poyhip_cm3;
int poyhip_set_cm3(poyhip_ctx *ctx, const poyhip_cm3 *cm3);

/* algn_CAML_simple_3 (src/algn.c:3194) = Align.cost_3: algn_nw_3d's cube
 * cost (algn_fill_cube, src/algn.c:2593, with its row-pointer arithmetic).
 * `uk` is accepted and unused, as algn_fill_cube ignores it.  Leaves the
 * directions in the context for poyhip_backtrack_3d. */
int poyhip_cost_3(poyhip_ctx *ctx, const uint8_t *s1, int l1, const uint8_t *s2, int l2,
                  const uint8_t *s3, int l3, int uk, int *cost);
/* algn_CAML_backtrack_3d (src/algn.c:3477) = Align.extract_edited_3: the
 * three aligned sequences of the last poyhip_cost_3 on the same inputs
 * (the walk stops at the origin cell, as backtrack_3d does);
 * cap >= l1 + l2 + l3. */
int poyhip_backtrack_3d(poyhip_ctx *ctx, const uint8_t *s1, int l1, const uint8_t *s2,
                        int l2, const uint8_t *s3, int l3, uint8_t *r1, uint8_t *r2,
                        uint8_t *r3, int cap, int *len);
/* algn_CAML_align_3d (src/algn.c:3521) = Align.c_align_3: fill + backtrack. */
int poyhip_align_3d(poyhip_ctx *ctx, const uint8_t *s1, int l1, const uint8_t *s2, int l2,
                    const uint8_t *s3, int l3, int uk, uint8_t *r1, uint8_t *r2,
                    uint8_t *r3, int cap, int *len, int *cost);
/* algn_CAML_median_3 (src/algn.c:3680) = Align.c_median_3: like
 * algn_get_median_3d, every position receives the median of the last
 * column; out_len = len. */
int poyhip_median_3(poyhip_ctx *ctx, const uint8_t *a, const uint8_t *b, const uint8_t *c,
                    int len, uint8_t *out, int cap, int *out_len);

/* ---- batches (no reference counterpart) --------------------------------- */
/* Operation of a batch. */
#define POYHIP_OP_ALIGN_AFFINE 1 /* align_affine_3 per pair (shorter first  */
                                 /* and swap flag handled inside)            */
#define POYHIP_OP_COST_AFFINE 2  /* cost_2_affine per pair                   */
#define POYHIP_OP_ALIGN_2D 3     /* Align.align_2 (linear) + ancestor_2 +    */
                                 /* median_2_with_gaps per pair              */
#define POYHIP_OP_COST_2D 4      /* Align.cost_2 (linear)                    */
#define POYHIP_OP_NEWKK_ALIGN 7  /* NewkkAlign.align_2 (affine) + ancestor_2 */
                                 /* + median_2_with_gaps (DOS.median's       */
                                 /* use_ukk branch, src/seqCS.ml:1044-1052)  */
#define POYHIP_OP_NEWKK_COST 8   /* NewkkAlign.cost_2 (affine)               */

typedef struct poyhip_batch poyhip_batch;

/* Plans a batch of n pairs whose lengths (leading gap included) are
 * len1[k], len2[k].  Planning sorts pairs largest-first for the persistent
 * kernels, sizes every per-pair region and splits the batch into chunks whose
 * direction stores fit `dm_budget_bytes` of device memory (0 = automatic).
 * deltaw: NULL for Align.cost_2's ?deltaw=None (what align_2 uses), else the
 * per-pair ~deltaw of DOS.distance. */
int poyhip_batch_create(poyhip_ctx *ctx, int op, int n, const int *len1,
                        const int *len2, const int *deltaw, size_t dm_budget_bytes,
                        poyhip_batch **out);
/* Waits for the context's stream and hands the batch's device blocks to
 * the context, which gives them to its next batches (up to 4 GiB in all;
 * poyhip_flush and poyhip_destroy free them).  Destroy every batch before
 * its context. */
int poyhip_batch_destroy(poyhip_batch *b);
/* Bytes of the output block of the batch and the offset of pair k in it.
 * Align ops write four consecutive regions of cap_k bytes per pair (affine:
 * resi, resj, median, medianwg with cap_k = l1+l2+2; linear and
 * diagonal-store: s1', s2', ancestor median, median with gaps with cap_k =
 * l1+l2+1).  Each region holds its sequence right-aligned: it starts at
 * cap_k - length. */
size_t poyhip_batch_out_bytes(const poyhip_batch *b);
int64_t poyhip_batch_out_offset(const poyhip_batch *b, int k);
int poyhip_batch_out_cap(const poyhip_batch *b, int k);
/* Both of the above for every pair of the batch (n entries each). */
int poyhip_batch_out_layout(const poyhip_batch *b, int64_t *off, int *cap);
/* Runs the batch on device-resident inputs.  d_blob holds all sequences;
 * d_off1/d_off2 give the offsets of the caller's first and second sequence
 * of each pair.  Results: d_cost[n]; for align ops d_out (out_bytes) and
 * d_outlen[4n] (region lengths in the order above, caller order for
 * the two aligned sequences).  Asynchronous on the context stream unless
 * `sync` is non-zero. */
int poyhip_batch_run(poyhip_batch *b, const uint8_t *d_blob, const int64_t *d_off1,
                     const int64_t *d_off2, int *d_cost, uint8_t *d_out,
                     int *d_outlen, int sync);
/* Per-pair statistics of the last run (host arrays of n; any may be NULL):
 * cells evaluated over all doubling rounds, lane-steps the fill issued for
 * them (cells plus idle lanes), rounds, final band half-width. */
int poyhip_batch_stats(poyhip_batch *b, int64_t *cells, int64_t *slots, int *rounds,
                       int *newk);
/* Captures the batch's launches into a HIP graph and replays it on later
 * runs with the same device pointers. */
int poyhip_batch_use_graph(poyhip_batch *b, int enable);

/* Many union merges in one launch (no reference counterpart), e.g. the
 * unions of one tree level: job k reads and writes the device blobs d_seq
 * (bytes) and d_ints (int32) at the offsets it names; d_c_len[k] receives
 * the new union's length and d_status[k] 0, or POYHIP_EINVAL where the
 * reference fails or a range of the job lies outside the blobs.  Jobs of
 * one launch must not write what another reads.  One lane walks one merge
 * and a few consecutive jobs share a wavefront, so a launch runs as long as
 * its longest merge; jobs ordered by aligned length (longest first) keep
 * similar walks together.  Asynchronous on the context stream unless `sync`
 * is non-zero. */
typedef struct {
  int64_t a, b, median;       /* aligned children and median (d_seq offsets) */
  int64_t ua_s, ub_s, c_s;    /* union sequence buffers (d_seq offsets) */
  int64_t ua_off, ub_off;     /* children's offset arrays (d_ints offsets) */
  int64_t c_off, c_c1, c_c2;  /* new union's arrays (d_ints offsets) */
  int32_t lena, ua_len, ua_cap, ub_len, ub_cap, c_cap;
} poyhip_union_job;
int poyhip_union_merge_batch(poyhip_ctx *ctx, int n, const poyhip_union_job *d_jobs,
                             uint8_t *d_seq, int64_t seq_bytes, int32_t *d_ints,
                             int64_t n_ints, int *d_c_len, int *d_status, int sync);

/* Tuning knobs of a context (no reference counterpart: the settings the
 * autotuner, poy5_amd/autotune.py, chooses per length bucket).  A value
 * applies to batches created and runs started after it is set.
 *   WAVES_PER_CU      resident wavefronts per CU of a pair batch's
 *                     persistent fill, 1..32 (default 16)
 *   AFFINE_ROTATING   1: gap-free affine pairs use the rotating stripe
 *                     schedule within its limits (default); 0: the plain
 *                     stripe loop
 *   LINEAR_WORKGROUP  when every cost entry fits 16 bits, 2: the linear
 *                     fill runs a workgroup of wavefronts per pair unless
 *                     the batch has at least 16 pairs per CU, which run one
 *                     wavefront each (default); 1: always a workgroup per
 *                     pair; 0: always one wavefront per pair
 *   CUBE_WORKGROUP    1: the 3D fill runs a workgroup per triple when its
 *                     keys fit (default); 0: one wavefront per triple
 * Results do not depend on the knobs.  get returns the value or
 * POYHIP_EINVAL. */
enum {
  POYHIP_TUNE_WAVES_PER_CU = 1,
  POYHIP_TUNE_AFFINE_ROTATING = 2,
  POYHIP_TUNE_LINEAR_WORKGROUP = 3,
  POYHIP_TUNE_CUBE_WORKGROUP = 4
};
int poyhip_set_tuning(poyhip_ctx *c, int knob, int value);
int poyhip_get_tuning(poyhip_ctx *c, int knob);
/* Timing of the last non-graph run with sync set, from HIP events recorded
 * on the context stream: last_fill_ms is the summed duration of the fill
 * stage over all chunks (for ALIGN_AFFINE a chunk's fill stage is the
 * gap-free kernel followed by the general kernel), last_total_ms the whole
 * run.  -1 after a graph replay. */
float poyhip_batch_last_fill_ms(const poyhip_batch *b);
float poyhip_batch_last_total_ms(const poyhip_batch *b);
/* Number of chunks the batch is split into (one fill stage each). */
int poyhip_batch_chunks(const poyhip_batch *b);

/* Batches of triples (no reference counterpart). */
#define POYHIP_OP_ALIGN_3D 5 /* algn_nw_3d + backtrack_3d + median_3 per triple */
#define POYHIP_OP_COST_3D 6  /* algn_nw_3d cost per triple                      */
typedef struct poyhip_batch3 poyhip_batch3;
/* Plans n triples (lengths with the leading gap); chunks keep each chunk's
 * direction stores within dm_budget_bytes (0 = automatic). */
int poyhip_batch3_create(poyhip_ctx *ctx, int op, int n, const int *len1, const int *len2,
                         const int *len3, size_t dm_budget_bytes, poyhip_batch3 **out);
int poyhip_batch3_destroy(poyhip_batch3 *b);
/* ALIGN_3D writes four regions of cap_k = l1+l2+l3 bytes per triple (r1,
 * r2, r3, median_3), each right-aligned; d_outlen[4k..4k+3] = lengths
 * (-1 when the walk left the cube). */
size_t poyhip_batch3_out_bytes(const poyhip_batch3 *b);
int64_t poyhip_batch3_out_offset(const poyhip_batch3 *b, int k);
int poyhip_batch3_out_cap(const poyhip_batch3 *b, int k);
int poyhip_batch3_run(poyhip_batch3 *b, const uint8_t *d_blob, const int64_t *d_off1,
                      const int64_t *d_off2, const int64_t *d_off3, int *d_cost,
                      uint8_t *d_out, int *d_outlen, int sync);
int poyhip_batch3_chunks(const poyhip_batch3 *b);
/* Summed fill duration of the last run with sync set (HIP events). */
float poyhip_batch3_last_fill_ms(const poyhip_batch3 *b);

/* ---- serving (no reference counterpart) --------------------------------- */
/* A continuous-batching server over one context: requests arrive one at a
 * time (a tree search's rejoin and join alignments), a worker thread admits
 * them into S slots (one HIP stream and preallocated buffers each) while
 * their direction stores fit a slot's budget, and launches each slot's
 * requests as one batch; slots overlap on the device.  Sequences live in a
 * device-resident pool; an align_2d request can keep its ancestor median in
 * the pool (written by the GPU) for later requests.  Install the cost model
 * on the context before creating the server and keep it while it runs. */
#define POYHIP_REQ_KEEP_MEDIAN 1
typedef struct poyhip_server poyhip_server;
typedef struct {
  int n_slots;          /* concurrent batches (default 3)                 */
  int max_batch_pairs;  /* requests per batch (default 4096)              */
  int max_len;          /* longest sequence, leading gap included (16384) */
  size_t slot_dm_bytes; /* direction-store budget per slot (4 GiB)        */
  size_t pool_bytes;    /* sequence pool (256 MiB)                        */
} poyhip_server_opts;
int poyhip_server_create(poyhip_ctx *ctx, const poyhip_server_opts *opts,
                         poyhip_server **out);
int poyhip_server_destroy(poyhip_server *s);
/* Appends a host sequence to the pool (uploaded with the next batch). */
int poyhip_server_put_seq(poyhip_server *s, const uint8_t *seq, int len, int64_t *id);
/* Copies a pool sequence (e.g. a kept median) back to the host. */
int poyhip_server_get_seq(poyhip_server *s, int64_t id, uint8_t *out, int cap, int *len);
/* Queues one request on pool sequences a and b (caller order), op one of the
 * POYHIP_OP_* pair ops, deltaw < 0 for Align.cost_2's ?deltaw=None.  With
 * POYHIP_REQ_KEEP_MEDIAN (align_2d) *median_id names the pool entry that
 * will hold the ancestor median; requests naming it wait until it exists. */
int poyhip_server_submit(poyhip_server *s, int op, int64_t a, int64_t b, int deltaw,
                         int flags, int64_t *req, int64_t *median_id);
/* Waits for a request (timeout_ms < 0: forever; POYHIP_EBUSY on timeout). */
int poyhip_server_wait(poyhip_server *s, int64_t req, int timeout_ms);
/* Result of a finished request: cost, aligned sequences (align ops, caller
 * order, cap >= their length), latency from submission to completion;
 * release != 0 forgets the request. */
int poyhip_server_result(poyhip_server *s, int64_t req, int *cost, uint8_t *r1,
                         uint8_t *r2, int cap, int *len, double *latency_us, int release);
int poyhip_server_stats(poyhip_server *s, int64_t *batches, int64_t *pairs,
                        double *cells_full, int *max_batch, int *max_inflight);
/* Open-loop load generator: submits request k at arrival_us[k] after the
 * call from a native thread, then waits for all and reports per-request
 * costs and latencies (any output array may be NULL) and the wall time. */
int poyhip_server_replay(poyhip_server *s, int n, const double *arrival_us, const int *ops,
                         const int64_t *a, const int64_t *b, const int *deltaw,
                         const int *flags, int *costs, double *latency_us,
                         int64_t *median_ids, double *wall_us);

/* ---- synchronisation helpers -------------------------------------------- */
int poyhip_sync(poyhip_ctx *ctx);
/* Opaque hipStream_t of the context (for callers that enqueue their own
 * copies on it). */
void *poyhip_stream(poyhip_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* POYHIP_H */

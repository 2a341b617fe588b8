/*
 * oracle.h -- CPU restatement of POY 5.1 direct-optimization (DO) pairwise
 * alignment, used ONLY as test infrastructure (parity checker, golden-vector
 * writer and the CPU baseline leg of bench.py).  Nothing in the product path
 * (poy5_amd/, include/) links or calls this code.
 *
 * Every function follows the reference C/OCaml functions named in its
 * comment (paths relative to the POY source tree).  The restatement is
 * memory-faithful where the reference reads buffers that persist between
 * Ukkonen doubling rounds (see orc_align_affine_3).
 *
 * Sequences are arrays of 5-bit nucleotide bitsets (A=1 C=2 G=4 T=8 gap=16),
 * leading gap included, stored left to right (index 0 = the leading gap).
 */
#ifndef POY_ORACLE_H
#define POY_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_HIGH_NUM 1000000 /* src/algn.c:37 */
#define ORC_GAP 16
#define ORC_LCM 5
#define ORC_TBL 2048 /* 2 * (1<<lcm)^2, src/cm.c:609-616 */

/* 2D flags, src/matrices.h:22-27 */
#define ORC_ALIGN 1
#define ORC_INSERT 2
#define ORC_DELETE 4

/* affine flags, src/algn.c:1185-1199 */
#define ORC_ALIGN_TO_ALIGN 1
#define ORC_ALIGN_TO_VERTICAL 2
#define ORC_ALIGN_TO_HORIZONTAL 4
#define ORC_ALIGN_TO_DIAGONAL 8
#define ORC_BEGIN_BLOCK 16
#define ORC_END_BLOCK 32
#define ORC_BEGIN_VERTICAL 64
#define ORC_END_VERTICAL 128
#define ORC_BEGIN_HORIZONTAL 256
#define ORC_END_HORIZONTAL 512
#define ORC_DO_ALIGN 1024
#define ORC_DO_VERTICAL 2048
#define ORC_DO_HORIZONTAL 4096
#define ORC_HORIZONTAL_EQ_VERTICAL 8192
#define ORC_DO_DIAGONAL 16384

/* 3D direction codes, src/matrices.h:34-40 */
#define ORC_P1 1
#define ORC_P2 2
#define ORC_P3 4
#define ORC_S1 8
#define ORC_S2 16
#define ORC_S3 32
#define ORC_SS 64

/* The fields of struct cm (src/cm.h:33-76) used on the DNA path. */
typedef struct {
    int lcm;          /* 5 for DNA with combinations */
    int gap;          /* 16 */
    int a_sz;         /* 31 */
    int combinations; /* 1 */
    int cost_model;   /* 0 linear, 1 affine (cost_model_type) */
    int gap_open;     /* affine gap opening */
    int cost[ORC_TBL];
    int worst[ORC_TBL];
    int prepend[64];
    int tail[64];
    uint8_t median[ORC_TBL];
} orc_cm;

/* struct cm_3d cost/median slice addressed by 5-bit codes (src/cm.h:253-280) */
typedef struct {
    int lcm, gap, a_sz;
    int cost[32 * 32 * 32];
    uint8_t median[32 * 32 * 32];
} orc_cm3;

typedef struct {
    int64_t cells;     /* cells evaluated, summed over all doubling rounds */
    int rounds;        /* fill passes (1 for the full plane) */
    int full_plane;    /* 1 when algn_fill_plane (no band) was used */
    int final_T;       /* last Ukkonen threshold */
    int final_k;       /* last newk */
    int gap_num;       /* last gap count used by the stop test */
    int min_gap_excess; /* min over band rounds of gap count - (L2 - L1);
                           INT_MAX when no band round ran */
} orc_stats;

/* ---- tables (src/cost_matrix.ml, src/cm.c) ---- */
/* of_transformations_and_gaps (cost_matrix.ml:1337) -> fill_cost_matrix
 * (1140) with the linear all-combinations fill (862) and default
 * prepend/tail (994); when gap_open >= 0 additionally set_cost_model
 * (Affine gap_open) (1003) which refills bitwise (721). */
void orc_cm_dna(orc_cm *cm, int trans, int gaps, int gap_open);
/* Three_D.of_two_dim_comb (cost_matrix.ml:1605) */
void orc_cm3_from_2d(orc_cm3 *c3, const orc_cm *cm);
/* cm_get_min_non0_cost (cm.c:1064) */
int orc_min_non0_cost(const orc_cm *cm);

/* ---- non-affine 2D (src/algn.c) ---- */
/* algn_nw (2972) -> algn_nw_limit -> algn_fill_plane_2 (1134).
 * L1 <= L2.  dm: u16 L1*L2 row-major (row i = position i of s1), filled
 * like the reference's direction matrix.  Returns the cost. */
int orc_nw(const uint8_t *s1, int L1, const uint8_t *s2, int L2,
           const orc_cm *cm, int deltawh, uint16_t *dm, orc_stats *st);
/* algn_nw_limit (2954) as algn_CAML_limit_2 (3161) calls it: the limit
 * lengths enter only the full-plane-vs-band decision's height. */
int orc_nw_limit(const uint8_t *s1, int L1, const uint8_t *s2, int L2,
                 const orc_cm *cm, int deltawh, int len_s1, int len_s2, uint16_t *dm,
                 orc_stats *st);
/* backtrace_2d (3278).  r1/r2 receive the aligned sequences left to right;
 * returns the aligned length. */
int orc_backtrace_2d(const uint8_t *s1, int L1, const uint8_t *s2, int L2,
                     const uint16_t *dm, int swaped, uint8_t *r1, uint8_t *r2);
/* backtrace_2d with the limits of algn_CAML_backtrace_2d_limit (3449). */
int orc_backtrace_2d_limit(const uint8_t *s1, int L1, const uint8_t *s2, int L2,
                           const uint16_t *dm, int st_s1, int st_s2, int algn_s1, int algn_s2,
                           int swaped, uint8_t *r1, uint8_t *r2);
/* backtrace_2d_gaps (978) */
int orc_backtrace_2d_gaps(const uint16_t *dm, int L1, int L2);
/* algn_ancestor_2 (3603); returns median length */
int orc_ancestor_2(const uint8_t *a, const uint8_t *b, int len,
                   const orc_cm *cm, uint8_t *out);
/* seq_get_median_2d_no_gaps / _with_gaps (seq.c:259/241) */
int orc_median_2_no_gaps(const uint8_t *a, const uint8_t *b, int len,
                         const orc_cm *cm, uint8_t *out);
int orc_median_2_with_gaps(const uint8_t *a, const uint8_t *b, int len,
                           const orc_cm *cm, uint8_t *out);
/* algn_calculate_from_2_aligned (3003) on the worst / cost tables */
int orc_worst_2(const uint8_t *a, const uint8_t *b, int len, const orc_cm *cm);
int orc_verify_2(const uint8_t *a, const uint8_t *b, int len, const orc_cm *cm);

/* ---- affine 2D ---- */
/* algn_CAML_align_affine_3 (algn.c:2359): banded affine fill with
 * T-doubling + backtrace_aff (1715).  Li <= Lj.  Outputs left to right,
 * each buffer must hold Li+Lj+2 bytes.  Returns cost. */
int orc_align_affine_3(const uint8_t *si, int Li, const uint8_t *sj, int Lj,
                       const orc_cm *cm, int swaped,
                       uint8_t *median, int *mlen, uint8_t *medianwg, int *mwglen,
                       uint8_t *resi, uint8_t *resj, int *alen, orc_stats *st);
/* One band round of the above at a given p (band half-width min(p, Li-2)),
 * as the first round algn_newkk_increaseT_aff (2311) runs it, then
 * backtrace_aff: the fill of algn_newkk_test_aff (2186) on freshly
 * initialised rows.  Test infrastructure for band-restricted pins. */
int orc_align_affine_round(const uint8_t *si, int Li, const uint8_t *sj, int Lj,
                           const orc_cm *cm, int swaped, int p,
                           uint8_t *median, int *mlen, uint8_t *medianwg, int *mwglen,
                           uint8_t *resi, uint8_t *resj, int *alen, orc_stats *st);
/* algn_CAML_cost_affine_3 (algn.c:2457): full affine DP without backtrace,
 * either order. */
int orc_cost_affine_3(const uint8_t *si, int Li, const uint8_t *sj, int Lj,
                      const orc_cm *cm);

/* ---- OCaml glue (src/sequence.ml, src/seq.c) ---- */
/* seq_CAML_count (seq.c:644) with startNO=0 == Align.count_gaps */
int orc_count_gaps(const uint8_t *s, int L);
/* Align.cost_2 (sequence.ml:868-936) deltawh: has_deltaw=0 for the
 * ?deltaw=None call (align_2), 1 for DOS.distance's ~deltaw. */
int orc_deltawh(const uint8_t *s1, int L1, const uint8_t *s2, int L2,
                int has_deltaw, int deltaw);
/* Align.cost_2 (linear): shorter first, deltawh as above; fills dm for
 * the shorter-first orientation. */
int orc_cost_2(const uint8_t *s1, int L1, const uint8_t *s2, int L2,
               const orc_cm *cm, int has_deltaw, int deltaw, uint16_t *dm,
               orc_stats *st);
/* Align.align_2 ~first_gap:true (sequence.ml:1061) for the linear model:
 * cost_2 + create_edited_2.  r1/r2 are aligned s1/s2 (caller order). */
int orc_align_2(const uint8_t *s1, int L1, const uint8_t *s2, int L2,
                const orc_cm *cm, uint8_t *r1, uint8_t *r2, int *alen,
                orc_stats *st);

/* ---- 3D (algn_nw_3d, algn.c:2977) ---- */
int orc_nw_3d(const uint8_t *s1, int L1, const uint8_t *s2, int L2,
              const uint8_t *s3, int L3, const orc_cm3 *c3, uint16_t *dm);
/* backtrack_3d (algn.c:3345); returns aligned length, -1 if the walk reads
 * outside the sequences (the reference would read out of bounds). */
int orc_backtrack_3d(const uint8_t *s1, int L1, const uint8_t *s2, int L2,
                     const uint8_t *s3, int L3, const uint16_t *dm, int gap,
                     uint8_t *r1, uint8_t *r2, uint8_t *r3);
/* algn_get_median_3d (algn.c:3641) */
int orc_median_3(const uint8_t *a, const uint8_t *b, const uint8_t *c, int len,
                 const orc_cm3 *c3, uint8_t *out);

/* ---- batch helper used by the CPU baseline ---- */
/* Runs orc_align_affine_3 over n pairs (shorter-first handled here);
 * returns the sum of costs. */
int64_t orc_batch_align_affine(int n, const uint8_t *blob, const int64_t *off1,
                               const int *len1, const int64_t *off2,
                               const int *len2, const orc_cm *cm,
                               int *costs, int64_t *cells_out);

/* Align.align_2 + ancestor_2 + median_2_with_gaps per pair (caller order);
 * returns the sum of costs. */
int64_t orc_batch_align_2(int n, const uint8_t *blob, const int64_t *off1, const int *len1,
                          const int64_t *off2, const int *len2, const orc_cm *cm, int *costs,
                          int64_t *cells_out);
/* algn_nw_3d + backtrack_3d + median_3 per triple; returns the sum of costs. */
int64_t orc_batch_align_3d(int n, const uint8_t *blob, const int64_t *off1, const int *len1,
                           const int64_t *off2, const int *len2, const int64_t *off3,
                           const int *len3, const orc_cm3 *c3, int *costs);

/* ---- diagonal-store Ukkonen aligner, affine (src/newkkonen.c; newkk.c) ---- */
/* direction flags, src/newkkonen.c:41-61 */
#define NK_START 0
#define NK_ALIGN 1
#define NK_DO_DELETE 16
#define NK_DO_INSERT 32
#define NK_END_INSERT 64
#define NK_END_DELETE 128
#define NK_END_DIAG 256
#define NK_DO_DIAG 512
#define NK_INS_EQ_DEL 1024

typedef struct {
    int64_t cells;  /* cell updates over all rounds */
    int rounds;
    int final_T;
    int final_k;
    int gap_num;    /* max of the two gap counts of the last cell, last round */
} orc_nk_stats;

/* newkkonen_CAML_algn_affine (newkkonen.c:1473) and, when r1 is non-NULL,
 * newkkonen_CAML_backtrace_affine (1847) on the same matrix: cost of s1
 * (lenX) against s2 (lenY >= lenX); r1/r2 (capacity lenX + lenY) receive the
 * aligned pair left to right, *rlen its length (-1 where the reference
 * fails).  The reference's incremental refill, band-only storage. */
int orc_newkk_align(const uint8_t *s1, int lenX, const uint8_t *s2, int lenY, const orc_cm *cm,
                    int swaped, uint8_t *r1, uint8_t *r2, int *rlen, orc_nk_stats *st);
/* The same with every band cell recomputed in every round. */
int orc_newkk_full(const uint8_t *s1, int lenX, const uint8_t *s2, int lenY, const orc_cm *cm,
                   int swaped, uint8_t *r1, uint8_t *r2, int *rlen, orc_nk_stats *st);
/* NewkkAlign.align_2 per pair of a blob (CPU baseline); returns the sum of
 * the costs. */
int64_t orc_batch_newkk(int n, const uint8_t *blob, const int64_t *off1, const int *len1,
                        const int64_t *off2, const int *len2, const orc_cm *cm, int *costs,
                        int64_t *cells_out);
/* ---- unions (src/union.c; union.c) ---- */
/* union_merge (union.c:170) behind union_CAML_make (318): merges the unions
 * ua and ub of two aligned children a, b (length lena, with their aligned
 * median) into a new union c of capacity c_cap (OCaml's Unions.union uses
 * len(ua) + len(ub) + 2).  A union is a sequence right-aligned in a buffer
 * of `cap` symbols with 32-bit offset arrays of `cap` entries; ua/ub need
 * only their offsets.  c's sequence, offsets and child offsets are written
 * right-aligned; entries the merge does not write are 0.  Returns 0, or -1
 * where the reference fails. */
int orc_union_merge(const uint8_t *a, const uint8_t *b, const uint8_t *median, int lena,
                    const uint8_t *ua_s, int ua_len, int ua_cap, const int32_t *ua_off,
                    const uint8_t *ub_s, int ub_len, int ub_cap, const int32_t *ub_off,
                    uint8_t *c_s, int c_cap, int32_t *c_off, int32_t *c_ca, int32_t *c_cb,
                    int *c_len);
/* ---- Align.closest's column map (src/sequence.ml:1180; closest.c) ---- */
/* Cost_matrix.Two_D.get_closest (src/cost_matrix.ml:1387); -1 on failure */
int orc_get_closest(const orc_cm *cm, int a, int b);
/* the column map over an aligned pair (or equal sequences, equal != 0) and
 * remove_gaps2; out holds len + 1 bytes; returns the length or -1 */
int orc_closest_map(const uint8_t *s1p, const uint8_t *s2p, int len, int equal, const orc_cm *cm,
                    uint8_t *out);

/* cost, P, Q, ED, CD of every cell of the full matrix (k = lenX - 1) */
void orc_newkk_matrix(const uint8_t *s1, int lenX, const uint8_t *s2, int lenY, const orc_cm *cm,
                      int *out);

#ifdef __cplusplus
}
#endif
#endif

/*
 * oracle.c -- CPU restatement of POY 5.1 DO pairwise alignment.
 *
 * TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker.  The product (poy5_amd/,
 * include/poyhip.h) never links or calls it.
 *
 * Each routine names the reference function (POY source tree path:line) it
 * restates.  The restatement keeps the reference's buffer model where that
 * model is observable: the banded affine fill keeps its two rolling rows and
 * u16 gap-count rows alive across Ukkonen doubling rounds exactly as
 * algn_newkk_test_aff does (src/algn.c:2186-2306), because row 0's
 * extend_block_diagonal and extend_horizontal[0] are not re-initialised
 * between rounds and later rounds read them.
 */
#include "oracle.h"

#include <limits.h>
#include <stdlib.h>
#include <string.h>

#define GAP ORC_GAP
#define HIGH ORC_HIGH_NUM

static inline int cidx(int a, int b) { return (a << ORC_LCM) + b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imax3(int a, int b, int c) { return imax(imax(a, b), c); }
static inline int popc(int x) { return __builtin_popcount((unsigned)x); }

/* ------------------------------------------------------------------------ */
/* Tables                                                                    */
/* ------------------------------------------------------------------------ */

/* of_transformations_and_gaps (src/cost_matrix.ml:1337-1348): element e of
 * the 5-letter alphabet (A C G T gap) is stored at code 1<<e
 * (store_input_list_in_cost_matrix_all_combinations, cost_matrix.ml:329). */
static void load_5x5(orc_cm *cm, int trans, int gaps) {
    for (int x = 0; x < 5; x++)
        for (int y = 0; y < 5; y++) {
            int v = (x == y) ? 0 : ((x == 4 || y == 4) ? gaps : trans);
            cm->cost[cidx(1 << x, 1 << y)] = v;
        }
}

/* fill_best_cost_and_median_for_all_combinations (cost_matrix.ml:862-897)
 * with test_combinations (479-507).  The model is linear when the tables are
 * created (use_cost_model = Linnear, cost_matrix.ml:99), so goa = 0. */
static void fill_all_combinations_linear(orc_cm *cm) {
    for (int i = 1; i <= 31; i++) {
        for (int j = 1; j <= 31; j++) {
            int best = 0, c = INT_MAX, w = 0;
            for (int a = 1; a <= 16; a <<= 1) {
                if (!(i & a)) continue;
                for (int b = 1; b <= 16; b <<= 1) {
                    if (!(j & b)) continue;
                    for (int v = 1; v <= 16; v <<= 1) {
                        int tc = cm->cost[cidx(a, v)] + cm->cost[cidx(v, b)];
                        if (tc < c) { c = tc; best = v; }
                        else if (tc == c) best |= v;
                    }
                    int cab = cm->cost[cidx(a, b)];
                    if (cab > w) w = cab;
                }
            }
            if (popc(i) == 1 && popc(j) == 1) {
                cm->median[cidx(i, j)] = (uint8_t)(i | j);
            } else {
                cm->cost[cidx(i, j)] = c;
                cm->median[cidx(i, j)] = (uint8_t)best; /* cleanup = id (linear) */
            }
            cm->worst[cidx(i, j)] = w;
        }
    }
}

/* fill_best_cost_and_median_for_all_combinations_bitwise
 * (cost_matrix.ml:721-804), create_original = false. */
static void fill_all_combinations_bitwise(orc_cm *cm) {
    int old[ORC_TBL];
    memcpy(old, cm->cost, sizeof old);
#define OC(a, b) old[cidx((a), (b))]
    for (int i = 1; i <= 31; i++) {
        for (int j = 1; j <= 31; j++) {
            int median, best, worst;
            if (popc(i) == 1 && popc(j) == 1) {
                if (i == j) {
                    int cii = OC(i, i) * 2;
                    median = i; best = cii; worst = cii;
                } else {
                    int cost1 = OC(i, j) + OC(j, j);
                    int cost2 = OC(i, i) + OC(i, j);
                    int cij, mij;
                    if (cost1 == cost2) { cij = cost1; mij = i | j; }
                    else if (cost1 > cost2) { cij = cost2; mij = i; }
                    else { cij = cost1; mij = j; }
                    median = mij; best = cij; worst = cij;
                }
            } else {
                /* find_best_combination: process (lj x li) then (li x lj) */
                int bc = INT_MAX, bm = 0, wc = 0;
                for (int pass = 0; pass < 2; pass++) {
                    int X = pass == 0 ? j : i, Y = pass == 0 ? i : j;
                    for (int x = 1; x <= 16; x <<= 1) {
                        if (!(X & x)) continue;
                        for (int y = 1; y <= 16; y <<= 1) {
                            if (!(Y & y)) continue;
                            int cost1 = OC(x, x) + OC(x, y);
                            int cost2 = OC(x, y) + OC(y, y);
                            int cij, mij;
                            if (cost1 == cost2) { cij = cost1; mij = x | y; }
                            else if (cost1 > cost2) { cij = cost2; mij = y; }
                            else { cij = cost1; mij = x; }
                            if (cij < bc) { bc = cij; bm = mij; }
                            else if (cij == bc) bm |= mij;
                            if (cij > wc) wc = cij;
                        }
                    }
                }
                median = bm; best = bc; worst = wc;
            }
            cm->median[cidx(i, j)] = (uint8_t)median;
            cm->cost[cidx(i, j)] = best;
            cm->worst[cidx(i, j)] = worst;
        }
    }
#undef OC
}

void orc_cm_dna(orc_cm *cm, int trans, int gaps, int gap_open) {
    memset(cm, 0, sizeof *cm);
    cm->lcm = ORC_LCM;
    cm->gap = GAP;
    cm->a_sz = 31;
    cm->combinations = 1;
    load_5x5(cm, trans, gaps);
    fill_all_combinations_linear(cm);
    /* fill_default_prepend_tail (cost_matrix.ml:994-1001) */
    for (int i = 1; i <= 31; i++) {
        cm->tail[i] = cm->cost[cidx(i, GAP)];
        cm->prepend[i] = cm->cost[cidx(GAP, i)];
    }
    if (gap_open >= 0) {
        /* set_cost_model (Affine go) (cost_matrix.ml:1003-1016): c_set_aff
         * then the bitwise refill; prepend/tail are not refilled. */
        cm->cost_model = 1;
        cm->gap_open = gap_open;
        fill_all_combinations_bitwise(cm);
    }
}

int orc_min_non0_cost(const orc_cm *cm) {
    int m = INT_MAX / 2;
    for (int i = 0; i < ORC_TBL; i++)
        if (cm->cost[i] > 0 && cm->cost[i] < m) m = cm->cost[i];
    return m;
}

/* Three_D.of_two_dim_comb (cost_matrix.ml:1605-1651), metric input: the
 * cheapest single-state median of three sets; ties keep the lowest bit
 * (pick_bit). */
void orc_cm3_from_2d(orc_cm3 *c3, const orc_cm *cm) {
    memset(c3, 0, sizeof *c3);
    c3->lcm = ORC_LCM;
    c3->gap = GAP;
    c3->a_sz = 31;
    for (int i = 1; i <= 31; i++)
        for (int j = 1; j <= 31; j++)
            for (int k = 1; k <= 31; k++) {
                int best = INT_MAX, med = 0;
                for (int l = 0; l < ORC_LCM; l++) {
                    int v = 1 << l;
                    int c = cm->cost[cidx(v, i)] + cm->cost[cidx(v, j)] +
                            cm->cost[cidx(v, k)];
                    if (c < best) { best = c; med = v; }
                    else if (c == best) med |= v;
                }
                int idx = (((i << 5) + j) << 5) + k;
                c3->cost[idx] = best;
                c3->median[idx] = (uint8_t)(med & -med);
            }
}

/* cm_precalc_4algn (src/cm.c:1334-1368): 33 rows of length L2. */
static int *precalc_2d(const orc_cm *cm, const uint8_t *s, int L) {
    int *to = (int *)calloc((size_t)33 * (L > 0 ? L : 1), sizeof(int));
    for (int m = 0; m < L; m++) to[m] = cm->prepend[s[m]];
    for (int a = 1; a <= cm->a_sz; a++) {
        int *row = to + (size_t)a * L;
        if (L > 0) row[0] = cm->tail[a];
        for (int m = 1; m < L; m++) row[m] = cm->cost[cidx(a, s[m])];
    }
    return to;
}

/* ------------------------------------------------------------------------ */
/* Non-affine 2D                                                             */
/* ------------------------------------------------------------------------ */

/* algn_fill_row (src/algn.c:458-533): the flag set is every tied argmin. */
static void fill_row(int *mm, const int *pm, const int *gap_row,
                     const int *alg_row, uint16_t *dm, int c, int st, int end) {
    for (int j = st; j <= end; j++) {
        int del = pm[j] + c;
        int ins = mm[j - 1] + gap_row[j];
        int aln = pm[j - 1] + alg_row[j];
        int m = del < ins ? del : ins;
        if (aln < m) m = aln;
        mm[j] = m;
        dm[j] = (uint16_t)((aln == m ? ORC_ALIGN : 0) | (ins == m ? ORC_INSERT : 0) |
                           (del == m ? ORC_DELETE : 0));
    }
}

/* algn_fill_ukk_right_cell (543): no DELETE candidate. */
static void right_cell(int *mm, const int *pm, const int *gap_row,
                       const int *alg_row, uint16_t *dm, int pos) {
    int ins = mm[pos - 1] + gap_row[pos];
    int aln = pm[pos - 1] + alg_row[pos];
    int m = ins < aln ? ins : aln;
    mm[pos] = m;
    dm[pos] = (uint16_t)((ins == m ? ORC_INSERT : 0) | (aln == m ? ORC_ALIGN : 0));
}

/* algn_fill_ukk_left_right_cell (588) */
static void left_right_cell(int *mm, const int *pm, const int *alg_row,
                            uint16_t *dm, int pos) {
    mm[pos] = pm[pos - 1] + alg_row[pos];
    dm[pos] = ORC_ALIGN;
}

/* algn_fill_ukk_left_cell (614): no INSERT candidate. */
static void left_cell(int *mm, const int *pm, const int *alg_row, uint16_t *dm,
                      int c, int pos) {
    int del = pm[pos] + c;
    int aln = pm[pos - 1] + alg_row[pos];
    int m = del < aln ? del : aln;
    mm[pos] = m;
    dm[pos] = (uint16_t)((del == m ? ORC_DELETE : 0) | (aln == m ? ORC_ALIGN : 0));
}

/* algn_fill_last_column (654) */
static void last_column(int *mm, const int *pm, int tlc, int l, uint16_t *dm) {
    if (l > 0) {
        int cst = tlc + pm[l];
        if (cst < mm[l]) { mm[l] = cst; dm[l] = ORC_DELETE; }
        else if (cst == mm[l]) dm[l] |= ORC_DELETE;
    }
}

/* algn_fill_plane (927-973): the full matrix. mm holds L1*L2 ints. */
static int fill_plane(const uint8_t *s1, int L1, int L2, const int *prec,
                      const orc_cm *cm, int *mm, uint16_t *dm) {
    const int *gap_row = prec + (size_t)GAP * L2;
    const int *first_gap_row = prec;
    mm[0] = 0;
    dm[0] = ORC_ALIGN;
    for (int j = 1; j < L2; j++) {
        mm[j] = mm[j - 1] + first_gap_row[j];
        dm[j] = ORC_INSERT;
    }
    for (int i = 1; i < L1; i++) {
        int *cur = mm + (size_t)i * L2;
        const int *pm = cur - L2;
        uint16_t *d = dm + (size_t)i * L2;
        int tlc = cm->tail[s1[i]];
        int c = cm->cost[cidx(s1[i], GAP)];
        const int *alg_row = prec + (size_t)s1[i] * L2;
        /* algn_fill_full_row (675) */
        cur[0] = c + pm[0];
        d[0] = ORC_DELETE;
        fill_row(cur, pm, gap_row, alg_row, d, c, 1, L2 - 1);
        last_column(cur, pm, tlc, L2 - 1, d);
    }
    return mm[(size_t)(L1 - 1) * L2 + L2 - 1];
}

int orc_backtrace_2d_gaps(const uint16_t *dm, int L1, int L2) {
    /* backtrace_2d_gaps (978-1005): ALIGN, then INSERT, then DELETE. */
    long end = (long)L2 * (L1 - 1) + L2 - 1;
    int nd = 0, ni = 0;
    while (end >= 0) {
        uint16_t f = dm[end];
        if (f & ORC_ALIGN) end -= L2 + 1;
        else if (f & ORC_INSERT) { ni++; end -= 1; }
        else { nd++; end -= L2; }
    }
    return imax(nd, ni);
}

/* One Ukkonen round: algn_newkk_test (src/algn.c:1063-1106) with
 * algn_newkk_fill_a_row (1008) and the four row fillers (746-923). */
static int newkk_test(const uint8_t *s1, int L1, int L2, const int *prec,
                      const orc_cm *cm, int *mm, uint16_t *dm, int p, int *cost,
                      int64_t *cells) {
    int newk = p;
    if (p >= L1) newk = L1 - 1;
    int len_first_row = L2 - L1 + 1 + p;
    if (len_first_row > L2) len_first_row = L2;
    const int *first_gap_row = prec;
    const int *gap_row = prec + (size_t)GAP * L2;
    /* algn_fill_first_row (693) */
    mm[0] = 0;
    dm[0] = ORC_ALIGN;
    for (int j = 1; j < len_first_row; j++) {
        mm[j] = mm[j - 1] + first_gap_row[j];
        dm[j] = ORC_INSERT;
    }
    *cells += len_first_row;
    for (int i = 1; i < L1; i++) {
        int startj, endj, hl, hr;
        if (i - newk > 0) { startj = i - newk; hl = 1; }
        else { startj = 0; hl = 0; }
        if (i + L2 - L1 + newk <= L2 - 1) { endj = i + L2 - L1 + newk; hr = 1; }
        else { endj = L2 - 1; hr = 0; }
        int *cur = mm + (size_t)i * L2;
        const int *pm = cur - L2;
        uint16_t *d = dm + (size_t)i * L2;
        int ch = s1[i];
        int c = cm->cost[cidx(ch, GAP)];
        int tlc = cm->tail[ch];
        const int *alg_row = prec + (size_t)ch * L2;
        int len = endj - startj + 1;
        *cells += len;
        if (hr && hl) {
            if (len == 1) {
                left_right_cell(cur, pm, alg_row, d, startj);
            } else {
                left_cell(cur, pm, alg_row, d, c, startj);
                fill_row(cur, pm, gap_row, alg_row, d, c, startj + 1, startj + len - 2);
                right_cell(cur, pm, gap_row, alg_row, d, startj + len - 1);
            }
        } else if (hr) {
            /* algn_fill_first_cell (720): pm[0] + alg_row[0] */
            cur[0] = pm[0] + alg_row[0];
            d[0] = ORC_DELETE;
            fill_row(cur, pm, gap_row, alg_row, d, c, 1, len - 2);
            right_cell(cur, pm, gap_row, alg_row, d, len - 1);
        } else if (hl) {
            left_cell(cur, pm, alg_row, d, c, startj);
            fill_row(cur, pm, gap_row, alg_row, d, c, startj + 1, startj + len - 1);
            last_column(cur, pm, tlc, startj + len - 1, d);
        } else {
            cur[0] = pm[0] + alg_row[0];
            d[0] = ORC_DELETE;
            fill_row(cur, pm, gap_row, alg_row, d, c, 1, L2 - 1);
            last_column(cur, pm, tlc, L2 - 1, d);
        }
    }
    *cost = mm[(size_t)(L1 - 1) * L2 + L2 - 1];
    return orc_backtrace_2d_gaps(dm, L1, L2);
}

int orc_nw(const uint8_t *s1, int L1, const uint8_t *s2, int L2,
           const orc_cm *cm, int deltawh, uint16_t *dm, orc_stats *st) {
    /* algn_nw (2972): the limits are the whole sequences */
    return orc_nw_limit(s1, L1, s2, L2, cm, deltawh, L1, L2, dm, st);
}

int orc_nw_limit(const uint8_t *s1, int L1, const uint8_t *s2, int L2,
                 const orc_cm *cm, int deltawh, int len_s1, int len_s2, uint16_t *dm,
                 orc_stats *st) {
    orc_stats local;
    if (!st) st = &local;
    memset(st, 0, sizeof *st);
    st->min_gap_excess = INT_MAX;
    int *prec = precalc_2d(cm, s2, L2);
    int *mm = (int *)malloc(sizeof(int) * (size_t)L1 * L2);
    /* algn_nw_limit (2954) -> algn_fill_plane_2 (1134): width 50, height
     * 50 + (len_s1 - len_s2), each widened by deltawh; the whole sequences
     * are filled whatever the limits */
    int width = 50 + deltawh;
    if (width > L2) width = L2;
    int height = 50 + (len_s1 - len_s2) + deltawh;
    if (height > L1) height = L1;
    (void)width;
    int delta = orc_min_non0_cost(cm);
    int T = (L2 - L1 + 1) * delta;
    int full;
    if ((float)L1 >= 1.5 * (float)L2) full = 1;
    else if (2 * height < L1) full = 0;
    else full = (8 >= (L1 - height));
    int cost;
    if (full) {
        cost = fill_plane(s1, L1, L2, prec, cm, mm, dm);
        st->cells = (int64_t)L1 * L2;
        st->rounds = 1;
        st->full_plane = 1;
    } else {
        /* algn_newkk_increaseT (1118-1130) */
        for (;;) {
            int p = (T - (L2 - L1)) / 2;
            int gap_num = newkk_test(s1, L1, L2, prec, cm, mm, dm, p, &cost, &st->cells);
            st->rounds++;
            st->final_T = T;
            st->final_k = p >= L1 ? L1 - 1 : p;
            st->gap_num = gap_num;
            if (gap_num - (L2 - L1) < st->min_gap_excess) st->min_gap_excess = gap_num - (L2 - L1);
            int newp = (2 * T - (L2 - L1)) / 2;
            if ((gap_num + 1) < p || newp - L2 + 1 >= 0) break;
            T *= 2;
        }
    }
    free(mm);
    free(prec);
    return cost;
}

/* Sequences built by prepending (seq_prepend, src/seq.c:155) are assembled
 * right to left in a scratch buffer and copied out left to right. */
typedef struct {
    uint8_t *buf;
    int cap, len;
} rseq;
static void rs_init(rseq *r, int cap) {
    r->buf = (uint8_t *)malloc(cap > 0 ? cap : 1);
    r->cap = cap;
    r->len = 0;
}
static inline void rs_prepend(rseq *r, uint8_t v) {
    r->len++;
    r->buf[r->cap - r->len] = v;
}
static inline uint8_t rs_first(const rseq *r) { return r->buf[r->cap - r->len]; }
static int rs_out(rseq *r, uint8_t *out) {
    if (out) memcpy(out, r->buf + r->cap - r->len, r->len);
    int n = r->len;
    free(r->buf);
    return n;
}

int orc_backtrace_2d(const uint8_t *s1, int L1, const uint8_t *s2, int L2,
                     const uint16_t *dm, int swaped, uint8_t *r1, uint8_t *r2) {
    /* algn_CAML_backtrace_2d (3424): no limits */
    return orc_backtrace_2d_limit(s1, L1, s2, L2, dm, 0, 0, L1, L2, swaped, r1, r2);
}

int orc_backtrace_2d_limit(const uint8_t *s1, int L1, const uint8_t *s2, int L2,
                           const uint16_t *dm, int st_s1, int st_s2, int algn_s1, int algn_s2,
                           int swaped, uint8_t *r1, uint8_t *r2) {
    /* backtrace_2d (3278-3327) + follow_* (79-124).  The walk's pointer
     * starts at beg + L2 (algn_s1 - 1) + algn_s2 - 1 with beg = dm + st_s2
     * and runs while it stays at or above beg; the symbol indices start at
     * algn_s1 + st_s1 and algn_s2 + st_s2 (st_s1 moves the symbols read,
     * not the row). */
    (void)L1;
    rseq a, b;
    rs_init(&a, L1 + L2);
    rs_init(&b, L1 + L2);
    long end = (long)st_s2 + (long)L2 * (algn_s1 - 1) + algn_s2 - 1;
    int i = algn_s1 + st_s1, j = algn_s2 + st_s2;
    while (end >= st_s2) {
        uint16_t f = dm[end];
        if (f & ORC_ALIGN) {
            i--; rs_prepend(&a, s1[i]);
            j--; rs_prepend(&b, s2[j]);
            end -= L2 + 1;
        } else {
            int do_ins = swaped ? ((f & ORC_INSERT) != 0) : !((f & ORC_DELETE) != 0);
            if (do_ins) {
                rs_prepend(&a, GAP);
                j--; rs_prepend(&b, s2[j]);
                end -= 1;
            } else {
                i--; rs_prepend(&a, s1[i]);
                rs_prepend(&b, GAP);
                end -= L2;
            }
        }
    }
    rs_out(&b, r2);
    return rs_out(&a, r1);
}

/* algn_remove_gaps (3539) */
static void remove_gaps(rseq *s) {
    int n = 0;
    uint8_t *tmp = (uint8_t *)malloc(s->len + 1);
    for (int k = 0; k < s->len; k++) {
        uint8_t v = s->buf[s->cap - s->len + k];
        if (v != GAP) tmp[n++] = v;
    }
    s->len = 0;
    for (int k = n - 1; k >= 0; k--) rs_prepend(s, tmp[k]);
    rs_prepend(s, GAP);
    free(tmp);
}

/* algn_correct_blocks_affine (3561-3598) */
static void correct_blocks_affine(rseq *s, const uint8_t *a, const uint8_t *b) {
    int len = s->len, extending_gap = 0, inside_block = 0, prev_block = 0;
    uint8_t *sb0 = s->buf + s->cap - s->len;
    for (int i = 0; i < len; i++) {
        int ab = a[i], bb = b[i], sb = sb0[i];
        if (!inside_block && (!(ab & GAP) || !(bb & GAP))) inside_block = 0;
        else if (inside_block && ((!(ab & GAP)) || (!(bb & GAP)))) inside_block = 0;
        else if (((ab & GAP) || (bb & GAP)) && ((ab != GAP) || (bb != GAP))) inside_block = 1;
        else inside_block = 0;
        if (((GAP & ab) || (GAP & bb)) && (!(sb & GAP)) && !extending_gap) {
            prev_block = inside_block;
            extending_gap = 1;
        } else if ((GAP & ab) && (GAP & bb) && (sb & GAP) && (sb != GAP) &&
                   extending_gap && inside_block && !prev_block) {
            sb = (~GAP) & sb;
            prev_block = 0;
        } else if ((GAP & ab) && (GAP & bb) && (1 == extending_gap)) {
            prev_block = inside_block;
            extending_gap = 0;
        }
        sb0[i] = (uint8_t)sb;
    }
    remove_gaps(s);
}

int orc_ancestor_2(const uint8_t *a, const uint8_t *b, int len,
                   const orc_cm *cm, uint8_t *out) {
    /* algn_ancestor_2 (3603-3626).  An empty result reads sm[0] out of
     * bounds in the reference; both of its branches then yield [gap]. */
    rseq sm;
    rs_init(&sm, len + 2);
    for (int i = len - 1; i >= 0; i--) {
        int interm = cm->median[cidx(a[i], b[i])];
        if (interm == 0) { free(sm.buf); return -1; } /* failwith */
        if (!cm->combinations || 1 != cm->cost_model) {
            if (interm != GAP) rs_prepend(&sm, (uint8_t)interm);
        } else {
            rs_prepend(&sm, (uint8_t)interm);
        }
    }
    if (sm.len == 0) {
        rs_prepend(&sm, GAP);
    } else if (!cm->combinations || (1 != cm->cost_model && GAP != rs_first(&sm))) {
        rs_prepend(&sm, GAP);
    } else if (cm->combinations) {
        correct_blocks_affine(&sm, a, b);
    }
    return rs_out(&sm, out);
}

int orc_median_2_no_gaps(const uint8_t *a, const uint8_t *b, int len,
                         const orc_cm *cm, uint8_t *out) {
    rseq sm;
    rs_init(&sm, len + 1);
    for (int i = len - 1; i >= 0; i--) {
        int interm = cm->median[cidx(a[i], b[i])];
        if (interm != GAP) rs_prepend(&sm, (uint8_t)interm);
    }
    rs_prepend(&sm, GAP);
    return rs_out(&sm, out);
}

int orc_median_2_with_gaps(const uint8_t *a, const uint8_t *b, int len,
                           const orc_cm *cm, uint8_t *out) {
    for (int i = 0; i < len; i++) out[i] = cm->median[cidx(a[i], b[i])];
    return len;
}

/* algn_calculate_from_2_aligned (3003-3089), uselevel 0 with combinations */
static int calc_from_aligned(const uint8_t *s1, const uint8_t *s2, int len,
                             const orc_cm *cm, const int *tbl) {
    int res = 0, state = 0, i;
    int go = cm->gap_open;
    if (len <= 0) return 0;
    i = ((GAP & s1[0]) && (GAP & s2[0])) ? 1 : 0;
    for (; i < len; i++) {
        int x = s1[i], y = s2[i];
        if (state == 0) {
            if ((x & GAP) && !(y & GAP)) { res += go; state = 1; }
            else if ((y & GAP) && !(x & GAP)) { res += go; state = 2; }
        } else if (state == 1) {
            if (!(x & GAP)) {
                if ((y & GAP) && !(x & GAP)) { res += go; state = 2; }
                else state = 0;
            }
        } else {
            if (!(y & GAP)) {
                if (x & GAP) { res += go; state = 1; }
                else state = 0;
            }
        }
        res += tbl[cidx(x, y)];
    }
    return res;
}

int orc_worst_2(const uint8_t *a, const uint8_t *b, int len, const orc_cm *cm) {
    /* algn_CAML_worst_2 (3108) -> algn_worst_2 (3092); the all-gap test of
     * Align.max_cost_2 (sequence.ml:652-658) is OCaml-side. */
    return calc_from_aligned(a, b, len, cm, cm->worst);
}

int orc_verify_2(const uint8_t *a, const uint8_t *b, int len, const orc_cm *cm) {
    return calc_from_aligned(a, b, len, cm, cm->cost);
}

/* ------------------------------------------------------------------------ */
/* Affine 2D                                                                 */
/* ------------------------------------------------------------------------ */

/* HAS_GAP_OPENING (1240-1253), gap_startNO = 0 */
static inline int has_gap_opening(int idx, int prev, int curr, int go) {
    if (idx == 1 && (GAP & curr)) return 0;
    if (idx > 1 && !(GAP & prev) && (GAP & curr)) return 0;
    return go;
}

/* algn_fill_gapnum (127-176): u16 rows, arithmetic in int, stored mod 2^16 */
static void fill_gapnum(int j, int al, int in, int de, const uint16_t *pg1,
                        uint16_t *g1, const uint16_t *pg2, uint16_t *g2) {
    int a, b;
    if (al) {
        if (de && in) {
            a = imax3(pg1[j - 1], g1[j - 1] + 1, pg1[j]);
            b = imax3(pg2[j - 1], g2[j - 1], pg2[j] + 1);
        } else if (de) {
            a = imax(pg1[j], pg1[j - 1]);
            b = imax(pg2[j - 1], pg2[j] + 1);
        } else if (in) {
            a = imax(g1[j - 1] + 1, pg1[j - 1]);
            b = imax(pg2[j - 1], g2[j - 1]);
        } else {
            a = pg1[j - 1];
            b = pg2[j - 1];
        }
    } else if (in) {
        if (de) {
            a = imax(g1[j - 1] + 1, pg1[j]);
            b = imax(pg2[j] + 1, g2[j - 1]);
        } else {
            a = g1[j - 1] + 1;
            b = g2[j - 1];
        }
    } else {
        a = pg1[j];
        b = pg2[j] + 1;
    }
    g1[j] = (uint16_t)a;
    g2[j] = (uint16_t)b;
}

typedef struct {
    const orc_cm *cm;
    const uint8_t *si, *sj;
    int leni, lenj; /* lengths minus one */
    int *eh, *ev, *ebd, *cbd, *fc, *gop, *shx;
    uint16_t *dm;
    uint16_t *g[4];
    const int *prec;
    int64_t cells;
} aff_ctx;

/* algn_newkk_fill_a_row_aff (2113-2172) with FILL_EXTEND_HORIZONTAL (1283),
 * FILL_EXTEND_VERTICAL (1327), FILL_EXTEND_BLOCK_DIAGONAL (1376),
 * FILL_CLOSE_BLOCK_DIAGONAL (1436) and ASSIGN_MINIMUM (1936). */
static void aff_fill_row(aff_ctx *x, int k, int baseband, int i, int start,
                         int end, const int *gap_row, int go, int ic, int ip,
                         int si_vext, int si_ge, int si_go, int si_ng,
                         const int *si_vec, const int *peh, const int *pev,
                         const int *pcbd, const int *pebd, uint16_t *dmrow,
                         const uint16_t *pg1, uint16_t *g1, const uint16_t *pg2,
                         uint16_t *g2, int *eh, int *ev, int *cbd, int *ebd) {
    const uint8_t *sj = x->sj;
    int jc = sj[start], jp;
    for (int j = start; j <= end; j++) {
        int at_right = ((j - i) - baseband + 1 == k);
        int at_left = (j == start);
        jp = jc;
        jc = sj[j];
        int dir = 0;
        int sj_ng = 15 & jc;
        int sj_ge = gap_row[j];
        int sj_go = x->gop[j];
        if (!at_left) {
            int ext = eh[j - 1] + x->shx[j];
            int opn = cbd[j - 1] + sj_go + sj_ge;
            if (ext < opn) { dir |= ORC_BEGIN_HORIZONTAL; eh[j] = ext; }
            else { dir |= ORC_END_HORIZONTAL; eh[j] = opn; }
        } else {
            eh[j] = HIGH;
        }
        if (!at_right) {
            int ext = pev[j] + si_vext;
            int opn = pcbd[j] + si_go + si_ge;
            if (ext < opn) { dir |= ORC_BEGIN_VERTICAL; ev[j] = ext; }
            else { dir |= ORC_END_VERTICAL; ev[j] = opn; }
        } else {
            ev[j] = HIGH;
        }
        if (j > 0) {
            int flag = (GAP & ic) && (GAP & jc);
            int flag2 = !(GAP & ip) && !(GAP & jp);
            int diag = flag ? 0 : HIGH;
            int open_diag = flag ? (flag2 ? 0 : 2 * go) : HIGH;
            int ext = pebd[j - 1] + diag;
            int opn = pcbd[j - 1] + open_diag;
            if (ext < opn) { dir |= ORC_BEGIN_BLOCK; ebd[j] = ext; }
            else { dir |= ORC_END_BLOCK; ebd[j] = opn; }

            int d = si_vec[sj_ng];
            int extra = (sj_go < si_go) ? si_go : sj_go;
            int algn = pcbd[j - 1] + d;
            int fv = (ic == si_ng) ? pev[j - 1] + d : pev[j - 1] + d + sj_go;
            int fh = (jc == sj_ng) ? peh[j - 1] + d : peh[j - 1] + d + si_go;
            int fd = pebd[j - 1] + d + extra;
            int mask = ORC_ALIGN_TO_ALIGN, cb = algn;
            if (cb >= fv) { if (cb > fv) { cb = fv; mask = ORC_ALIGN_TO_VERTICAL; } else mask |= ORC_ALIGN_TO_VERTICAL; }
            if (cb >= fh) { if (cb > fh) { cb = fh; mask = ORC_ALIGN_TO_HORIZONTAL; } else mask |= ORC_ALIGN_TO_HORIZONTAL; }
            if (cb >= fd) { if (cb > fd) { cb = fd; mask = ORC_ALIGN_TO_DIAGONAL; } else mask |= ORC_ALIGN_TO_DIAGONAL; }
            cbd[j] = cb;
            dir |= mask;
        } else {
            cbd[j] = HIGH;
            ebd[j] = HIGH;
        }
        /* ASSIGN_MINIMUM */
        int H = eh[j], V = ev[j], E = ebd[j], C = cbd[j];
        int f = H, mask = ORC_DO_HORIZONTAL;
        if (f >= V) { if (f > V) { f = V; mask = ORC_DO_VERTICAL; } else mask |= ORC_DO_VERTICAL; }
        if (f >= E) { if (f > E) { f = E; mask = ORC_DO_DIAGONAL; } else mask |= ORC_DO_DIAGONAL; }
        if (f >= C) { if (f > C) { f = C; mask = ORC_DO_ALIGN; } else mask |= ORC_DO_ALIGN; }
        if (f == H && H == V) mask |= ORC_HORIZONTAL_EQ_VERTICAL;
        x->fc[j] = f;
        dir |= mask;
        fill_gapnum(j, (dir & (ORC_DO_ALIGN | ORC_DO_DIAGONAL)) != 0,
                    (dir & ORC_DO_HORIZONTAL) != 0, (dir & ORC_DO_VERTICAL) != 0, pg1,
                    g1, pg2, g2);
        dmrow[j] = (uint16_t)dir;
    }
    x->cells += end - start + 1;
}

/* algn_newkk_test_aff (2186-2306) */
static int aff_test(aff_ctx *x, int p, int *res_cost) {
    const orc_cm *cm = x->cm;
    int lenY = x->lenj, lenX = x->leni;
    int newk = p;
    if (p >= lenX) newk = lenX - 1;
    int go = cm->gap_open;
    int offset = x->lenj + 2;
    int *eh = x->eh, *ev = x->ev, *ebd = x->ebd, *cbd = x->cbd;
    int *peh = eh + offset, *pev = ev + offset, *pebd = ebd + offset, *pcbd = cbd + offset;
    uint16_t *gm1 = x->g[0], *ngm1 = x->g[1], *gm2 = x->g[2], *ngm2 = x->g[3];
    const int *gap_row = x->prec;
    const uint8_t *si = x->si, *sj = x->sj;
    int ic = si[0], ip;
    ev[0] = HIGH;
    cbd[0] = 0;
    x->shx[0] = 0;
    gm1[0] = 0;
    gm2[0] = 0;
    for (int j = 1; j <= x->lenj; j++) {
        x->gop[j] = has_gap_opening(j, sj[j - 1], sj[j], go);
        if ((sj[j - 1] & GAP) && !(sj[j] & GAP)) x->shx[j] = x->gop[j] + gap_row[j];
        else x->shx[j] = gap_row[j];
        gm1[j] = (uint16_t)j;
        gm2[j] = 0;
        ngm1[j] = 0xFFFF;
        ngm2[j] = 0xFFFF;
        x->dm[j] = ORC_DO_HORIZONTAL;
        ev[j] = HIGH;
        cbd[j] = cbd[j - 1] + x->shx[j];
        eh[j] = cbd[j];
    }
    x->shx[1] = gap_row[1];
    for (int i = 1; i <= x->leni; i++) {
        uint16_t *cdm = x->dm + (size_t)i * offset;
        int startj = (i - newk > 0) ? i - newk : 0;
        int endj = (i + lenY - lenX + newk <= lenY - 1) ? i + lenY - lenX + newk : lenY;
        int *t;
        t = peh; peh = eh; eh = t;
        t = pev; pev = ev; ev = t;
        t = pebd; pebd = ebd; ebd = t;
        t = pcbd; pcbd = cbd; cbd = t;
        ip = ic;
        ic = si[i];
        int si_ge = cm->cost[cidx(ic, GAP)];
        int si_go = has_gap_opening(i, ip, ic, go);
        int si_ng = 15 & ic;
        const int *si_vec = cm->cost + (si_ng << ORC_LCM);
        int si_vext = (i > 1 && (ip & GAP) && !(ic & GAP)) ? si_go + si_ge : si_ge;
        aff_fill_row(x, newk, lenY - lenX + 1, i, startj, endj, gap_row, go, ic, ip,
                     si_vext, si_ge, si_go, si_ng, si_vec, peh, pev, pcbd, pebd, cdm,
                     gm1, ngm1, gm2, ngm2, eh, ev, cbd, ebd);
        if (i <= x->leni - 1) {
            uint16_t *u = gm1; gm1 = ngm1; ngm1 = u;
            u = gm2; gm2 = ngm2; ngm2 = u;
        }
    }
    *res_cost = x->fc[x->lenj];
    return imax(ngm1[x->lenj], ngm2[x->lenj]);
}

enum { M_TODO, M_VERT, M_HORZ, M_DIAG, M_ALGN };

/* Grow-only scratch, like the reference's realloc-grown Matrix.default
 * (src/matrices.c:74-137): contents persist between calls. */
static void scratch_aff(size_t n, int **matrix, uint16_t **dm) {
    static int *m = NULL;
    static uint16_t *d = NULL;
    static size_t cap = 0;
    if (n > cap) {
        free(m);
        free(d);
        m = (int *)calloc(n, sizeof(int));
        d = (uint16_t *)calloc(n, sizeof(uint16_t));
        cap = n;
    }
    *matrix = m;
    *dm = d;
}

/* choose_dir (1594-1624) */
static int choose_dir(int has_diag, int has_algn, int has_vert, int has_hrzn, int sw) {
    (void)has_algn;
    if (!sw) {
        if (has_vert) return M_VERT;
        if (has_hrzn) return M_HORZ;
        if (has_diag) return M_DIAG;
        return M_ALGN;
    }
    if (has_hrzn) return M_HORZ;
    if (has_vert) return M_VERT;
    if (has_diag) return M_DIAG;
    return M_ALGN;
}

/* backtrace_aff (1715-1819) with follow_*_affine (1531-1592) */
static void backtrace_aff(const uint16_t *dmbase, const uint8_t *si, int Li,
                          const uint8_t *sj, int Lj, const orc_cm *cm, int sw,
                          rseq *median, rseq *medianwg, rseq *resi, rseq *resj) {
    int offset = Lj + 1;
    int i = Li - 1, j = Lj - 1;
    int ic = si[i], jc = sj[j];
    long p = (long)Li * (Lj + 1) - 2;
    int mode = M_TODO;
    while (i != 0 && j != 0) {
        uint16_t d = dmbase[p];
        if (mode == M_TODO) {
            mode = choose_dir(d & ORC_DO_DIAGONAL, d & ORC_DO_ALIGN, d & ORC_DO_VERTICAL,
                              d & ORC_DO_HORIZONTAL, sw);
        } else if (mode == M_VERT || mode == M_HORZ) {
            int vert = sw ? (mode != M_HORZ) : (mode == M_VERT);
            if (vert) {
                if ((d & ORC_END_VERTICAL) || (d & ORC_HORIZONTAL_EQ_VERTICAL)) mode = M_TODO;
                if (!(ic & GAP)) {
                    rs_prepend(median, (uint8_t)(ic | GAP));
                    rs_prepend(medianwg, (uint8_t)(ic | GAP));
                } else {
                    rs_prepend(medianwg, GAP);
                }
                rs_prepend(resi, (uint8_t)ic);
                rs_prepend(resj, GAP);
                i--;
                p -= offset;
                ic = si[i];
            } else {
                if ((d & ORC_END_HORIZONTAL) || (d & ORC_HORIZONTAL_EQ_VERTICAL)) mode = M_TODO;
                if (!(jc & GAP)) {
                    rs_prepend(median, (uint8_t)(jc | GAP));
                    rs_prepend(medianwg, (uint8_t)(jc | GAP));
                } else {
                    rs_prepend(medianwg, GAP);
                }
                rs_prepend(resi, GAP);
                rs_prepend(resj, (uint8_t)jc);
                j--;
                p -= 1;
                jc = sj[j];
            }
        } else if (mode == M_DIAG) {
            if (d & ORC_END_BLOCK) mode = M_TODO;
            rs_prepend(resi, (uint8_t)ic);
            rs_prepend(resj, (uint8_t)jc);
            rs_prepend(medianwg, GAP);
            i--;
            j--;
            p -= offset + 1;
            jc = sj[j];
            ic = si[i];
        } else {
            mode = choose_dir(d & ORC_ALIGN_TO_DIAGONAL, d & ORC_ALIGN_TO_ALIGN,
                              d & ORC_ALIGN_TO_VERTICAL, d & ORC_ALIGN_TO_HORIZONTAL, sw);
            uint8_t prep = cm->median[cidx(ic & 15, jc & 15)];
            rs_prepend(median, prep);
            rs_prepend(medianwg, prep);
            rs_prepend(resi, (uint8_t)ic);
            rs_prepend(resj, (uint8_t)jc);
            i--;
            j--;
            p -= offset + 1;
            jc = sj[j];
            ic = si[i];
        }
    }
    while (i != 0) {
        if (!(ic & GAP)) {
            rs_prepend(median, (uint8_t)(ic | GAP));
            rs_prepend(medianwg, (uint8_t)(ic | GAP));
        } else {
            rs_prepend(medianwg, GAP);
        }
        rs_prepend(resi, (uint8_t)ic);
        rs_prepend(resj, GAP);
        i--;
        ic = si[i];
    }
    while (j != 0) {
        if (!(jc & GAP)) {
            rs_prepend(median, (uint8_t)(jc | GAP));
            rs_prepend(medianwg, (uint8_t)(jc | GAP));
        } else {
            rs_prepend(medianwg, GAP);
        }
        rs_prepend(resi, GAP);
        rs_prepend(resj, (uint8_t)jc);
        j--;
        jc = sj[j];
    }
    rs_prepend(resi, GAP);
    rs_prepend(resj, GAP);
    rs_prepend(medianwg, GAP);
    /* an empty median reads past its buffer in the reference; the leading gap
     * is prepended in that case */
    if (median->len == 0 || GAP != rs_first(median)) rs_prepend(median, GAP);
}

static int align_affine_impl(const uint8_t *si, int Li, const uint8_t *sj, int Lj,
                             const orc_cm *cm, int swaped, uint8_t *median, int *mlen,
                             uint8_t *medianwg, int *mwglen, uint8_t *resi, uint8_t *resj,
                             int *alen, orc_stats *st, int fixed_p);

int orc_align_affine_3(const uint8_t *si, int Li, const uint8_t *sj, int Lj,
                       const orc_cm *cm, int swaped, uint8_t *median, int *mlen,
                       uint8_t *medianwg, int *mwglen, uint8_t *resi, uint8_t *resj,
                       int *alen, orc_stats *st) {
    return align_affine_impl(si, Li, sj, Lj, cm, swaped, median, mlen, medianwg, mwglen,
                             resi, resj, alen, st, -1);
}

int orc_align_affine_round(const uint8_t *si, int Li, const uint8_t *sj, int Lj,
                           const orc_cm *cm, int swaped, int p, uint8_t *median, int *mlen,
                           uint8_t *medianwg, int *mwglen, uint8_t *resi, uint8_t *resj,
                           int *alen, orc_stats *st) {
    if (p < 0) return -1;
    return align_affine_impl(si, Li, sj, Lj, cm, swaped, median, mlen, medianwg, mwglen,
                             resi, resj, alen, st, p);
}

static int align_affine_impl(const uint8_t *si, int Li, const uint8_t *sj, int Lj,
                             const orc_cm *cm, int swaped, uint8_t *median, int *mlen,
                             uint8_t *medianwg, int *mwglen, uint8_t *resi, uint8_t *resj,
                             int *alen, orc_stats *st, int fixed_p) {
    orc_stats local;
    if (!st) st = &local;
    memset(st, 0, sizeof *st);
    st->min_gap_excess = INT_MAX;
    if (Lj < Li) return -1; /* "pass the shorter one as first" */
    int largest = Lj + 1;
    size_t msz = (size_t)(Li + 12) * (Lj + 1);
    int *matrix;
    uint16_t *dm;
    scratch_aff(msz, &matrix, &dm);
    aff_ctx x;
    memset(&x, 0, sizeof x);
    x.cm = cm;
    x.si = si;
    x.sj = sj;
    x.leni = Li - 1;
    x.lenj = Lj - 1;
    x.cbd = matrix;
    x.ebd = matrix + 2 * largest;
    x.ev = matrix + 4 * largest;
    x.eh = matrix + 6 * largest;
    x.fc = matrix + 8 * largest;
    x.gop = matrix + 10 * largest;
    x.shx = matrix + 11 * largest;
    x.dm = dm;
    for (int g = 0; g < 4; g++) x.g[g] = (uint16_t *)calloc(Lj + 1, sizeof(uint16_t));
    int *prec = precalc_2d(cm, sj, Lj);
    x.prec = prec;
    /* initialize_matrices_affine (1866-1933) */
    {
        int go = cm->gap_open;
        x.cbd[0] = 0;
        x.ebd[0] = HIGH;
        x.eh[0] = go;
        x.ev[0] = go;
        x.fc[0] = 0;
        dm[0] = 0xFFFF;
        for (int j = 1; j <= x.lenj; j++) {
            int r = x.eh[j - 1] + prec[j];
            x.eh[j] = r;
            x.cbd[j] = HIGH;
            x.ebd[j] = HIGH;
            x.ev[j] = HIGH;
            x.fc[j] = r;
            dm[j] = ORC_DO_HORIZONTAL | ORC_END_HORIZONTAL;
        }
    }
    /* algn_fill_plane_3_aff (2341) -> algn_newkk_increaseT_aff (2311) */
    int delta = orc_min_non0_cost(cm);
    int T = (x.lenj - x.leni + 1) * delta;
    int res = 0;
    for (;;) {
        int p = fixed_p >= 0 ? fixed_p : (T - (x.lenj - x.leni)) / 2;
        int gap_num = aff_test(&x, p, &res);
        st->rounds++;
        st->final_T = T;
        st->final_k = p >= x.leni ? x.leni - 1 : p;
        st->gap_num = gap_num;
        if (gap_num - (x.lenj - x.leni) < st->min_gap_excess)
            st->min_gap_excess = gap_num - (x.lenj - x.leni);
        int newp = (2 * T - (x.lenj - x.leni)) / 2;
        if (fixed_p >= 0 || gap_num < p || newp - x.lenj + 1 >= 0) break;
        T *= 2;
    }
    st->cells = x.cells;
    rseq rm, rmwg, ri, rj;
    int cap = Li + Lj + 2;
    rs_init(&rm, cap);
    rs_init(&rmwg, cap);
    rs_init(&ri, cap);
    rs_init(&rj, cap);
    backtrace_aff(dm, si, Li, sj, Lj, cm, swaped, &rm, &rmwg, &ri, &rj);
    int n;
    n = rs_out(&rm, median);
    if (mlen) *mlen = n;
    n = rs_out(&rmwg, medianwg);
    if (mwglen) *mwglen = n;
    rs_out(&rj, resj);
    n = rs_out(&ri, resi);
    if (alen) *alen = n;
    for (int g = 0; g < 4; g++) free(x.g[g]);
    free(prec);
    return res;
}

/* algn_CAML_cost_affine_3 (2457-2515) -> initialize_matrices_affine_nobt
 * (1822) -> algn_fill_plane_3_aff_nobt (1987-2110) */
int orc_cost_affine_3(const uint8_t *si, int Li, const uint8_t *sj, int Lj,
                      const orc_cm *cm) {
    const uint8_t *A = si, *B = sj;
    int la = Li, lb = Lj;
    if (Li > Lj) { A = sj; B = si; la = Lj; lb = Li; }
    int largest = Li > Lj ? Li : Lj;
    /* the reference carves 12 rows of 2*largest from a (largest+1)^2 block
     * that only ever grows; size it for the carve-out directly */
    size_t msz = (size_t)(largest + 1) * (largest + 1);
    if (msz < (size_t)13 * (largest + 1)) msz = (size_t)13 * (largest + 1);
    int *matrix = (int *)calloc(msz, sizeof(int));
    int *cbd = matrix, *ebd = matrix + 2 * largest, *ev = matrix + 4 * largest;
    int *eh = matrix + 6 * largest, *gop = matrix + 10 * largest, *shx = matrix + 11 * largest;
    int *prec = precalc_2d(cm, B, lb);
    const int *gap_row = prec;
    int go = cm->gap_open;
    int leni = la - 1, lenj = lb - 1;
    cbd[0] = 0;
    ebd[0] = HIGH;
    eh[0] = go;
    ev[0] = go;
    for (int j = 1; j <= lenj; j++) {
        eh[j] = eh[j - 1] + gap_row[j];
        cbd[j] = HIGH;
        ebd[j] = HIGH;
        ev[j] = HIGH;
    }
    int offset = lenj + 2;
    int *peh = eh + offset, *pev = ev + offset, *pebd = ebd + offset, *pcbd = cbd + offset;
    int ic = A[0], ip;
    shx[0] = 0;
    for (int j = 1; j <= lenj; j++) {
        gop[j] = has_gap_opening(j, B[j - 1], B[j], go);
        if ((B[j - 1] & GAP) && !(B[j] & GAP)) shx[j] = gop[j] + gap_row[j];
        else shx[j] = gap_row[j];
    }
    shx[1] = gap_row[1];
    for (int i = 1; i <= leni; i++) {
        int *t;
        t = peh; peh = eh; eh = t;
        t = pev; pev = ev; ev = t;
        t = pebd; pebd = ebd; ebd = t;
        t = pcbd; pcbd = cbd; cbd = t;
        ip = ic;
        ic = A[i];
        int si_ge = cm->cost[cidx(ic, GAP)];
        int si_go = has_gap_opening(i, ip, ic, go);
        int si_ng = 15 & ic;
        int si_vext = (i > 1 && (ip & GAP) && !(ic & GAP)) ? si_go + si_ge : si_ge;
        int r = pev[0] + si_vext;
        eh[0] = HIGH;
        ebd[0] = HIGH;
        ev[0] = r;
        cbd[0] = HIGH;
        const int *si_vec = cm->cost + (si_ng << ORC_LCM);
        int jc = B[0], jp;
        for (int j = 1; j <= lenj; j++) {
            jp = jc;
            jc = B[j];
            int sj_ng = 15 & jc, sj_ge = gap_row[j], sj_go = gop[j];
            int ext = eh[j - 1] + shx[j];
            int opn = cbd[j - 1] + sj_go + sj_ge;
            eh[j] = ext < opn ? ext : opn;
            ext = pev[j] + si_vext;
            opn = pcbd[j] + si_go + si_ge;
            ev[j] = ext < opn ? ext : opn;
            int flag = (GAP & ic) && (GAP & jc);
            int flag2 = !(GAP & ip) && !(GAP & jp);
            int diag = flag ? 0 : HIGH;
            int open_diag = flag ? (flag2 ? 0 : 2 * go) : HIGH;
            ext = pebd[j - 1] + diag;
            opn = pcbd[j - 1] + open_diag;
            ebd[j] = ext < opn ? ext : opn;
            int d = si_vec[sj_ng];
            int extra = (sj_go < si_go) ? si_go : sj_go;
            int cb = pcbd[j - 1] + d;
            int fv = (ic == si_ng) ? pev[j - 1] + d : pev[j - 1] + d + sj_go;
            int fh = (jc == sj_ng) ? peh[j - 1] + d : peh[j - 1] + d + si_go;
            int fd = pebd[j - 1] + d + extra;
            if (cb > fv) cb = fv;
            if (cb > fh) cb = fh;
            if (cb > fd) cb = fd;
            cbd[j] = cb;
        }
    }
    int res = eh[lenj];
    if (res > ev[lenj]) res = ev[lenj];
    if (res > ebd[lenj]) res = ebd[lenj];
    if (res > cbd[lenj]) res = cbd[lenj];
    free(prec);
    free(matrix);
    return res;
}

/* ------------------------------------------------------------------------ */
/* OCaml glue                                                                */
/* ------------------------------------------------------------------------ */

int orc_count_gaps(const uint8_t *s, int L) {
    int n = 0;
    for (int i = 0; i < L; i++)
        if (s[i] >= GAP && (s[i] & GAP)) n++;
    return n;
}

/* Align.cost_2's deltaw_calc (sequence.ml:874-884); int_of_float (n * 0.10)
 * equals n / 10 for every non-negative int n. */
static int deltaw_calc(int s1len, int s2len, int has_deltaw, int v) {
    int dif = s1len - s2len;
    int lower = s1len / 10;
    if (!has_deltaw) return dif < lower ? lower / 2 : 2;
    return dif < lower ? lower : v;
}

int orc_deltawh(const uint8_t *s1, int L1, const uint8_t *s2, int L2,
                int has_deltaw, int deltaw) {
    int g = imax(orc_count_gaps(s1, L1), orc_count_gaps(s2, L2));
    if (L1 <= L2) return g + deltaw_calc(L1, L2, has_deltaw, deltaw);
    return g + deltaw_calc(L2, L1, has_deltaw, deltaw);
}

int orc_cost_2(const uint8_t *s1, int L1, const uint8_t *s2, int L2,
               const orc_cm *cm, int has_deltaw, int deltaw, uint16_t *dm,
               orc_stats *st) {
    int dw = orc_deltawh(s1, L1, s2, L2, has_deltaw, deltaw);
    if (L1 <= L2) return orc_nw(s1, L1, s2, L2, cm, dw, dm, st);
    return orc_nw(s2, L2, s1, L1, cm, dw, dm, st);
}

int orc_align_2(const uint8_t *s1, int L1, const uint8_t *s2, int L2,
                const orc_cm *cm, uint8_t *r1, uint8_t *r2, int *alen,
                orc_stats *st) {
    uint16_t *dm = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)L1 * L2);
    int cost = orc_cost_2(s1, L1, s2, L2, cm, 0, 0, dm, st);
    int n;
    if (L1 <= L2) n = orc_backtrace_2d(s1, L1, s2, L2, dm, 0, r1, r2);
    else n = orc_backtrace_2d(s2, L2, s1, L1, dm, 1, r2, r1);
    if (alen) *alen = n;
    free(dm);
    return cost;
}

/* ------------------------------------------------------------------------ */
/* 3D cube                                                                   */
/* ------------------------------------------------------------------------ */

/* algn_nw_3d (2977) -> cm_precalc_4algn_3d (cm.c:1390) -> algn_fill_cube
 * (2593-2763).  The cube's row pointers are restated with the reference's
 * exact pointer arithmetic: at the start of plane i >= 1 the "diag" and
 * "upper" rows lag their geometric neighbours by i-1 rows and "prev" by i
 * rows, because the per-plane row loop advances them s2_len-1 times while
 * the plane advances s2_len rows. */
int orc_nw_3d(const uint8_t *s1, int L1, const uint8_t *s2, int L2,
              const uint8_t *s3, int L3, const orc_cm3 *c3, uint16_t *dm) {
    size_t n = (size_t)L1 * L2 * L3;
    int *mm = (int *)calloc(n ? n : 1, sizeof(int));
    const int asz = 32; /* a_sz + 1 */
    int *prec = (int *)calloc((size_t)33 * 33 * L3, sizeof(int));
    for (int a = 1; a <= 31; a++)
        for (int b = 1; b <= 31; b++)
            for (int k = 0; k < L3; k++)
                prec[(size_t)a * asz * L3 + (size_t)b * L3 + k] =
                    c3->cost[(((a << 5) + b) << 5) + s3[k]];
#define P3D(a, b) (prec + (size_t)(a) * asz * L3 + (size_t)(b) * L3)
    const int gap = GAP;
    const int *ggs3 = P3D(gap, gap);
    mm[0] = 0;
    dm[0] = ORC_S2;
    for (int k = 1; k < L3; k++) {
        mm[k] = mm[k - 1] + ggs3[k];
        dm[k] = ORC_SS;
    }
    long prev = 0, cur = L3;
    for (int r = 1; r < L2; r++, prev += L3, cur += L3) {
        const int *gs2s3 = P3D(gap, s2[r]);
        mm[cur] = mm[prev] + gs2s3[0];
        dm[cur] = ORC_P1;
        for (int k = 1; k < L3; k++) {
            mm[cur + k] = mm[prev + k] + gs2s3[0];
            dm[cur + k] = ORC_P1;
            int t = mm[prev + k - 1] + gs2s3[k];
            if (t < mm[cur + k]) { mm[cur + k] = t; dm[cur + k] = ORC_S1; }
            t = mm[cur + k - 1] + ggs3[k];
            if (t < mm[cur + k]) { mm[cur + k] = t; dm[cur + k] = ORC_SS; }
        }
    }
    cur = (long)L3 * L2;
    long diag = 0, upper = L3;
    prev = cur - L3;
    for (int i = 1; i < L1; i++) {
        int s1i = s1[i];
        const int *s1gs3 = P3D(s1i, gap);
        mm[cur] = mm[diag] + s1gs3[0];
        dm[cur] = ORC_P3;
        for (int k = 1; k < L3; k++) {
            mm[cur + k] = mm[diag + k] + s1gs3[0];
            dm[cur + k] = ORC_P3;
            int t = mm[diag + k - 1] + s1gs3[k];
            if (t < mm[cur + k]) { mm[cur + k] = t; dm[cur + k] = ORC_S3; }
            t = ggs3[k] + mm[cur + k - 1];
            if (t < mm[cur + k]) { mm[cur + k] = t; dm[cur + k] = ORC_SS; }
        }
        cur += L3;
        for (int r = 1; r < L2; r++, diag += L3, upper += L3, prev += L3, cur += L3) {
            int s2r = s2[r];
            const int *gs2s3 = P3D(gap, s2r);
            const int *s1s2s3 = P3D(s1i, s2r);
            /* fill_parallel (2552) */
            for (int k = 0; k < L3; k++) {
                mm[cur + k] = mm[upper + k] + s1gs3[0];
                dm[cur + k] = ORC_P3;
                int t = mm[prev + k] + gs2s3[0];
                if (t < mm[cur + k]) { mm[cur + k] = t; dm[cur + k] = ORC_P1; }
                t = mm[diag + k] + s1s2s3[0];
                if (t < mm[cur + k]) { mm[cur + k] = t; dm[cur + k] = ORC_P2; }
            }
            /* fill_moved (2524) on rows shifted one cell back */
            for (int k = 1; k < L3; k++) {
                int t = mm[upper - 1 + k] + s1gs3[k];
                if (t < mm[cur + k]) { mm[cur + k] = t; dm[cur + k] = ORC_S3; }
                t = mm[prev - 1 + k] + gs2s3[k];
                if (t < mm[cur + k]) { mm[cur + k] = t; dm[cur + k] = ORC_S1; }
                t = mm[diag - 1 + k] + s1s2s3[k];
                if (t < mm[cur + k]) { mm[cur + k] = t; dm[cur + k] = ORC_S2; }
            }
            for (int k = 1; k < L3; k++) {
                int t = mm[cur + k - 1] + ggs3[k];
                if (t < mm[cur + k]) { mm[cur + k] = t; dm[cur + k] = ORC_SS; }
            }
        }
    }
#undef P3D
    int res = mm[cur - 1];
    free(prec);
    free(mm);
    return res;
}

int orc_backtrack_3d(const uint8_t *s1, int L1, const uint8_t *s2, int L2,
                     const uint8_t *s3, int L3, const uint16_t *dm, int gap,
                     uint8_t *r1, uint8_t *r2, uint8_t *r3) {
    rseq a, b, c;
    int cap = L1 + L2 + L3;
    rs_init(&a, cap);
    rs_init(&b, cap);
    rs_init(&c, cap);
    long plane = (long)L2 * L3, line = L3;
    long end = (long)L1 * L2 * L3 - 1;
    int l1 = L1 - 1, l2 = L2 - 1, l3 = L3 - 1;
    int bad = 0;
#define TAKE(r, s, idx) do { if ((idx) < 0) { bad = 1; break; } rs_prepend(&(r), (s)[(idx)--]); } while (0)
    while (end > 0 && !bad && a.len < cap) {
        uint16_t d = dm[end];
        if (d & ORC_S2) {
            TAKE(a, s1, l1); TAKE(b, s2, l2); TAKE(c, s3, l3);
            end -= plane + line + 1;
        } else if (d & ORC_S3) {
            TAKE(a, s1, l1); rs_prepend(&b, (uint8_t)gap); TAKE(c, s3, l3);
            end -= plane + 1;
        } else if (d & ORC_S1) {
            rs_prepend(&a, (uint8_t)gap); TAKE(b, s2, l2); TAKE(c, s3, l3);
            end -= line + 1;
        } else if (d & ORC_P3) {
            TAKE(a, s1, l1); rs_prepend(&b, (uint8_t)gap); rs_prepend(&c, (uint8_t)gap);
            end -= plane;
        } else if (d & ORC_SS) {
            rs_prepend(&a, (uint8_t)gap); rs_prepend(&b, (uint8_t)gap); TAKE(c, s3, l3);
            end -= 1;
        } else if (d & ORC_P1) {
            rs_prepend(&a, (uint8_t)gap); TAKE(b, s2, l2); rs_prepend(&c, (uint8_t)gap);
            end -= line;
        } else if (d & ORC_P2) {
            TAKE(a, s1, l1); TAKE(b, s2, l2); rs_prepend(&c, (uint8_t)gap);
            end -= plane + line;
        } else {
            bad = 1;
        }
    }
#undef TAKE
    rs_out(&c, r3);
    rs_out(&b, r2);
    int n = rs_out(&a, r1);
    return bad ? -1 : n;
}

int orc_median_3(const uint8_t *a, const uint8_t *b, const uint8_t *c, int len,
                 const orc_cm3 *c3, uint8_t *out) {
    /* algn_get_median_3d (3641-3654): the loop reads the last column of each
     * input on every iteration (its end pointers never move). */
    if (len <= 0) return 0;
    uint8_t m = c3->median[(((a[len - 1] << 5) + b[len - 1]) << 5) + c[len - 1]];
    for (int i = 0; i < len; i++) out[i] = m;
    return len;
}

/* ------------------------------------------------------------------------ */

int64_t orc_batch_align_affine(int n, const uint8_t *blob, const int64_t *off1,
                               const int *len1, const int64_t *off2,
                               const int *len2, const orc_cm *cm, int *costs,
                               int64_t *cells_out) {
    int64_t total = 0, cells = 0;
    for (int k = 0; k < n; k++) {
        const uint8_t *a = blob + off1[k], *b = blob + off2[k];
        int la = len1[k], lb = len2[k];
        int sw = 0;
        if (la > lb) {
            const uint8_t *t = a; a = b; b = t;
            int tl = la; la = lb; lb = tl;
            sw = 1;
        }
        int cap = la + lb + 2;
        uint8_t *buf = (uint8_t *)malloc((size_t)cap * 4);
        int ml, mwl, al;
        orc_stats st;
        int c = orc_align_affine_3(a, la, b, lb, cm, sw, buf, &ml, buf + cap, &mwl,
                                   buf + 2 * cap, buf + 3 * cap, &al, &st);
        free(buf);
        if (costs) costs[k] = c;
        total += c;
        cells += st.cells;
    }
    if (cells_out) *cells_out = cells;
    return total;
}

/* Align.align_2 + ancestor_2 + median_2_with_gaps per pair (the linear batch
 * op); returns the sum of costs.  CPU baseline helper. */
int64_t orc_batch_align_2(int n, const uint8_t *blob, const int64_t *off1, const int *len1,
                          const int64_t *off2, const int *len2, const orc_cm *cm, int *costs,
                          int64_t *cells_out) {
    int64_t total = 0, cells = 0;
    for (int k = 0; k < n; k++) {
        const uint8_t *a = blob + off1[k], *b = blob + off2[k];
        int la = len1[k], lb = len2[k];
        int cap = la + lb + 1;
        uint8_t *buf = (uint8_t *)malloc((size_t)cap * 4);
        int al = 0;
        orc_stats st;
        int c = orc_align_2(a, la, b, lb, cm, buf, buf + cap, &al, &st);
        orc_ancestor_2(buf, buf + cap, al, cm, buf + 2 * cap);
        orc_median_2_with_gaps(buf, buf + cap, al, cm, buf + 3 * cap);
        free(buf);
        if (costs) costs[k] = c;
        total += c;
        cells += st.cells;
    }
    if (cells_out) *cells_out = cells;
    return total;
}

/* algn_nw_3d + backtrack_3d + median_3 per triple; returns the sum of costs.
 * CPU baseline helper. */
int64_t orc_batch_align_3d(int n, const uint8_t *blob, const int64_t *off1, const int *len1,
                           const int64_t *off2, const int *len2, const int64_t *off3,
                           const int *len3, const orc_cm3 *c3, int *costs) {
    int64_t total = 0;
    for (int k = 0; k < n; k++) {
        const uint8_t *a = blob + off1[k], *b = blob + off2[k], *c = blob + off3[k];
        int l1 = len1[k], l2 = len2[k], l3 = len3[k];
        size_t cells = (size_t)l1 * l2 * l3;
        uint16_t *dm = (uint16_t *)malloc(sizeof(uint16_t) * (cells ? cells : 1));
        int cost = orc_nw_3d(a, l1, b, l2, c, l3, c3, dm);
        int cap = l1 + l2 + l3;
        uint8_t *buf = (uint8_t *)malloc((size_t)cap * 4);
        int len = orc_backtrack_3d(a, l1, b, l2, c, l3, dm, GAP, buf, buf + cap, buf + 2 * cap);
        if (len > 0) orc_median_3(buf, buf + cap, buf + 2 * cap, len, c3, buf + 3 * cap);
        free(buf);
        free(dm);
        if (costs) costs[k] = cost;
        total += cost;
    }
    return total;
}

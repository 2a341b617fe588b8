/*
 * newkk.c -- CPU restatement of POY 5.1's diagonal-store Ukkonen aligner
 * (src/newkkonen.c), affine cost model.
 *
 * TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker.  The product (poy5_amd/,
 * include/poyhip.h) never links or calls it.
 *
 * Two restatements of the same fill:
 *  - orc_newkk_align: the reference's procedure as written.  Band-only
 *    storage, one array per diagonal in the reference's order (0..bb-1, then
 *    k=1 right, k=1 left, k=2 right, ...; newkkonen.h:77-79, get_idx
 *    newkkonen.c:453), T doubling (increaseT 1155), and after the first round
 *    the queue-driven refill that revisits only cells whose inputs changed
 *    (update_a_row 964, update_a_cell 875, ukktest 1077).
 *  - orc_newkk_full: every band cell recomputed in every round (dense
 *    store).  The two agreeing on every input is what pins the queue logic.
 * Both share the cell (update_internal_cell 680 + assign_best_cost_and_
 * direction 600) and the backtrace (backtrace_affine 1666).
 *
 * Sequences: 5-bit bitsets, leading gap at index 0, s1 the shorter
 * (lenX <= lenY), as newkk_algn requires (newkkonen.c:1361).
 */
#include "oracle.h"

#include <limits.h>
#include <stdlib.h>
#include <string.h>

#define NK_INF (INT_MAX / 2)
#define NK_MUCH_LONGER 100 /* newkkonen.c:35 */

typedef struct {
    int cost;
    uint16_t dir, g1, g2;
    int P, Q, ED, CD;
} nk_cell;

static inline int nk_min(int a, int b) { return a < b ? a : b; }
static inline int nk_max(int a, int b) { return a > b ? a : b; }
static inline int cidx(int a, int b) { return (a << ORC_LCM) + b; }

/* my_add (newkkonen.c:65-71) */
static inline int my_add(int a, int b) {
    if (a >= NK_INF) return NK_INF;
    if (b >= NK_INF) return NK_INF;
    return a + b;
}

/* get_go / get_extgo (newkkonen.c:189-215) */
static inline int get_go(int base, int prebase, int idx, int go) {
    if (idx == 1 && (base & ORC_GAP)) return 0;
    if (idx > 1 && !(prebase & ORC_GAP) && (base & ORC_GAP)) return 0;
    return go;
}
static inline int get_extgo(int base, int prebase, int go) {
    return ((prebase & ORC_GAP) && !(base & ORC_GAP)) ? go : 0;
}

/* cm_calc_cost with level 0 (get_cmcost, newkkonen.c:583-592) */
static inline int cmc(const orc_cm *c, int a, int b) { return c->cost[cidx(a, b)]; }

/* update_internal_cell (newkkonen.c:680-870) followed by
 * assign_best_cost_and_direction (600-677).  L = (i, j-1), R = (i-1, j),
 * M = (i-1, j-1); a NULL neighbour is never read (the border / first row /
 * first column cases of the reference). */
static nk_cell nk_cell_update(const uint8_t *s1, const uint8_t *s2, const orc_cm *c, int go,
                              int i, int j, int at_left, int at_right, const nk_cell *L,
                              const nk_cell *R, const nk_cell *M) {
    const int base1 = s1[i], base2 = s2[j];
    const int prev1 = i > 0 ? s1[i - 1] : 0, prev2 = j > 0 ? s2[j - 1] : 0;
    const int go1 = get_go(base1, prev1, i, go), go2 = get_go(base2, prev2, j, go);
    const int xg1 = get_extgo(base1, prev1, go), xg2 = get_extgo(base2, prev2, go);
    int thisP = NK_INF, thisQ = NK_INF, thisED = NK_INF, thisCD = NK_INF;
    /* from the left (insert) */
    int costL, extL, openL, gL1 = 0, gL2 = 0;
    if (at_left || j == 0) {
        costL = extL = openL = NK_INF;
    } else {
        const int add = cmc(c, base2, ORC_GAP);
        extL = my_add(L->Q, add + xg2);
        openL = my_add(L->CD, add + go2);
        costL = nk_min(openL, extL);
        thisQ = costL;
        gL1 = L->g1;
        gL2 = L->g2;
    }
    /* from above (delete) */
    int costR, extR, openR, gR1 = 0, gR2 = 0;
    if (at_right || i == 0) {
        costR = extR = openR = NK_INF;
    } else {
        const int add = cmc(c, base1, ORC_GAP);
        extR = my_add(R->P, add + xg1);
        openR = my_add(R->CD, add + go1);
        costR = nk_min(openR, extR);
        thisP = costR;
        gR1 = R->g1;
        gR2 = R->g2;
    }
    /* from the diagonal (align, gap-gap block) */
    int costM, costDiag, extD, openD, gM1 = 0, gM2 = 0;
    if (i * j == 0) {
        costM = costDiag = extD = openD = NK_INF;
    } else {
        const int add = cmc(c, base1 & 15, base2 & 15);
        const int b1n = base1 & 15, b2n = base2 & 15;
        const int fromCD = my_add(M->CD, add);
        /* plain additions, as written (newkkonen.c:822-829) */
        const int fromR = M->P + cmc(c, b1n, base2) + (b1n == base1 ? 0 : go);
        const int fromL = M->Q + cmc(c, base1, b2n) + (b2n == base2 ? 0 : go);
        thisCD = nk_min(nk_min(fromCD, fromR), fromL);
        const int fromED = my_add(M->ED, go1 + go2 + add);
        thisCD = nk_min(thisCD, fromED);
        const int diag = ((base1 & ORC_GAP) && (base2 & ORC_GAP)) ? 0 : NK_INF;
        openD = my_add(M->CD, go1 + go2);
        extD = my_add(M->ED, diag);
        thisED = nk_min(extD, openD);
        costM = thisCD;
        costDiag = thisED;
        gM1 = M->g1;
        gM2 = M->g2;
    }
    /* assign_best_cost_and_direction: insert, then delete, align, gap-gap */
    int dir = NK_DO_INSERT, best = costL, r1 = gL1 + 1, r2 = gL2;
    if (costR <= best) {
        if (costR < best) {
            best = costR; dir = NK_DO_DELETE; r1 = gR1; r2 = gR2 + 1;
        } else {
            dir |= NK_DO_DELETE; r1 = nk_max(r1, gR1); r2 = nk_max(r2, gR2 + 1);
        }
    }
    if (costM <= best) {
        if (costM < best) {
            best = thisCD; dir = NK_ALIGN; r1 = gM1; r2 = gM2;
        } else {
            dir |= NK_ALIGN; r1 = nk_max(r1, gM1); r2 = nk_max(r2, gM2);
        }
    }
    if (costDiag <= best) {
        if (costDiag < best) {
            best = costDiag; dir = NK_DO_DIAG; r1 = gM1; r2 = gM2;
        } else {
            dir |= NK_DO_DIAG; r1 = nk_max(r1, gM1); r2 = nk_max(r2, gM2);
        }
    }
    if (extR >= openR) dir |= NK_END_DELETE;
    if (extL >= openL) dir |= NK_END_INSERT;
    if (extD >= openD) dir |= NK_END_DIAG;
    if (extR == extL && extR == best) dir |= NK_INS_EQ_DEL;
    nk_cell out;
    out.cost = best;
    out.dir = (uint16_t)dir;
    out.g1 = (uint16_t)r1; /* DIRECTION_MATRIX stores wrap at 2^16 */
    out.g2 = (uint16_t)r2;
    out.P = thisP;
    out.Q = thisQ;
    out.ED = thisED;
    out.CD = thisCD;
    return out;
}

/* Row 0 as newkk_algn initialises it (newkkonen.c:1390-1426): (0,0) and
 * the baseband columns 1..bb-1. */
static nk_cell nk_origin(int go) {
    nk_cell o;
    o.cost = 0; o.dir = NK_START; o.g1 = 0; o.g2 = 0;
    o.P = go; o.Q = go; o.ED = NK_INF; o.CD = 0;
    return o;
}
static nk_cell nk_first_row(const uint8_t *s2, const orc_cm *c, int go, int j,
                            const nk_cell *L) {
    const int thisbase = s2[j], prevbase = s2[j - 1];
    const int flag = ORC_GAP & prevbase, flag2 = ORC_GAP & thisbase;
    const int addcost = cmc(c, thisbase, ORC_GAP);
    int gocost;
    if (j == 1) gocost = flag2 ? 0 : go;
    else gocost = (!flag && flag2) ? 0 : go;
    const int extL = (flag && !flag2) ? L->Q + addcost + go : L->Q + addcost;
    const int openL = L->cost + addcost + gocost;
    const int costL = nk_min(openL, extL);
    nk_cell o;
    o.cost = costL;
    o.dir = NK_DO_INSERT | NK_END_INSERT | NK_END_DELETE | NK_END_DIAG;
    o.g1 = (uint16_t)(L->g1 + 1);
    o.g2 = L->g2;
    o.P = NK_INF; o.Q = costL; o.ED = NK_INF; o.CD = NK_INF;
    return o;
}

/* trivial_algn (newkkonen.c:1344-1358): indel costs of both sequences,
 * gap bits stripped from the symbol. */
static int nk_trivial_cost(const uint8_t *s1, int l1, const uint8_t *s2, int l2,
                           const orc_cm *c) {
    int t = 0;
    for (int i = 0; i < l1; i++) t += cmc(c, s1[i] & 15, ORC_GAP);
    for (int j = 0; j < l2; j++) t += cmc(c, s2[j] & 15, ORC_GAP);
    return t;
}

/* ---------------- the reference's procedure ---------------------------- */

typedef struct {
    int lenX, lenY, bb, k, diag_in_use, diag_max;
    int64_t *doff;  /* offset of every diagonal that can exist */
    nk_cell *pool;
    int64_t used;
} nk_mat;

/* get_idx (newkkonen.c:453-495): diagonal and index of (i, j); border flags
 * for the current k; first-time flag against oldk when oldk >= 0. */
static nk_cell *nk_at(nk_mat *m, int i, int j) {
    int d;
    if (i > j) d = 2 * (i - j) + m->bb - 1;
    else if (j - i < m->bb) d = j - i;
    else d = m->bb + 2 * (j - i + 1 - m->bb) - 2;
    return m->pool + m->doff[d] + (i <= j ? i : j);
}
static int nk_first_time(const nk_mat *m, int oldk, int i, int j) {
    return (j - i - m->bb + 1) > oldk || (i - j) > oldk;
}

/* expand_mat (newkkonen.c:288-440): two more diagonals per unit of k, their
 * costs set to -1 so update_a_cell sees every new cell as changed. */
static void nk_expand(nk_mat *m, int newk, int oldk) {
    const int newsize = m->diag_in_use + 2 * (newk - oldk);
    const int64_t end = m->doff[newsize];
    if (end > m->used) {
        m->pool = realloc(m->pool, sizeof(nk_cell) * (size_t)end);
        m->used = end;
    }
    for (int64_t q = m->doff[m->diag_in_use]; q < end; q++) {
        memset(&m->pool[q], 0xff, sizeof(nk_cell));
    }
    m->diag_in_use = newsize;
    m->k = newk;
}

/* update_a_cell (newkkonen.c:875-946): returns 1 when the cell is new or
 * any of cost, gap counts, P, Q, ED, CD changed. */
static int nk_update_a_cell(nk_mat *m, const uint8_t *s1, const uint8_t *s2, const orc_cm *c,
                            int go, int i, int j, int oldk, int getcost, int64_t *cells) {
    nk_cell *x = nk_at(m, i, j);
    const int first = nk_first_time(m, oldk, i, j);
    if (first) getcost = 0;
    const nk_cell old = *x;
    if (i == 0 && j == 0) return 0;
    const int at_left = (i - j) == m->k;
    const int at_right = (j - i - m->bb + 1) == m->k;
    const nk_cell *L = (at_left || j == 0) ? NULL : nk_at(m, i, j - 1);
    const nk_cell *R = (at_right || i == 0) ? NULL : nk_at(m, i - 1, j);
    const nk_cell *M = (i * j == 0) ? NULL : nk_at(m, i - 1, j - 1);
    *x = nk_cell_update(s1, s2, c, go, i, j, at_left, at_right, L, R, M);
    (*cells)++;
    if (getcost) {
        return !(x->CD == old.CD && x->ED == old.ED && x->Q == old.Q && x->P == old.P &&
                 x->cost == old.cost && x->g1 == old.g1 && x->g2 == old.g2);
    }
    return first ? 1 : 0;
}

/* queue_with_linkedlist.c as a FIFO of column indices */
typedef struct {
    int *v;
    int head, tail;
} nk_q;
static void q_push(nk_q *q, int x) { q->v[q->tail++] = x; }
static int q_empty(const nk_q *q) { return q->head == q->tail; }
static int q_pop(nk_q *q) { return q->v[q->head++]; }
static int q_peek(const nk_q *q) { return q->v[q->head]; }

/* in_new_left_band (newkkonen.c:951-956) */
static int in_new_left_band(int i, int j, int newk, int oldk) {
    return i >= j && (i - j) <= newk && (i - j) >= oldk;
}

/* update_a_row (newkkonen.c:964-1073) */
static void nk_update_a_row(nk_mat *m, const uint8_t *s1, const uint8_t *s2, const orc_cm *c,
                            int go, int i, int startj, int endj, int newk, int oldk, nk_q *prevq,
                            nk_q *thisq, int64_t *cells) {
    if (newk == oldk) {
        for (int j = startj; j <= endj; j++)
            nk_update_a_cell(m, s1, s2, c, go, i, j, oldk, 0, cells);
        return;
    }
    int next_j = startj, follow_prevq = 0, last_updated_j = startj - 1, msb;
    do {
        const int changed = nk_update_a_cell(m, s1, s2, c, go, i, next_j, oldk, 1, cells);
        if (changed && !in_new_left_band(i, next_j, newk, oldk)) q_push(thisq, next_j);
        if (changed) last_updated_j = next_j;
        else if (follow_prevq) last_updated_j = next_j;
        else last_updated_j = -1;
        if (last_updated_j == -1) {
            if (q_empty(prevq)) {
                msb = endj + 1;
                follow_prevq = 0;
            } else {
                msb = q_pop(prevq);
                follow_prevq = 1;
            }
            next_j = msb;
        } else {
            if (q_empty(prevq)) {
                msb = endj + 1;
                follow_prevq = 0;
            } else {
                msb = q_peek(prevq);
                follow_prevq = 1;
            }
            if (last_updated_j + 1 < msb) {
                next_j = last_updated_j + 1;
                follow_prevq = 0;
            } else {
                if (follow_prevq) (void)q_pop(prevq);
                next_j = msb;
            }
        }
    } while (next_j != endj + 1 && last_updated_j != endj);
}

/* ukktest (newkkonen.c:1077-1152) */
static void nk_ukktest(nk_mat *m, const uint8_t *s1, const uint8_t *s2, const orc_cm *c, int go,
                       int p, int *res_cost, int *res_gap, int64_t *cells) {
    const int lenX = m->lenX, lenY = m->lenY;
    const int newk = p >= lenX ? lenX - 1 : p;
    const int oldk = (m->diag_in_use - m->bb) / 2;
    nk_expand(m, newk, oldk);
    /* each row pushes at most its band once, plus the row-0 seed */
    const int qcap = 2 * (lenY + 2);
    nk_q prevq = {malloc(sizeof(int) * qcap), 0, 0};
    nk_q thisq = {malloc(sizeof(int) * qcap), 0, 0};
    for (int i = 0; i < lenX; i++) {
        const int startj = (i - newk) > 0 ? i - newk : 0;
        const int endj = (i + lenY - lenX + newk) < lenY - 1 ? i + lenY - lenX + newk : lenY - 1;
        const int oldendj = (i + lenY - lenX + oldk) < lenY - 1 ? i + lenY - lenX + oldk : lenY - 1;
        if (i == 0) q_push(&prevq, oldendj + 1);
        nk_update_a_row(m, s1, s2, c, go, i, startj, endj, newk, oldk, &prevq, &thisq, cells);
        /* transfer_queue: prevq := thisq, thisq := empty */
        nk_q t = prevq;
        prevq = thisq;
        thisq = t;
        thisq.head = thisq.tail = 0;
    }
    free(prevq.v);
    free(thisq.v);
    const nk_cell *f = nk_at(m, lenX - 1, lenY - 1);
    *res_cost = f->cost;
    *res_gap = f->g1 > f->g2 ? f->g1 : f->g2;
}

/* newkk_algn (newkkonen.c:1361-1444) with init_mat (1219), reset_mat and the
 * row-0 initialisation, then increaseT (1155) iterated. */
static int nk_fill(nk_mat *m, const uint8_t *s1, int lenX, const uint8_t *s2, int lenY,
                   const orc_cm *c, orc_nk_stats *st) {
    const int go = c->gap_open > 0 ? c->gap_open : 0;
    m->lenX = lenX;
    m->lenY = lenY;
    m->bb = lenY - lenX + 1;
    m->k = 0;
    m->diag_in_use = m->bb;
    m->diag_max = m->bb + 2 * (lenX - 1);
    m->doff = malloc(sizeof(int64_t) * (m->diag_max + 1));
    int64_t o = 0;
    for (int d = 0; d < m->diag_max; d++) {
        m->doff[d] = o;
        o += d < m->bb ? lenX : lenX - ((d - m->bb) / 2 + 1);
    }
    m->doff[m->diag_max] = o;
    /* band-only storage: the baseband now, more diagonals as k grows */
    m->used = m->doff[m->bb];
    m->pool = malloc(sizeof(nk_cell) * (size_t)(m->used > 0 ? m->used : 1));
    /* reset_mat: every cost -1 */
    memset(m->pool, 0xff, sizeof(nk_cell) * (size_t)m->doff[m->bb]);
    *nk_at(m, 0, 0) = nk_origin(go);
    for (int j = 1; j < m->bb; j++) *nk_at(m, 0, j) = nk_first_row(s2, c, go, j, nk_at(m, 0, j - 1));
    const int delta = orc_min_non0_cost(c);
    long long T = (long long)m->bb * delta;
    int rounds = 0, cost = 0, gap = 0;
    int64_t cells = 0;
    for (;;) {
        const int p = (int)((T - (lenY - lenX)) / 2);
        nk_ukktest(m, s1, s2, c, go, p, &cost, &gap, &cells);
        rounds++;
        const long long newp = (T * 2 - (lenY - lenX)) / 2;
        if (p > gap || newp - lenY + 1 >= 0) break;
        T *= 2;
    }
    if (st) {
        st->rounds = rounds;
        st->final_T = (int)T;
        st->final_k = m->k;
        st->gap_num = gap;
        st->cells = cells;
    }
    return cost;
}

/* ---------------- every band cell, every round -------------------------- */

typedef struct {
    int lenX, lenY, bb, k;
    nk_cell *a; /* dense lenX x lenY */
} nk_dense;

static nk_cell *nd_at(nk_dense *d, int i, int j) { return d->a + (size_t)i * d->lenY + j; }

static int nk_fill_full(nk_dense *d, const uint8_t *s1, int lenX, const uint8_t *s2, int lenY,
                        const orc_cm *c, orc_nk_stats *st) {
    const int go = c->gap_open > 0 ? c->gap_open : 0;
    d->lenX = lenX;
    d->lenY = lenY;
    d->bb = lenY - lenX + 1;
    d->a = malloc(sizeof(nk_cell) * (size_t)lenX * lenY);
    const int D = lenY - lenX;
    const int delta = orc_min_non0_cost(c);
    long long T = (long long)d->bb * delta;
    int rounds = 0, cost = 0, gap = 0;
    int64_t cells = 0;
    *nd_at(d, 0, 0) = nk_origin(go);
    for (int j = 1; j < d->bb; j++) *nd_at(d, 0, j) = nk_first_row(s2, c, go, j, nd_at(d, 0, j - 1));
    for (;;) {
        const int p = (int)((T - D) / 2);
        const int k = p >= lenX ? lenX - 1 : p;
        d->k = k;
        for (int i = 0; i < lenX; i++) {
            const int lo = i - k > 0 ? i - k : 0;
            const int hi = i + D + k < lenY - 1 ? i + D + k : lenY - 1;
            for (int j = lo; j <= hi; j++) {
                if (i == 0 && j < d->bb) continue; /* initialised above */
                const int at_left = (i - j) == k, at_right = (j - i - d->bb + 1) == k;
                const nk_cell *L = (at_left || j == 0) ? NULL : nd_at(d, i, j - 1);
                const nk_cell *R = (at_right || i == 0) ? NULL : nd_at(d, i - 1, j);
                const nk_cell *M = (i * j == 0) ? NULL : nd_at(d, i - 1, j - 1);
                *nd_at(d, i, j) = nk_cell_update(s1, s2, c, go, i, j, at_left, at_right, L, R, M);
                cells++;
            }
        }
        const nk_cell *f = nd_at(d, lenX - 1, lenY - 1);
        cost = f->cost;
        gap = f->g1 > f->g2 ? f->g1 : f->g2;
        rounds++;
        const long long newp = (T * 2 - D) / 2;
        if (p > gap || newp - lenY + 1 >= 0) break;
        T *= 2;
    }
    if (st) {
        st->rounds = rounds;
        st->final_T = (int)T;
        st->final_k = d->k;
        st->gap_num = gap;
        st->cells = cells;
    }
    return cost;
}

/* ---------------- backtrace_affine -------------------------------------- */

enum { M_TODO, M_DELETE, M_INSERT, M_DIAGONAL, M_ALIGN };

/* backtrace_affine (newkkonen.c:1666-1775) with copy_dir_to_mode (1639) and
 * newkk_follow_deletion_or_insertion_affine (1563); trivial_backtrace (1498)
 * when lenY > 100 lenX.  Writes the aligned pair left to right into r1, r2
 * and returns its length; -1 where the reference fails (an invalid
 * direction). */
typedef const nk_cell *(*nk_get)(void *, int, int);

static int nk_backtrace(nk_get get, void *st, int k, const uint8_t *s1, int lenX,
                        const uint8_t *s2, int lenY, int swaped, uint8_t *r1, uint8_t *r2) {
    const int cap = lenX + lenY;
    uint8_t *t1 = malloc(cap), *t2 = malloc(cap);
    int n = 0;
    if (lenY > NK_MUCH_LONGER * lenX) {
        /* prepends s1[0..], then gaps against s2[0..] */
        for (int i = 0; i < lenX; i++) { t1[n] = s1[i]; t2[n] = ORC_GAP; n++; }
        for (int j = 0; j < lenY; j++) { t1[n] = ORC_GAP; t2[n] = s2[j]; n++; }
    } else {
        int i = lenX - 1, j = lenY - 1, mode = M_TODO;
        while (i >= 0 && j >= 0) {
            /* a cell outside the band is get_ukkcost's failure */
            if (j < i - k || j > i + (lenY - lenX) + k) { n = -1; break; }
            const int dir = get(st, i, j)->dir;
            if (dir == NK_START) {
                if (i != 0 || j != 0) { n = -1; break; }
                t1[n] = ORC_GAP; t2[n] = ORC_GAP; n++;
                i--; j--;
                continue;
            }
            if (mode == M_TODO) {
                const int hi = dir & NK_DO_INSERT, hd = dir & NK_DO_DELETE;
                const int ha = dir & NK_ALIGN, hg = dir & NK_DO_DIAG;
                if (!swaped) {
                    if (hd) mode = M_DELETE;
                    else if (hi) mode = M_INSERT;
                    else if (hg) mode = M_DIAGONAL;
                    else if (ha) mode = M_ALIGN;
                    else { n = -1; break; }
                } else {
                    if (hi) mode = M_INSERT;
                    else if (hd) mode = M_DELETE;
                    else if (hg) mode = M_DIAGONAL;
                    else if (ha) mode = M_ALIGN;
                    else { n = -1; break; }
                }
            } else if (mode == M_DELETE || mode == M_INSERT) {
                const int eq = dir & NK_INS_EQ_DEL;
                if (mode == M_DELETE) {
                    t1[n] = s1[i]; t2[n] = ORC_GAP; n++;
                    i--;
                    if ((dir & NK_END_DELETE) || eq) mode = M_TODO;
                } else {
                    t1[n] = ORC_GAP; t2[n] = s2[j]; n++;
                    j--;
                    if ((dir & NK_END_INSERT) || eq) mode = M_TODO;
                }
            } else if (mode == M_DIAGONAL) {
                if (dir & NK_END_DIAG) mode = M_TODO;
                t1[n] = s1[i]; t2[n] = s2[j]; n++;
                i--; j--;
            } else {
                t1[n] = s1[i] & 15; t2[n] = s2[j] & 15; n++;
                i--; j--;
                mode = M_TODO;
            }
        }
    }
    if (n >= 0)
        for (int q = 0; q < n; q++) {
            r1[q] = t1[n - 1 - q];
            r2[q] = t2[n - 1 - q];
        }
    free(t1);
    free(t2);
    return n;
}

static const nk_cell *get_mat(void *st, int i, int j) { return nk_at((nk_mat *)st, i, j); }
static const nk_cell *get_dense(void *st, int i, int j) { return nd_at((nk_dense *)st, i, j); }

int orc_newkk_align(const uint8_t *s1, int lenX, const uint8_t *s2, int lenY, const orc_cm *cm,
                    int swaped, uint8_t *r1, uint8_t *r2, int *rlen, orc_nk_stats *st) {
    if (lenX < 1 || lenX > lenY) return -1;
    if (st) memset(st, 0, sizeof *st);
    if (lenX * NK_MUCH_LONGER < lenY) {
        if (r1) *rlen = nk_backtrace(NULL, NULL, 0, s1, lenX, s2, lenY, swaped, r1, r2);
        return nk_trivial_cost(s1, lenX, s2, lenY, cm);
    }
    nk_mat m;
    memset(&m, 0, sizeof m);
    const int cost = nk_fill(&m, s1, lenX, s2, lenY, cm, st);
    if (r1) *rlen = nk_backtrace(get_mat, &m, m.k, s1, lenX, s2, lenY, swaped, r1, r2);
    free(m.pool);
    free(m.doff);
    return cost;
}

int orc_newkk_full(const uint8_t *s1, int lenX, const uint8_t *s2, int lenY, const orc_cm *cm,
                   int swaped, uint8_t *r1, uint8_t *r2, int *rlen, orc_nk_stats *st) {
    if (lenX < 1 || lenX > lenY) return -1;
    if (st) memset(st, 0, sizeof *st);
    if (lenX * NK_MUCH_LONGER < lenY) {
        if (r1) *rlen = nk_backtrace(NULL, NULL, 0, s1, lenX, s2, lenY, swaped, r1, r2);
        return nk_trivial_cost(s1, lenX, s2, lenY, cm);
    }
    nk_dense d;
    memset(&d, 0, sizeof d);
    const int cost = nk_fill_full(&d, s1, lenX, s2, lenY, cm, st);
    if (r1) *rlen = nk_backtrace(get_dense, &d, d.k, s1, lenX, s2, lenY, swaped, r1, r2);
    free(d.a);
    return cost;
}

/* The full-matrix values (k = lenX - 1 in one round) of every state at
 * every cell, for tests: out[5 * (i * lenY + j) + {0..4}] = cost, P, Q, ED,
 * CD. */
void orc_newkk_matrix(const uint8_t *s1, int lenX, const uint8_t *s2, int lenY, const orc_cm *cm,
                      int *out) {
    const int go = cm->gap_open > 0 ? cm->gap_open : 0;
    nk_dense d;
    d.lenX = lenX;
    d.lenY = lenY;
    d.bb = lenY - lenX + 1;
    d.k = lenX - 1;
    d.a = malloc(sizeof(nk_cell) * (size_t)lenX * lenY);
    *nd_at(&d, 0, 0) = nk_origin(go);
    for (int j = 1; j < d.bb; j++) *nd_at(&d, 0, j) = nk_first_row(s2, cm, go, j, nd_at(&d, 0, j - 1));
    const int k = d.k, D = lenY - lenX;
    for (int i = 0; i < lenX; i++)
        for (int j = (i - k > 0 ? i - k : 0); j <= (i + D + k < lenY - 1 ? i + D + k : lenY - 1); j++) {
            if (i == 0 && j < d.bb) continue;
            const int at_left = (i - j) == k, at_right = (j - i - d.bb + 1) == k;
            const nk_cell *L = (at_left || j == 0) ? NULL : nd_at(&d, i, j - 1);
            const nk_cell *R = (at_right || i == 0) ? NULL : nd_at(&d, i - 1, j);
            const nk_cell *M = (i * j == 0) ? NULL : nd_at(&d, i - 1, j - 1);
            *nd_at(&d, i, j) = nk_cell_update(s1, s2, cm, go, i, j, at_left, at_right, L, R, M);
        }
    for (int i = 0; i < lenX; i++)
        for (int j = 0; j < lenY; j++) {
            const nk_cell *x = nd_at(&d, i, j);
            int *o = out + 5 * ((size_t)i * lenY + j);
            o[0] = x->cost; o[1] = x->P; o[2] = x->Q; o[3] = x->ED; o[4] = x->CD;
        }
    free(d.a);
}

/* CPU baseline driver: NewkkAlign.align_2 per pair (shorter sequence first,
 * swaped when exchanged, src/sequence.ml:1885-1888) over a blob; returns the
 * sum of the costs, costs[k] per pair and *cells_out the cells evaluated. */
int64_t orc_batch_newkk(int n, const uint8_t *blob, const int64_t *off1, const int *len1,
                        const int64_t *off2, const int *len2, const orc_cm *cm, int *costs,
                        int64_t *cells_out) {
    int64_t sum = 0, cells = 0;
    for (int k = 0; k < n; k++) {
        const uint8_t *a = blob + off1[k], *b = blob + off2[k];
        int la = len1[k], lb = len2[k], sw = 0;
        if (la > lb) {
            const uint8_t *t = a; a = b; b = t;
            const int tl = la; la = lb; lb = tl;
            sw = 1;
        }
        uint8_t *r1 = malloc((size_t)la + lb), *r2 = malloc((size_t)la + lb);
        int rl = 0;
        orc_nk_stats st;
        costs[k] = orc_newkk_align(a, la, b, lb, cm, sw, r1, r2, &rl, &st);
        sum += costs[k];
        cells += st.cells;
        free(r1);
        free(r2);
    }
    if (cells_out) *cells_out = cells;
    return sum;
}

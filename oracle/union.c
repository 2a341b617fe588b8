/*
 * union.c -- CPU restatement of POY 5.1's union merge (src/union.c) and the
 * OCaml Unions constructors around it (src/sequence.ml:2759-2912).
 *
 * TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
 * (poy5_amd/, include/poyhip.h) never links or calls it.
 *
 * A union (struct unionoff, src/union.h) is a sequence right-aligned in a
 * buffer of `cap` symbols plus three offset arrays of `cap` entries (offset,
 * union_c1 / ca_offsets, union_c2 / cb_offsets), their live part also at the
 * right end.  union_CAML_unwrap (src/union.c:297) turns the OCaml record into
 * cursors: `begin` / `end` index the offset array, `position` the sequence,
 * `ca` / `cb` the two child-offset arrays; the merge walks all of them
 * leftwards.  The offsets are 32-bit (the reference's USE_LONG_SEQUENCES
 * layout).
 */
#include "oracle.h"

#include <string.h>

typedef struct {
    uint8_t *s;      /* cap symbols; the sequence is s[cap - len .. cap) */
    int32_t *off, *ca, *cb;
    int cap, len;
    /* cursors (union_CAML_produce, src/union.c:276) */
    int begin, end, pca, pcb, position;
} un_t;

static void un_unwrap(un_t *u) {
    if (u->len == 0) {
        u->begin = u->end = u->cap - 1;
        u->pca = u->pcb = u->cap - 1;
    } else {
        u->begin = u->cap - u->len;
        u->end = u->begin + u->len - 1;
        u->pca = u->pcb = u->cap - 1;
    }
    u->position = u->len - 1;
}

/* seq_get on the union's sequence: index from the sequence's first symbol */
static inline int un_get(const un_t *u, int pos) { return u->s[u->cap - u->len + pos]; }

/* seq_prepend (src/seq.c:155) */
static int un_prepend(un_t *u, int v) {
    if (u->len >= u->cap) return -1;
    u->len++;
    u->s[u->cap - u->len] = (uint8_t)v;
    return 0;
}

/* union_prepend_pair (src/union.c:38) */
static int un_prepend_pair(un_t *a, un_t *b, un_t *c, int or_) {
    int v = or_;
    if (a->position >= 0) {
        v |= un_get(a, a->position);
        a->position--;
        a->end--;
    }
    if (b->position >= 0) {
        v |= un_get(b, b->position);
        b->position--;
        b->end--;
    }
    return un_prepend(c, v);
}

/* union_prepend_item (src/union.c:79) */
static int un_prepend_item(un_t *src, un_t *tgt, int or_) {
    int v = or_;
    if (src->position >= 0) {
        v |= un_get(src, src->position);
        src->position--;
        src->end--;
    }
    return un_prepend(tgt, v);
}

/* union_copy_single (src/union.c:104): the count stored at the source's end
 * cursor is the number of its items to copy across */
static int un_copy_single(un_t *a, un_t *c, int it, int or_, int *err) {
    int i = a->off[a->end];
    it += i;
    while (i > 0) {
        if (un_prepend_item(a, c, or_)) *err = 1;
        i--;
    }
    return it;
}

/* union_copy_non_homologous (src/union.c:118) */
static int un_copy_non_homologous(un_t *au, un_t *bu, un_t *c, int items, int gap, int *err) {
    if (au->position > -1) items = un_copy_single(au, c, items, gap, err);
    if (bu->position > -1) items = un_copy_single(bu, c, items, gap, err);
    return items;
}

/* union_prepend_and_fill_items (src/union.c:136) */
static int un_prepend_and_fill(int interm, un_t *c, int prep, int apos, int bpos) {
    if (c->end - prep < 0 || c->pca - prep < 0 || c->pcb - prep < 0) return -1;
    for (int i = prep; i > 0; i--) {
        c->off[c->end] = i;
        c->end--;
        c->ca[c->pca] = 0;
        c->pca--;
        c->cb[c->pcb] = 0;
        c->pcb--;
    }
    c->ca[c->pca] = apos;
    c->cb[c->pcb] = bpos;
    c->pca--;
    c->pcb--;
    const int r = un_prepend(c, interm);
    c->off[c->end] = 0;
    c->end--;
    return r;
}

/* union_move_left (src/union.c:161) */
static int un_move_left(un_t *a) {
    const int r = un_get(a, a->position);
    a->position--;
    a->end--;
    return r;
}

/* union_merge (src/union.c:170-274).  Returns 0, or -1 where the reference
 * fails ("infinit loop": the aligned sequences end while a union still has
 * items) or would overrun the new union's capacity. */
static int un_merge(const uint8_t *a, const uint8_t *b, const uint8_t *median, int lena,
                    un_t *au, un_t *bu, un_t *c) {
    const int gap = ORC_GAP;
    int items = 0, err = 0;
    int i = lena - 1;
    while (au->begin <= au->end || bu->begin <= bu->end) {
        /* a cursor that left its array reads nothing the reference defines */
        if (au->end < -1 || bu->end < -1 || c->end < 0 || c->pca < 0 || c->pcb < 0) return -1;
        items = un_copy_non_homologous(au, bu, c, items, gap, &err);
        if (err) return -1;
        if (i >= 0) {
            int interm = median[i];
            if (gap == interm) {
                if (i != 0) {
                    if (gap == a[i]) err |= un_prepend_item(bu, c, gap) != 0;
                    else if (gap == b[i]) err |= un_prepend_item(au, c, gap) != 0;
                    else err |= un_prepend_pair(au, bu, c, gap) != 0;
                } else {
                    if (un_prepend_and_fill(gap, c, items, 0, 0)) return -1;
                    au->end--;
                    bu->end--;
                    return 0;
                }
                items += 1;
            } else {
                int apos, bpos;
                if (gap == a[i]) {
                    apos = 0;
                    bpos = bu->position;
                    if (bpos >= 0) interm |= un_move_left(bu);
                } else if (gap == b[i]) {
                    bpos = 0;
                    apos = au->position;
                    if (apos >= 0) interm |= un_move_left(au);
                } else {
                    /* as written: each guard tests the other union's cursor */
                    bpos = bu->position;
                    apos = au->position;
                    if (bpos >= 0) interm |= un_move_left(au);
                    if (apos >= 0) interm |= un_move_left(bu);
                }
                if (un_prepend_and_fill(interm, c, items, apos, bpos)) return -1;
                items = 0;
            }
            if (err) return -1;
            i--;
        } else {
            return -1;
        }
    }
    if (c->len == 0 || gap != c->s[c->cap - c->len]) {
        if (c->end < 0 || c->pca < 0 || c->pcb < 0 || un_prepend(c, gap)) return -1;
        c->ca[c->pca] = 0;
        c->cb[c->pcb] = 0;
        c->off[c->end] = 0;
    }
    return 0;
}

int orc_union_merge(const uint8_t *a, const uint8_t *b, const uint8_t *median, int lena,
                    const uint8_t *ua_s, int ua_len, int ua_cap, const int32_t *ua_off,
                    const uint8_t *ub_s, int ub_len, int ub_cap, const int32_t *ub_off,
                    uint8_t *c_s, int c_cap, int32_t *c_off, int32_t *c_ca, int32_t *c_cb,
                    int *c_len) {
    /* the cursors only read the sources' sequences and offsets */
    un_t au, bu, c;
    memset(&au, 0, sizeof au);
    memset(&bu, 0, sizeof bu);
    memset(&c, 0, sizeof c);
    au.s = (uint8_t *)ua_s;
    au.off = (int32_t *)ua_off;
    au.cap = ua_cap;
    au.len = ua_len;
    bu.s = (uint8_t *)ub_s;
    bu.off = (int32_t *)ub_off;
    bu.cap = ub_cap;
    bu.len = ub_len;
    c.s = c_s;
    c.off = c_off;
    c.ca = c_ca;
    c.cb = c_cb;
    c.cap = c_cap;
    c.len = 0;
    memset(c_s, 0, (size_t)c_cap);
    memset(c_off, 0, sizeof(int32_t) * (size_t)c_cap);
    memset(c_ca, 0, sizeof(int32_t) * (size_t)c_cap);
    memset(c_cb, 0, sizeof(int32_t) * (size_t)c_cap);
    un_unwrap(&au);
    un_unwrap(&bu);
    un_unwrap(&c);
    const int r = un_merge(a, b, median, lena, &au, &bu, &c);
    *c_len = c.len;
    return r;
}

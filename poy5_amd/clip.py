"""``Sequence.Clip`` (src/sequence.ml:3441-3575) and ``Sequence.Unions``'
position lookup (``get_positions``, src/sequence.ml:2869-2904) on the MI355X
library.

Clipped sequences are the terminal fragments the implied alignment aligns
(``ImpliedAlignment.align_2``, src/impliedAlignment.ml:323) and
``DOS.distance`` prices over fragments (src/seqCS.ml:1515): a sequence is
``("DO", s)``, ``("First", s)`` (its left end is a clip) or ``("Last", s)``
(its right end is).  The longer of two clipped sequences loses the symbols
that make it longer, from its clipped end, before the alignment; the
aligned pair gets them back afterwards, and the cost leaves out the clipped
end's leading gap columns.  Alignments run on the GPU through ``Align``;
the rest is the OCaml module's bookkeeping over the aligned columns.
"""
from __future__ import annotations

import numpy as np

from .cost_matrix import GAP
from .sequence import Align, Context, NewkkAlign

KINDS = ("DO", "First", "Last")


def is_empty(s) -> bool:
    s = np.asarray(s, np.uint8)
    return len(s) == 0 or bool(np.all(s == GAP))


def safe_reverse(s):
    """Sequence.safe_reverse (sequence.ml:186): the leading gap stays in
    front, the rest is reversed."""
    s = np.asarray(s, np.uint8)
    if len(s) == 0:
        return s.copy()
    return np.concatenate([s[:1], s[1:][::-1]]).astype(np.uint8)


def count_tip_gaps(seqa, seqb, gap: int = GAP) -> int:
    """Clip.count_tip_gaps (sequence.ml:3443): over the leading columns that
    hold a gap, those where only one side does."""
    acc = 0
    for x, y in zip(np.asarray(seqa), np.asarray(seqb)):
        if x != gap and y != gap:
            return acc
        if not (x == gap and y == gap):
            acc += 1
    raise IndexError("Clip.count_tip_gaps: every column holds a gap")


def correct_distance(ctx: Context, seqa, seqb, cost: int, gap: int = GAP) -> int:
    """Clip.correct_distance (sequence.ml:3455): the cost less the leading
    columns that hold a gap (their substitution-table costs) and one gap
    opening, when those columns cost anything."""
    cm = ctx.cm
    subst = 0
    for x, y in zip(np.asarray(seqa), np.asarray(seqb)):
        if x != gap and y != gap:
            break
        subst += cm.c(int(x), int(y))
    else:
        raise IndexError("Clip.correct_distance: every column holds a gap")
    if subst == 0:
        return cost
    return cost - (subst + cm.gap_open)


def corrected_distance(ctx: Context, missing: int, a, b, use_ukk: bool = False) -> int:
    """Clip.corrected_distance (sequence.ml:3469)."""
    if is_empty(a) or is_empty(b):
        return missing
    if use_ukk:
        sa, sb, cost = NewkkAlign(ctx).align_2(a, b)
    else:
        sa, sb, cost = Align(ctx).align_2(a, b)
    return correct_distance(ctx, sa, sb, cost)


def clip_n_fix(s1, s2):
    """Clip.clip_n_fix (sequence.ml:3489-3539) with fraction 1: the longer
    sequence keeps its first symbol and its last len(shorter) - 1 (its
    `dif` symbols after the first go), and the two fix functions put them
    back into an aligned pair (the longer side's own symbols, gaps on the
    other).  Returns (s1', s2', fix1, fix2, dif)."""
    s1 = np.asarray(s1, np.uint8)
    s2 = np.asarray(s2, np.uint8)

    def aux(a, b):
        if len(a) > len(b):
            maxlen = len(b)
            dif = len(a) - maxlen
            # Sequence.init calls its function from the last index down, so
            # the counter walks a from its end: element 0 is a[0], the rest
            # are a's last maxlen - 1 symbols
            a2 = np.concatenate([a[:1], a[len(a) - maxlen + 1:]]).astype(np.uint8)

            def fix_a(x):
                x = np.asarray(x, np.uint8)
                return np.concatenate([a[:dif + 1], x[1:]]).astype(np.uint8)

            def fix_b(x):
                x = np.asarray(x, np.uint8)
                return np.concatenate([np.full(dif + 1, a[0], np.uint8), x[1:]]).astype(np.uint8)

            return a2, b, fix_a, fix_b, dif

        def ident(x):
            return np.asarray(x, np.uint8)
        return a, b, ident, ident, 0

    if len(s1) >= len(s2):
        return aux(s1, s2)
    a, b, fa, fb, dif = aux(s2, s1)
    return b, a, fb, fa, dif


def align_2(ctx: Context, s1, s2, first_gap: bool = True):
    """Clip.Align.align_2 (sequence.ml:3541-3573) on (kind, sequence) pairs:
    ((kind, s1'), (kind, s2'), cost, clipped length, (kind, s1 aligned),
    (kind, s1 aligned)) -- the last two as the reference returns them."""
    k1, a = s1
    k2, b = s2
    if k1 not in KINDS or k2 not in KINDS:
        raise ValueError("kind must be DO, First or Last")
    al = Align(ctx)

    def algn_and_fix(x, y):
        x2, y2, f1, f2, dif = clip_n_fix(x, y)
        p1, p2, cost = al.align_2(x2, y2, first_gap=first_gap)
        cost = correct_distance(ctx, p1, p2, cost)
        tip = count_tip_gaps(p1, p2)
        return f1(p1), f2(p2), cost, dif + tip, p1, p1

    if k1 == "DO" and k2 == "DO":
        p1, p2, cost = al.align_2(a, b, first_gap=first_gap)
        return ("DO", p1), ("DO", p2), cost, 0, ("DO", p1), ("DO", p2)
    if "Last" in (k1, k2) and "First" not in (k1, k2):
        x, y, cost, dif, xo, yo = algn_and_fix(safe_reverse(a), safe_reverse(b))
        return (("Last", safe_reverse(x)), ("Last", safe_reverse(y)), cost, dif,
                ("Last", safe_reverse(xo)), ("Last", safe_reverse(yo)))
    if "First" in (k1, k2) and "Last" not in (k1, k2):
        x, y, cost, dif, xo, yo = algn_and_fix(a, b)
        return ("First", x), ("First", y), cost, dif, ("First", xo), ("First", yo)
    raise AssertionError("a First and a Last fragment cannot be aligned")


def get_positions(u, positions):
    """Unions.get_positions (sequence.ml:2869-2904): for each (x, y) pair of
    positions in the union's sequence, the positions its children hold
    there (union_c1 and union_c2), searching outwards past entries of 0;
    returns (list for c1, list for c2)."""
    ln = len(u.offset)
    seq_len = u.length
    last = len(positions) - 1

    def finder(step, arr, x):
        while True:
            if x == 0:
                return int(arr[0])
            if x == ln:
                step = -1
                x -= 1
                continue
            if x < 0 or (x > ln and step > 0):
                # the reference's bigarray access fails, or its search
                # never ends
                raise IndexError(f"Unions.get_positions: position {x} outside {ln}")
            if x < ln and arr[x] != 0:
                return int(arr[x])
            x += step

    def one(arr, k, xy):
        x, y = xy
        xs = 1 if k == last else -1
        ys = -1 if k == 0 else 1
        return (finder(xs, arr, ln - seq_len + x), finder(ys, arr, ln - seq_len + y))

    return ([one(u.c1, k, p) for k, p in enumerate(positions)],
            [one(u.c2, k, p) for k, p in enumerate(positions)])

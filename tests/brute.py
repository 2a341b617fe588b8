"""Brute-force enumeration of pairwise alignments (tiny inputs only).

Independent of the oracle's recurrences: every alignment of two sequences is
enumerated as a column list and priced column by column, so a dropped term,
a transposed operand or a wrong tie order in the DP fails these checks.

Sequences carry POY's leading gap (code 16) at index 0; the enumeration aligns
s[1:] with t[1:] after the fixed (16, 16) column.
"""
from __future__ import annotations

from functools import lru_cache

GAP = 16


def _idx(a, b):
    return (int(a) << 5) + int(b)


def all_alignments(n1, n2):
    """Yield move strings over {'A','D','I'} aligning n1 and n2 symbols.
    D consumes s1 only, I consumes s2 only."""
    @lru_cache(None)
    def rec(i, j):
        if i == 0 and j == 0:
            return [""]
        out = []
        if i > 0 and j > 0:
            out += [m + "A" for m in rec(i - 1, j - 1)]
        if i > 0:
            out += [m + "D" for m in rec(i - 1, j)]
        if j > 0:
            out += [m + "I" for m in rec(i, j - 1)]
        return out
    return rec(n1, n2)


def columns(s1, s2, moves):
    """Aligned sequences (leading (16,16) column included)."""
    r1, r2 = [s1[0]], [s2[0]]
    i = j = 1
    for m in moves:
        if m == "A":
            r1.append(s1[i]); r2.append(s2[j]); i += 1; j += 1
        elif m == "D":
            r1.append(s1[i]); r2.append(GAP); i += 1
        else:
            r1.append(GAP); r2.append(s2[j]); j += 1
    return r1, r2


def linear_cost(s1, s2, moves, cost):
    r1, r2 = columns(s1, s2, moves)
    return sum(cost[_idx(a, b)] for a, b in zip(r1[1:], r2[1:]))


def in_band(moves, band):
    """band(i) -> (startj, endj) for rows i >= 1 of the DP grid."""
    i = j = 0
    for m in moves:
        if m in "AD":
            i += 1
        if m in "AI":
            j += 1
        if i >= 1:
            lo, hi = band(i)
            if not (lo <= j <= hi):
                return False
    return True


def best_linear(s1, s2, cost, swaped, band=None):
    """Optimal cost and the optimal alignment POY's backtrace_2d selects:
    the lexicographically smallest reversed move string, ALIGN first, then
    DELETE before INSERT (unswapped) or INSERT before DELETE (swapped)."""
    order = {"A": 0, "D": 1, "I": 2} if not swaped else {"A": 0, "I": 1, "D": 2}
    best = None
    pick = None
    for mv in all_alignments(len(s1) - 1, len(s2) - 1):
        if band is not None and not in_band(mv, band):
            continue
        c = linear_cost(s1, s2, mv, cost)
        key = [order[m] for m in reversed(mv)]
        if best is None or c < best or (c == best and key < pick[0]):
            best, pick = c, (key, mv)
    return best, columns(s1, s2, pick[1])


def affine_cost(s1, s2, moves, sub, go, ge_of):
    """Textbook affine: substitution per aligned column plus go + sum of
    per-symbol extension costs for each maximal run of one gap type."""
    r1, r2 = columns(s1, s2, moves)
    total = 0
    prev = None
    for m, a, b in zip(moves, r1[1:], r2[1:]):
        if m == "A":
            total += sub(a, b)
        elif m == "D":
            total += ge_of(a) + (go if prev != "D" else 0)
        else:
            total += ge_of(b) + (go if prev != "I" else 0)
        prev = m
    return total


def best_affine(s1, s2, sub, go, ge_of):
    best = None
    for mv in all_alignments(len(s1) - 1, len(s2) - 1):
        c = affine_cost(s1, s2, mv, sub, go, ge_of)
        if best is None or c < best:
            best = c
    return best


def affine_cost_of_aligned(r1, r2, sub, go, ge_of):
    """Price an aligned pair (leading column skipped) with the textbook model."""
    r1 = [int(x) for x in r1]
    r2 = [int(x) for x in r2]
    moves = []
    for a, b in zip(r1[1:], r2[1:]):
        if a == GAP and b != GAP:
            moves.append("I")
        elif b == GAP and a != GAP:
            moves.append("D")
        else:
            moves.append("A")
    total, prev = 0, None
    for m, a, b in zip(moves, r1[1:], r2[1:]):
        if m == "A":
            total += sub(a, b)
        elif m == "D":
            total += ge_of(a) + (go if prev != "D" else 0)
        else:
            total += ge_of(b) + (go if prev != "I" else 0)
        prev = m
    return total

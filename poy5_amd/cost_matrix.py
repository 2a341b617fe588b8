"""Host-side mirror of POY's ``Cost_matrix.Two_D`` for DNA with combinations.

POY builds its 31x31 ambiguity-aware tables in OCaml and pushes them into the
C ``struct cm`` one entry at a time (``src/cost_matrix.ml:1140-1189``).  This
module builds the same tables and installs them on a device context through
``poyhip_set_cm``.  Codes are 5-bit sets over (A, C, G, T, gap) = bits 0..4;
a table entry for codes (a, b) lives at ``(a << 5) + b``.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib

GAP = 16
LCM = 5
SINGLES = (1, 2, 4, 8, 16)


def _bits(code: int):
    return [s for s in SINGLES if code & s]


@dataclass
class TwoD:
    """Tables of one 2D cost model (the struct cm fields on the DNA path)."""
    cost: np.ndarray      # int32[2048]
    worst: np.ndarray     # int32[2048]
    median: np.ndarray    # uint8[2048]
    prepend: np.ndarray   # int32[64]
    tail: np.ndarray      # int32[64]
    cost_model: int       # 0 linear, 1 affine
    gap_open: int

    def c(self, a: int, b: int) -> int:
        return int(self.cost[(a << LCM) + b])

    def med(self, a: int, b: int) -> int:
        return int(self.median[(a << LCM) + b])

    def min_non0_cost(self) -> int:
        """cm_get_min_non0_cost (src/cm.c:1064)."""
        pos = self.cost[self.cost > 0]
        return int(pos.min()) if pos.size else (2 ** 31 - 1) // 2

    def to_c(self) -> _lib.PoyhipCM:
        s = _lib.PoyhipCM()
        s.lcm, s.gap, s.a_sz, s.combinations = LCM, GAP, 31, 1
        s.cost_model, s.gap_open = self.cost_model, self.gap_open
        s.cost[:] = [int(x) for x in self.cost]
        s.worst[:] = [int(x) for x in self.worst]
        s.median[:] = [int(x) for x in self.median]
        s.prepend[:] = [int(x) for x in self.prepend]
        s.tail[:] = [int(x) for x in self.tail]
        return s


def of_transformations_and_gaps(trans: int, gaps: int, gap_open: int | None = None) -> TwoD:
    """``Cost_matrix.Two_D.of_transformations_and_gaps`` (cost_matrix.ml:1337)
    for the 5-letter DNA alphabet, followed by ``set_cost_model (Affine
    gap_open)`` (cost_matrix.ml:1003) when gap_open is given."""
    base = np.zeros((5, 5), dtype=np.int64)
    for x in range(5):
        for y in range(5):
            if x != y:
                base[x, y] = gaps if 4 in (x, y) else trans
    single = {1 << x: x for x in range(5)}

    def c0(a, b):  # cost between single states
        return int(base[single[a], single[b]])

    cost = np.zeros(2048, dtype=np.int32)
    worst = np.zeros(2048, dtype=np.int32)
    median = np.zeros(2048, dtype=np.uint8)
    for a in SINGLES:
        for b in SINGLES:
            cost[(a << LCM) + b] = c0(a, b)
    # fill_best_cost_and_median_for_all_combinations (cost_matrix.ml:862):
    # sets take the cheapest path a -> v -> b over single states v, and the
    # median is the union of every v on a cheapest path; singleton pairs keep
    # their cost and take the union of the two states as median.
    for i in range(1, 32):
        li = _bits(i)
        for j in range(1, 32):
            lj = _bits(j)
            trip = [(c0(a, v) + c0(v, b), v) for a in li for b in lj for v in SINGLES]
            best = min(t for t, _ in trip)
            med = 0
            for t, v in trip:
                if t == best:
                    med |= v
            w = max(c0(a, b) for a in li for b in lj)
            k = (i << LCM) + j
            if len(li) == 1 and len(lj) == 1:
                median[k] = i | j
            else:
                cost[k] = best
                median[k] = med
            worst[k] = w
    # fill_default_prepend_tail (cost_matrix.ml:994)
    prepend = np.zeros(64, dtype=np.int32)
    tail = np.zeros(64, dtype=np.int32)
    for i in range(1, 32):
        tail[i] = cost[(i << LCM) + GAP]
        prepend[i] = cost[(GAP << LCM) + i]
    if gap_open is None:
        return TwoD(cost, worst, median, prepend, tail, 0, 0)
    # fill_best_cost_and_median_for_all_combinations_bitwise
    # (cost_matrix.ml:721) on a snapshot of the linear tables
    old = cost.copy()

    def oc(a, b):
        return int(old[(a << LCM) + b])

    def pair_choice(x, y):
        """aux_find_best_combination's price and median of one state pair."""
        c1 = oc(x, x) + oc(x, y)
        c2 = oc(x, y) + oc(y, y)
        if c1 == c2:
            return c1, x | y
        return (c2, y) if c1 > c2 else (c1, x)

    for i in range(1, 32):
        li = _bits(i)
        for j in range(1, 32):
            lj = _bits(j)
            k = (i << LCM) + j
            if len(li) == 1 and len(lj) == 1:
                if i == j:
                    v = oc(i, i) * 2
                    median[k], cost[k], worst[k] = i, v, v
                else:
                    c1 = oc(i, j) + oc(j, j)
                    c2 = oc(i, i) + oc(i, j)
                    if c1 == c2:
                        v, m = c1, i | j
                    elif c1 > c2:
                        v, m = c2, i
                    else:
                        v, m = c1, j
                    median[k], cost[k], worst[k] = m, v, v
            else:
                prices = [pair_choice(x, y) for x in lj for y in li] + \
                         [pair_choice(x, y) for x in li for y in lj]
                best = min(p for p, _ in prices)
                m = 0
                for p, mm in prices:
                    if p == best:
                        m |= mm
                median[k], cost[k], worst[k] = m, best, max(p for p, _ in prices)
    return TwoD(cost, worst, median, prepend, tail, 1, int(gap_open))


def default_nucleotides() -> TwoD:
    """``Cost_matrix.Two_D.default`` (cost_matrix.ml:1297): unit costs."""
    return of_transformations_and_gaps(1, 1)


@dataclass
class ThreeD:
    """Tables of ``Cost_matrix.Three_D`` for 5-bit codes: entry (a, b, c) at
    ``((a << 5) + b) << 5 + c``."""
    cost: np.ndarray      # int32[32768]
    median: np.ndarray    # uint8[32768]

    def to_c(self) -> _lib.PoyhipCM3:
        s = _lib.PoyhipCM3()
        s.cost[:] = [int(x) for x in self.cost]
        s.median[:] = [int(x) for x in self.median]
        return s


def is_metric(m: TwoD) -> bool:
    """Triangle inequality over the five single states."""
    for a in SINGLES:
        for b in SINGLES:
            for c in SINGLES:
                if m.c(a, c) > m.c(a, b) + m.c(b, c):
                    return False
    return True


def three_d_of_two_dim_comb(m: TwoD) -> ThreeD:
    """``Three_D.of_two_dim_comb`` (cost_matrix.ml:1605-1651): for every
    triple of codes, the cheapest single state l (a "Steiner" state of the
    three) prices c(l, i) + c(l, j) + c(l, k); the median is the lowest tied
    state (pick_bit).  For a non-metric 2D model the gap state is only
    allowed when at least two of the three codes contain it."""
    metric = is_metric(m)
    cost = np.zeros(32768, dtype=np.int32)
    median = np.zeros(32768, dtype=np.uint8)
    big = 2 ** 62
    for i in range(1, 32):
        for j in range(1, 32):
            for k in range(1, 32):
                best, med = big, 0
                for inter in SINGLES:
                    shared = (i & inter != 0) + (j & inter != 0) + (k & inter != 0)
                    if metric or shared >= 2 or inter != GAP:
                        c = m.c(inter, i) + m.c(inter, j) + m.c(inter, k)
                    else:
                        c = big
                    if c < best:
                        best, med = c, inter
                    elif c == best:
                        med |= inter
                idx = (((i << LCM) + j) << LCM) + k
                cost[idx] = best if best < big else np.iinfo(np.int32).max
                median[idx] = med & -med
    return ThreeD(cost, median)

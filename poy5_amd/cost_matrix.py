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
    # set_metric (cost_matrix.ml:1166): the input list was positive,
    # symmetric and zero on the diagonal (input_is_metric, 1117); the 3D
    # builder reads this flag, not the triangle inequality
    metric: bool = True

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


@dataclass
class NoComb:
    """Tables of an alphabet without combinations (``use_comb`` false): codes
    1..a_sz index single states (gap last), entry (a, b) at a * (a_sz + 1)
    + b.  Host only: the device kernels take 5-bit state sets."""
    a_sz: int
    cost: np.ndarray      # int64[(a_sz + 1)^2]
    worst: np.ndarray
    median: np.ndarray
    prepend: np.ndarray   # int64[a_sz + 1]
    tail: np.ndarray

    def c(self, a: int, b: int) -> int:
        return int(self.cost[a * (self.a_sz + 1) + b])

    def med(self, a: int, b: int) -> int:
        return int(self.median[a * (self.a_sz + 1) + b])


def input_is_metric(l) -> tuple[bool, bool]:
    """input_is_metric (cost_matrix.ml:1117-1138): (positive, symmetric and
    zero on the diagonal; zero on the diagonal)."""
    a = np.asarray(l, dtype=np.int64)
    if np.any(a < 0):
        raise ValueError("a cost matrix has no negative entries")  # assert ispos
    iside = bool(np.all(np.diag(a) == 0))
    return bool(np.array_equal(a, a.T)) and iside, iside


def _prepend_tail(cost):
    """fill_default_prepend_tail (cost_matrix.ml:994)."""
    prepend = np.zeros(64, dtype=np.int32)
    tail = np.zeros(64, dtype=np.int32)
    for i in range(1, 32):
        tail[i] = cost[(i << LCM) + GAP]
        prepend[i] = cost[(GAP << LCM) + i]
    return prepend, tail


def _all_combinations(base):
    """fill_best_cost_and_median_for_all_combinations (cost_matrix.ml:862):
    sets take the cheapest path a -> v -> b over single states v
    (test_combinations, 479-505), the median is the union of every v on a
    cheapest path; singleton pairs keep their cost and take the union of the
    two states as median.  base[x][y]: the input cost of states 1<<x, 1<<y."""
    single = {1 << x: x for x in range(5)}

    def c0(a, b):
        return int(base[single[a]][single[b]])

    cost = np.zeros(2048, dtype=np.int32)
    worst = np.zeros(2048, dtype=np.int32)
    median = np.zeros(2048, dtype=np.uint8)
    for a in SINGLES:
        for b in SINGLES:
            cost[(a << LCM) + b] = c0(a, b)
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
    return cost, worst, median


def _bitwise(old, create_original):
    """fill_best_cost_and_median_for_all_combinations_bitwise
    (cost_matrix.ml:721-805) on a snapshot `old` of the singleton costs.
    A pair of states (x, y) costs old(x, x) + old(x, y) through x or
    old(x, y) + old(y, y) through y (create_original takes the diagonal
    entry back off); a set pair takes the cheapest of its state pairs, both
    orders, and the union of the tied medians."""
    cost = np.zeros(2048, dtype=np.int32)
    worst = np.zeros(2048, dtype=np.int32)
    median = np.zeros(2048, dtype=np.uint8)

    def oc(a, b):
        return int(old[(a << LCM) + b])

    def pair_choice(x, y):
        c1 = oc(x, x) + oc(x, y)
        c2 = oc(x, y) + oc(y, y)
        if c1 == c2:
            return (c1 - oc(x, x) if create_original else c1), x | y
        if c1 > c2:
            return (c2 - oc(y, y) if create_original else c2), y
        return (c1 - oc(x, x) if create_original else c1), x

    for i in range(1, 32):
        li = _bits(i)
        for j in range(1, 32):
            lj = _bits(j)
            k = (i << LCM) + j
            if len(li) == 1 and len(lj) == 1:
                if i == j:
                    v = oc(i, i) if create_original else oc(i, i) * 2
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
                    if create_original:
                        v = oc(i, j)
                    median[k], cost[k], worst[k] = m, v, v
            else:
                # process l2 l1 first, then l1 l2 (find_best_combination)
                prices = [pair_choice(x, y) for x in lj for y in li] + \
                         [pair_choice(x, y) for x in li for y in lj]
                best = min(p for p, _ in prices)
                m = 0
                for p, mm in prices:
                    if p == best:
                        m |= mm
                median[k], cost[k], worst[k] = m, best, max(p for p, _ in prices)
    return cost, worst, median


def _fill_medians(a_sz, cost, all_elements, tie_breaker):
    """fill_medians (cost_matrix.ml:899-991) for an alphabet without
    combinations: the median of (i, j) is a state k minimising cost(i, k) +
    cost(k, j), from the tied list built by prepending (so `first` is the
    largest tied k); the all-elements row and column come from the other
    entries as the reference writes them (its column scan never lowers its
    running minimum, so it keeps the last row's list)."""
    n = a_sz + 1
    mat = [[[] for _ in range(n)] for _ in range(n)]

    def c(a, b):
        return int(cost[a * n + b])

    for i in range(1, a_sz + 1):
        all_min, all_cst = [], None
        for j in range(1, a_sz + 1):
            if i == all_elements or j == all_elements:
                continue
            best, res = None, []
            for k in range(1, a_sz + 1):
                if k == all_elements:
                    continue
                v = c(k, j) + c(i, k)
                if best is None or v < best:
                    best, res = v, [k]
                elif v == best:
                    res = [k] + res
            if all_cst is None or best < all_cst:
                all_cst, all_min = best, res
            elif best == all_cst:
                all_min = res + all_min
            mat[i][j] = res
        if all_elements >= 0:
            mat[i][all_elements] = all_min
    if all_elements >= 0:
        for j in range(1, a_sz + 1):
            all_min = []
            for i in range(1, a_sz + 1):
                if i != all_elements:
                    all_min = mat[i][j]  # cst < max_int always holds
            mat[all_elements][j] = all_min
        all_min = []
        for i in range(1, a_sz + 1):
            if i != all_elements:
                all_min = mat[i][all_elements]
        mat[all_elements][all_elements] = all_min
    median = np.zeros(n * n, dtype=np.int64)
    for a in range(n):
        for b in range(n):
            x = mat[a][b]
            if not x:
                continue
            if tie_breaker == "first":
                median[a * n + b] = x[0]
            elif tie_breaker == "last":
                median[a * n + b] = x[-1]
            else:
                raise ValueError("tie_breaker is 'first' or 'last' (Keep_Random draws from "
                                 "OCaml's Random)")
    return median


def of_list(l, use_comb: bool = True, all_elements: int = -1, level: int = 0,
            tie_breaker: str = "first"):
    """``Cost_matrix.Two_D.of_list`` (cost_matrix.ml:1259-1272) ->
    ``fill_cost_matrix ~create_original:true`` (1140-1189).

    l: the a_sz x a_sz input costs, gap state last.  With combinations
    (a_sz = 5, DNA): the singleton costs are stored (329-344); a metric input
    (positive, symmetric, zero diagonal) fills the sets through
    test_combinations, any other through the bitwise builder with
    create_original.  Without combinations (use_comb false, or level 1 as
    of_channel maps it): the input costs, the all-elements row and column
    (363-395) and fill_medians' tie_breaker choice (`first` or `last`).
    Combination levels above 1 (fill_best_cost_and_median_for_some_
    combinations, 807-860) renumber the states and are not built."""
    a = np.asarray(l, dtype=np.int64)
    if a.ndim == 1:
        w = int(round(len(a) ** 0.5))
        a = a.reshape(w, w)
    a_sz = a.shape[0]
    if level == 1:
        use_comb = False  # of_channel (1191)
    if use_comb and level > 1:
        raise NotImplementedError("combination levels renumber the states; not built")
    metric, _ = input_is_metric(a)
    if use_comb:
        if a_sz != 5:
            raise ValueError("combinations are built for the five DNA states")
        if metric:
            cost, worst, median = _all_combinations(a)
        else:
            old = np.zeros(2048, dtype=np.int32)
            for x in range(5):
                for y in range(5):
                    old[((1 << x) << LCM) + (1 << y)] = a[x, y]
            cost, worst, median = _bitwise(old, True)
        prepend, tail = _prepend_tail(cost)
        return TwoD(cost, worst, median, prepend, tail, 0, 0, metric)
    n = a_sz + 1
    cost = np.zeros(n * n, dtype=np.int64)
    for i in range(1, a_sz + 1):
        for j in range(1, a_sz + 1):
            cost[i * n + j] = 0 if all_elements in (i, j) else a[i - 1, j - 1]
    worst = cost.copy()
    if all_elements >= 0:
        for i in range(1, a_sz + 1):
            if i != all_elements:
                cost[i * n + all_elements] = min(cost[i * n + j] for j in range(1, a_sz + 1)
                                                 if j != all_elements)
        all_min = None
        for j in range(1, a_sz + 1):
            if j == all_elements:
                continue
            mc = min(cost[i * n + j] for i in range(1, a_sz + 1) if i != all_elements)
            all_min = mc if all_min is None else min(all_min, mc)
            cost[all_elements * n + j] = mc
        cost[all_elements * n + all_elements] = all_min
    median = _fill_medians(a_sz, cost, all_elements, tie_breaker)
    gap = a_sz
    prepend = np.zeros(n, dtype=np.int64)
    tail = np.zeros(n, dtype=np.int64)
    for i in range(1, a_sz + 1):
        tail[i] = cost[i * n + gap]
        prepend[i] = cost[gap * n + i]
    return NoComb(a_sz, cost, worst, median, prepend, tail)


def of_transformations_and_gaps(trans: int, gaps: int, gap_open: int | None = None) -> TwoD:
    """``Cost_matrix.Two_D.of_transformations_and_gaps`` (cost_matrix.ml:1337)
    for the 5-letter DNA alphabet, followed by ``set_cost_model (Affine
    gap_open)`` (cost_matrix.ml:1003) when gap_open is given."""
    base = [[0 if x == y else (gaps if 4 in (x, y) else trans) for y in range(5)]
            for x in range(5)]
    m = of_list(base)
    if gap_open is None:
        return m
    # set_cost_model (Affine g) refills the sets bitwise on a snapshot of the
    # linear tables (1003-1016), prepend and tail staying
    cost, worst, median = _bitwise(m.cost.copy(), False)
    return TwoD(cost, worst, median, m.prepend, m.tail, 1, int(gap_open), m.metric)


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
    """``Two_D.is_metric`` (cost_matrix.ml:83): the builder's metric flag."""
    return bool(m.metric)


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

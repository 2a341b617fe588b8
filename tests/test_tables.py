"""Cost tables (CPU): the oracle's and the product's tables against closed
forms the text of src/cost_matrix.ml fixes, and against each other.

Closed forms used (src/cost_matrix.ml:1337-1348 builds the 5x5 tcm;
862-897 fills sets for the linear model; 721-804 refills them bitwise after
set_cost_model (Affine g), 1003-1016; 994 fills prepend/tail; 1605-1651
builds the 3D tables):
  * singletons: cost(x, x) = 0, cost(base, base') = trans, cost(x, gap) =
    gaps, median(x, y) = x | y when trans < 2 * gaps;
  * metric sets: cost(I, J) = min over a in I, b in J of cost(a, b);
  * prepend[x] = cost(gap, x), tail[x] = cost(x, gap), delta = min positive;
  * 3D: cost3 is symmetric in its three codes, cost3(a, a, b) = cost(a, b)
    for singletons, and the hand-derived entries below.
"""
import itertools

import numpy as np
import pytest

import oracle
from poy5_amd import cost_matrix

BASES = (1, 2, 4, 8)
SINGLES = (1, 2, 4, 8, 16)
A, C, G, T, GAP = 1, 2, 4, 8, 16


def bits(x):
    return [s for s in SINGLES if x & s]


def arr(cm, name):
    return np.array(getattr(cm, name))


@pytest.mark.parametrize("trans,gaps", [(1, 2), (1, 1), (2, 3), (3, 2)])
def test_linear_singletons(trans, gaps):
    cm = oracle.cm_dna(trans, gaps)
    cost, med = arr(cm, "cost"), arr(cm, "median")
    for a in SINGLES:
        for b in SINGLES:
            k = (a << 5) + b
            want = 0 if a == b else (gaps if GAP in (a, b) else trans)
            assert cost[k] == want, (a, b)
            if a == b:
                assert med[k] == a
            elif a != GAP and b != GAP and trans < 2 * gaps:
                assert med[k] == a | b


@pytest.mark.parametrize("trans,gaps", [(1, 2), (1, 1), (2, 3)])
def test_linear_metric_sets(trans, gaps):
    cm = oracle.cm_dna(trans, gaps)
    cost = arr(cm, "cost")
    for i in range(1, 32):
        for j in range(1, 32):
            want = min(cost[(a << 5) + b] for a in bits(i) for b in bits(j))
            assert cost[(i << 5) + j] == want, (i, j)


@pytest.mark.parametrize("args", [(1, 2, None), (1, 1, None), (1, 1, 3), (2, 1, 2)])
def test_prepend_tail_delta(args):
    cm = oracle.cm_dna(args[0], args[1], args[2])
    cost, pre, tail = arr(cm, "cost"), arr(cm, "prepend"), arr(cm, "tail")
    for x in range(1, 32):
        assert pre[x] == cost[(GAP << 5) + x]
        assert tail[x] == cost[(x << 5) + GAP]
    pos = cost[cost > 0]
    assert oracle.min_non0_cost(cm) == pos.min()


def test_affine_unit_hand_entries():
    """Bitwise refill with tcm (1,1), gap_open 3 (the affine workload)."""
    cm = oracle.cm_dna(1, 1, 3)
    cost, med = arr(cm, "cost"), arr(cm, "median")
    assert cm.cost_model == 1 and cm.gap_open == 3
    assert cost[(A << 5) + A] == 0 and med[(A << 5) + A] == A
    assert cost[(A << 5) + C] == 1 and med[(A << 5) + C] == A | C
    assert cost[(A << 5) + GAP] == 1 and med[(A << 5) + GAP] == A | GAP
    # A vs the set {A,C}: the shared state costs nothing
    assert cost[(A << 5) + (A | C)] == 0 and med[(A << 5) + (A | C)] == A
    # {A,C} vs {G,T}: one substitution, every state pair ties
    assert cost[((A | C) << 5) + (G | T)] == 1
    assert med[((A | C) << 5) + (G | T)] == A | C | G | T


def test_linear_hand_set_medians_tcm12():
    """Hand-derived set entries of tcm (1, 2), linear model: for sets I, J
    the cost is the least cost(a, v) + cost(v, b) over a in I, b in J and a
    state v, and the median is the union of every v reaching it
    (test_combinations, src/cost_matrix.ml:479-505, filled at 862-897; no
    cleanup for the linear model, 700-704).  With substitution 1 and gap 2
    a base pair (a, b) reaches 1 through v = a or v = b, a shared state
    reaches 0 through itself, and (x, gap) reaches 2 through x or gap."""
    entries = [
        (A, C | G, 1, A | C | G),
        (A | C, A | G, 0, A),
        (A | GAP, C, 1, A | C),
        (A | GAP, C | GAP, 0, GAP),
        (A | C, G | T, 1, A | C | G | T),
        (A | C | G, T, 1, A | C | G | T),
        (C, G | T | GAP, 1, C | G | T),
        (A | C, GAP, 2, A | C | GAP),
    ]
    for cm in (oracle.cm_dna(1, 2), cost_matrix.of_transformations_and_gaps(1, 2)):
        cost, med = np.asarray(cm.cost), np.asarray(cm.median)
        for i, j, c, m in entries:
            assert cost[(i << 5) + j] == c, (i, j)
            assert med[(i << 5) + j] == m, (i, j)


@pytest.mark.parametrize("args", [(1, 2, None), (1, 1, None), (1, 1, 3), (2, 3, None), (2, 1, 2),
                                  (5, 1, None)])
def test_product_tables_match_oracle(args):
    """poy5_amd.cost_matrix (Python) and the oracle (C) are separate
    restatements of the same OCaml table builders."""
    o = oracle.cm_dna(args[0], args[1], args[2])
    p = cost_matrix.of_transformations_and_gaps(*args)
    for name in ("cost", "median", "worst", "prepend", "tail"):
        po = np.asarray(getattr(p, name)).astype(np.int64)
        oo = arr(o, name).astype(np.int64)
        n = min(len(po), len(oo))
        assert np.array_equal(po[:n], oo[:n]), name
    assert p.cost_model == o.cost_model
    assert p.min_non0_cost() == oracle.min_non0_cost(o)


def test_3d_tables():
    c2 = oracle.cm_dna(1, 1)
    c3 = oracle.cm3_from_2d(c2)
    cost, med = np.array(c3.cost), np.array(c3.median)

    def e(a, b, c):
        k = (((a << 5) + b) << 5) + c
        return int(cost[k]), int(med[k])

    # hand-derived (uniform 1/1): the Steiner state of three leaves
    assert e(A, A, A) == (0, A)
    assert e(A, A, C) == (1, A)
    assert e(A, C, G)[0] == 2
    assert e(A, GAP, GAP) == (1, GAP)
    assert e(GAP, GAP, GAP)[0] == 0
    c2c = np.array(c2.cost)
    for a, b, c in itertools.product(range(1, 32), repeat=3):
        if (a * 7 + b * 3 + c) % 5:  # a deterministic subset keeps this quick
            continue
        v = e(a, b, c)[0]
        for p in set(itertools.permutations((a, b, c))):
            assert e(*p)[0] == v
    for a in SINGLES:
        for b in SINGLES:
            assert e(a, a, b)[0] == c2c[(a << 5) + b]


@pytest.mark.parametrize("tcm", [(1, 1), (1, 2), (2, 1), (5, 1)])
def test_product_3d_tables_match_oracle(tcm):
    o = oracle.cm3_from_2d(oracle.cm_dna(*tcm))
    p = cost_matrix.three_d_of_two_dim_comb(cost_matrix.of_transformations_and_gaps(*tcm))
    # the oracle restates the metric branch of of_two_dim_comb only
    assert cost_matrix.is_metric(cost_matrix.of_transformations_and_gaps(*tcm))
    idx = np.array([(((i << 5) + j) << 5) + k for i in range(1, 32) for j in range(1, 32)
                    for k in range(1, 32)])
    assert np.array_equal(np.array(o.cost)[idx], p.cost[idx])
    assert np.array_equal(np.array(o.median)[idx], p.median[idx])


def test_metric_flag_follows_the_input_not_the_triangle_inequality():
    """input_is_metric (src/cost_matrix.ml:1117): positive, symmetric and
    zero on the diagonal.  Substitution 5 / gap 1 breaks the triangle
    inequality (5 > 1 + 1) and is still metric for POY; an asymmetric or a
    non-zero diagonal input is not."""
    assert cost_matrix.of_transformations_and_gaps(5, 1).metric
    base = [[0 if x == y else (2 if 4 in (x, y) else 2) for y in range(5)] for x in range(5)]
    base[0][1], base[1][0] = 1, 3
    assert not cost_matrix.of_list(base).metric
    base = [[0 if x == y else 2 for y in range(5)] for x in range(5)]
    base[0][0] = 1
    assert not cost_matrix.of_list(base).metric


def test_non_metric_bitwise_hand_entries():
    """A non-metric input (A->C 1, C->A 3, every other change 2) fills the
    sets with the bitwise builder and create_original (src/cost_matrix.ml:
    721-805, 1180-1183): a state pair (x, y) is priced old(x, x) + old(x, y)
    through x or old(x, y) + old(y, y) through y, the diagonal entry taken
    back off; singleton pairs keep their input cost."""
    base = [[0 if x == y else 2 for y in range(5)] for x in range(5)]
    base[0][1], base[1][0] = 1, 3
    m = cost_matrix.of_list(base)
    k = lambda a, b: (a << 5) + b  # noqa: E731
    assert m.cost[k(A, C)] == 1 and m.median[k(A, C)] == A | C
    assert m.cost[k(C, A)] == 3 and m.median[k(C, A)] == A | C
    assert m.cost[k(A | C, G)] == 2 and m.median[k(A | C, G)] == A | C | G
    assert m.worst[k(A | C, G)] == 2
    # {A, C} vs A: the pair (A, A) costs 0 through A; the worst pair is C -> A
    assert m.cost[k(A | C, A)] == 0 and m.median[k(A | C, A)] == A
    assert m.worst[k(A | C, A)] == 3
    base = [[0 if x == y else 2 for y in range(5)] for x in range(5)]
    base[0][0] = 1  # a cost between equal states
    m = cost_matrix.of_list(base)
    assert m.cost[k(A, A)] == 1 and m.median[k(A, A)] == A


def test_no_comb_medians_and_tie_breaker():
    """fill_medians (src/cost_matrix.ml:899-991) without combinations: the
    tied states of cost(i, k) + cost(k, j) are collected by prepending, so
    tie_breaker `first` takes the largest tied state and `last` the
    smallest.  States 1, 2 and gap 3 with 1-2 costing 1 and gaps 2: for
    (1, 2) both k = 1 and k = 2 cost 1."""
    base = [[0, 1, 2], [1, 0, 2], [2, 2, 0]]
    first = cost_matrix.of_list(base, use_comb=False, tie_breaker="first")
    last = cost_matrix.of_list(base, use_comb=False, tie_breaker="last")
    assert first.med(1, 2) == 2 and last.med(1, 2) == 1
    assert first.med(1, 1) == 1 and first.med(3, 3) == 3
    assert first.c(1, 3) == 2 and first.tail[1] == 2 and first.prepend[2] == 2
    # level 1 is read as no combinations (of_channel)
    assert cost_matrix.of_list(base, level=1).med(1, 2) == 2


def test_no_comb_all_elements_row():
    """The all-elements state (4 here): its input costs are zeroed, then
    cost(i, 4) = min_j cost(i, j) and cost(4, j) = min_i cost(i, j)
    (src/cost_matrix.ml:363-395); median(i, 4) is the list of i's cheapest
    column; the column scan for median(4, j) never lowers its running
    minimum, so it keeps the last state's list: median(4, 1) is the first
    of mat[3][1], whose ties k = 3 and k = 1 (cost 2 each) were prepended."""
    base = [[0, 1, 2, 7], [1, 0, 2, 7], [2, 2, 0, 7], [7, 7, 7, 7]]
    m = cost_matrix.of_list(base, use_comb=False, all_elements=4)
    assert m.c(1, 4) == 0 and m.c(4, 3) == 0 and m.c(4, 4) == 0
    assert m.med(1, 4) == 1
    assert m.med(4, 1) == 3


def test_combination_levels_are_not_built():
    with pytest.raises(NotImplementedError):
        cost_matrix.of_list([[0] * 5] * 5, level=2)

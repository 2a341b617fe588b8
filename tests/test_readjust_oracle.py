"""Pins of the oracle's Sequence.readjust / DOS.readjust restatement
(src/sequence.ml:2097-2146, src/seqCS.ml:776-952) and of the column map of
Align.closest (Cost_matrix.Two_D.get_closest, src/cost_matrix.ml:1387;
remove_gaps2, src/sequence.ml:209): entries derived by hand from the
definitions, and properties the construction fixes."""
import numpy as np
import pytest

import oracle

A, Cc, G, T, GAP = 1, 2, 4, 8, 16


def test_get_closest_hand_entries():
    cm = oracle.cm_dna(1, 1, 3)  # every substitution 1, every indel 1
    # a member of b costs 0 against a
    assert oracle.get_closest(cm, A, A | G) == A
    # a tie between two states: the lower bit wins (the fold keeps the
    # first strictly cheaper state)
    assert oracle.get_closest(cm, A, Cc | G) == Cc
    # the cheaper state is the higher bit: A|G against C|G picks G (0), not C
    assert oracle.get_closest(cm, A | G, Cc | G) == G
    # a is the gap: b keeps its gap bit, and the gap is cheapest
    assert oracle.get_closest(cm, GAP, A | GAP) == GAP
    # both carry a gap: the gap
    assert oracle.get_closest(cm, A | GAP, G | GAP) == GAP
    # only b carries one: it is dropped
    assert oracle.get_closest(cm, A | Cc, A | GAP) == A
    # b is the gap itself
    assert oracle.get_closest(cm, A, GAP) == GAP
    # the reference's asserts
    assert oracle.get_closest(cm, 0, A) == -1
    assert oracle.get_closest(cm, A, 0) == -1


def test_closest_map_hand_cases():
    cm = oracle.cm_dna(1, 1, 3)
    s1 = np.array([GAP, A, GAP, A | G], np.uint8)
    s2 = np.array([GAP, A | Cc, G, Cc | G], np.uint8)
    # (gap, gap) -> gap (dropped); (A, A|C) -> A; (gap, G) -> G;
    # (A|G, C|G) -> G; remove_gaps2 prepends one gap
    assert list(oracle.closest_map(s1, s2, cm)) == [GAP, A, G, G]
    # equal sequences: columns after the first lose their gap bit, then each
    # symbol's cheapest state against itself, the lowest on ties
    s = np.array([GAP, A | GAP, A | Cc], np.uint8)
    assert list(oracle.closest_map(s, s, cm, equal=True)) == [GAP, A, A]
    # an interior lone gap becomes 0 under the mask: the reference's assert
    with pytest.raises(ValueError):
        oracle.closest_map(np.array([GAP, GAP, A], np.uint8),
                           np.array([GAP, GAP, A], np.uint8), cm, equal=True)


def test_closest_special_branches():
    cm = oracle.cm_dna(1, 2)
    x = np.array([GAP, A, Cc], np.uint8)
    # an all-gap s2 comes back as it is, cost 0
    r, c = oracle.closest(x, np.array([GAP, GAP], np.uint8), cm)
    assert list(r) == [GAP, GAP] and c == 0
    # equal sequences: no alignment, cost 0
    r, c = oracle.closest(x, x, cm)
    assert list(r) == [GAP, A, Cc] and c == 0
    # otherwise the cost of s1 against the result
    r, c = oracle.closest(np.array([GAP, Cc], np.uint8), np.array([GAP, A | Cc], np.uint8), cm)
    assert list(r) == [GAP, Cc] and c == 0


def test_select_one_lowest_state():
    assert list(oracle.select_one([GAP, A | Cc, G | T, A | GAP])) == [GAP, A, G, A]
    with pytest.raises(ValueError):
        oracle.select_one([GAP, 0])


@pytest.mark.parametrize("cm", [oracle.cm_dna(1, 2), oracle.cm_dna(1, 2, 3)])
def test_readjust_hand_case(cm):
    # a = b = [-, A], parent [-, C]; one substitution costs 1, an indel 2:
    # ab = [-, A] (0), bc = ac = [-, A|C] (1); against the third each costs
    # 1, 0, 0, so the sums tie at 1 and make_center takes closest(parent, ab)
    # = the closest states of ab to [-, C] along their alignment = [-, A];
    # the children reach it at 0, the parent at 1
    a = np.array([GAP, A], np.uint8)
    par = np.array([GAP, Cc], np.uint8)
    c3, c2, m, p1, p2, p3 = oracle.seq_readjust(a, a, a, par, cm)
    assert (c3, c2) == (1, 0)
    assert list(m) == [GAP, A] and list(p1) == [GAP, A]
    assert p1 is p2 is p3


def test_readjust_of_three_equal_sequences_is_the_sequence():
    cm = oracle.cm_dna(1, 1, 3)
    x = np.array([GAP, A, Cc, G, T, A, A], np.uint8)
    c3, c2, m, *_ = oracle.seq_readjust(x, x, x, x, cm)
    assert (c3, c2) == (0, 0) and np.array_equal(m, x)


def test_readjust_costs_are_the_alignments_of_the_result():
    # the returned costs are Align.align_2 of each child and the parent
    # against the new median, and no worse than the median's own use
    rng = np.random.default_rng(3)
    for model in ((1, 2), (1, 1, 3)):
        cm = oracle.cm_dna(*model)
        for _ in range(10):
            root = rng.choice([A, Cc, G, T], 30)
            tri = []
            for _ in range(3):
                s = root.copy()
                s[rng.random(30) < 0.2] = rng.choice([A, Cc, G, T])
                keep = rng.random(30) > 0.1
                tri.append(np.concatenate([[GAP], s[keep]]).astype(np.uint8))
            c3, c2, m, *_ = oracle.seq_readjust(tri[0], tri[1], tri[0], tri[2], cm)
            costs = [oracle.align_2_any(x, m, cm)[2] for x in tri]
            assert c2 == costs[0] + costs[1] and c3 == sum(costs)
            assert m[0] == GAP and np.all(m[1:] != GAP)


def _node(seq, c2=0.0, c3=0.0, s=0.0):
    return {"seq": np.asarray(seq, np.uint8), "cost2": c2, "cost2_max": 0.0, "cost3": c3,
            "sum_cost": s}


def test_dos_readjust_branches():
    cm = oracle.cm_dna(1, 2)
    e = [GAP]
    x = [GAP, A, Cc, G]
    y = [GAP, A, G]
    mine = _node(x, 1.0, 2.0, 5.0)
    # two empty neighbours: the remaining one is returned with the
    # children's summed subtree cost; all empty: the parent
    ch, node, c3, c2, s = oracle.dos_readjust(_node(e, s=1.0), _node(y, s=2.0), _node(e), mine, cm)
    assert list(node["seq"]) == y and (c3, c2, s) == (0, 0, 3) and node["sum_cost"] == 3.0
    ch, node, *_ = oracle.dos_readjust(_node(e), _node(e), _node(y), mine, cm)
    assert list(node["seq"]) == y
    ch, node, *_ = oracle.dos_readjust(_node(x), _node(e), _node(e), mine, cm)
    assert list(node["seq"]) == x and ch  # costs differ from mine's
    # empty parent: the lowest states of the children's median, cost2 = cost3
    ch, node, c3, c2, s = oracle.dos_readjust(_node(x), _node(y), _node(e), mine, cm)
    assert c2 == c3 == oracle.align_2_any(x, y, cm)[2]
    assert np.all((node["seq"] & (node["seq"] - 1)) == 0)  # one state each
    # an empty child: the other child's and the parent's median, distances
    # to both
    ch, node, c3, c2, s = oracle.dos_readjust(_node(e), _node(y), _node(x), mine, cm)
    med = node["seq"]
    assert c2 == oracle.dos_distance(y, med, cm, False)
    assert c3 == c2 + oracle.dos_distance(med, x, cm, False)
    # all present: Sequence.readjust's result, unchanged when mine already
    # holds it with its costs
    c3, c2, m, *_ = oracle.seq_readjust(np.array(x, np.uint8), np.array(y, np.uint8), None,
                                        np.array(x, np.uint8), cm)
    again = _node(m, c2, c3, c2)
    ch, node, *_ = oracle.dos_readjust(_node(x), _node(y), _node(x), again, cm)
    assert not ch and np.array_equal(node["seq"], m)

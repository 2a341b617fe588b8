"""Pins of the oracle's union merge (oracle/union.c, restating src/union.c
and Sequence.Unions, src/sequence.ml:2759-2912) by hand-derived cases: two
leaves aligned without gaps, with an inserted column, with a column whose
median is the gap (a non-homologous item, counted in the new union's
offsets), and that union merged again (its count copies the item across)."""
import numpy as np

import oracle


def right(u, k):
    """The last k entries of each array of a union."""
    return (list(u.sequence()), list(u.offset[-k:]), list(u.c1[-k:]), list(u.c2[-k:]))


def test_leaf():
    u = oracle.union_leaf([16, 1, 4])
    assert u.cap == 3 and u.length == 3
    assert list(u.offset) == [0, 0, 0] and list(u.c1) == [0, 1, 2] and list(u.c2) == [0, 1, 2]


def test_two_leaves_without_gaps():
    # column 1: both children move left, the union symbol is the median OR'd
    # with both children's symbols, and the child offsets record positions 1, 1;
    # column 0 (the leading gap, median 16) closes the merge at i = 0
    u = oracle.union_merge([16, 1], [16, 2], [16, 3], oracle.union_leaf([16, 1]),
                           oracle.union_leaf([16, 2]))
    assert u.cap == 6
    assert right(u, 2) == ([16, 3], [0, 0], [0, 1], [0, 1])


def test_inserted_column():
    # a' = -A-, b' = -AG: the last column has a gap in a, so only b moves
    # (position 2, a's offset 0); median 20 | G = 20
    u = oracle.union_merge([16, 1, 16], [16, 1, 4], [16, 1, 20], oracle.union_leaf([16, 1]),
                           oracle.union_leaf([16, 1, 4]))
    assert u.cap == 7
    assert right(u, 3) == ([16, 1, 20], [0, 0, 0], [0, 1, 0], [0, 1, 2])


def test_gap_median_is_a_counted_item_and_is_copied_again():
    # column 1 has median 16 (the gap) with b a gap: a's symbol 17 is
    # prepended as a non-homologous item, and the fill at column 0 writes its
    # count 1 into the offsets
    u = oracle.union_merge([16, 17, 1], [16, 16, 1], [16, 16, 1],
                           oracle.union_leaf([16, 17, 1]), oracle.union_leaf([16, 1]))
    assert right(u, 3) == ([16, 17, 1], [0, 1, 0], [0, 0, 2], [0, 0, 1])
    # merged again with a leaf: the count 1 at the union's cursor copies the
    # item across before column 0 closes the merge
    v = oracle.union_merge([16, 1], [16, 1], [16, 1], u, oracle.union_leaf([16, 1]))
    assert right(v, 3) == ([16, 17, 1], [0, 1, 0], [0, 0, 2], [0, 0, 1])

"""Pins of the oracle's Sequence.Clip restatement (src/sequence.ml:
3441-3575) and Unions.get_positions (2869-2904): cases worked by hand from
the OCaml, chosen so that keeping the longer fragment's head instead of its
tail, forgetting the gap opening, counting two-sided gap columns as tips or
reading the wrong child array would each fail one."""
import numpy as np
import pytest

import oracle

A, Cc, G, T, GAP = 1, 2, 4, 8, 16


def arr(*x):
    return np.array(x, np.uint8)


def test_count_tip_gaps_by_hand():
    # columns (-,-) (-,A) (C,-) (G,G): one-sided gaps 2, the (-,-) skipped
    assert oracle.clip_count_tip_gaps(arr(GAP, GAP, Cc, G), arr(GAP, A, GAP, G)) == 2
    assert oracle.clip_count_tip_gaps(arr(GAP, A), arr(GAP, A)) == 0
    with pytest.raises(IndexError):
        oracle.clip_count_tip_gaps(arr(GAP, GAP), arr(GAP, A))


def test_correct_distance_by_hand():
    lin = oracle.cm_dna(1, 2)
    aff = oracle.cm_dna(1, 1, 3)
    # leading gap columns cost 0 + cost(gap, A) + cost(C, gap) = 2 + 2 under
    # (1, 2); the linear table opens no gap: 10 - 4
    a, b = arr(GAP, GAP, Cc, G), arr(GAP, A, GAP, G)
    assert oracle.clip_correct_distance(a, b, 10, lin) == 6
    # affine (1, 1, 3): 1 + 1 and one gap opening of 3: 10 - 5
    assert oracle.clip_correct_distance(a, b, 10, aff) == 5
    # nothing in the leading columns costs anything: unchanged
    assert oracle.clip_correct_distance(arr(GAP, A), arr(GAP, A), 7, aff) == 7


def test_clip_n_fix_keeps_the_tail_by_hand():
    a = arr(GAP, A, Cc, G, T, A, Cc)  # 7
    b = arr(GAP, A, G, A)             # 4
    s1, s2, f1, f2, dif = oracle.clip_n_fix(a, b)
    # Sequence.init's counter runs from the end: first symbol, then the last 3
    assert list(s1) == [GAP, T, A, Cc] and list(s2) == list(b) and dif == 3
    x, y = arr(GAP, T, A, Cc), arr(GAP, A, G, A)
    assert list(f1(x)) == [GAP, A, Cc, G, T, A, Cc]
    assert list(f2(y)) == [GAP, GAP, GAP, GAP, A, G, A]
    # the shorter first: the results come back in the caller's order
    s1, s2, f1, f2, dif = oracle.clip_n_fix(b, a)
    assert list(s1) == list(b) and list(s2) == [GAP, T, A, Cc] and dif == 3
    assert list(f2(x)) == [GAP, A, Cc, G, T, A, Cc]
    # equal lengths: untouched
    s1, s2, f1, f2, dif = oracle.clip_n_fix(b, b)
    assert dif == 0 and list(f1(b)) == list(b)


def test_clip_align_by_hand():
    cm = oracle.cm_dna(1, 2)
    a = arr(GAP, A, Cc, G, T, A, Cc)
    b = arr(GAP, A, G, A)
    # First: [-, T, A, C] against [-, A, G, A] costs 3 (three substitutions
    # beat any gap at 2), no leading gap column, clipped length 3
    r1, r2, cost, clen, o1, o2 = oracle.clip_align_2(("First", a), ("DO", b), cm)
    assert cost == 3 and clen == 3
    assert r1[0] == "First" and list(r1[1]) == list(a)
    assert list(r2[1]) == [GAP, GAP, GAP, GAP, A, G, A]
    assert list(o1[1]) == [GAP, T, A, Cc] and o2[1] is o1[1]
    # Last: reversed, clipped from the (reversed) front, reversed back
    r1, r2, cost, clen, _, _ = oracle.clip_align_2(("Last", a), ("Last", b), cm)
    assert list(r1[1]) == list(a) and list(r2[1]) == [GAP, A, G, A, GAP, GAP, GAP]
    assert cost == 2 and clen == 3  # [-, A, C, G] vs [-, A, G, A]: A=A, C/G, G/A
    with pytest.raises(AssertionError):
        oracle.clip_align_2(("First", a), ("Last", b), cm)


def test_union_get_positions_by_hand():
    # a union of capacity 6 holding 4 symbols; c1 / c2 right-aligned with
    # zeros where the child has no position
    u = oracle.Union(np.array([0, 0, GAP, A, Cc, G], np.uint8), 4,
                     np.zeros(6, np.int32),
                     np.array([0, 0, 7, 0, 9, 11], np.int32),
                     np.array([0, 0, 5, 6, 0, 8], np.int32))
    # positions in the union's sequence map to offset len - seq_len + x = 2 + x
    a, b = oracle.union_get_positions(u, [(0, 1), (1, 2), (3, 3)])
    # first pair: x = 0 -> index 2 (7); y = 1 -> index 3 holds 0, and the
    # first pair's y searches downwards: index 2 (7)
    assert a[0] == (7, 7)
    # middle pair: x moves down from index 3 (0) to 2 (7); y up from 4 (9)
    assert a[1] == (7, 9)
    # last pair: x moves up from index 5 (11); y up from 5 (11)
    assert a[2] == (11, 11)
    # c2: index 2 (5), index 3 (6) / index 3 (6), index 4 (0) -> up to 5 (8)
    assert b[0] == (5, 6) and b[1] == (6, 8) and b[2] == (8, 8)

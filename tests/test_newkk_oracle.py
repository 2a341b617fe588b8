"""Pins of the oracle's diagonal-store affine aligner (oracle/newkk.c,
restating src/newkkonen.c).

  * every state of every cell of the full matrix equals the cheapest path
    into it, enumerated move by move (tests/brute_newkk.py), with gap-bit
    symbols and several cost models;
  * the reference's incremental refill (only cells whose inputs changed are
    revisited after a band widens) gives exactly what recomputing every band
    cell gives: cost, alignment, rounds, final band and gap count;
  * with (1,1) and gap opening 3 on inputs without gap-bit symbols the full
    band is the textbook affine-gap optimum, and a returned alignment
    re-prices to the returned cost;
  * hand-derived cases fix the backtrace's tie choices under both swap
    flags, the stop rule's round count and the trivial case's layout.
"""
import numpy as np
import pytest

import brute
import brute_newkk as bn
import oracle

BASES = (1, 2, 4, 8)
IUPAC = (3, 5, 9, 6, 10, 12, 7, 11, 13, 14, 15)
GAPPY = (17, 18, 20, 24, 31, 19, 28)
INF = (2 ** 31 - 1) // 2


def rseq(rng, n, amb=0.0, gapbit=0.0):
    s = [16]
    for _ in range(n):
        r = rng.random()
        if r < amb:
            s.append(int(rng.choice(IUPAC)))
        elif r < amb + gapbit:
            s.append(int(rng.choice(GAPPY)))
        else:
            s.append(int(rng.choice(BASES)))
    return np.array(s, np.uint8)


def mutate(rng, s, k, gapbit=0.0):
    b = [int(x) for x in s[1:]]
    for _ in range(k):
        pos = int(rng.integers(0, len(b) + 1))
        op = int(rng.integers(0, 3))
        sym = int(rng.choice(GAPPY)) if rng.random() < gapbit else int(rng.choice(BASES))
        if op == 0 and pos < len(b):
            b[pos] = sym
        elif op == 1:
            b.insert(pos, sym)
        elif pos < len(b):
            del b[pos]
    return np.array([16] + b, np.uint8)


MODELS = [(1, 1, 3), (2, 1, 4), (1, 3, 2), (1, 2, 0), (3, 1, 1)]


def table(cm):
    t = np.ctypeslib.as_array(cm.cost).copy()
    return lambda a, b: int(t[(int(a) << 5) + int(b)])


@pytest.mark.parametrize("model", MODELS)
def test_newkk_matrix_vs_paths(model):
    cm = oracle.cm_dna(*model)
    cost = table(cm)
    rng = np.random.default_rng(100 + model[2])
    for t in range(40):
        n1 = int(rng.integers(0, 5))
        n2 = int(rng.integers(n1, 6))
        a = rseq(rng, n1, 0.2, 0.3)
        b = rseq(rng, n2, 0.2, 0.3)
        m = oracle.newkk_matrix(a, b, cm)
        want = bn.state_minima(a, b, cost, model[2])
        for i in range(len(a)):
            for j in range(len(b)):
                vals = []
                for st, col in ((bn.ALIGN, 4), (bn.GG, 3), (bn.VERT, 1), (bn.HORZ, 2)):
                    v = int(m[i, j, col])
                    w = want.get((i, j, st))
                    if w is None:
                        assert v >= INF, (a, b, i, j, st)
                    else:
                        assert v == w, (list(a), list(b), i, j, st)
                        vals.append(w)
                if (i, j) != (0, 0):
                    assert int(m[i, j, 0]) == min(vals)


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("gapbit", [0.0, 0.25])
def test_newkk_incremental_refill_equals_full_recompute(model, gapbit):
    cm = oracle.cm_dna(*model)
    rng = np.random.default_rng(7 + model[0] + 10 * model[2] + int(gapbit * 100))
    multi = 0
    for t in range(60):
        n1 = int(rng.integers(0, 50))
        a = rseq(rng, n1, 0.1, gapbit)
        b = mutate(rng, a, int(rng.integers(0, 12)), gapbit)
        if len(a) > len(b):
            a, b = b, a
        sw = t & 1
        x = oracle.newkk_align(a, b, cm, sw)
        y = oracle.newkk_full(a, b, cm, sw)
        assert x["cost"] == y["cost"], (list(a), list(b))
        assert np.array_equal(x["r1"], y["r1"]) and np.array_equal(x["r2"], y["r2"])
        for f in ("rounds", "final_T", "final_k", "gap_num"):
            assert getattr(x["stats"], f) == getattr(y["stats"], f), f
        multi += x["stats"].rounds > 1 and x["stats"].final_k > 0
    assert multi > 10  # the incremental path (a band that widened) ran


def _sub(a, b):
    return int(_T[(int(a) << 5) + int(b)])


def _ge(x):
    return int(_T[(int(x) << 5) + 16])


_AFF = oracle.cm_dna(1, 1, 3)
_T = np.ctypeslib.as_array(_AFF.cost).copy()


def test_newkk_full_band_is_textbook_affine():
    rng = np.random.default_rng(21)
    full = 0
    for t in range(160):
        n1 = int(rng.integers(0, 6))
        n2 = int(rng.integers(n1, 7))
        a, b = rseq(rng, n1, 0.2), rseq(rng, n2, 0.2)
        r = oracle.newkk_align(a, b, _AFF, t & 1)
        best = brute.best_affine([int(x) for x in a], [int(x) for x in b], _sub, 3, _ge)
        assert r["cost"] >= best
        if r["stats"].final_k == len(a) - 1:
            full += 1
            assert r["cost"] == best, (list(a), list(b))
        # the returned alignment is a valid alignment priced at the cost
        assert brute.affine_cost_of_aligned(r["r1"], r["r2"], _sub, 3, _ge) == r["cost"]
        back1 = [int(x) for x in r["r1"][1:] if x != 16]
        back2 = [int(x) for x in r["r2"][1:] if x != 16]
        assert back1 == [int(x) for x in a[1:]] and back2 == [int(x) for x in b[1:]]
    assert full > 60


def test_newkk_band_never_beats_full_matrix():
    rng = np.random.default_rng(5)
    for model in MODELS:
        cm = oracle.cm_dna(*model)
        for t in range(30):
            a = rseq(rng, int(rng.integers(1, 40)), 0.1, 0.1)
            b = mutate(rng, a, int(rng.integers(0, 10)), 0.1)
            if len(a) > len(b):
                a, b = b, a
            r = oracle.newkk_align(a, b, cm, 0, backtrace=False)
            m = oracle.newkk_matrix(a, b, cm)
            assert r["cost"] >= int(m[-1, -1, 0])
            if r["stats"].final_k == len(a) - 1 and len(b) <= 100 * len(a):
                assert r["cost"] == int(m[-1, -1, 0])


def test_newkk_hand_cases():
    cm = oracle.cm_dna(1, 1, 3)
    # -A vs -AA: the first doubling step (k = 0) and the second (k = 1)
    # both price 4 (one aligned A, one gap of go 3 + 1); at (1, 2) insert
    # and align tie and
    # both swap orders take the insertion first, so the gap is the last
    # column.
    for sw in (0, 1):
        r = oracle.newkk_align([16, 1], [16, 1, 1], cm, sw)
        assert r["cost"] == 4
        assert list(r["r1"]) == [16, 1, 16] and list(r["r2"]) == [16, 1, 1]
        assert r["stats"].rounds == 2 and r["stats"].final_k == 1
    # -AC vs -CA with (2,1) and gap opening 0: two optimal alignments of
    # cost 2 (delete A, match C, insert A / insert C, match A, delete C); at
    # the last cell insert and delete tie, unswapped takes the deletion,
    # swapped the insertion.  The diagonal-only first round prices 4 and
    # cannot stop.
    cm2 = oracle.cm_dna(2, 1, 0)
    r0 = oracle.newkk_align([16, 1, 2], [16, 2, 1], cm2, 0)
    r1 = oracle.newkk_align([16, 1, 2], [16, 2, 1], cm2, 1)
    assert r0["cost"] == r1["cost"] == 2
    assert list(r0["r1"]) == [16, 16, 1, 2] and list(r0["r2"]) == [16, 2, 1, 16]
    assert list(r1["r1"]) == [16, 1, 2, 16] and list(r1["r2"]) == [16, 16, 2, 1]
    assert r0["stats"].rounds == 2 and r0["stats"].final_k == 1


def test_newkk_identical_sequences_stop_after_two_rounds():
    # D = 0: T = delta = 1, p = 0 (diagonal only, gap count 0, no stop since
    # p > 0 fails and newp = 1 < L - 1); then T = 2, p = 1 > 0 stops.
    cm = oracle.cm_dna(1, 1, 3)
    rng = np.random.default_rng(3)
    for L in (3, 5, 20, 64):
        a = rseq(rng, L - 1)
        r = oracle.newkk_align(a, a.copy(), cm, 0)
        assert r["cost"] == 0
        assert r["stats"].rounds == 2 and r["stats"].final_k == 1
        assert list(r["r1"]) == list(a) and list(r["r2"]) == list(a)


def test_newkk_trivial_case_layout():
    # lenY > 100 lenX: the cost is every symbol's indel cost (gap bit
    # stripped) and the backtrace prepends s1 then gaps (trivial_backtrace),
    # so the aligned pair is gaps+reversed(s1) over reversed(s2)+gaps.
    cm = oracle.cm_dna(1, 2, 3)
    rng = np.random.default_rng(9)
    a = np.array([16], np.uint8)
    b = rseq(rng, 150, 0.1, 0.1)
    r = oracle.newkk_align(a, b, cm, 0)
    t = table(cm)
    assert r["cost"] == sum(t(int(x) & 15, 16) for x in a) + sum(t(int(x) & 15, 16) for x in b)
    assert list(r["r1"]) == [16] * len(b) + [int(x) for x in a[::-1]]
    assert list(r["r2"]) == [int(x) for x in b[::-1]] + [16] * len(a)


def test_newkk_align_2_and_cost_2_wrappers():
    cm = oracle.cm_dna(1, 1, 3)
    rng = np.random.default_rng(4)
    for t in range(20):
        a = rseq(rng, int(rng.integers(1, 30)), 0.1)
        b = mutate(rng, a, 5)
        p1, p2, c = oracle.newkk_align_2(a, b, cm)
        if len(a) > len(b):
            r = oracle.newkk_align(b, a, cm, 1)
            assert list(p1) == list(r["r2"]) and list(p2) == list(r["r1"])
        else:
            r = oracle.newkk_align(a, b, cm, 0)
            assert list(p1) == list(r["r1"]) and list(p2) == list(r["r2"])
        assert c == r["cost"] == oracle.newkk_cost_2(a, b, cm) == oracle.newkk_cost_2(b, a, cm)
        q1, q2, c2 = oracle.newkk_align_2(a[1:], b[1:], cm, first_gap=False)
        assert c2 == c and list(q1) == list(p1[1:]) and list(q2) == list(p2[1:])

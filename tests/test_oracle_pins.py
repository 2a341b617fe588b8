"""Pins of the oracle (CPU) against what the mathematics fixes.

Every check here is independent of the oracle's recurrences:
  * brute-force enumeration of all alignments of tiny inputs (tests/brute.py)
    gives the optimal linear cost and, with POY's tie order (ALIGN first, then
    DELETE before INSERT unswapped, INSERT before DELETE swapped,
    src/algn.c:3278-3327), the exact alignment backtrace_2d must return;
  * unit costs without ambiguity reduce to the Levenshtein distance, for the
    full plane and for Ukkonen bands (src/algn.c:1007-1130);
  * re-pricing a returned alignment column by column gives the reported cost;
  * affine (tcm (1,1), gap_open 3) on inputs without gap-bit symbols is the
    textbook affine-gap model (one opening per maximal gap run);
  * the 3D cube's first plane (s1 = the leading gap only) is an exact 2D
    alignment of s2 and s3 priced by the 3D table (src/algn.c:2622-2668).
A plausible mistake in the oracle (a dropped term, a wrong sign, a swapped
operand or tie order, an off-by-one band edge) fails one of them.
"""
import itertools

import numpy as np
import pytest

import brute
import oracle

BASES = (1, 2, 4, 8)
IUPAC = (3, 5, 9, 6, 10, 12, 7, 11, 13, 14, 15)
GAPPY = (17, 18, 20, 24, 31)


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


def ints(a):
    return [int(x) for x in a]


def mutate(rng, s, k):
    b = ints(s[1:])
    for _ in range(k):
        pos = int(rng.integers(0, len(b) + 1))
        op = int(rng.integers(0, 3))
        if op == 0 and pos < len(b):
            b[pos] = int(rng.choice(BASES))
        elif op == 1:
            b.insert(pos, int(rng.choice(BASES)))
        elif pos < len(b):
            del b[pos]
    return np.array([16] + b, np.uint8)


def levenshtein(a, b):
    a, b = ints(a[1:]), ints(b[1:])
    d = list(range(len(b) + 1))
    for i, x in enumerate(a, 1):
        p = d[:]
        d[0] = i
        for j, y in enumerate(b, 1):
            d[j] = min(p[j] + 1, d[j - 1] + 1, p[j - 1] + (x != y))
    return d[-1]


@pytest.mark.parametrize("tcm", [(1, 2), (1, 1), (2, 1), (3, 2)])
@pytest.mark.parametrize("gapbit", [0.0, 0.3])
def test_linear_full_plane_vs_brute(tcm, gapbit):
    rng = np.random.default_rng(100 + tcm[0] * 10 + tcm[1] + int(gapbit * 10))
    cm = oracle.cm_dna(*tcm)
    C = np.array(cm.cost)
    for _ in range(60):
        n1 = int(rng.integers(0, 6))
        n2 = int(rng.integers(n1, 7))
        a, b = rseq(rng, n1, 0.15, gapbit), rseq(rng, n2, 0.15, gapbit)
        c, dm, st = oracle.nw(a, b, cm, 100)
        for sw in (0, 1):
            bc, (r1, r2) = brute.best_linear(ints(a), ints(b), C, sw)
            assert c == bc
            x1, x2 = oracle.backtrace_2d(a, b, dm, sw)
            assert ints(x1) == ints(r1) and ints(x2) == ints(r2), (ints(a), ints(b), sw)


def test_linear_levenshtein_full_and_band():
    rng = np.random.default_rng(5)
    cm = oracle.cm_dna(1, 1)
    C = np.array(cm.cost)
    bands = 0
    for _ in range(120):
        a = rseq(rng, int(rng.integers(20, 160)))
        b = mutate(rng, a, int(rng.integers(0, 25)))
        if len(b) < len(a):
            a, b = b, a
        want = levenshtein(a, b)
        for dw in (0, 2, 60):
            c, dm, st = oracle.nw(a, b, cm, dw)
            bands += 0 if st.full_plane else 1
            assert c == want
            r1, r2 = oracle.backtrace_2d(a, b, dm, dw & 1)
            assert sum(int(C[(int(x) << 5) + int(y)]) for x, y in zip(r1[1:], r2[1:])) == c
            assert oracle.verify_2(r1, r2, cm) == c
    assert bands > 100  # the band path was exercised


def test_linear_edge_cases():
    cm = oracle.cm_dna(1, 2)
    s = np.array([16, 1, 2, 4, 8, 3], np.uint8)
    c, _, _ = oracle.nw(s, s, cm, 0)
    assert c == 0
    c, dm, _ = oracle.nw(np.array([16], np.uint8), s, cm, 0)
    assert c == 2 * 5  # every symbol against a gap
    r1, r2 = oracle.backtrace_2d(np.array([16], np.uint8), s, dm, 0)
    assert ints(r1) == [16] * 6 and ints(r2) == ints(s)


AFF = oracle.cm_dna(1, 1, 3)
_AC = np.array(AFF.cost)


def _sub(a, b):
    return int(_AC[(int(a) << 5) + int(b)])


def _ge(x):
    return int(_AC[(int(x) << 5) + 16])


def test_affine_vs_textbook():
    rng = np.random.default_rng(11)
    for t in range(150):
        n1 = int(rng.integers(0, 6))
        n2 = int(rng.integers(n1, 7))
        a, b = rseq(rng, n1, 0.2), rseq(rng, n2, 0.2)
        want = brute.best_affine(ints(a), ints(b), _sub, 3, _ge)
        r = oracle.align_affine_3(a, b, AFF, t & 1)
        assert r["cost"] == want, (ints(a), ints(b))
        assert oracle.cost_affine_3(a, b, AFF) == want
        assert oracle.cost_affine_3(b, a, AFF) == want
        assert brute.affine_cost_of_aligned(r["resi"], r["resj"], _sub, 3, _ge) == want
        assert oracle.verify_2(r["resi"], r["resj"], AFF) == want


def test_affine_band_rounds_agree_with_full():
    """Banded doubling (align) and the full DP (cost-only) agree on inputs
    without gap-bit symbols, at lengths where the band is narrow."""
    rng = np.random.default_rng(12)
    multi = 0
    for t in range(40):
        a = rseq(rng, int(rng.integers(50, 400)), 0.05)
        b = mutate(rng, a, int(rng.integers(0, 40)))
        if len(b) < len(a):
            a, b = b, a
        r = oracle.align_affine_3(a, b, AFF, t & 1)
        multi += r["stats"].rounds > 1
        assert r["cost"] == oracle.cost_affine_3(a, b, AFF)
        assert oracle.verify_2(r["resi"], r["resj"], AFF) == r["cost"]
    assert multi > 5


@pytest.mark.parametrize("model", [(1, 1, 3), (1, 2, 1), (2, 3, 4), (1, 1, 0)])
@pytest.mark.parametrize("gapbit", [0.0, 0.2])
def test_gap_count_at_least_length_difference(model, gapbit):
    """Every band round's gap count is at least D = len2 - len1: each path
    makes D more horizontal than vertical moves, and the count keeps the
    most moves of a kind over tied paths.  The fill kernels skip rounds
    whose stop test therefore cannot pass (affine gap_num < p with p <= D,
    linear gap_num + 1 < p with p <= D + 1); this pins the premise on the
    oracle's rounds, affine and linear, with and without gap-bit symbols."""
    rng = np.random.default_rng(21 + model[2] + int(10 * gapbit))
    acm, lcm = oracle.cm_dna(*model), oracle.cm_dna(model[0], model[1])
    rounds = 0
    for t in range(30):
        a = rseq(rng, int(rng.integers(1, 300)), 0.05, gapbit)
        b = mutate(rng, a, int(rng.integers(0, 60)))
        if len(b) < len(a):
            a, b = b, a
        st = oracle.align_affine_3(a, b, acm, t & 1)["stats"]
        assert st.min_gap_excess >= 0, (t, st.min_gap_excess)
        rounds += st.rounds
        st2 = oracle.align_2(a, b, lcm)[3]
        assert st2.full_plane or st2.min_gap_excess >= 0, (t, st2.min_gap_excess)
    assert rounds > 60


def test_affine_medians_follow_alignment():
    """median_with_gaps is the column-wise median table of the alignment;
    the median drops the gap columns (inputs without gap-bit symbols)."""
    rng = np.random.default_rng(13)
    med = np.array(AFF.median)
    for t in range(40):
        a = rseq(rng, int(rng.integers(1, 60)), 0.1)
        b = mutate(rng, a, int(rng.integers(0, 8)))
        if len(b) < len(a):
            a, b = b, a
        r = oracle.align_affine_3(a, b, AFF, 0)
        ri, rj, mw, m = ints(r["resi"]), ints(r["resj"]), ints(r["medianwg"]), ints(r["median"])
        assert len(ri) == len(rj) == len(mw)
        for x, y, z in zip(ri[1:], rj[1:], mw[1:]):
            if x != 16 and y != 16:
                assert z == int(med[(x << 5) + y])
            else:
                assert z == (x | y)
        assert m[0] == 16


def test_linear_medians_closed_forms():
    rng = np.random.default_rng(14)
    cm = oracle.cm_dna(1, 2)
    med = np.array(cm.median)
    for _ in range(40):
        n = int(rng.integers(1, 30))
        a = rseq(rng, n, 0.2, 0.2)
        b = rseq(rng, n, 0.2, 0.2)
        cols = [int(med[(int(x) << 5) + int(y)]) for x, y in zip(a, b)]
        assert ints(oracle.median_2_with_gaps(a, b, cm)) == cols
        kept = [v for v in cols if v != 16]
        assert ints(oracle.median_2_no_gaps(a, b, cm)) == [16] + kept
        anc = ints(oracle.ancestor_2(a, b, cm))
        assert anc == (kept if kept and kept[0] == 16 else [16] + kept)


def _brute_3d_plane0(s2, s3, c3):
    """Min over all 2D alignments of s2 and s3 with columns priced by the
    3D table with the first code fixed to the gap."""
    cost = np.array(c3.cost)

    def e(b, c):
        return int(cost[(((16 << 5) + b) << 5) + c])

    best = None
    for mv in brute.all_alignments(len(s2) - 1, len(s3) - 1):
        i = j = 1
        tot = 0
        for m in mv:
            if m == "A":
                tot += e(s2[i], s3[j]); i += 1; j += 1
            elif m == "D":
                tot += e(s2[i], 16); i += 1
            else:
                tot += e(16, s3[j]); j += 1
        best = tot if best is None else min(best, tot)
    return best


def test_3d_first_plane_vs_brute():
    rng = np.random.default_rng(15)
    c3 = oracle.cm3_from_2d(oracle.cm_dna(1, 1))
    for _ in range(50):
        s1 = np.array([16], np.uint8)
        s2 = rseq(rng, int(rng.integers(0, 5)), 0.2)
        s3 = rseq(rng, int(rng.integers(0, 5)), 0.2)
        cost, dm = oracle.nw_3d(s1, s2, s3, c3)
        assert cost == _brute_3d_plane0(ints(s2), ints(s3), c3)
        out = oracle.backtrack_3d(s1, s2, s3, dm)
        assert out is not None
        r1, r2, r3 = (ints(x) for x in out)
        assert all(x == 16 for x in r1)
        assert [x for x in r2 if x != 16] == [x for x in ints(s2) if x != 16]
        assert [x for x in r3 if x != 16] == [x for x in ints(s3) if x != 16]



# ---- four-state affine model by path enumeration (tests/brute_affine.py) ----
import brute_affine  # noqa: E402

AFF_MODELS = [(1, 1, 3), (2, 2, 4), (1, 3, 2)]


def _bmodel(a, b, cm, mode, newk=None):
    return brute_affine.Model(a, b, list(cm.cost), list(cm.prepend), list(cm.median),
                              cm.gap_open, mode, newk)


@pytest.mark.parametrize("model", AFF_MODELS)
@pytest.mark.parametrize("gapbit", [0.0, 0.3])
def test_affine_band_round_tie_exact_vs_paths(model, gapbit):
    """One band round at every half-width, both swap flags: cost and the
    exact aligned pair, median and median with gaps equal the walk of
    backtrace_aff's mode machine over flags derived from path-enumerated
    state values (todo priority V>H>block>align, swapped H>V; runs end on
    END_* or HORIZONTAL_EQ_VERTICAL; ALIGN_TO_* priority likewise).  With
    gap-bit symbols this covers POY's gap-opening rules and the gap-gap block
    state, which the textbook model does not have."""
    cm = oracle.cm_dna(*model)
    rng = np.random.default_rng(300 + model[0] + 10 * model[2] + int(gapbit * 10))
    n = 0
    for _ in range(60):
        la = int(rng.integers(0, 6))
        lb = int(rng.integers(la, 7))
        a, b = rseq(rng, la, 0.1, gapbit), rseq(rng, lb, 0.1, gapbit)
        for p in range(0, 6):
            newk = p if p < len(a) - 1 else len(a) - 2
            bm = _bmodel(a, b, cm, "band", newk)
            for sw in (0, 1):
                r = oracle.align_affine_round(a, b, cm, sw, p)
                med, mwg, ri, rj = bm.walk(sw)
                assert r["cost"] == bm.cost_at_end(), (ints(a), ints(b), p, sw)
                assert ints(r["resi"]) == ri and ints(r["resj"]) == rj, (ints(a), ints(b), p, sw)
                assert ints(r["median"]) == med and ints(r["medianwg"]) == mwg
                n += 1
    assert n > 600


@pytest.mark.parametrize("model", AFF_MODELS)
def test_affine_doubling_result_tie_exact_vs_paths(model):
    """The full Ukkonen doubling (algn_CAML_align_affine_3) on inputs without
    gap-bit symbols: its result is the path-enumerated model's walk at the
    final round's band (rows carried over from earlier rounds only touch the
    gap-gap block state and row 0's H, which cannot win without gap bits)."""
    cm = oracle.cm_dna(*model)
    rng = np.random.default_rng(400 + model[0] + 10 * model[2])
    for _ in range(120):
        la = int(rng.integers(0, 6))
        lb = int(rng.integers(la, 7))
        a, b = rseq(rng, la, 0.15), rseq(rng, lb, 0.15)
        for sw in (0, 1):
            r = oracle.align_affine_3(a, b, cm, sw)
            bm = _bmodel(a, b, cm, "band", r["stats"].final_k)
            med, mwg, ri, rj = bm.walk(sw)
            assert r["cost"] == bm.cost_at_end(), (ints(a), ints(b), sw)
            assert ints(r["resi"]) == ri and ints(r["resj"]) == rj, (ints(a), ints(b), sw)
            assert ints(r["median"]) == med and ints(r["medianwg"]) == mwg


@pytest.mark.parametrize("model", AFF_MODELS)
@pytest.mark.parametrize("gapbit", [0.0, 0.3])
def test_affine_cost_only_vs_paths(model, gapbit):
    """algn_CAML_cost_affine_3 (full DP, no directions): the minimum over
    every four-state path under POY's move costs, including the effect of
    the stub's overlapping row buffers; either argument order."""
    cm = oracle.cm_dna(*model)
    rng = np.random.default_rng(500 + model[0] + 10 * model[2] + int(gapbit * 10))
    overlap_mattered = 0
    for _ in range(300):
        la = int(rng.integers(0, 6))
        lb = int(rng.integers(la, 7))
        a, b = rseq(rng, la, 0.1, gapbit), rseq(rng, lb, 0.1, gapbit)
        # the stub takes the shorter sequence as rows, the first on a tie
        want = _bmodel(a, b, cm, "full").cost_at_end()
        assert oracle.cost_affine_3(a, b, cm) == want, (ints(a), ints(b))
        rows, cols = (b, a) if len(b) == len(a) else (a, b)
        assert oracle.cost_affine_3(b, a, cm) == _bmodel(rows, cols, cm, "full").cost_at_end()
        clean = brute_affine.Model(a, b, list(cm.cost), list(cm.prepend), list(cm.median),
                                   cm.gap_open, "full", overlap=False).cost_at_end()
        assert clean <= want  # the overlap only removes candidates
        overlap_mattered += clean < want
    if gapbit:
        assert overlap_mattered > 0  # the overlap is exercised, not vacuous


def test_affine_cost_only_overlap_known_cases():
    """Pairs where the stub's overlapping row buffers change the cost-only
    result (found by the path enumeration): the clean four-state minimum is
    lower than what algn_CAML_cost_affine_3 computes."""
    cases = [((1, 1, 3), [16, 4, 1, 8, 18], [16, 4, 24, 20, 17, 2], 4, 5),
             ((1, 3, 2), [16, 4, 24, 1, 24], [16, 2, 20, 17, 17, 24, 1], 1, 4),
             ((1, 3, 2), [16, 3, 18], [16, 24, 17, 20, 24, 4], 3, 4)]
    for model, a, b, clean, poy in cases:
        cm = oracle.cm_dna(*model)
        a, b = np.array(a, np.uint8), np.array(b, np.uint8)
        assert oracle.cost_affine_3(a, b, cm) == poy
        assert _bmodel(a, b, cm, "full").cost_at_end() == poy
        assert brute_affine.Model(a, b, list(cm.cost), list(cm.prepend), list(cm.median),
                                  cm.gap_open, "full", overlap=False).cost_at_end() == clean


# ---- three-way cube by path enumeration (tests/brute_cube.py) -------------
import brute_cube  # noqa: E402


@pytest.mark.parametrize("tcm", [(1, 1), (1, 2), (2, 1)])
def test_3d_cube_vs_path_enumeration(tcm):
    """Every plane of algn_fill_cube, drifted row pointers included: the
    cost, the whole direction cube (first minimising move in the
    reference's candidate order) and backtrack_3d's walk equal the
    path-enumerated model, for cubes up to 4 x 5 x 5 (planes >= 1 with their
    lags of i-1 and i rows, L1 = 2 included)."""
    c3 = oracle.cm3_from_2d(oracle.cm_dna(*tcm))
    cost3 = list(c3.cost)
    rng = np.random.default_rng(600 + 10 * tcm[0] + tcm[1])
    planes = 0
    for t in range(150):
        n1 = int(rng.integers(0, 4)) if t % 3 else 1  # every third cube has L1 = 2
        s1 = rseq(rng, n1, 0.2)
        s2 = rseq(rng, int(rng.integers(0, 5)), 0.2)
        s3 = rseq(rng, int(rng.integers(0, 5)), 0.2)
        cube = brute_cube.Cube(s1, s2, s3, cost3)
        cost, dm = oracle.nw_3d(s1, s2, s3, c3)
        assert cost == cube.cost(), (ints(s1), ints(s2), ints(s3))
        L1, L2, L3 = len(s1), len(s2), len(s3)
        flat = np.asarray(dm).reshape(-1)
        for i in range(L1):
            for r in range(L2):
                for k in range(L3):
                    if (i, r, k) != (0, 0, 0):
                        assert int(flat[(i * L2 + r) * L3 + k]) == cube.dm_code(i, r, k), \
                            (ints(s1), ints(s2), ints(s3), i, r, k)
        want = cube.walk()
        got = oracle.backtrack_3d(s1, s2, s3, dm)
        assert (want is None) == (got is None)
        if want is not None:
            assert [ints(x) for x in got] == [list(x) for x in want]
        planes += L1 - 1
    assert planes > 100


def test_3d_one_row_planes_all_read_plane_zero():
    """With L2 = 1 the drifted pointers make every plane read plane 0, so
    the cube's cost depends on s1 only through its last symbol."""
    c3 = oracle.cm3_from_2d(oracle.cm_dna(1, 2))
    rng = np.random.default_rng(17)
    s2 = np.array([16], np.uint8)
    for _ in range(40):
        s3 = rseq(rng, int(rng.integers(0, 8)), 0.2)
        last = int(rng.choice(BASES))
        ref, _ = oracle.nw_3d(np.array([16, last], np.uint8), s2, s3, c3)
        for n in (2, 3, 6):
            s1 = np.concatenate([rseq(rng, n - 1, 0.3), [last]]).astype(np.uint8)
            assert oracle.nw_3d(s1, s2, s3, c3)[0] == ref


def test_median_3_repeats_the_last_column():
    """algn_get_median_3d (src/algn.c:3641-3654) never moves its end
    pointers: the median has len(s1) entries, each the 3D median of the
    last column."""
    c3 = oracle.cm3_from_2d(oracle.cm_dna(1, 2))
    med = np.array(c3.median)
    rng = np.random.default_rng(18)
    for _ in range(60):
        n = int(rng.integers(1, 12))
        a, b, c = (rng.integers(1, 32, n).astype(np.uint8) for _ in range(3))
        want = med[(int(a[-1]) << 10) + (int(b[-1]) << 5) + int(c[-1])]
        got = oracle.median_3(a, b, c, c3)
        assert len(got) == n and np.all(got == want)


def _moves_of(r1, r2):
    """Walk moves of an aligned pair without gap-bit symbols, last column
    first: 0 align (both advance), 1 row only, 2 column only."""
    out = []
    for x, y in zip(reversed(ints(r1)), reversed(ints(r2))):
        if x == 16 and y == 16:
            out.append(0)  # the leading gaps, cell (0, 0)
        elif y == 16:
            out.append(1)
        elif x == 16:
            out.append(2)
        else:
            out.append(0)
    return out


@pytest.mark.parametrize("tcm", [(1, 2), (2, 1), (1, 1)])
def test_limit_walk_from_any_cell_is_the_prefix_alignment(tcm):
    """backtrace_2d with the limits of algn_CAML_backtrace_2d_limit: from
    cell (r, c) of a full-plane matrix (st_s1 = st_s2 = 0, algn = r+1, c+1)
    the walk is the tie-exact alignment of the prefixes s1[:r+1], s2[:c+1]
    (the fill is prefix-closed), for both swap orders."""
    cm = oracle.cm_dna(*tcm)
    C = np.array(cm.cost)
    rng = np.random.default_rng(900 + 10 * tcm[0] + tcm[1])
    for _ in range(50):
        n1 = int(rng.integers(1, 6))
        n2 = int(rng.integers(n1, 7))
        a, b = rseq(rng, n1, 0.15), rseq(rng, n2, 0.15)
        _, dm, st = oracle.nw(a, b, cm, 100)
        assert st.full_plane
        for r in range(len(a)):
            for c in range(len(b)):
                for sw in (0, 1):
                    _, (p1, p2) = brute.best_linear(ints(a[:r + 1]), ints(b[:c + 1]), C, sw)
                    x1, x2 = oracle.backtrace_2d_limit(a, b, dm, 0, 0, r + 1, c + 1, sw)
                    assert ints(x1) == ints(p1) and ints(x2) == ints(p2), (r, c, sw)


def test_limit_walk_stops_left_of_st_s2_and_st_s1_moves_symbols_only():
    """With st_s2 = x the walk from column x + algn_s2 - 1 is the unlimited
    walk's columns up to its first cell of row 0 left of column x; st_s1
    changes the row symbols read, not the moves."""
    cm = oracle.cm_dna(1, 2)
    rng = np.random.default_rng(910)
    for _ in range(60):
        n1 = int(rng.integers(2, 12))
        n2 = int(rng.integers(n1, 14))
        a, b = rseq(rng, n1, 0.1), rseq(rng, n2, 0.1)
        _, dm, _ = oracle.nw(a, b, cm, 100)
        L2 = len(b)
        r = int(rng.integers(1, len(a)))
        c = int(rng.integers(1, len(b)))
        sw = int(rng.integers(0, 2))
        f1, f2 = oracle.backtrace_2d_limit(a, b, dm, 0, 0, r + 1, c + 1, sw)
        for x in range(0, c + 1):
            # columns of the unlimited walk before it passes below column x of row 0
            i, j, keep = r, c, 0
            for m in _moves_of(f1, f2):
                if i * L2 + j < x:
                    break
                keep += 1
                i, j = (i - 1, j - 1) if m == 0 else ((i - 1, j) if m == 1 else (i, j - 1))
            g1, g2 = oracle.backtrace_2d_limit(a, b, dm, 0, x, r + 1, c + 1 - x, sw)
            assert ints(g1) == ints(f1[len(f1) - keep:]) and ints(g2) == ints(f2[len(f2) - keep:])
        # st_s1: the same moves, row symbols from s1[y:]
        y = int(rng.integers(0, len(a) - r))
        h1, h2 = oracle.backtrace_2d_limit(a, b, dm, y, 0, r + 1, c + 1, sw)
        assert _moves_of(h1, h2) == _moves_of(f1, f2) or y > 0
        assert ints(h2) == ints(f2)
        rows = [v for v, m in zip(reversed(ints(h1)), _moves_of(f1, f2)) if m != 2]
        assert rows == ints(a[y:y + r + 1])[::-1][:len(rows)]


def test_limit_2_moves_only_the_decision():
    """algn_nw_limit's height is 50 + (s1_len - s2_len) + w: limits that
    raise it turn a banded fill into a full plane; the cost, optimal either
    way, does not change."""
    cm = oracle.cm_dna(1, 2)
    rng = np.random.default_rng(920)
    for _ in range(10):
        b = rseq(rng, 320)
        a = mutate(rng, b, 40)[:301]
        c0, _, st0 = oracle.nw(a, b, cm, 0)
        assert not st0.full_plane  # height 50 + (301 - 321) = 30 < 301 / 2
        c1, _, st1 = oracle.limit_2(a, b, cm, 0, 900, 100)  # height 850, clamped to 301
        assert st1.full_plane and c1 == c0
        c2, _, st2 = oracle.limit_2(a, b, cm, 0, len(a), len(b))
        assert not st2.full_plane and c2 == c0

"""GPU parity: every drop-in and batch entry point against the CPU oracle,
bit-exact (costs, aligned sequences, medians are integer/byte data)."""
import sys

import numpy as np
import pytest

import oracle
from poy5_amd import cost_matrix, workloads
from poy5_amd.batch import run_pairs
from poy5_amd.sequence import Align, Context

pytestmark = pytest.mark.gpu

PLAIN = np.array([1, 2, 4, 8], np.uint8)
IUPAC = np.array([3, 5, 9, 6, 10, 12, 7, 11, 13, 14, 15], np.uint8)
GAPPY = np.array([17, 18, 20, 24, 31, 19, 28], np.uint8)  # codes carrying the gap bit


def rand_seq(rng, n, gappy=0.0, iupac=0.05):
    s = rng.choice(PLAIN, n)
    m = rng.random(n)
    s = np.where(m < iupac, rng.choice(IUPAC, n), s)
    s = np.where(m > 1 - gappy, rng.choice(GAPPY, n), s)
    return np.concatenate([[16], s]).astype(np.uint8)


def mutate(rng, s, d):
    out = [16]
    for c in s[1:]:
        u = rng.random()
        if u < d / 3:
            continue
        if u < 2 * d / 3:
            out.append(int(rng.choice(PLAIN)))
        out.append(int(rng.choice(PLAIN)) if rng.random() < d else int(c))
    return np.array(out, np.uint8)


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def pair_cases(seed, count, maxlen, gappy=0.0):
    rng = np.random.default_rng(seed)
    cases = []
    for k in range(count):
        la = int(rng.integers(0, maxlen))
        root = rand_seq(rng, la, gappy)
        a = mutate(rng, root, rng.uniform(0.02, 0.3))
        b = mutate(rng, root, rng.uniform(0.02, 0.3))
        cases.append((a, b))
    # edge cases: leading gap only, equal, single symbol
    cases += [(np.array([16], np.uint8), np.array([16], np.uint8)),
              (np.array([16], np.uint8), rand_seq(rng, 5)),
              (rand_seq(rng, 1), rand_seq(rng, 1)),
              (rand_seq(rng, 7), rand_seq(rng, 7))]
    return cases


# (substitution, gap extension, gap opening): the minimum non-zero cost
# (the T-doubling unit delta) is 1 for the first and third model and 2 for
# the second, whose extension also differs from its substitution
AFFINE_MODELS = [(1, 1, 3), (2, 2, 4), (1, 3, 2)]


@pytest.mark.parametrize("model", AFFINE_MODELS)
@pytest.mark.parametrize("gappy", [0.0, 0.15])
def test_align_affine_3_dropin(ctx, gappy, model):
    say(f"test_align_affine_3_dropin gappy={gappy} model={model}")
    cm = cost_matrix.of_transformations_and_gaps(*model)
    ctx.set_cm(cm)
    ocm = oracle.cm_dna(*model)
    al = Align(ctx)
    n = 60 if model == (1, 1, 3) else 30
    for k, (a, b) in enumerate(pair_cases(11 + int(gappy * 100), n, 400, gappy)):
        if len(a) > len(b):
            a, b = b, a
        for sw in (0, 1):
            m, ra, rb, c, mwg = al.c_align_affine_3(a, b, sw)
            ref = oracle.align_affine_3(a, b, ocm, sw)
            assert c == ref["cost"], (k, sw, c, ref["cost"])
            assert np.array_equal(ra, ref["resi"]), (k, sw)
            assert np.array_equal(rb, ref["resj"]), (k, sw)
            assert np.array_equal(m, ref["median"]), (k, sw)
            assert np.array_equal(mwg, ref["medianwg"]), (k, sw)


@pytest.mark.parametrize("model", AFFINE_MODELS)
def test_cost_affine_3_dropin(ctx, model):
    say(f"test_cost_affine_3_dropin model={model}")
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(*model))
    ocm = oracle.cm_dna(*model)
    al = Align(ctx)
    for k, (a, b) in enumerate(pair_cases(5, 40, 300, 0.1)):
        assert al.cost_2_affine(a, b) == oracle.cost_affine_3(a, b, ocm), k


def test_cost_affine_3_row_overlap(ctx):
    """algn_CAML_cost_affine_3's overlapping row buffers (the last column's V
    of even rows has only HIGH sources; with one column the aligned state of
    odd rows reads the cell's own V and E): known pairs where this changes
    the cost, then thousands of tiny pairs with gap-bit symbols, through
    the batch and the drop-in, against the oracle."""
    say("test_cost_affine_3_row_overlap")
    rng = np.random.default_rng(0x0E7A)
    for model in AFFINE_MODELS:
        ctx.set_cm(cost_matrix.of_transformations_and_gaps(*model))
        ocm = oracle.cm_dna(*model)
        known = [([16, 4, 1, 8, 18], [16, 4, 24, 20, 17, 2]),
                 ([16, 4, 24, 1, 24], [16, 2, 20, 17, 17, 24, 1]),
                 ([16, 3, 18], [16, 24, 17, 20, 24, 4])]
        pairs = [(np.array(a, np.uint8), np.array(b, np.uint8)) for a, b in known]
        for _ in range(3000):
            a = rand_seq(rng, int(rng.integers(0, 7)), gappy=0.3, iupac=0.1)
            b = rand_seq(rng, int(rng.integers(0, 7)), gappy=0.3, iupac=0.1)
            pairs.append((a, b))
        costs, _ = run_pairs(ctx, "cost_affine", pairs)
        want = np.array([oracle.cost_affine_3(a, b, ocm) for a, b in pairs])
        bad = np.nonzero(costs != want)[0]
        assert len(bad) == 0, (model, [(list(pairs[k][0]), list(pairs[k][1])) for k in bad[:5]])
        al = Align(ctx)
        for k in range(0, len(pairs), 60):
            a, b = pairs[k]
            assert al.cost_2_affine(a, b) == want[k], (model, k)


@pytest.mark.parametrize("tcm", [(1, 2), (1, 1)])
def test_align_2_linear(ctx, tcm):
    say(f"test_align_2_linear tcm={tcm}")
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(*tcm))
    ocm = oracle.cm_dna(*tcm)
    al = Align(ctx)
    for k, (a, b) in enumerate(pair_cases(21 + tcm[1], 60, 500, 0.1)):
        r1, r2, c = al.align_2(a, b)
        o1, o2, oc, _ = oracle.align_2(a, b, ocm)
        assert c == oc, (k, c, oc)
        assert np.array_equal(r1, o1) and np.array_equal(r2, o2), k
        # the deltaw form used by DOS.distance
        assert al.cost_2(a, b, deltaw=8) == oracle.cost_2(a, b, ocm, deltaw=8)[0], k


def test_medians_dropin(ctx):
    say("test_medians_dropin")
    for model in [(1, 2, None), (1, 1, 3)]:
        ctx.set_cm(cost_matrix.of_transformations_and_gaps(*model))
        ocm = oracle.cm_dna(*model)
        al = Align(ctx)
        for k, (a, b) in enumerate(pair_cases(31, 20, 200, 0.2)):
            o1, o2, _, _ = oracle.align_2(a, b, oracle.cm_dna(1, 2))
            assert np.array_equal(al.ancestor_2(o1, o2), oracle.ancestor_2(o1, o2, ocm)), k
            assert np.array_equal(al.median_2(o1, o2), oracle.median_2_no_gaps(o1, o2, ocm)), k
            assert np.array_equal(al.median_2_with_gaps(o1, o2),
                                  oracle.median_2_with_gaps(o1, o2, ocm)), k
            # Align.max_cost_2 (sequence.ml:652): 0 when a side is all gaps,
            # else the C stub algn_CAML_worst_2
            empty = bool(np.all(o1 == 16) or np.all(o2 == 16))
            want = 0 if empty else oracle.worst_2(o1, o2, ocm)
            assert al.max_cost_2(o1, o2) == want, k


@pytest.mark.parametrize("model", AFFINE_MODELS)
def test_batch_affine_config2_sample(ctx, model):
    say(f"test_batch_affine_config2_sample model={model}")
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(*model))
    ocm = oracle.cm_dna(*model)
    pairs = workloads.pairs_list(2, 96)
    pairs = [p for p in pairs if len(p[0]) * len(p[1]) < 1_200_000][:48]
    costs, outs = run_pairs(ctx, "align_affine", pairs)
    for k, (a, b) in enumerate(pairs):
        sw = len(a) > len(b)
        si, sj = (b, a) if sw else (a, b)
        ref = oracle.align_affine_3(si, sj, ocm, int(sw))
        ra, rb = (ref["resj"], ref["resi"]) if sw else (ref["resi"], ref["resj"])
        assert costs[k] == ref["cost"], (k, costs[k], ref["cost"])
        assert np.array_equal(outs[k][0], ra) and np.array_equal(outs[k][1], rb), k
        assert np.array_equal(outs[k][2], ref["median"]), k
        assert np.array_equal(outs[k][3], ref["medianwg"]), k


def test_batch_affine_config2_long(ctx):
    """The longest pairs of a 256-pair draw (many 64-row stripes, several
    doubling rounds, wide and narrow bands) together with short ones."""
    say("test_batch_affine_config2_long")
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1, 3))
    ocm = oracle.cm_dna(1, 1, 3)
    pairs = workloads.pairs_list(2, 256, seed=0x5EED)
    pairs.sort(key=lambda p: -len(p[0]) * len(p[1]))
    pairs = pairs[:6] + pairs[-10:]
    costs, outs = run_pairs(ctx, "align_affine", pairs)
    for k, (a, b) in enumerate(pairs):
        sw = len(a) > len(b)
        si, sj = (b, a) if sw else (a, b)
        ref = oracle.align_affine_3(si, sj, ocm, int(sw))
        ra, rb = (ref["resj"], ref["resi"]) if sw else (ref["resi"], ref["resj"])
        assert costs[k] == ref["cost"], (k, costs[k], ref["cost"])
        assert np.array_equal(outs[k][0], ra) and np.array_equal(outs[k][1], rb), k
        assert np.array_equal(outs[k][2], ref["median"]), k
        assert np.array_equal(outs[k][3], ref["medianwg"]), k


def wide_band_pairs(seed):
    """Gap-free pairs whose doubling rounds reach bands of 600-1400 columns:
    long-vs-short random pairs (one wide round) and block-indel pairs whose
    narrow round is followed by a wide one, in both orders."""
    rng = np.random.default_rng(seed)
    pairs = []
    for la, lb in [(200, 800), (130, 700), (66, 650), (300, 950), (257, 1000), (129, 720),
                   (400, 1100), (640, 1300)]:
        a = np.concatenate([[16], rng.choice(PLAIN, la - 1)]).astype(np.uint8)
        b = np.concatenate([[16], rng.choice(PLAIN, lb - 1)]).astype(np.uint8)
        pairs.append((a, b))
    for la, ins_at, x, del_at, y in [(800, 200, 200, 500, 100), (1000, 300, 230, 700, 130),
                                     (700, 50, 300, 600, 200)]:
        a = np.concatenate([[16], rng.choice(PLAIN, la - 1)]).astype(np.uint8)
        b = np.concatenate([a[:ins_at], rng.choice(PLAIN, x), a[ins_at:del_at],
                            a[del_at + y:]]).astype(np.uint8)
        pairs.append((a, b))
    return pairs + [(b, a) for a, b in pairs[::3]]


def test_batch_affine_wide_bands(ctx):
    """Bands wider than the shared-memory FIFO of the rotating schedule: the
    fill spills it to global memory."""
    say("test_batch_affine_wide_bands")
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1, 3))
    ocm = oracle.cm_dna(1, 1, 3)
    pairs = wide_band_pairs(0x71DE)
    costs, outs = run_pairs(ctx, "align_affine", pairs)
    for k, (a, b) in enumerate(pairs):
        sw = len(a) > len(b)
        si, sj = (b, a) if sw else (a, b)
        ref = oracle.align_affine_3(si, sj, ocm, int(sw))
        ra, rb = (ref["resj"], ref["resi"]) if sw else (ref["resi"], ref["resj"])
        assert costs[k] == ref["cost"], (k, costs[k], ref["cost"])
        assert np.array_equal(outs[k][0], ra) and np.array_equal(outs[k][1], rb), k
        assert np.array_equal(outs[k][2], ref["median"]), k
        assert np.array_equal(outs[k][3], ref["medianwg"]), k


def test_batch_affine_mixed_classes(ctx):
    """One batch of gap-free pairs and pairs with gap-bit symbols: the fast
    kernel fills and backtraces the first class, the general kernel and the
    backtrace kernel the second; a small direction budget splits the batch
    into several chunks."""
    say("test_batch_affine_mixed_classes")
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1, 3))
    ocm = oracle.cm_dna(1, 1, 3)
    pairs = pair_cases(0xC1A5, 24, 160) + pair_cases(0xC1A6, 24, 160, gappy=0.15)
    for budget in (0, 1 << 18):
        costs, outs = run_pairs(ctx, "align_affine", pairs, dm_budget=budget)
        for k, (a, b) in enumerate(pairs):
            sw = len(a) > len(b)
            si, sj = (b, a) if sw else (a, b)
            ref = oracle.align_affine_3(si, sj, ocm, int(sw))
            ra, rb = (ref["resj"], ref["resi"]) if sw else (ref["resi"], ref["resj"])
            assert costs[k] == ref["cost"], (budget, k, costs[k], ref["cost"])
            assert np.array_equal(outs[k][0], ra) and np.array_equal(outs[k][1], rb), (budget, k)
            assert np.array_equal(outs[k][2], ref["median"]), (budget, k)
            assert np.array_equal(outs[k][3], ref["medianwg"]), (budget, k)


def test_batch_affine_plain_stripe_fallback(ctx):
    """Gap-free pairs outside the rotating schedule's limits run the plain
    stripe loop: more than 128 stripes (8,300 rows) and more than 65,535
    columns."""
    say("test_batch_affine_plain_stripe_fallback")
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1, 3))
    ocm = oracle.cm_dna(1, 1, 3)
    rng = np.random.default_rng(0xFA11)
    root = np.concatenate([[16], rng.choice(PLAIN, 8300)]).astype(np.uint8)
    pairs = [(mutate(rng, root, 0.02), mutate(rng, root, 0.02)),
             (np.concatenate([[16], rng.choice(PLAIN, 40)]).astype(np.uint8),
              np.concatenate([[16], rng.choice(PLAIN, 65_600)]).astype(np.uint8))]
    costs, outs = run_pairs(ctx, "align_affine", pairs)
    for k, (a, b) in enumerate(pairs):
        sw = len(a) > len(b)
        si, sj = (b, a) if sw else (a, b)
        ref = oracle.align_affine_3(si, sj, ocm, int(sw))
        ra, rb = (ref["resj"], ref["resi"]) if sw else (ref["resi"], ref["resj"])
        assert costs[k] == ref["cost"], (k, costs[k], ref["cost"])
        assert np.array_equal(outs[k][0], ra) and np.array_equal(outs[k][1], rb), k
        assert np.array_equal(outs[k][2], ref["median"]), k
        assert np.array_equal(outs[k][3], ref["medianwg"]), k


def test_batch_linear_config1_sample(ctx):
    say("test_batch_linear_config1_sample")
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 2))
    ocm = oracle.cm_dna(1, 2)
    pairs = workloads.pairs_list(1, 32)
    costs, outs = run_pairs(ctx, "align_2d", pairs)
    for k, (a, b) in enumerate(pairs):
        o1, o2, oc, _ = oracle.align_2(a, b, ocm)
        assert costs[k] == oc, (k, costs[k], oc)
        assert np.array_equal(outs[k][0], o1) and np.array_equal(outs[k][1], o2), k
        assert np.array_equal(outs[k][2], oracle.ancestor_2(o1, o2, ocm)), k
        assert np.array_equal(outs[k][3], oracle.median_2_with_gaps(o1, o2, ocm)), k
    costs2, _ = run_pairs(ctx, "cost_2d", pairs)
    assert np.array_equal(costs2, costs)


def test_batch_linear_config4_long(ctx):
    """Long banded pairs of the config-4 process (16k+ symbols, 10% IUPAC,
    sub 1 / indel 1): the six shortest of 24 draws, bit-exact in cost,
    aligned pair, ancestor median and median with gaps."""
    say("test_batch_linear_config4_long")
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1))
    ocm = oracle.cm_dna(1, 1)
    pairs = workloads.pairs_list(4, 24)
    pairs.sort(key=lambda p: len(p[0]) * len(p[1]))
    pairs = pairs[:6]
    costs, outs = run_pairs(ctx, "align_2d", pairs)
    for k, (a, b) in enumerate(pairs):
        o1, o2, oc, st = oracle.align_2(a, b, ocm)
        assert not st.full_plane
        assert costs[k] == oc, (k, costs[k], oc)
        assert np.array_equal(outs[k][0], o1) and np.array_equal(outs[k][1], o2), k
        assert np.array_equal(outs[k][2], oracle.ancestor_2(o1, o2, ocm)), k
        assert np.array_equal(outs[k][3], oracle.median_2_with_gaps(o1, o2, ocm)), k


def test_batch_linear_config4_beyond_int_indexing(ctx):
    """The three largest of 24 config-4 draws (L1 * L2 above 2^31, where
    the reference's int products at src/algn.c:985 and :1019 overflow):
    size-independent checks.  align_2 and cost_2 agree, the aligned pair
    re-prices to the cost, drops back to the inputs, and its ancestor and
    median with gaps equal the oracle's medians of that aligned pair."""
    say("test_batch_linear_config4_beyond_int_indexing")
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1))
    ocm = oracle.cm_dna(1, 1)
    pairs = workloads.pairs_list(4, 24)
    pairs.sort(key=lambda p: len(p[0]) * len(p[1]))
    pairs = pairs[-3:]
    assert all(len(a) * len(b) > 2 ** 31 for a, b in pairs)
    costs, outs = run_pairs(ctx, "align_2d", pairs)
    c2, _ = run_pairs(ctx, "cost_2d", pairs)
    assert np.array_equal(costs, c2)
    for k, (a, b) in enumerate(pairs):
        r1, r2, anc, mwg = outs[k]
        assert len(r1) == len(r2)
        assert oracle.verify_2(r1, r2, ocm) == costs[k], k
        assert np.array_equal(r1[r1 != 16], a[a != 16]) and np.array_equal(r2[r2 != 16], b[b != 16])
        assert np.array_equal(anc, oracle.ancestor_2(r1, r2, ocm)), k
        assert np.array_equal(mwg, oracle.median_2_with_gaps(r1, r2, ocm)), k


def test_batch_linear_two_phase(ctx):
    """A direction-store budget far below the full-width estimate: cost-only
    rounds for every pair, then final rounds in chunks sized to the bands.
    Same results as the one-phase plan and the oracle."""
    say("test_batch_linear_two_phase")
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 2))
    ocm = oracle.cm_dna(1, 2)
    pairs = workloads.pairs_list(1, 48)
    c1, o1 = run_pairs(ctx, "align_2d", pairs)
    c2, o2 = run_pairs(ctx, "align_2d", pairs, dm_budget=1 << 20)
    assert np.array_equal(c1, c2)
    for k, (a, b) in enumerate(pairs):
        for r in range(4):
            assert np.array_equal(o1[k][r], o2[k][r]), (k, r)
        r1, r2, oc, _ = oracle.align_2(a, b, ocm)
        assert c2[k] == oc and np.array_equal(o2[k][0], r1) and np.array_equal(o2[k][1], r2), k


def test_limit_2_and_backtrace_2d_limit_dropins(ctx):
    """algn_CAML_limit_2 and algn_CAML_backtrace_2d_limit through the C ABI
    against the oracle: costs for limits on both sides of the
    full-plane-vs-band decision, and windowed walks (start cell, st_s2 stop,
    st_s1 symbol shift) on full-plane fills, both swap orders; the unlimited
    window equals backtrace_2d on a banded fill too."""
    say("test_limit_2_and_backtrace_2d_limit_dropins")
    from poy5_amd.sequence import Align
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 2))
    ocm = oracle.cm_dna(1, 2)
    al = Align(ctx)
    rng = np.random.default_rng(0x11A)
    for t in range(40):
        n1 = int(rng.integers(2, 40))
        n2 = int(rng.integers(n1, 48))
        a = np.concatenate([[16], rng.choice(PLAIN, n1 - 1)]).astype(np.uint8)
        b = np.concatenate([[16], rng.choice(PLAIN, n2 - 1)]).astype(np.uint8)
        s1_len, s2_len = int(rng.integers(0, 400)), int(rng.integers(0, 400))
        w = int(rng.integers(0, 30))
        c = al.c_cost_2_limit(a, b, w, 7, 1, 2, s1_len, s2_len)
        oc, odm, ost = oracle.limit_2(a, b, ocm, w, s1_len, s2_len)
        assert c == oc, t
        if not ost.full_plane:
            continue
        for _ in range(6):
            st1 = int(rng.integers(0, n1))
            a1 = int(rng.integers(0, n1 - st1 + 1))
            st2 = int(rng.integers(0, n2))
            a2 = int(rng.integers(0, n2 - st2 + 1))
            sw = int(rng.integers(0, 2))
            r1, r2 = al.extract_edited_2_limit(a, b, st1, st2, a1, a2, sw)
            o1, o2 = oracle.backtrace_2d_limit(a, b, odm, st1, st2, a1, a2, sw)
            assert np.array_equal(r1, o1) and np.array_equal(r2, o2), (t, st1, st2, a1, a2, sw)
    # banded fill, whole window = backtrace_2d
    root = np.concatenate([[16], rng.choice(PLAIN, 600)]).astype(np.uint8)
    a, b = mutate(rng, root, 0.03), mutate(rng, root, 0.03)
    if len(a) > len(b):
        a, b = b, a
    c = al.c_cost_2_limit(a, b, 0, 0, 0, 0, len(a), len(b))
    oc, odm, ost = oracle.limit_2(a, b, ocm, 0, len(a), len(b))
    assert c == oc and not ost.full_plane
    for sw in (0, 1):
        r1, r2 = al.extract_edited_2_limit(a, b, 0, 0, len(a), len(b), sw)
        o1, o2 = oracle.backtrace_2d(a, b, odm, sw)
        assert np.array_equal(r1, o1) and np.array_equal(r2, o2), sw
    from poy5_amd._lib import PoyhipError
    with pytest.raises(PoyhipError):
        al.extract_edited_2_limit(a, b, 1, 0, len(a), len(b), 0)


def _orc_cm_of(m):
    """The oracle's table struct filled from a product table set."""
    o = oracle.OrcCM()
    o.lcm, o.gap, o.a_sz, o.combinations = 5, 16, 31, 1
    o.cost_model, o.gap_open = m.cost_model, m.gap_open
    o.cost[:] = [int(x) for x in m.cost]
    o.worst[:] = [int(x) for x in m.worst]
    o.median[:] = [int(x) for x in m.median]
    o.prepend[:] = [int(x) for x in m.prepend]
    o.tail[:] = [int(x) for x in m.tail]
    return o


def test_non_metric_tcm_batch_and_dropin(ctx):
    """An asymmetric, non-metric TCM (A->C 1, C->A 3, gaps 2, other changes
    2; sets from the bitwise builder with create_original): the linear batch
    and the drop-in against the oracle on the same tables, so the kernels'
    orientation of an asymmetric cost matches the reference's."""
    say("test_non_metric_tcm_batch_and_dropin")
    from poy5_amd.sequence import Align
    base = [[0 if x == y else 2 for y in range(5)] for x in range(5)]
    base[0][1], base[1][0] = 1, 3
    m = cost_matrix.of_list(base)
    assert not m.metric
    ctx.set_cm(m)
    ocm = _orc_cm_of(m)
    pairs = workloads.pairs_list(1, 24)
    costs, outs = run_pairs(ctx, "align_2d", pairs)
    al = Align(ctx)
    for k, (a, b) in enumerate(pairs):
        o1, o2, oc, _ = oracle.align_2(a, b, ocm)
        assert costs[k] == oc, k
        assert np.array_equal(outs[k][0], o1) and np.array_equal(outs[k][1], o2), k
        assert np.array_equal(outs[k][2], oracle.ancestor_2(o1, o2, ocm)), k
        if k < 6:
            assert al.cost_2(a, b) == oracle.cost_2(a, b, ocm)[0], k

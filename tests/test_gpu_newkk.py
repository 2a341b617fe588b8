"""GPU parity of the diagonal-store affine aligner (src/newkkonen.c) against
the oracle, bit-exact: the drop-ins (newkkonen_CAML_algn_affine and the
affine backtrace), NewkkAlign.align_2 / cost_2, and the batch ops (align +
ancestor_2 + median_2_with_gaps, cost only), including the two-phase plan,
both gap-count representations and a longer sequence past 2^16."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from poy5_amd import cost_matrix
from poy5_amd._lib import PoyhipError
from poy5_amd.batch import run_pairs
from poy5_amd.sequence import Context, NewkkAlign

pytestmark = pytest.mark.gpu

PLAIN = np.array([1, 2, 4, 8], np.uint8)
IUPAC = np.array([3, 5, 9, 6, 10, 12, 7, 11, 13, 14, 15], np.uint8)
GAPPY = np.array([17, 18, 20, 24, 31, 19, 28], np.uint8)

MODELS = [(1, 1, 3), (2, 1, 4), (1, 3, 2), (1, 2, 0)]


def rand_seq(rng, n, gappy=0.0, iupac=0.05):
    s = rng.choice(PLAIN, n)
    m = rng.random(n)
    s = np.where(m < iupac, rng.choice(IUPAC, n), s)
    s = np.where(m > 1 - gappy, rng.choice(GAPPY, n), s)
    return np.concatenate([[16], s]).astype(np.uint8)


def mutate(rng, s, d, gappy=0.0):
    out = [16]
    for c in s[1:]:
        u = rng.random()
        if u < d / 3:
            continue
        if u < 2 * d / 3:
            out.append(int(rng.choice(GAPPY if rng.random() < gappy else PLAIN)))
        out.append(int(rng.choice(PLAIN)) if rng.random() < d else int(c))
    return np.array(out, np.uint8)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def cases(seed, count, maxlen, gappy=0.0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        root = rand_seq(rng, int(rng.integers(0, maxlen)), gappy)
        a = mutate(rng, root, rng.uniform(0.02, 0.3), gappy)
        b = mutate(rng, root, rng.uniform(0.02, 0.3), gappy)
        out.append((a, b))
    out += [(np.array([16], np.uint8), np.array([16], np.uint8)),
            (np.array([16], np.uint8), rand_seq(rng, 5)),
            (rand_seq(rng, 1), rand_seq(rng, 1)),
            (rand_seq(rng, 7), rand_seq(rng, 7)),
            (rand_seq(rng, 2), rand_seq(rng, 250)),   # trivial case (250 > 100 x 2)
            (rand_seq(rng, 40), rand_seq(rng, 40))]
    return out


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("gappy", [0.0, 0.2])
def test_newkk_dropins_match_oracle(ctx, model, gappy):
    say(f"newkk drop-ins model={model} gappy={gappy}")
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(*model))
    ocm = oracle.cm_dna(*model)
    nk = NewkkAlign(ctx)
    for t, (a, b) in enumerate(cases(31 + model[2], 60, 300, gappy)):
        x, y = (a, b) if len(a) <= len(b) else (b, a)
        sw = t & 1
        want = oracle.newkk_align(x, y, ocm, sw)
        assert nk.newkk_cost2_affine(x, y, sw) == want["cost"], (t, list(x), list(y))
        r1, r2, c = nk.get_alignment(x, y, sw)
        assert c == want["cost"]
        assert np.array_equal(r1, want["r1"]) and np.array_equal(r2, want["r2"]), t


def test_newkk_ocaml_wrappers_match_oracle(ctx):
    say("newkk align_2 / cost_2 wrappers")
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1, 3))
    ocm = oracle.cm_dna(1, 1, 3)
    nk = NewkkAlign(ctx)
    for t, (a, b) in enumerate(cases(5, 30, 200, 0.1)):
        for fg in (True, False):
            a_, b_ = (a, b) if fg else (a[1:], b[1:])
            p1, p2, c = nk.align_2(a_, b_, first_gap=fg)
            q1, q2, d = oracle.newkk_align_2(a_, b_, ocm, first_gap=fg)
            assert c == d and np.array_equal(p1, q1) and np.array_equal(p2, q2), t
        assert nk.cost_2(a, b) == oracle.newkk_cost_2(a, b, ocm)


def test_newkk_refuses_the_linear_model(ctx):
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1))
    nk = NewkkAlign(ctx)
    with pytest.raises(PoyhipError):
        nk.newkk_cost2_affine(np.array([16, 1], np.uint8), np.array([16, 1, 2], np.uint8))


def _batch_check(ctx, model, pairs, dm_budget=0):
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(*model))
    ocm = oracle.cm_dna(*model)
    costs, outs = run_pairs(ctx, "newkk_align", pairs, dm_budget=dm_budget)
    ccosts, _ = run_pairs(ctx, "newkk_cost", pairs)
    for k, (a, b) in enumerate(pairs):
        p1, p2, c = oracle.newkk_align_2(a, b, ocm)
        assert int(costs[k]) == c and int(ccosts[k]) == c, k
        s1p, s2p, anc, mwg = outs[k]
        assert np.array_equal(s1p, p1) and np.array_equal(s2p, p2), k
        # DOS.median's use_ukk branch: Align.ancestor_2 and
        # median_2_with_gaps over the aligned pair
        assert np.array_equal(anc, oracle.ancestor_2(p1, p2, ocm)), k
        assert np.array_equal(mwg, oracle.median_2_with_gaps(p1, p2, ocm)), k
    return costs


@pytest.mark.parametrize("model", [(1, 1, 3), (2, 1, 4)])
def test_newkk_batch_matches_oracle(ctx, model):
    say(f"newkk batch model={model}")
    pairs = cases(77 + model[0], 300, 500, 0.05)
    _batch_check(ctx, model, pairs)


def test_newkk_batch_two_phase_plan(ctx):
    say("newkk batch two-phase")
    pairs = cases(91, 60, 400, 0.05)
    _batch_check(ctx, (1, 1, 3), pairs, dm_budget=1 << 20)


def _child(code: str, env: dict) -> str:
    return subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                          timeout=600).stdout


def test_newkk_both_gap_counts_give_the_same_results():
    say("newkk both gap counts (child process)")
    code = r'''
import sys
sys.path.insert(0, "tests"); sys.path.insert(0, ".")
import numpy as np
from poy5_amd import cost_matrix
from poy5_amd.batch import run_pairs
from poy5_amd.sequence import Context
import test_gpu_newkk as T
ctx = Context(0)
ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1, 3))
pairs = T.cases(13, 120, 400, 0.1)
costs, outs = run_pairs(ctx, "newkk_align", pairs)
h = 0
for k, o in enumerate(outs):
    for x in o:
        h = (h * 1000003 + int(np.asarray(x, np.int64).sum()) * 7 + len(x)) % (1 << 61)
print("RESULT", h, int(np.asarray(costs, np.int64).sum()))
'''
    env = dict(os.environ)
    env.pop("POYHIP_NEWKK_COUNTS", None)
    one = _child(code, env)
    env["POYHIP_NEWKK_COUNTS"] = "both"
    two = _child(code, env)
    r1 = [ln for ln in one.splitlines() if ln.startswith("RESULT")]
    r2 = [ln for ln in two.splitlines() if ln.startswith("RESULT")]
    assert r1 and r1 == r2, (one[-500:], two[-500:])


def test_newkk_longer_sequence_past_16_bit_counts(ctx):
    # 700 x 70,000: the insertion count passes 65,535 and wraps as the
    # reference's u16 stores wrap it
    say("newkk 700 x 70000")
    rng = np.random.default_rng(8)
    a = rand_seq(rng, 699)
    b = np.concatenate([a, rand_seq(rng, 69300)[1:]]).astype(np.uint8)
    rng.shuffle(b[1:800])
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1, 3))
    ocm = oracle.cm_dna(1, 1, 3)
    want = oracle.newkk_align(a, b, ocm, 0)
    r1, r2, c = NewkkAlign(ctx).get_alignment(a, b, 0)
    assert c == want["cost"]
    assert np.array_equal(r1, want["r1"]) and np.array_equal(r2, want["r2"])


def test_dropins_refuse_bytes_beyond_5_bit_codes(ctx):
    # every per-pair entry point checks its sequences before touching the
    # device: a byte of 32 or more would index past the 32 x 32 tables
    from poy5_amd.sequence import Align
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1, 3))
    bad = np.array([16, 1, 40, 2], np.uint8)
    good = np.array([16, 1, 2, 4, 8], np.uint8)
    with pytest.raises(PoyhipError):
        NewkkAlign(ctx).newkk_cost2_affine(bad, good)
    with pytest.raises(PoyhipError):
        Align(ctx).cost_2_affine(bad, good)
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1))
    with pytest.raises(PoyhipError):
        Align(ctx).cost_2(good, bad)


def test_newkk_config2_sample_and_long_pairs_match_oracle(ctx):
    # 200 pairs drawn from the config-2 workload (50-5,000 symbols) and two
    # long pairs of 16k-20k (the long-sequence use of NewkkAlign) through the
    # batch, against the oracle's incremental refill
    from poy5_amd import workloads
    say("newkk config-2 sample and long pairs")
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1, 3))
    ocm = oracle.cm_dna(1, 1, 3)
    pairs = workloads.pairs_list(2, 65536)[::328][:200]
    rng = np.random.default_rng(12)
    for n in (16000, 20000):
        root = rand_seq(rng, n)
        pairs.append((mutate(rng, root, 0.01), mutate(rng, root, 0.02)))
    costs, outs = run_pairs(ctx, "newkk_align", pairs)
    for k, (a, b) in enumerate(pairs):
        p1, p2, c = oracle.newkk_align_2(a, b, ocm)
        assert int(costs[k]) == c, k
        assert np.array_equal(outs[k][0], p1) and np.array_equal(outs[k][1], p2), k

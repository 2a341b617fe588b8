"""GPU parity of Sequence.Clip (src/sequence.ml:3441-3575) and
Unions.get_positions (2869-2904) against the oracle: clipped alignments of
every kind pair, corrected distances, under linear and affine tables; the
alignments run through the library."""
import sys

import numpy as np
import pytest

import oracle
from poy5_amd import clip, cost_matrix, unions
from poy5_amd.sequence import Context

pytestmark = pytest.mark.gpu

PLAIN = np.array([1, 2, 4, 8], np.uint8)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def rand_seq(rng, n):
    return np.concatenate([[16], rng.choice(PLAIN, n)]).astype(np.uint8)


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


KIND_PAIRS = [("DO", "DO"), ("First", "DO"), ("DO", "First"), ("First", "First"),
              ("Last", "DO"), ("DO", "Last"), ("Last", "Last")]


@pytest.mark.parametrize("model", [(1, 2), (1, 1, 3)])
def test_clip_align_and_distance_match_oracle(ctx, model):
    say(f"clip model={model}")
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(*model))
    ocm = oracle.cm_dna(*model)
    rng = np.random.default_rng(3 + len(model))
    for t in range(40):
        root = rand_seq(rng, int(rng.integers(5, 200)))
        a = mutate(rng, root, 0.1)
        b = mutate(rng, root, 0.1)
        # fragments: one side cut short at an end
        cut = int(rng.integers(0, max(1, len(b) // 2)))
        b = np.concatenate([b[:1], b[1 + cut:]]).astype(np.uint8) if t % 2 else b[:len(b) - cut]
        if len(b) < 2:
            continue
        for k1, k2 in KIND_PAIRS:
            got = clip.align_2(ctx, (k1, a), (k2, b))
            want = oracle.clip_align_2((k1, a), (k2, b), ocm)
            assert got[2] == want[2] and got[3] == want[3], (t, k1, k2)
            for g, w in zip(got[:2] + got[4:], want[:2] + want[4:]):
                assert g[0] == w[0] and np.array_equal(g[1], w[1]), (t, k1, k2)
        assert clip.corrected_distance(ctx, 0, a, b) == oracle.clip_corrected_distance(a, b, ocm)
    if len(model) == 3:
        # the diagonal-store aligner's branch
        assert clip.corrected_distance(ctx, 0, a, b, True) == \
            oracle.clip_corrected_distance(a, b, ocm, use_ukk=True)


def test_get_positions_on_merged_unions(ctx):
    say("union positions")
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1, 3))
    cm = oracle.cm_dna(1, 1, 3)
    rng = np.random.default_rng(9)
    for t in range(20):
        root = rand_seq(rng, int(rng.integers(10, 120)))
        sa, sb = mutate(rng, root, 0.15), mutate(rng, root, 0.15)
        x, y, sw = (sa, sb, 0) if len(sa) <= len(sb) else (sb, sa, 1)
        r = oracle.align_affine_3(x, y, cm, sw)
        ra, rb = (r["resi"], r["resj"]) if not sw else (r["resj"], r["resi"])
        u = unions.union(ctx, ra, rb, r["medianwg"], unions.leaf(sa), unions.leaf(sb))
        ou = oracle.union_merge(ra, rb, r["medianwg"], oracle.union_leaf(sa),
                                oracle.union_leaf(sb))
        n = u.length
        pos = [(0, 0)] + [tuple(sorted(rng.integers(1, n, 2).tolist())) for _ in range(5)] + \
              [(n - 1, n - 1)]
        assert clip.get_positions(u, pos) == oracle.union_get_positions(ou, pos), t

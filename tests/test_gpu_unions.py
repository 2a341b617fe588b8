"""GPU parity of the union merge (src/union.c union_merge behind
union_CAML_make) against the oracle: the hand-derived cases through the
drop-in, and unions built bottom-up over random trees, each level merged in
one launch, the children aligned as DOS.median aligns them under the affine
model (align_affine_3: aligned children and median with gaps)."""
import sys

import numpy as np
import pytest

import oracle
from poy5_amd import unions
from poy5_amd.sequence import Context

pytestmark = pytest.mark.gpu

PLAIN = np.array([1, 2, 4, 8], np.uint8)
GAPPY = np.array([17, 18, 20, 24, 31], np.uint8)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def same(u, v):
    assert u.length == v.length and u.cap == v.cap
    assert np.array_equal(u.seq[u.cap - u.length:], v.seq[v.cap - v.length:])
    for f in ("offset", "c1", "c2"):
        assert np.array_equal(getattr(u, f), getattr(v, f)), f


def test_union_hand_cases_through_the_dropin(ctx):
    say("union hand cases")
    cases = [([16, 1], [16, 2], [16, 3], [16, 1], [16, 2]),
             ([16, 1, 16], [16, 1, 4], [16, 1, 20], [16, 1], [16, 1, 4]),
             ([16, 17, 1], [16, 16, 1], [16, 16, 1], [16, 17, 1], [16, 1])]
    for a, b, m, sa, sb in cases:
        want = oracle.union_merge(a, b, m, oracle.union_leaf(sa), oracle.union_leaf(sb))
        got = unions.union(ctx, a, b, m, unions.leaf(sa), unions.leaf(sb))
        same(got, want)
    # a union with a counted item merged again
    u = oracle.union_merge([16, 17, 1], [16, 16, 1], [16, 16, 1], oracle.union_leaf([16, 17, 1]),
                           oracle.union_leaf([16, 1]))
    want = oracle.union_merge([16, 1], [16, 1], [16, 1], u, oracle.union_leaf([16, 1]))
    gu = unions.Union(u.seq, u.length, u.offset, u.c1, u.c2)
    same(unions.union(ctx, [16, 1], [16, 1], [16, 1], gu, unions.leaf([16, 1])), want)


def rand_seq(rng, n, gappy):
    s = rng.choice(PLAIN, n)
    s = np.where(rng.random(n) < gappy, rng.choice(GAPPY, n), s)
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


@pytest.mark.parametrize("gappy", [0.0, 0.1])
def test_union_trees_match_oracle_level_by_level(ctx, gappy):
    say(f"union trees gappy={gappy}")
    rng = np.random.default_rng(5 + int(gappy * 10))
    cm = oracle.cm_dna(1, 1, 3)
    for tree in range(6):
        root = rand_seq(rng, int(rng.integers(20, 300)), gappy)
        # 16 tips: a balanced tree, merged level by level
        tips = [mutate(rng, root, 0.1) for _ in range(16)]
        seqs = tips
        ounions = [oracle.union_leaf(s) for s in tips]
        gunions = [unions.leaf(s) for s in tips]
        while len(seqs) > 1:
            merges, nseqs, onew = [], [], []
            for k in range(0, len(seqs), 2):
                a, b = seqs[k], seqs[k + 1]
                x, y, sw = (a, b, 0) if len(a) <= len(b) else (b, a, 1)
                r = oracle.align_affine_3(x, y, cm, sw)
                ra, rb = (r["resi"], r["resj"]) if not sw else (r["resj"], r["resi"])
                mwg = r["medianwg"]
                onew.append(oracle.union_merge(ra, rb, mwg, ounions[k], ounions[k + 1]))
                merges.append((ra, rb, mwg, gunions[k], gunions[k + 1]))
                nseqs.append(r["median"])
            gnew = unions.union_level(ctx, merges)
            for o, g in zip(onew, gnew):
                assert g is not None
                same(g, o)
            # the drop-in on the first merge of the level
            a, b, m, ua, ub = merges[0]
            same(unions.union(ctx, a, b, m, ua, ub), onew[0])
            seqs, ounions, gunions = nseqs, onew, gnew

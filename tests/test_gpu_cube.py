"""GPU parity of the 3D cube (algn_nw_3d, backtrack_3d, median_3) against
the oracle and the golden fixtures, through the C ABI.  Bit-exact."""
import json
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def L(x):
    return [int(v) for v in x]


def rseq(rng, n, amb=0.1):
    s = [16]
    for _ in range(n):
        s.append(int(rng.choice([3, 12, 15])) if rng.random() < amb else int(rng.choice([1, 2, 4, 8])))
    return np.array(s, np.uint8)


@pytest.fixture(scope="module")
def env():
    from poy5_amd import cost_matrix
    from poy5_amd.sequence import Align, Context
    ctx = Context(0)
    two = cost_matrix.of_transformations_and_gaps(1, 1)
    ctx.set_cm(two)
    ctx.set_cm3(cost_matrix.three_d_of_two_dim_comb(two))
    yield ctx, Align(ctx), oracle.cm3_from_2d(oracle.cm_dna(1, 1))
    ctx.close()


def check_triple(al, c3, s1, s2, s3):
    want, dm = oracle.nw_3d(s1, s2, s3, c3)
    bt = oracle.backtrack_3d(s1, s2, s3, dm)
    (r1, r2, r3), cost = al.align_3(s1, s2, s3)
    assert cost == want, (len(s1), len(s2), len(s3))
    assert bt is not None
    assert [L(r1), L(r2), L(r3)] == [L(x) for x in bt]
    return bt


def test_align_3d_dropin_random(env):
    ctx, al, c3 = env
    rng = np.random.default_rng(31)
    shapes = [(1, 1, 1), (1, 5, 7), (6, 1, 4), (5, 6, 1), (2, 2, 2)]
    shapes += [tuple(int(x) for x in rng.integers(1, 40, 3)) for _ in range(25)]
    for n1, n2, n3 in shapes:
        s1, s2, s3 = rseq(rng, n1 - 1), rseq(rng, n2 - 1), rseq(rng, n3 - 1)
        bt = check_triple(al, c3, s1, s2, s3)
        m = al.median_3(*bt)
        assert L(m) == L(oracle.median_3(*bt, c3))
        assert al.cost_3(s1, s2, s3) == oracle.nw_3d(s1, s2, s3, c3)[0]
        assert [L(x) for x in al.extract_edited_3(s1, s2, s3)] == [L(x) for x in bt]


def test_align_3d_long_rows(env):
    """Rows longer than one 1024-cell segment (s3 of 1,030 and 2,100)."""
    ctx, al, c3 = env
    rng = np.random.default_rng(32)
    for n1, n2, n3 in [(3, 4, 1030), (2, 3, 2100), (4, 1, 1500)]:
        check_triple(al, c3, rseq(rng, n1 - 1), rseq(rng, n2 - 1), rseq(rng, n3 - 1))


def test_backtrack_requires_fill(env):
    ctx, al, c3 = env
    from poy5_amd._lib import PoyhipError
    a = np.array([16, 1, 2], np.uint8)
    al.cost_3(a, a, a)
    with pytest.raises(PoyhipError):
        al.extract_edited_3(a, a, np.array([16, 4], np.uint8))


def test_cube_golden(env):
    ctx, al, c3 = env
    with open(os.path.join(GOLD, "cube_3d_tcm11.json")) as f:
        g = json.load(f)
    u8 = lambda x: np.array(x, np.uint8)  # noqa: E731
    for c in g["cases"]:
        (r1, r2, r3), cost = al.align_3(u8(c["s1"]), u8(c["s2"]), u8(c["s3"]))
        assert cost == c["cost"]
        assert [L(r1), L(r2), L(r3)] == [c["r1"], c["r2"], c["r3"]]
        assert L(al.median_3(r1, r2, r3)) == c["median3"]


def test_batch_3d_config5_sample(env):
    """Smallest triples of a 512-triple draw of the config-5 generator plus
    random short ones, as one batch; cost-only batch agrees."""
    from poy5_amd import batch, workloads
    ctx, al, c3 = env
    blob, off, ln = workloads.gen_triples(512)
    vol = ln[0::3].astype(np.float64) * ln[1::3] * ln[2::3]
    pick = np.argsort(vol)[:6]
    trip = [tuple(blob[off[3 * k + u]: off[3 * k + u] + ln[3 * k + u]] for u in range(3))
            for k in pick]
    rng = np.random.default_rng(33)
    for _ in range(40):
        n = rng.integers(1, 30, 3)
        trip.append(tuple(rseq(rng, int(x) - 1) for x in n))
    costs, outs = batch.run_triples(ctx, "align_3d", trip)
    for k, (s1, s2, s3) in enumerate(trip):
        want, dm = oracle.nw_3d(s1, s2, s3, c3)
        assert costs[k] == want, k
        bt = oracle.backtrack_3d(s1, s2, s3, dm)
        assert [L(x) for x in outs[k][:3]] == [L(x) for x in bt]
        assert L(outs[k][3]) == L(oracle.median_3(*bt, c3))
    c2, _ = batch.run_triples(ctx, "cost_3d", trip)
    assert np.array_equal(c2, costs)


def test_batch_3d_workgroup_fill_equals_wavefront_fill(env):
    """The workgroup-per-triple fill (one segment rows) against the
    wavefront-per-triple fill on config-5 triples, and the oracle on the two
    smallest: costs, aligned triples and medians identical."""
    from poy5_amd import batch, workloads
    from poy5_amd import cost_matrix
    from poy5_amd.sequence import Context
    ctx, al, c3 = env
    blob, off, ln = workloads.gen_triples(256)
    vol = ln[0::3].astype(np.float64) * ln[1::3] * ln[2::3]
    pick = np.argsort(vol)[:24]
    trip = [tuple(blob[off[3 * k + u]: off[3 * k + u] + ln[3 * k + u]] for u in range(3))
            for k in pick]
    rng = np.random.default_rng(34)
    for _ in range(24):
        n = rng.integers(1, 90, 3)
        trip.append(tuple(rseq(rng, int(x) - 1) for x in n))
    # rows of two and three segments
    for n in [(6, 9, 1030), (12, 5, 2100), (3, 17, 3072), (9, 9, 1025)]:
        trip.append(tuple(rseq(rng, x - 1) for x in n))
    costs, outs = batch.run_triples(ctx, "align_3d", trip)
    os.environ["POYHIP_CUBE_FILL"] = "wave"
    try:
        wctx = Context(0)
    finally:
        del os.environ["POYHIP_CUBE_FILL"]
    try:
        two = cost_matrix.of_transformations_and_gaps(1, 1)
        wctx.set_cm(two)
        wctx.set_cm3(cost_matrix.three_d_of_two_dim_comb(two))
        wcosts, wouts = batch.run_triples(wctx, "align_3d", trip)
        wc2, _ = batch.run_triples(wctx, "cost_3d", trip)
    finally:
        wctx.close()
    assert np.array_equal(costs, wcosts)
    assert np.array_equal(wc2, costs)
    for k in range(len(trip)):
        assert [L(x) for x in outs[k]] == [L(x) for x in wouts[k]], k
    for k in [0, 1, len(trip) - 4, len(trip) - 3]:
        s1, s2, s3 = trip[k]
        want, dm = oracle.nw_3d(s1, s2, s3, c3)
        assert costs[k] == want
        bt = oracle.backtrack_3d(s1, s2, s3, dm)
        assert [L(x) for x in outs[k][:3]] == [L(x) for x in bt]


@pytest.mark.parametrize("tcm", [(5, 7), (1000000, 1500000)])
def test_cube_ring_widths_and_fallback(tcm):
    """Cost models whose cube values need 32-bit ring rows or keys too wide
    for the workgroup fill (which then falls back to the wavefront fill):
    sub 5 / indel 7 keeps 16-bit rows only while L1+L2+L3 < 666, sub 10^6 /
    indel 1.5*10^6 leaves the workgroup fill below L1+L2+L3 = 52.  Drop-in
    calls decide per triple, batches per batch; all against the oracle."""
    from poy5_amd import batch, cost_matrix
    from poy5_amd.sequence import Align, Context
    ctx = Context(0)
    try:
        two = cost_matrix.of_transformations_and_gaps(*tcm)
        ctx.set_cm(two)
        ctx.set_cm3(cost_matrix.three_d_of_two_dim_comb(two))
        al = Align(ctx)
        c3 = oracle.cm3_from_2d(oracle.cm_dna(*tcm))
        rng = np.random.default_rng(35)
        if tcm == (5, 7):
            small = [tuple(int(x) for x in rng.integers(5, 200, 3)) for _ in range(10)]
            large = [(300, 250, 330), (180, 420, 260)]
        else:
            small = [tuple(int(x) for x in rng.integers(3, 17, 3)) for _ in range(10)]
            large = [(20, 30, 25), (33, 12, 40)]
        for shapes in (small, small + large):
            trip = [tuple(rseq(rng, n - 1) for n in sh) for sh in shapes]
            costs, outs = batch.run_triples(ctx, "align_3d", trip)
            for k, (s1, s2, s3) in enumerate(trip):
                want, dm = oracle.nw_3d(s1, s2, s3, c3)
                assert costs[k] == want, (tcm, shapes[k])
                bt = oracle.backtrack_3d(s1, s2, s3, dm)
                assert [L(x) for x in outs[k][:3]] == [L(x) for x in bt], (tcm, shapes[k])
        for s1, s2, s3 in trip[-3:]:
            want, dm = oracle.nw_3d(s1, s2, s3, c3)
            (r1, r2, r3), cost = al.align_3(s1, s2, s3)
            assert cost == want
            assert [L(r1), L(r2), L(r3)] == [L(x) for x in oracle.backtrack_3d(s1, s2, s3, dm)]
    finally:
        ctx.close()

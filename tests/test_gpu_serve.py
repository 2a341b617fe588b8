"""Continuous-batching server: results equal the oracle and do not depend on
what else shares a batch (alone vs in the mix, small vs large slots)."""
import numpy as np
import pytest

import oracle
from poy5_amd import cost_matrix, spr
from poy5_amd.sequence import Context
from poy5_amd.serve import Server

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    c.set_cm(cost_matrix.of_transformations_and_gaps(1, 2))
    yield c
    c.close()


def oracle_medians(children, leaves, ocm):
    med = dict(leaves)
    for v in spr.post_order(children):
        if children[v]:
            a, b = children[v]
            r1, r2, _, _ = oracle.align_2(med[a], med[b], ocm)
            med[v] = oracle.ancestor_2(r1, r2, ocm)
    return med


def test_downpass_and_rejoins_match_oracle(ctx):
    print("test_downpass_and_rejoins_match_oracle", flush=True)
    ocm = oracle.cm_dna(1, 2)
    parent, children, leaves = spr.simulate(n_taxa=16, root_len=300, seed=11)
    srv = Server(ctx, n_slots=3, max_batch_pairs=256, max_len=2048, slot_dm_bytes=64 << 20,
                 pool_bytes=16 << 20)
    try:
        mid = spr.downpass(srv, parent, children, leaves)
        want = oracle_medians(children, leaves, ocm)
        for v in mid:
            assert np.array_equal(srv.get(mid[v]), want[v]), v
        rng = np.random.default_rng(3)
        a, b = spr.rejoin_requests(children, mid, 300, rng)
        inv = {m: v for v, m in mid.items()}
        costs, lat, _, wall = srv.replay(np.zeros(len(a)), ["cost_2d"] * len(a), a, b)
        for k in range(len(a)):
            oc = oracle.cost_2(want[inv[a[k]]], want[inv[b[k]]], ocm)[0]
            assert costs[k] == oc, k
        assert np.all(lat > 0)
    finally:
        srv.close()


def test_alone_equals_mix(ctx):
    """Each request alone (submit, wait) and the same requests in one burst
    from the native load generator give identical results."""
    print("test_alone_equals_mix", flush=True)
    parent, children, leaves = spr.simulate(n_taxa=24, root_len=400, seed=12)
    srv = Server(ctx, n_slots=3, max_batch_pairs=512, max_len=2048, slot_dm_bytes=32 << 20,
                 pool_bytes=16 << 20)
    try:
        ids = [srv.put(s) for s in leaves.values()]
        rng = np.random.default_rng(5)
        pairs = [(int(rng.choice(ids)), int(rng.choice(ids))) for _ in range(120)]
        alone = []
        for k, (x, y) in enumerate(pairs):
            op = "align_2d" if k % 3 == 0 else "cost_2d"
            r, _ = srv.submit(op, x, y, deltaw=None if k % 2 else 8)
            srv.wait(r)
            alone.append(srv.result(r)[:3])
        # the mix: all at once, several batches in flight
        reqs = []
        for k, (x, y) in enumerate(pairs):
            op = "align_2d" if k % 3 == 0 else "cost_2d"
            reqs.append(srv.submit(op, x, y, deltaw=None if k % 2 else 8)[0])
        for k, r in enumerate(reqs):
            srv.wait(r)
            c, r1, r2, _ = srv.result(r)
            assert c == alone[k][0], k
            assert np.array_equal(r1, alone[k][1]) and np.array_equal(r2, alone[k][2]), k
        st = srv.stats()
        assert st["pairs"] == 2 * len(pairs)
        assert st["batches"] < 2 * len(pairs)  # the mix was batched
    finally:
        srv.close()


def test_affine_requests_match_oracle(ctx):
    print("test_affine_requests_match_oracle", flush=True)
    c2 = Context(0)
    c2.set_cm(cost_matrix.of_transformations_and_gaps(1, 1, 3))
    ocm = oracle.cm_dna(1, 1, 3)
    srv = Server(c2, n_slots=2, max_batch_pairs=128, max_len=4096, slot_dm_bytes=64 << 20,
                 pool_bytes=8 << 20)
    try:
        parent, children, leaves = spr.simulate(n_taxa=12, root_len=500, seed=13)
        seqs = list(leaves.values())
        ids = [srv.put(s) for s in seqs]
        reqs = []
        for k in range(len(ids) - 1):
            reqs.append((k, srv.submit("align_affine", ids[k], ids[k + 1])[0],
                         srv.submit("cost_affine", ids[k], ids[k + 1])[0]))
        for k, ra, rc in reqs:
            a, b = seqs[k], seqs[k + 1]
            srv.wait(ra)
            c, r1, r2, _ = srv.result(ra)
            sw = len(a) > len(b)
            si, sj = (b, a) if sw else (a, b)
            ref = oracle.align_affine_3(si, sj, ocm, int(sw))
            ra_, rb_ = (ref["resj"], ref["resi"]) if sw else (ref["resi"], ref["resj"])
            assert c == ref["cost"] and np.array_equal(r1, ra_) and np.array_equal(r2, rb_), k
            srv.wait(rc)
            assert srv.result(rc)[0] == oracle.cost_affine_3(a, b, ocm), k
    finally:
        srv.close()
        c2.close()

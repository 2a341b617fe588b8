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


def check_spr_workload(ctx, n_taxa, root_len, seed, n_req, slot_dm_bytes):
    """Down pass, edge medians and rejoin costs of a simulated tree on the
    server, each against the oracle: node medians (align_2 + ancestor_2),
    edge medians (the same on the edge's two node medians) and
    cost_2(clade median, edge median) with DOS.distance's deltaw."""
    ocm = oracle.cm_dna(1, 2)
    parent, children, leaves = spr.simulate(n_taxa=n_taxa, root_len=root_len, seed=seed)
    srv = Server(ctx, n_slots=3, max_batch_pairs=2048, max_len=4096,
                 slot_dm_bytes=slot_dm_bytes, pool_bytes=64 << 20)
    try:
        mid = spr.downpass(srv, parent, children, leaves)
        want = oracle_medians(children, leaves, ocm)
        for v in mid:
            assert np.array_equal(srv.get(mid[v]), want[v]), v
        emid = spr.edge_medians(srv, parent, mid)
        ewant = {}
        for v in emid:
            r1, r2, _, _ = oracle.align_2(want[v], want[parent[v]], ocm)
            ewant[v] = oracle.ancestor_2(r1, r2, ocm)
            assert np.array_equal(srv.get(emid[v]), ewant[v]), ("edge", v)
        rng = np.random.default_rng(3)
        a, b, cv, ev = spr.rejoin_requests(parent, children, mid, emid, n_req, rng)
        dw, _ = spr.rejoin_deltaw(srv, a, b)
        costs, lat, _, wall = srv.replay(np.zeros(len(a)), ["cost_2d"] * len(a), a, b, deltaw=dw)
        for k in range(len(a)):
            x, y = want[int(cv[k])], ewant[int(ev[k])]
            assert dw[k] == max(abs(len(x) - len(y)), 8)
            oc = oracle.cost_2(x, y, ocm, deltaw=int(dw[k]))[0]
            assert costs[k] == oc, k
        assert np.all(lat > 0)
    finally:
        srv.close()


def test_downpass_and_rejoins_match_oracle(ctx):
    print("test_downpass_and_rejoins_match_oracle", flush=True)
    check_spr_workload(ctx, 16, 300, 11, 300, 64 << 20)


def test_spr_config3_full_size(ctx):
    """The serving workload at its configuration: the seeded 128-taxon tree
    with root length 1,500 (spr.simulate defaults), all 127 down-pass
    medians, all 254 edge medians and 1,200 rejoin costs."""
    print("test_spr_config3_full_size", flush=True)
    check_spr_workload(ctx, 128, 1500, 0x504F59003, 1200, 1 << 30)


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


def test_admission_refuses_over_budget_and_serves_the_rest(ctx):
    """A request whose direction store exceeds a slot's budget is refused
    with POYHIP_EBUSY at submission; the requests around it still match the
    oracle."""
    print("test_admission_refuses_over_budget_and_serves_the_rest", flush=True)
    from poy5_amd import _lib
    ocm = oracle.cm_dna(1, 2)
    rng = np.random.default_rng(21)
    small = [np.concatenate([[16], rng.choice(spr.BASES, n)]).astype(np.uint8)
             for n in (150, 170, 190, 210)]
    big = np.concatenate([[16], rng.choice(spr.BASES, 3000)]).astype(np.uint8)
    # ~0.25 B per cell: 3001 x 3001 cells need ~2.3 MB, over the 1 MiB budget
    srv = Server(ctx, n_slots=2, max_batch_pairs=64, max_len=4096, slot_dm_bytes=1 << 20,
                 pool_bytes=4 << 20)
    try:
        ids = [srv.put(s) for s in small]
        bid = srv.put(big)
        ok = []
        for k in range(len(ids) - 1):
            ok.append((k, srv.submit("align_2d", ids[k], ids[k + 1])[0]))
            if k == 1:
                with pytest.raises(_lib.PoyhipError) as e:
                    srv.submit("align_2d", bid, bid)
                assert e.value.code == -6  # POYHIP_EBUSY
        for k, r in ok:
            srv.wait(r)
            c, r1, r2, _ = srv.result(r)
            o1, o2, oc, _ = oracle.align_2(small[k], small[k + 1], ocm)
            assert c == oc and np.array_equal(r1, o1) and np.array_equal(r2, o2), k
    finally:
        srv.close()


def test_failed_producer_releases_waiting_submissions(ctx):
    """A request that keeps its median fails at launch (forced from the host,
    no kernel runs); a later submission naming that median returns an error
    instead of blocking, and so does every submission after it."""
    print("test_failed_producer_releases_waiting_submissions", flush=True)
    import threading
    from poy5_amd import _lib
    rng = np.random.default_rng(22)
    seqs = [np.concatenate([[16], rng.choice(spr.BASES, 120)]).astype(np.uint8)
            for _ in range(3)]
    srv = Server(ctx, n_slots=1, max_batch_pairs=16, max_len=1024, slot_dm_bytes=8 << 20,
                 pool_bytes=1 << 20)
    try:
        ids = [srv.put(s) for s in seqs]
        ctx.chk(srv.L.poyhip_server_fail_next_launch(srv.h))
        r, m = srv.submit("align_2d", ids[0], ids[1], keep_median=True)
        out = {}

        def dependent():
            try:
                srv.submit("cost_2d", m, ids[2])
                out["rc"] = 0
            except _lib.PoyhipError as e:
                out["rc"] = e.code

        t = threading.Thread(target=dependent, daemon=True)
        t.start()
        t.join(timeout=30)
        assert not t.is_alive(), "submission naming a failed median blocked"
        assert out["rc"] == -3  # POYHIP_EDEVICE
        with pytest.raises(_lib.PoyhipError) as e:
            srv.wait(r)
        assert e.value.code == -3
    finally:
        srv.close()


@pytest.mark.parametrize("model", [(1, 1, 3), (2, 1, 4), (1, 3, 2)])
def test_affine_gap_bit_and_gap_free_mix_match_oracle(model):
    """Serving affine requests: pairs with gap-bit symbols (codes 17-31) and
    gap-free pairs in the same slot batch; cost, aligned pair and, for align
    requests, the kept alignment match oracle.align_affine_3 under several
    cost models (minimum non-zero cost 1 and 2)."""
    print(f"test_affine_gap_bit_and_gap_free_mix_match_oracle {model}", flush=True)
    c2 = Context(0)
    c2.set_cm(cost_matrix.of_transformations_and_gaps(*model))
    ocm = oracle.cm_dna(*model)
    rng = np.random.default_rng(23 + model[0] + 10 * model[2])
    codes = np.array([1, 2, 4, 8, 1, 2, 4, 8, 3, 5, 15], np.uint8)
    gapc = np.arange(17, 32, dtype=np.uint8)
    seqs = []
    for k in range(16):
        s = rng.choice(codes, int(rng.integers(80, 400)))
        if k % 4 >= 2:  # pairs (0,1), (4,5), ... stay gap-free
            m = rng.random(len(s)) < 0.15
            s[m] = rng.choice(gapc, int(m.sum()))
        seqs.append(np.concatenate([[16], s]).astype(np.uint8))
    srv = Server(c2, n_slots=1, max_batch_pairs=64, max_len=1024, slot_dm_bytes=64 << 20,
                 pool_bytes=2 << 20)
    try:
        ids = [srv.put(s) for s in seqs]
        reqs = [(k, srv.submit("align_affine", ids[k], ids[k + 1])[0],
                 srv.submit("cost_affine", ids[k], ids[k + 1])[0]) for k in range(len(ids) - 1)]
        for k, ra, rc in reqs:
            a, b = seqs[k], seqs[k + 1]
            srv.wait(ra)
            c, r1, r2, _ = srv.result(ra)
            sw = len(a) > len(b)
            si, sj = (b, a) if sw else (a, b)
            ref = oracle.align_affine_3(si, sj, ocm, int(sw))
            ra_, rb_ = (ref["resj"], ref["resi"]) if sw else (ref["resi"], ref["resj"])
            assert c == ref["cost"], k
            assert np.array_equal(r1, ra_) and np.array_equal(r2, rb_), k
            srv.wait(rc)
            assert srv.result(rc)[0] == oracle.cost_affine_3(a, b, ocm), k
    finally:
        srv.close()
        c2.close()

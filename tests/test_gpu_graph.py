"""HIP graph replay of a planned batch: every replay gives the eager run's
costs and output sequences bit for bit, with eager work of the same
context, torch work and host synchronisations between replays."""
import numpy as np
import pytest
import torch

import oracle
from poy5_amd import cost_matrix, workloads
from poy5_amd.batch import Batch, pack_pairs
from poy5_amd.sequence import Align, Context

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def used_outputs(b, out, outlen):
    """Costs aside, only the written part of each output region counts."""
    oh = out.cpu().numpy()
    lh = outlen.cpu().numpy()
    return [b.unpack(oh, lh, k) for k in range(b.n)]


@pytest.mark.parametrize("op,model", [("align_affine", (1, 1, 3)), ("align_2d", (1, 2, None)),
                                      ("cost_affine", (1, 1, 3))])
def test_graph_replays_equal_eager(ctx, op, model):
    print(f"test_graph_replays_equal_eager {op}", flush=True)
    tr, ga, go = model
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(tr, ga, go))
    pairs = workloads.pairs_list(2, 160)
    pairs = [p for p in pairs if len(p[0]) * len(p[1]) < 2_000_000][:96]
    blob, off1, off2, len1, len2 = pack_pairs(pairs)
    d_blob = torch.from_numpy(blob).cuda()
    d1, d2 = torch.from_numpy(off1).cuda(), torch.from_numpy(off2).cuda()
    b = Batch(ctx, op, len1, len2)
    try:
        cost, out, outlen = b.alloc_outputs("cuda")
        b.run(d_blob, d1, d2, cost, out, outlen, sync=True)
        want_cost = cost.cpu().numpy().copy()
        want_out = used_outputs(b, out, outlen) if op != "cost_affine" else None
        # a few costs against the oracle, so the eager run is itself right
        ocm = oracle.cm_dna(tr, ga, go)
        for k in range(0, len(pairs), 24):
            a, c = pairs[k]
            if go is None:
                assert want_cost[k] == oracle.align_2(a, c, ocm)[2]
            else:
                sw = len(a) > len(c)
                si, sj = (c, a) if sw else (a, c)
                assert want_cost[k] == oracle.align_affine_3(si, sj, ocm, int(sw))["cost"]
        b.use_graph(True)
        al = Align(ctx)
        for rep in range(4):
            cost.fill_(-7)
            out.zero_()
            outlen.fill_(-1)
            b.run(d_blob, d1, d2, cost, out, outlen, sync=True)
            assert np.array_equal(cost.cpu().numpy(), want_cost), rep
            if want_out is not None:
                got = used_outputs(b, out, outlen)
                for k in range(b.n):
                    for r in range(4):
                        assert np.array_equal(got[k][r], want_out[k][r]), (rep, k, r)
            # eager work between replays: a drop-in call, torch kernels, syncs
            x, y = pairs[rep]
            al.cost_2(x, y) if go is None else al.align_affine_3(x, y)
            torch.cuda.synchronize()
            _ = (torch.arange(1 << 20, device="cuda") * 3).sum().item()
        # new output buffers: the graph is captured again for them
        cost2, out2, outlen2 = b.alloc_outputs("cuda")
        b.run(d_blob, d1, d2, cost2, out2, outlen2, sync=True)
        assert np.array_equal(cost2.cpu().numpy(), want_cost)
    finally:
        b.use_graph(False)
        b.close()

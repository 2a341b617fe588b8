"""Tuning knobs and the autotuner: every setting gives the same costs and
output sequences; the autotuner picks one of its candidates and leaves it on
the context."""
import numpy as np
import pytest

from poy5_amd import autotune, cost_matrix, workloads
from poy5_amd._lib import PoyhipError
from poy5_amd.batch import pack_pairs, run_pairs, run_triples
from poy5_amd.sequence import Context

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def same_outputs(x, y):
    return all(np.array_equal(p, q) for a, b in zip(x, y) for p, q in zip(a, b))


@pytest.mark.parametrize("op,model", [("align_affine", (1, 1, 3)), ("align_2d", (1, 2, None))])
def test_results_do_not_depend_on_knobs(ctx, op, model):
    print(f"test_results_do_not_depend_on_knobs {op}", flush=True)
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(*model))
    pairs = workloads.pairs_list(2, 128)
    pairs = [p for p in pairs if len(p[0]) * len(p[1]) < 3_000_000][:64]
    default = autotune.get_knobs(ctx)
    try:
        base_c, base_o = run_pairs(ctx, op, pairs)
        for cand in autotune.candidates(op):
            autotune.set_knobs(ctx, cand)
            c, o = run_pairs(ctx, op, pairs)
            assert np.array_equal(c, base_c), cand
            assert same_outputs(o, base_o), cand
    finally:
        autotune.set_knobs(ctx, default)


def test_cube_knob(ctx):
    print("test_cube_knob", flush=True)
    two = cost_matrix.of_transformations_and_gaps(1, 1)
    ctx.set_cm(two)
    ctx.set_cm3(cost_matrix.three_d_of_two_dim_comb(two))
    rng = np.random.default_rng(9)
    trip = [tuple(np.concatenate([[16], rng.choice([1, 2, 4, 8], int(n))]).astype(np.uint8)
                  for n in rng.integers(20, 120, 3)) for _ in range(12)]
    c1, o1 = run_triples(ctx, "align_3d", trip)
    autotune.set_knobs(ctx, {"cube_workgroup": 0})
    try:
        c0, o0 = run_triples(ctx, "align_3d", trip)
    finally:
        autotune.set_knobs(ctx, {"cube_workgroup": 1})
    assert np.array_equal(c0, c1) and same_outputs(o0, o1)


def test_autotune_picks_a_candidate(ctx):
    print("test_autotune_picks_a_candidate", flush=True)
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1, 3))
    blob, off, ln = workloads.gen_pairs(2, 600)
    l1, l2 = ln[0::2].astype(np.int32), ln[1::2].astype(np.int32)
    default = autotune.get_knobs(ctx)
    try:
        res = autotune.tune_pairs(ctx, "align_affine", blob, off[0::2], off[1::2], l1, l2,
                                  sample=48, repeats=1)
        assert res["choice"] in autotune.candidates("align_affine")
        got = autotune.get_knobs(ctx)
        assert all(got[k] == v for k, v in res["choice"].items())
        assert len(res["buckets"]) >= 2
    finally:
        autotune.set_knobs(ctx, default)


def test_set_tuning_rejects_bad_values(ctx):
    with pytest.raises(PoyhipError):
        autotune.set_knobs(ctx, {"waves_per_cu": 0})
    with pytest.raises(PoyhipError):
        ctx.chk(ctx._lib.poyhip_set_tuning(ctx.h, 99, 1))
    assert ctx._lib.poyhip_get_tuning(ctx.h, 99) < 0


@pytest.mark.parametrize("n", [300, 700, 1500, 5000])
def test_linear_batch_sizes_match_oracle(ctx, n):
    # the linear fill follows the batch size (a workgroup per pair, or one
    # wavefront per pair from 16 pairs per CU): sizes on both sides against
    # the oracle's Align.align_2
    import oracle
    print(f"test_linear_batch_sizes_match_oracle n={n}", flush=True)
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 2))
    ocm = oracle.cm_dna(1, 2)
    rng = np.random.default_rng(n)
    plain = np.array([1, 2, 4, 8], np.uint8)
    pairs = []
    for _ in range(n):
        root = rng.choice(plain, int(rng.integers(10, 120)))
        sides = []
        for _ in range(2):
            s = root.copy()
            s[rng.random(len(s)) < 0.1] = rng.choice(plain)
            s = s[rng.random(len(s)) > 0.03]
            sides.append(np.concatenate([[16], s]).astype(np.uint8))
        pairs.append(tuple(sides))
    costs, outs = run_pairs(ctx, "align_2d", pairs)
    for k in range(0, n, max(1, n // 200)):
        a, b = pairs[k]
        p1, p2, c, _ = oracle.align_2(a, b, ocm)
        assert int(costs[k]) == c, k
        assert np.array_equal(outs[k][0], p1) and np.array_equal(outs[k][1], p2), k



def test_batches_reuse_device_blocks(ctx):
    # a destroyed batch's device blocks go to the context's next batches with
    # whatever they last held: a large affine batch, then a smaller one that
    # takes its blocks, then one after poyhip_flush freed them, each against
    # the oracle's align_affine_3
    import oracle
    print("test_batches_reuse_device_blocks", flush=True)
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1, 3))
    ocm = oracle.cm_dna(1, 1, 3)
    pairs = [p for p in workloads.pairs_list(2, 900, seed=0xB10C)
             if len(p[0]) * len(p[1]) < 1_200_000]
    for label, part in (("large", pairs), ("reused", pairs[:150]), ("flushed", pairs[150:300])):
        if label == "flushed":
            ctx.chk(ctx._lib.poyhip_flush(ctx.h))
        print(f"  {label}: {len(part)} pairs", flush=True)
        costs, outs = run_pairs(ctx, "align_affine", part)
        for k in range(0, len(part), 9):
            a, b = part[k]
            sw = len(a) > len(b)
            si, sj = (b, a) if sw else (a, b)
            ref = oracle.align_affine_3(si, sj, ocm, int(sw))
            ra, rb = (ref["resj"], ref["resi"]) if sw else (ref["resi"], ref["resj"])
            assert costs[k] == ref["cost"], (label, k)
            assert np.array_equal(outs[k][0], ra) and np.array_equal(outs[k][1], rb), (label, k)
            assert np.array_equal(outs[k][2], ref["median"]), (label, k)

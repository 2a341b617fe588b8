"""poyhip_align_affine_3_noalloc: the allocation-free form of
poyhip_align_affine_3 (fixed device scratch) gives the same results as the
allocating entry point and the oracle, call after call, with the second call
of each sequence smaller than the first (the scratch keeps the previous
call's contents)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from poy5_amd import _lib, cost_matrix
from poy5_amd.sequence import Context

pytestmark = pytest.mark.gpu

PLAIN = np.array([1, 2, 4, 8], np.uint8)
GAPPY = np.array([17, 18, 20, 24, 31], np.uint8)


def seq(rng, n, gappy):
    s = rng.choice(PLAIN, n)
    if gappy:
        m = rng.random(n) < 0.15
        s[m] = rng.choice(GAPPY, int(m.sum()))
    return np.concatenate([[16], s]).astype(np.uint8)


def call(ctx, fn, si, sj, sw):
    cap = len(si) + len(sj) + 2
    bufs = [np.zeros(cap, np.uint8) for _ in range(4)]
    ml, mwl, rl, cost = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    P = lambda x: x.ctypes.data_as(C.POINTER(C.c_uint8))  # noqa: E731
    rc = fn(ctx.h, P(si), len(si), P(sj), len(sj), sw, P(bufs[0]), C.byref(ml), P(bufs[1]),
            C.byref(mwl), P(bufs[2]), P(bufs[3]), C.byref(rl), cap, C.byref(cost))
    if rc != 0:
        return rc, None
    return 0, (cost.value, bufs[2][:rl.value].copy(), bufs[3][:rl.value].copy(),
               bufs[0][:ml.value].copy(), bufs[1][:mwl.value].copy())


@pytest.mark.parametrize("gappy", [False, True])
def test_noalloc_matches_allocating_entry_point(gappy):
    print(f"test_noalloc_matches_allocating_entry_point gappy={gappy}", flush=True)
    ctx = Context(0)
    try:
        ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1, 3))
        ocm = oracle.cm_dna(1, 1, 3)
        L = _lib.lib()
        rng = np.random.default_rng(77 + int(gappy))
        # two calls in a row, the second smaller, then again with a swap
        for sizes in [(2900, 3100), (400, 450), (1200, 1500), (60, 61)]:
            si, sj = seq(rng, sizes[0], gappy), seq(rng, sizes[1], gappy)
            for sw in (0, 1):
                rc, got = call(ctx, L.poyhip_align_affine_3_noalloc, si, sj, sw)
                assert rc == 0, (sizes, sw, rc, L.poyhip_last_error(ctx.h))
                rc, want = call(ctx, L.poyhip_align_affine_3, si, sj, sw)
                assert rc == 0
                assert got[0] == want[0], (sizes, sw, got[0], want[0])
                for q in range(1, 5):
                    assert np.array_equal(got[q], want[q]), (sizes, sw, q)
                ref = oracle.align_affine_3(si, sj, ocm, sw)
                assert got[0] == ref["cost"]
                assert np.array_equal(got[1], ref["resi"]) and np.array_equal(got[2], ref["resj"])
                assert np.array_equal(got[3], ref["median"])
                assert np.array_equal(got[4], ref["medianwg"])
    finally:
        ctx.close()


def test_noalloc_rejects_pairs_beyond_the_scratch():
    print("test_noalloc_rejects_pairs_beyond_the_scratch", flush=True)
    ctx = Context(0)
    try:
        ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1, 3))
        L = _lib.lib()
        rng = np.random.default_rng(5)
        si, sj = seq(rng, 9000, False), seq(rng, 9000, False)
        rc, _ = call(ctx, L.poyhip_align_affine_3_noalloc, si, sj, 0)
        assert rc == -5  # POYHIP_ECAP
        assert b"scratch" in L.poyhip_last_error(ctx.h)
        # the scratch is still usable after the refusal
        si, sj = seq(rng, 300, False), seq(rng, 330, False)
        rc, got = call(ctx, L.poyhip_align_affine_3_noalloc, si, sj, 0)
        assert rc == 0
        rc, want = call(ctx, L.poyhip_align_affine_3, si, sj, 0)
        assert got[0] == want[0] and all(np.array_equal(got[q], want[q]) for q in range(1, 5))
    finally:
        ctx.close()

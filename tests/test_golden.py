"""Golden fixtures (tests/golden/*.json, written by make_golden.py).

CPU: the oracle still reproduces every committed vector.
GPU: the drop-in entry points of libpoyhip reproduce them through the C ABI.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def u8(x):
    return np.array(x, np.uint8)


def L(x):
    return [int(v) for v in x]


# ---------------------------------------------------------------- oracle --

def test_oracle_affine_golden():
    import oracle
    g = load("affine_align_3.json")
    cm = oracle.cm_dna(1, 1, 3)
    for c in g["cases"]:
        r = oracle.align_affine_3(u8(c["si"]), u8(c["sj"]), cm, c["swaped"])
        assert r["cost"] == c["cost"]
        assert L(r["resi"]) == c["resi"] and L(r["resj"]) == c["resj"]
        assert L(r["median"]) == c["median"] and L(r["medianwg"]) == c["medianwg"]
        assert oracle.cost_affine_3(u8(c["si"]), u8(c["sj"]), cm) == c["cost_full"]


@pytest.mark.parametrize("name", ["linear_align_2_tcm12.json", "linear_align_2_tcm11.json"])
def test_oracle_linear_golden(name):
    import oracle
    g = load(name)
    cm = oracle.cm_dna(*g["tcm"])
    for c in g["cases"]:
        r1, r2, cost, _ = oracle.align_2(u8(c["s1"]), u8(c["s2"]), cm)
        assert (L(r1), L(r2), cost) == (c["r1"], c["r2"], c["cost"])
        cd, _, _ = oracle.cost_2(u8(c["s1"]), u8(c["s2"]), cm, c["deltaw"])
        assert cd == c["cost_deltaw"]


def test_oracle_medians_golden():
    import oracle
    g = load("medians_tcm12.json")
    cm = oracle.cm_dna(1, 2)
    for c in g["cases"]:
        a, b = u8(c["a"]), u8(c["b"])
        assert L(oracle.ancestor_2(a, b, cm)) == c["ancestor"]
        assert L(oracle.median_2_no_gaps(a, b, cm)) == c["median"]
        assert L(oracle.median_2_with_gaps(a, b, cm)) == c["median_wg"]
        assert oracle.worst_2(a, b, cm) == c["worst"]
        assert oracle.verify_2(a, b, cm) == c["verify"]


def test_oracle_cube_golden():
    import oracle
    g = load("cube_3d_tcm11.json")
    c3 = oracle.cm3_from_2d(oracle.cm_dna(1, 1))
    for c in g["cases"]:
        cost, dm = oracle.nw_3d(u8(c["s1"]), u8(c["s2"]), u8(c["s3"]), c3)
        assert cost == c["cost"]
        bt = oracle.backtrack_3d(u8(c["s1"]), u8(c["s2"]), u8(c["s3"]), dm)
        assert (bt is not None) == c["walk_ok"]
        if bt is not None:
            assert [L(x) for x in bt] == [c["r1"], c["r2"], c["r3"]]
            assert L(oracle.median_3(*bt, c3)) == c["median3"]


# ------------------------------------------------------------------- GPU --

@pytest.fixture(scope="module")
def ctx():
    from poy5_amd import cost_matrix
    from poy5_amd.sequence import Context
    c = Context(0)
    yield c, cost_matrix
    c.close()


@pytest.mark.gpu
def test_gpu_affine_golden(ctx):
    from poy5_amd.sequence import Align
    c0, cmod = ctx
    c0.set_cm(cmod.of_transformations_and_gaps(1, 1, 3))
    al = Align(c0)
    g = load("affine_align_3.json")
    for c in g["cases"]:
        m, ri, rj, cost, mwg = al.c_align_affine_3(u8(c["si"]), u8(c["sj"]), c["swaped"])
        assert cost == c["cost"], c
        assert L(ri) == c["resi"] and L(rj) == c["resj"]
        assert L(m) == c["median"] and L(mwg) == c["medianwg"]
        assert al.cost_2_affine(u8(c["si"]), u8(c["sj"])) == c["cost_full"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["linear_align_2_tcm12.json", "linear_align_2_tcm11.json"])
def test_gpu_linear_golden(ctx, name):
    from poy5_amd.sequence import Align
    c0, cmod = ctx
    g = load(name)
    c0.set_cm(cmod.of_transformations_and_gaps(*g["tcm"]))
    al = Align(c0)
    for c in g["cases"]:
        r1, r2, cost = al.align_2(u8(c["s1"]), u8(c["s2"]))
        assert (L(r1), L(r2), cost) == (c["r1"], c["r2"], c["cost"]), c
        assert al.cost_2(u8(c["s1"]), u8(c["s2"]), c["deltaw"]) == c["cost_deltaw"]


@pytest.mark.gpu
def test_gpu_medians_golden(ctx):
    from poy5_amd.sequence import Align
    c0, cmod = ctx
    c0.set_cm(cmod.of_transformations_and_gaps(1, 2))
    al = Align(c0)
    g = load("medians_tcm12.json")
    for c in g["cases"]:
        a, b = u8(c["a"]), u8(c["b"])
        assert L(al.ancestor_2(a, b)) == c["ancestor"]
        assert L(al.median_2(a, b)) == c["median"]
        assert L(al.median_2_with_gaps(a, b)) == c["median_wg"]
        out = C.c_int()
        P = lambda v: v.ctypes.data_as(C.POINTER(C.c_uint8))  # noqa: E731
        c0.chk(c0._lib.poyhip_worst_2(c0.h, P(a), P(b), len(a), C.byref(out)))
        assert out.value == c["worst"]

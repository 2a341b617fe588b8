#!/usr/bin/env python3
"""Writes the golden fixtures of this directory from the oracle.

The reference holds no golden vectors or known-answer tests for the DO
alignment path, so these fixtures are produced by the oracle after it has
been pinned by tests/test_oracle_pins.py and tests/test_tables.py.  They
freeze its outputs on small seeded inputs (including ambiguity codes,
gap-bit symbols, empty sides and both swap orders) so that the oracle and
the GPU path are both checked against the same committed data.

Run from the repository root:  python3 tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

BASES = (1, 2, 4, 8)
IUPAC = (3, 5, 9, 6, 10, 12, 7, 11, 13, 14, 15)
GAPPY = (17, 18, 20, 24, 31)


def rseq(rng, n, amb, gapbit):
    s = [16]
    for _ in range(n):
        r = rng.random()
        if r < amb:
            s.append(int(rng.choice(IUPAC)))
        elif r < amb + gapbit:
            s.append(int(rng.choice(GAPPY)))
        else:
            s.append(int(rng.choice(BASES)))
    return s


def related(rng, n, amb, gapbit, edits):
    a = rseq(rng, n, amb, gapbit)
    b = a[1:]
    for _ in range(edits):
        pos = int(rng.integers(0, len(b) + 1))
        op = int(rng.integers(0, 3))
        if op == 0 and pos < len(b):
            b[pos] = int(rng.choice(BASES))
        elif op == 1:
            b.insert(pos, int(rng.choice(BASES + IUPAC[:2])))
        elif pos < len(b):
            del b[pos]
    return a, [16] + b


def L(x):
    return [int(v) for v in x]


def u8(x):
    return np.array(x, np.uint8)


def affine_cases(rng):
    cm = oracle.cm_dna(1, 1, 3)
    out = []
    shapes = [(0, 0, 0), (0, 5, 0), (3, 3, 1)] + [
        (int(rng.integers(1, 90)), 0, int(rng.integers(0, 15))) for _ in range(40)]
    for k, (n, extra, ed) in enumerate(shapes):
        gapbit = 0.0 if k % 3 else 0.2
        a, b = related(rng, n, 0.1, gapbit, ed)
        b = b + rseq(rng, extra, 0.1, gapbit)[1:]
        if len(a) > len(b):
            a, b = b, a
        for sw in (0, 1):
            r = oracle.align_affine_3(u8(a), u8(b), cm, sw)
            out.append(dict(si=a, sj=b, swaped=sw, cost=int(r["cost"]), median=L(r["median"]),
                            medianwg=L(r["medianwg"]), resi=L(r["resi"]), resj=L(r["resj"]),
                            cost_full=int(oracle.cost_affine_3(u8(a), u8(b), cm))))
    return dict(tcm=[1, 1], gap_open=3, cases=out)


def linear_cases(rng, tcm):
    cm = oracle.cm_dna(*tcm)
    out = []
    for k in range(40):
        n = int(rng.integers(0, 120))
        a, b = related(rng, n, 0.1, 0.0 if k % 2 else 0.15, int(rng.integers(0, 20)))
        if k % 5 == 0:
            a, b = b, a  # longer first: Align.align_2 swaps
        r1, r2, c, _ = oracle.align_2(u8(a), u8(b), cm)
        dw = int(rng.integers(0, 12))
        cd, _, st = oracle.cost_2(u8(a), u8(b), cm, dw)
        out.append(dict(s1=a, s2=b, r1=L(r1), r2=L(r2), cost=int(c), deltaw=dw,
                        cost_deltaw=int(cd), full_plane=int(st.full_plane)))
    return dict(tcm=list(tcm), cases=out)


def median_cases(rng):
    cm = oracle.cm_dna(1, 2)
    out = []
    for k in range(30):
        n = int(rng.integers(1, 60))
        a = rseq(rng, n, 0.15, 0.25)
        b = rseq(rng, n, 0.15, 0.25)
        if k == 0:
            a = [16] * (n + 1)
        out.append(dict(a=a, b=b, ancestor=L(oracle.ancestor_2(u8(a), u8(b), cm)),
                        median=L(oracle.median_2_no_gaps(u8(a), u8(b), cm)),
                        median_wg=L(oracle.median_2_with_gaps(u8(a), u8(b), cm)),
                        worst=int(oracle.worst_2(u8(a), u8(b), cm)),
                        verify=int(oracle.verify_2(u8(a), u8(b), cm))))
    return dict(tcm=[1, 2], cases=out)


def cube_cases(rng):
    c3 = oracle.cm3_from_2d(oracle.cm_dna(1, 1))
    out = []
    for k in range(24):
        n = int(rng.integers(0, 25))
        a, b = related(rng, n, 0.0, 0.0, int(rng.integers(0, 5)))
        _, c = related(rng, n, 0.0, 0.0, int(rng.integers(0, 5)))
        cost, dm = oracle.nw_3d(u8(a), u8(b), u8(c), c3)
        bt = oracle.backtrack_3d(u8(a), u8(b), u8(c), dm)
        case = dict(s1=a, s2=b, s3=c, cost=int(cost), walk_ok=bt is not None)
        if bt is not None:
            r1, r2, r3 = bt
            case.update(r1=L(r1), r2=L(r2), r3=L(r3),
                        median3=L(oracle.median_3(r1, r2, r3, c3)))
        out.append(case)
    return dict(tcm=[1, 1], cases=out)


def main():
    rng = np.random.default_rng(0x504F59)
    files = {
        "affine_align_3.json": affine_cases(rng),
        "linear_align_2_tcm12.json": linear_cases(rng, (1, 2)),
        "linear_align_2_tcm11.json": linear_cases(rng, (1, 1)),
        "medians_tcm12.json": median_cases(rng),
        "cube_3d_tcm11.json": cube_cases(rng),
    }
    for name, data in files.items():
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(data, f, separators=(",", ":"))
        print(name, len(data["cases"]), "cases")


if __name__ == "__main__":
    main()

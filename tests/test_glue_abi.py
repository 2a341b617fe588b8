"""Host logic (CPU): OCaml glue restated in the product and the oracle, the
C ABI's exported symbols, the workload generator and the bench contract.

Glue pins are hand-computed from the text of src/sequence.ml:868-936
(Align.cost_2's deltaw_calc, shorter-first swap) and src/seq.c:644-669
(seq_CAML_count with startNO = 0, i.e. codes >= 16 carrying the gap bit).
"""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle
from poy5_amd import _lib, batch, workloads
from poy5_amd.sequence import Align

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def u8(x):
    return np.array(x, np.uint8)


def test_count_gaps_hand_cases():
    s = u8([16, 1, 17, 2, 31, 16, 15])
    assert oracle.count_gaps(s) == 4
    assert Align.count_gaps(s) == 4
    assert oracle.count_gaps(u8([16])) == 1
    assert Align.count_gaps(u8([1, 2, 15])) == 0


def test_deltawh_hand_cases():
    a = u8([16] + [1] * 99)    # 100 symbols, leading gap only
    b = u8([16] + [2] * 119)   # 120
    # gaps = 1; lower = 10; dif = -20 < 10: no deltaw -> 10 / 2, deltaw -> 10
    assert oracle.deltawh(a, b) == 1 + 5
    assert oracle.deltawh(b, a) == 1 + 5
    assert oracle.deltawh(a, b, 7) == 1 + 10
    c = u8([16, 1, 2, 4, 8])   # equal short lengths: lower = 0, dif = 0
    assert oracle.deltawh(c, c) == 1 + 2
    assert oracle.deltawh(c, c, 7) == 1 + 7
    g = u8([16, 17, 18, 1, 1])
    assert oracle.deltawh(g, c) == 3 + 2


def test_deltaw_calc_product_matches_oracle():
    rng = np.random.default_rng(3)
    for _ in range(400):
        l1 = int(rng.integers(1, 5000))
        l2 = int(rng.integers(l1, l1 + 300))
        dw = None if rng.random() < 0.5 else int(rng.integers(0, 40))
        a = u8([16] + [1] * (l1 - 1))
        b = u8([16] + [2] * (l2 - 1))
        want = oracle.deltawh(a, b, dw)
        got = max(Align.count_gaps(a), Align.count_gaps(b)) + Align.deltaw_calc(l1, l2, dw)
        assert got == want, (l1, l2, dw)


def _header_functions(*names):
    found = set()
    for name in names or ("poyhip.h", "poyhip_tools.h"):
        text = open(os.path.join(ROOT, "include", name)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        found |= set(re.findall(r"\b(poyhip_\w+)\s*\(", text))
    return found


def test_abi_exports_every_declared_symbol():
    declared = _header_functions()
    assert len(declared) >= 20
    L = C.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), name
    # the ctypes binding covers exactly what the headers declare
    assert declared == set(_lib.SIGNATURES), declared ^ set(_lib.SIGNATURES)
    # benchmark generators and test hooks stay out of the drop-in header
    assert not any(n.startswith("poyhip_gen_") for n in _header_functions("poyhip.h"))


def test_library_loads_without_gpu():
    L = _lib.lib()
    h = C.c_void_p()
    # no device in this container: creation must fail with an error code,
    # never fall back to a CPU path
    import torch
    if not torch.cuda.is_available():
        rc = L.poyhip_create(0, C.byref(h))
        assert rc != _lib.POYHIP_OK


def test_workloads_deterministic_and_valid():
    for cfg, n in ((1, 64), (2, 256), (4, 4)):
        b1, o1, l1 = workloads.gen_pairs(cfg, n)
        b2, o2, l2 = workloads.gen_pairs(cfg, n)
        assert np.array_equal(b1, b2) and np.array_equal(o1, o2) and np.array_equal(l1, l2)
        assert len(l1) == 2 * n
        for k in range(2 * n):
            s = b1[o1[k]: o1[k] + l1[k]]
            assert s[0] == 16
            assert np.all((s[1:] >= 1) & (s[1:] <= 15))
        b3, _, _ = workloads.gen_pairs(cfg, n, seed=12345)
        assert not np.array_equal(b1[:64], b3[:64])
    _, _, ln = workloads.gen_pairs(2, 4096)
    assert ln.min() >= 20 and ln.max() <= 8000  # root 50-5000 after indels


def test_pack_pairs_roundtrip():
    pairs = [(u8([16, 1, 2]), u8([16, 4])), (u8([16]), u8([16, 8, 8, 8]))]
    blob, off1, off2, len1, len2 = batch.pack_pairs(pairs)
    for k, (a, b) in enumerate(pairs):
        assert np.array_equal(blob[off1[k]: off1[k] + len1[k]], a)
        assert np.array_equal(blob[off2[k]: off2[k] + len2[k]], b)


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "single MI355X" in r.stderr


def test_spr_workload_deterministic():
    from poy5_amd import spr
    p1, c1, l1 = spr.simulate(n_taxa=32, root_len=200, seed=9)
    p2, c2, l2 = spr.simulate(n_taxa=32, root_len=200, seed=9)
    assert np.array_equal(p1, p2) and c1 == c2
    assert len(l1) == 32 and len(p1) == 63
    for v in l1:
        assert np.array_equal(l1[v], l2[v]) and l1[v][0] == 16
        assert np.all((l1[v][1:] >= 1) & (l1[v][1:] <= 8))
    order = spr.post_order(c1)
    assert order[-1] == 0 and sorted(order) == list(range(63))

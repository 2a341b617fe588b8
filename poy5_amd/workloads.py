"""Synthetic inputs of the benchmark configurations.

The pair generators live in libpoyhip (``poyhip_gen_pairs``, host code) so
that 65,536-pair workloads are produced in seconds.  Configuration numbers:

1. 1,024 DNA pairs, root U[300,700], 10% substitutions, 3% indel events
   (geometric lengths, mean 3), 5% IUPAC codes; linear tcm sub 1 / indel 2.
2. 65,536 pairs, root log-uniform [50,5000], divergence U[2%,25%] at the
   10:3 substitution:indel ratio, 5% IUPAC; affine sub 1, ext 1, open 3.
4. 2,048 pairs, root U[16k,100k], divergence U[1%,8%], 10% IUPAC; linear
   sub 1 / indel 1.
5. 16,384 triples, root U[200,1000], three branches each with divergence
   U[3%,15%] at the 10:3 ratio, no IUPAC; uniform 1/1 tables, 3D cube.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

SEED_BASE = 0x504F59000


def gen_pairs(config: int, n: int, seed: int | None = None):
    """Returns (blob, off[2n], len[2n]) with pair k = (off[2k], off[2k+1])."""
    L = _lib.lib()
    seed = SEED_BASE + config if seed is None else seed
    total = C.c_int64()
    off = np.zeros(2 * n, np.int64)
    ln = np.zeros(2 * n, np.int32)
    _lib.check(L.poyhip_gen_pairs(config, n, seed, None, off.ctypes.data_as(C.POINTER(C.c_int64)),
                                  ln.ctypes.data_as(C.POINTER(C.c_int)), C.byref(total)))
    blob = np.zeros(max(total.value, 1), np.uint8)
    _lib.check(L.poyhip_gen_pairs(config, n, seed, blob.ctypes.data_as(C.POINTER(C.c_uint8)),
                                  off.ctypes.data_as(C.POINTER(C.c_int64)),
                                  ln.ctypes.data_as(C.POINTER(C.c_int)), C.byref(total)))
    return blob, off, ln


def pairs_list(config: int, n: int, seed: int | None = None):
    blob, off, ln = gen_pairs(config, n, seed)
    return [(blob[off[2 * k]:off[2 * k] + ln[2 * k]].copy(),
             blob[off[2 * k + 1]:off[2 * k + 1] + ln[2 * k + 1]].copy()) for k in range(n)]


def gen_triples(n: int, seed: int | None = None, config: int = 5):
    """Returns (blob, off[3n], len[3n]) with triple k at entries 3k..3k+2."""
    L = _lib.lib()
    seed = SEED_BASE + config if seed is None else seed
    total = C.c_int64()
    off = np.zeros(3 * n, np.int64)
    ln = np.zeros(3 * n, np.int32)
    _lib.check(L.poyhip_gen_triples(config, n, seed, None,
                                    off.ctypes.data_as(C.POINTER(C.c_int64)),
                                    ln.ctypes.data_as(C.POINTER(C.c_int)), C.byref(total)))
    blob = np.zeros(max(total.value, 1), np.uint8)
    _lib.check(L.poyhip_gen_triples(config, n, seed, blob.ctypes.data_as(C.POINTER(C.c_uint8)),
                                    off.ctypes.data_as(C.POINTER(C.c_int64)),
                                    ln.ctypes.data_as(C.POINTER(C.c_int)), C.byref(total)))
    return blob, off, ln

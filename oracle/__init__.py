"""CPU oracle for POY's DO pairwise alignment -- TEST INFRASTRUCTURE ONLY.

ctypes binding of ``oracle/liboracle.so`` (built from ``oracle/oracle.c`` by
``oracle/Makefile``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker.  The product package ``poy5_amd`` never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

TBL = 2048
GAP = 16


class OrcCM(C.Structure):
    _fields_ = [
        ("lcm", C.c_int), ("gap", C.c_int), ("a_sz", C.c_int),
        ("combinations", C.c_int), ("cost_model", C.c_int), ("gap_open", C.c_int),
        ("cost", C.c_int * TBL), ("worst", C.c_int * TBL),
        ("prepend", C.c_int * 64), ("tail", C.c_int * 64),
        ("median", C.c_uint8 * TBL),
    ]


class OrcCM3(C.Structure):
    _fields_ = [
        ("lcm", C.c_int), ("gap", C.c_int), ("a_sz", C.c_int),
        ("cost", C.c_int * 32768), ("median", C.c_uint8 * 32768),
    ]


class OrcStats(C.Structure):
    _fields_ = [
        ("cells", C.c_int64), ("rounds", C.c_int), ("full_plane", C.c_int),
        ("final_T", C.c_int), ("final_k", C.c_int), ("gap_num", C.c_int),
        ("min_gap_excess", C.c_int),
    ]


class OrcNkStats(C.Structure):
    _fields_ = [
        ("cells", C.c_int64), ("rounds", C.c_int), ("final_T", C.c_int),
        ("final_k", C.c_int), ("gap_num", C.c_int),
    ]


# direction flags of the diagonal-store aligner (src/newkkonen.c:41-61)
NK_ALIGN, NK_DO_DELETE, NK_DO_INSERT = 1, 16, 32
NK_END_INSERT, NK_END_DELETE, NK_END_DIAG, NK_DO_DIAG, NK_INS_EQ_DEL = 64, 128, 256, 512, 1024


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        u8p = C.POINTER(C.c_uint8)
        ip = C.POINTER(C.c_int)
        u16p = C.POINTER(C.c_uint16)
        L.orc_cm_dna.argtypes = [C.POINTER(OrcCM), C.c_int, C.c_int, C.c_int]
        L.orc_cm3_from_2d.argtypes = [C.POINTER(OrcCM3), C.POINTER(OrcCM)]
        L.orc_min_non0_cost.argtypes = [C.POINTER(OrcCM)]
        L.orc_nw.argtypes = [u8p, C.c_int, u8p, C.c_int, C.POINTER(OrcCM), C.c_int,
                             u16p, C.POINTER(OrcStats)]
        L.orc_backtrace_2d.argtypes = [u8p, C.c_int, u8p, C.c_int, u16p, C.c_int, u8p, u8p]
        L.orc_nw_limit.argtypes = [u8p, C.c_int, u8p, C.c_int, C.POINTER(OrcCM), C.c_int,
                                   C.c_int, C.c_int, u16p, C.POINTER(OrcStats)]
        L.orc_backtrace_2d_limit.argtypes = [u8p, C.c_int, u8p, C.c_int, u16p, C.c_int, C.c_int,
                                             C.c_int, C.c_int, C.c_int, u8p, u8p]
        L.orc_backtrace_2d_gaps.argtypes = [u16p, C.c_int, C.c_int]
        for f in ("orc_ancestor_2", "orc_median_2_no_gaps", "orc_median_2_with_gaps"):
            getattr(L, f).argtypes = [u8p, u8p, C.c_int, C.POINTER(OrcCM), u8p]
        for f in ("orc_worst_2", "orc_verify_2"):
            getattr(L, f).argtypes = [u8p, u8p, C.c_int, C.POINTER(OrcCM)]
        L.orc_align_affine_3.argtypes = [u8p, C.c_int, u8p, C.c_int, C.POINTER(OrcCM), C.c_int,
                                         u8p, ip, u8p, ip, u8p, u8p, ip, C.POINTER(OrcStats)]
        L.orc_align_affine_round.argtypes = [u8p, C.c_int, u8p, C.c_int, C.POINTER(OrcCM),
                                             C.c_int, C.c_int, u8p, ip, u8p, ip, u8p, u8p, ip,
                                             C.POINTER(OrcStats)]
        L.orc_cost_affine_3.argtypes = [u8p, C.c_int, u8p, C.c_int, C.POINTER(OrcCM)]
        L.orc_count_gaps.argtypes = [u8p, C.c_int]
        L.orc_deltawh.argtypes = [u8p, C.c_int, u8p, C.c_int, C.c_int, C.c_int]
        L.orc_cost_2.argtypes = [u8p, C.c_int, u8p, C.c_int, C.POINTER(OrcCM), C.c_int, C.c_int,
                                 u16p, C.POINTER(OrcStats)]
        L.orc_align_2.argtypes = [u8p, C.c_int, u8p, C.c_int, C.POINTER(OrcCM), u8p, u8p, ip,
                                  C.POINTER(OrcStats)]
        L.orc_nw_3d.argtypes = [u8p, C.c_int, u8p, C.c_int, u8p, C.c_int, C.POINTER(OrcCM3), u16p]
        L.orc_backtrack_3d.argtypes = [u8p, C.c_int, u8p, C.c_int, u8p, C.c_int, u16p, C.c_int,
                                       u8p, u8p, u8p]
        L.orc_median_3.argtypes = [u8p, u8p, u8p, C.c_int, C.POINTER(OrcCM3), u8p]
        L.orc_batch_align_affine.argtypes = [C.c_int, u8p, C.POINTER(C.c_int64), ip,
                                             C.POINTER(C.c_int64), ip, C.POINTER(OrcCM), ip,
                                             C.POINTER(C.c_int64)]
        L.orc_batch_align_affine.restype = C.c_int64
        L.orc_batch_align_2.argtypes = [C.c_int, u8p, C.POINTER(C.c_int64), ip,
                                        C.POINTER(C.c_int64), ip, C.POINTER(OrcCM), ip,
                                        C.POINTER(C.c_int64)]
        L.orc_batch_align_2.restype = C.c_int64
        L.orc_batch_align_3d.argtypes = [C.c_int, u8p, C.POINTER(C.c_int64), ip,
                                         C.POINTER(C.c_int64), ip, C.POINTER(C.c_int64), ip,
                                         C.POINTER(OrcCM3), ip]
        L.orc_batch_align_3d.restype = C.c_int64
        for f in ("orc_newkk_align", "orc_newkk_full"):
            getattr(L, f).argtypes = [u8p, C.c_int, u8p, C.c_int, C.POINTER(OrcCM), C.c_int,
                                      u8p, u8p, ip, C.POINTER(OrcNkStats)]
        L.orc_batch_newkk.argtypes = [C.c_int, u8p, C.POINTER(C.c_int64), ip,
                                      C.POINTER(C.c_int64), ip, C.POINTER(OrcCM), ip,
                                      C.POINTER(C.c_int64)]
        L.orc_batch_newkk.restype = C.c_int64
        i32p = C.POINTER(C.c_int32)
        L.orc_union_merge.argtypes = [u8p, u8p, u8p, C.c_int, u8p, C.c_int, C.c_int, i32p,
                                      u8p, C.c_int, C.c_int, i32p, u8p, C.c_int, i32p, i32p,
                                      i32p, ip]
        L.orc_newkk_matrix.argtypes = [u8p, C.c_int, u8p, C.c_int, C.POINTER(OrcCM), ip]
        L.orc_get_closest.argtypes = [C.POINTER(OrcCM), C.c_int, C.c_int]
        L.orc_closest_map.argtypes = [u8p, u8p, C.c_int, C.c_int, C.POINTER(OrcCM), u8p]
        _lib = L
    return _lib


def _u8(a):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    return a, a.ctypes.data_as(C.POINTER(C.c_uint8))


def cm_dna(trans: int, gaps: int, gap_open: int | None = None) -> OrcCM:
    cm = OrcCM()
    lib().orc_cm_dna(C.byref(cm), trans, gaps, -1 if gap_open is None else gap_open)
    return cm


def cm3_from_2d(cm: OrcCM) -> OrcCM3:
    c3 = OrcCM3()
    lib().orc_cm3_from_2d(C.byref(c3), C.byref(cm))
    return c3


def min_non0_cost(cm) -> int:
    return lib().orc_min_non0_cost(C.byref(cm))


def nw(s1, s2, cm, deltawh):
    """algn_nw: returns (cost, dm[L1,L2] uint16, stats)."""
    a, ap = _u8(s1)
    b, bp = _u8(s2)
    dm = np.zeros((len(a), len(b)), dtype=np.uint16)
    st = OrcStats()
    c = lib().orc_nw(ap, len(a), bp, len(b), C.byref(cm), deltawh,
                     dm.ctypes.data_as(C.POINTER(C.c_uint16)), C.byref(st))
    return c, dm, st


def limit_2(s1, s2, cm, w, s1_len, s2_len):
    """algn_CAML_limit_2 (h, s1_st and s2_st are read nowhere): returns
    (cost, dm[L1,L2] uint16, stats)."""
    a, ap = _u8(s1)
    b, bp = _u8(s2)
    dm = np.zeros((len(a), len(b)), dtype=np.uint16)
    st = OrcStats()
    c = lib().orc_nw_limit(ap, len(a), bp, len(b), C.byref(cm), w, s1_len, s2_len,
                           dm.ctypes.data_as(C.POINTER(C.c_uint16)), C.byref(st))
    return c, dm, st


def backtrace_2d_limit(s1, s2, dm, st_s1, st_s2, algn_s1, algn_s2, swaped):
    """algn_CAML_backtrace_2d_limit: the aligned pair of the limited walk."""
    a, ap = _u8(s1)
    b, bp = _u8(s2)
    cap = len(a) + len(b)
    r1 = np.zeros(cap, np.uint8)
    r2 = np.zeros(cap, np.uint8)
    n = lib().orc_backtrace_2d_limit(ap, len(a), bp, len(b),
                                     np.ascontiguousarray(dm).ctypes.data_as(C.POINTER(C.c_uint16)),
                                     st_s1, st_s2, algn_s1, algn_s2, int(swaped),
                                     r1.ctypes.data_as(C.POINTER(C.c_uint8)),
                                     r2.ctypes.data_as(C.POINTER(C.c_uint8)))
    return r1[:n].copy(), r2[:n].copy()


def backtrace_2d(s1, s2, dm, swaped):
    a, ap = _u8(s1)
    b, bp = _u8(s2)
    cap = len(a) + len(b)
    r1 = np.zeros(cap, np.uint8)
    r2 = np.zeros(cap, np.uint8)
    n = lib().orc_backtrace_2d(ap, len(a), bp, len(b),
                               np.ascontiguousarray(dm).ctypes.data_as(C.POINTER(C.c_uint16)),
                               int(swaped), r1.ctypes.data_as(C.POINTER(C.c_uint8)),
                               r2.ctypes.data_as(C.POINTER(C.c_uint8)))
    return r1[:n].copy(), r2[:n].copy()


def _pair_out(fname, a, b, cm, extra=2):
    a_, ap = _u8(a)
    b_, bp = _u8(b)
    out = np.zeros(len(a_) + extra, np.uint8)
    n = getattr(lib(), fname)(ap, bp, len(a_), C.byref(cm), out.ctypes.data_as(C.POINTER(C.c_uint8)))
    if n < 0:
        raise ValueError("median should not be 0")
    return out[:n].copy()


def ancestor_2(a, b, cm):
    return _pair_out("orc_ancestor_2", a, b, cm)


def median_2_no_gaps(a, b, cm):
    return _pair_out("orc_median_2_no_gaps", a, b, cm)


def median_2_with_gaps(a, b, cm):
    return _pair_out("orc_median_2_with_gaps", a, b, cm)


def worst_2(a, b, cm):
    a_, ap = _u8(a)
    b_, bp = _u8(b)
    return lib().orc_worst_2(ap, bp, len(a_), C.byref(cm))


def verify_2(a, b, cm):
    a_, ap = _u8(a)
    b_, bp = _u8(b)
    return lib().orc_verify_2(ap, bp, len(a_), C.byref(cm))


def align_affine_3(si, sj, cm, swaped):
    """algn_CAML_align_affine_3 body: si must not be longer than sj.
    Returns dict(cost, median, medianwg, resi, resj, stats)."""
    a, ap = _u8(si)
    b, bp = _u8(sj)
    cap = len(a) + len(b) + 2
    bufs = [np.zeros(cap, np.uint8) for _ in range(4)]
    ml, mwl, al = C.c_int(), C.c_int(), C.c_int()
    st = OrcStats()
    P = lambda x: x.ctypes.data_as(C.POINTER(C.c_uint8))
    c = lib().orc_align_affine_3(ap, len(a), bp, len(b), C.byref(cm), int(swaped),
                                 P(bufs[0]), C.byref(ml), P(bufs[1]), C.byref(mwl),
                                 P(bufs[2]), P(bufs[3]), C.byref(al), C.byref(st))
    if c == -1 and len(b) < len(a):
        raise ValueError("pass the shorter one as first")
    return dict(cost=c, median=bufs[0][:ml.value].copy(), medianwg=bufs[1][:mwl.value].copy(),
                resi=bufs[2][:al.value].copy(), resj=bufs[3][:al.value].copy(), stats=st)


def align_affine_round(si, sj, cm, swaped, p):
    """One band round at half-width parameter p on fresh rows (the first
    round of align_affine_3's doubling) plus backtrace_aff."""
    a, ap = _u8(si)
    b, bp = _u8(sj)
    cap = len(a) + len(b) + 2
    bufs = [np.zeros(cap, np.uint8) for _ in range(4)]
    ml, mwl, al = C.c_int(), C.c_int(), C.c_int()
    st = OrcStats()
    P = lambda x: x.ctypes.data_as(C.POINTER(C.c_uint8))  # noqa: E731
    c = lib().orc_align_affine_round(ap, len(a), bp, len(b), C.byref(cm), int(swaped), int(p),
                                     P(bufs[0]), C.byref(ml), P(bufs[1]), C.byref(mwl),
                                     P(bufs[2]), P(bufs[3]), C.byref(al), C.byref(st))
    return dict(cost=c, median=bufs[0][:ml.value].copy(), medianwg=bufs[1][:mwl.value].copy(),
                resi=bufs[2][:al.value].copy(), resj=bufs[3][:al.value].copy(), stats=st)


def cost_affine_3(si, sj, cm):
    a, ap = _u8(si)
    b, bp = _u8(sj)
    return lib().orc_cost_affine_3(ap, len(a), bp, len(b), C.byref(cm))


def count_gaps(s):
    a, ap = _u8(s)
    return lib().orc_count_gaps(ap, len(a))


def deltawh(s1, s2, deltaw=None):
    a, ap = _u8(s1)
    b, bp = _u8(s2)
    return lib().orc_deltawh(ap, len(a), bp, len(b), 0 if deltaw is None else 1,
                             0 if deltaw is None else deltaw)


def cost_2(s1, s2, cm, deltaw=None):
    """Align.cost_2 (linear): returns (cost, dm, stats) for the shorter-first
    orientation."""
    a, ap = _u8(s1)
    b, bp = _u8(s2)
    n1, n2 = (len(a), len(b)) if len(a) <= len(b) else (len(b), len(a))
    dm = np.zeros((n1, n2), np.uint16)
    st = OrcStats()
    c = lib().orc_cost_2(ap, len(a), bp, len(b), C.byref(cm), 0 if deltaw is None else 1,
                         0 if deltaw is None else deltaw,
                         dm.ctypes.data_as(C.POINTER(C.c_uint16)), C.byref(st))
    return c, dm, st


def align_2(s1, s2, cm):
    """Align.align_2 ~first_gap:true, linear model: (s1p, s2p, cost, stats)."""
    a, ap = _u8(s1)
    b, bp = _u8(s2)
    cap = len(a) + len(b)
    r1 = np.zeros(cap, np.uint8)
    r2 = np.zeros(cap, np.uint8)
    n = C.c_int()
    st = OrcStats()
    c = lib().orc_align_2(ap, len(a), bp, len(b), C.byref(cm),
                          r1.ctypes.data_as(C.POINTER(C.c_uint8)),
                          r2.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(n), C.byref(st))
    return r1[:n.value].copy(), r2[:n.value].copy(), c, st


def nw_3d(s1, s2, s3, c3):
    a, ap = _u8(s1)
    b, bp = _u8(s2)
    c, cp = _u8(s3)
    dm = np.zeros(len(a) * len(b) * len(c), np.uint16)
    cost = lib().orc_nw_3d(ap, len(a), bp, len(b), cp, len(c), C.byref(c3),
                           dm.ctypes.data_as(C.POINTER(C.c_uint16)))
    return cost, dm


def backtrack_3d(s1, s2, s3, dm):
    a, ap = _u8(s1)
    b, bp = _u8(s2)
    c, cp = _u8(s3)
    cap = len(a) + len(b) + len(c)
    outs = [np.zeros(cap, np.uint8) for _ in range(3)]
    P = lambda x: x.ctypes.data_as(C.POINTER(C.c_uint8))
    n = lib().orc_backtrack_3d(ap, len(a), bp, len(b), cp, len(c),
                               np.ascontiguousarray(dm).ctypes.data_as(C.POINTER(C.c_uint16)),
                               GAP, P(outs[0]), P(outs[1]), P(outs[2]))
    if n < 0:
        return None
    return tuple(o[:n].copy() for o in outs)


def median_3(a, b, c, c3):
    a_, ap = _u8(a)
    b_, bp = _u8(b)
    c_, cp = _u8(c)
    out = np.zeros(len(a_) + 1, np.uint8)
    n = lib().orc_median_3(ap, bp, cp, len(a_), C.byref(c3), out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out[:n].copy()


def _newkk(fname, s1, s2, cm, swaped, backtrace):
    a, ap = _u8(s1)
    b, bp = _u8(s2)
    if len(a) > len(b):
        raise ValueError("pass the shorter one as first")
    cap = len(a) + len(b)
    r1 = np.zeros(cap, np.uint8)
    r2 = np.zeros(cap, np.uint8)
    n = C.c_int(0)
    st = OrcNkStats()
    P = lambda x: x.ctypes.data_as(C.POINTER(C.c_uint8))  # noqa: E731
    c = getattr(lib(), fname)(ap, len(a), bp, len(b), C.byref(cm), int(swaped),
                              P(r1) if backtrace else None, P(r2) if backtrace else None,
                              C.byref(n), C.byref(st))
    out = dict(cost=c, stats=st)
    if backtrace:
        if n.value < 0:
            raise RuntimeError("the reference backtrace fails on this input")
        out["r1"] = r1[:n.value].copy()
        out["r2"] = r2[:n.value].copy()
    return out


def newkk_align(s1, s2, cm, swaped=0, backtrace=True):
    """newkkonen_CAML_algn_affine + newkkonen_CAML_backtrace_affine (the
    reference's incremental refill): s1 must not be longer than s2.  Returns
    dict(cost, stats[, r1, r2])."""
    return _newkk("orc_newkk_align", s1, s2, cm, swaped, backtrace)


def newkk_full(s1, s2, cm, swaped=0, backtrace=True):
    """The same aligner with every band cell recomputed in every round."""
    return _newkk("orc_newkk_full", s1, s2, cm, swaped, backtrace)


def newkk_matrix(s1, s2, cm):
    """Full-matrix states: array (len1, len2, 5) of cost, P, Q, ED, CD."""
    a, ap = _u8(s1)
    b, bp = _u8(s2)
    out = np.zeros(len(a) * len(b) * 5, np.int32)
    lib().orc_newkk_matrix(ap, len(a), bp, len(b), C.byref(cm),
                           out.ctypes.data_as(C.POINTER(C.c_int)))
    return out.reshape(len(a), len(b), 5)


def newkk_align_2(s1, s2, cm, first_gap=True):
    """NewkkAlign.align_2 (src/sequence.ml:1883-1913) on the oracle: shorter
    first with swaped = 1 when exchanged, results swapped back; first_gap
    False prepends a gap to both and strips the first column.  Returns
    (s1', s2', cost)."""
    a = np.asarray(s1, np.uint8)
    b = np.asarray(s2, np.uint8)
    if not first_gap:
        a = np.concatenate([[GAP], a]).astype(np.uint8)
        b = np.concatenate([[GAP], b]).astype(np.uint8)
    exchange = len(a) > len(b)
    x, y = (b, a) if exchange else (a, b)
    r = newkk_align(x, y, cm, 1 if exchange else 0)
    p1, p2 = (r["r2"], r["r1"]) if exchange else (r["r1"], r["r2"])
    if not first_gap:
        p1, p2 = p1[1:], p2[1:]
    return p1.copy(), p2.copy(), r["cost"]


def newkk_cost_2(s1, s2, cm):
    """NewkkAlign.cost_2 (src/sequence.ml:1915-1963), affine model: shorter
    first, an empty sequence replaced by a lone gap."""
    a = np.asarray(s1, np.uint8)
    b = np.asarray(s2, np.uint8)
    if len(a) > len(b):
        a, b = b, a
    if len(a) == 0:
        a = np.array([GAP], np.uint8)
    if len(b) == 0:
        b = np.array([GAP], np.uint8)
    return newkk_align(a, b, cm, 0, backtrace=False)["cost"]


class Union:
    """Sequence.Unions.u (src/sequence.ml:2782): a sequence right-aligned in
    `cap` symbols and three int32 offset arrays of `cap` entries (offset,
    union_c1, union_c2)."""

    def __init__(self, seq, length, offset, c1, c2):
        self.seq = np.asarray(seq, np.uint8)
        self.length = int(length)
        self.offset = np.asarray(offset, np.int32)
        self.c1 = np.asarray(c1, np.int32)
        self.c2 = np.asarray(c2, np.int32)

    @property
    def cap(self):
        return len(self.seq)

    def sequence(self):
        return self.seq[self.cap - self.length:].copy()


def union_leaf(s):
    """Unions.leaf (src/sequence.ml:2833): offsets 0, union_c1 = union_c2 =
    0..len-1 right-aligned (the sequence fills its capacity)."""
    s = np.asarray(s, np.uint8).copy()
    n = len(s)
    return Union(s, n, np.zeros(n, np.int32), np.arange(n, dtype=np.int32),
                 np.arange(n, dtype=np.int32))


def union_merge(a, b, median, ua, ub):
    """Unions.union (src/sequence.ml:2849) = union_CAML_make (src/union.c:318):
    a, b the aligned children, median their aligned median.  Returns the new
    Union (capacity len(ua) + len(ub) + 2; entries the merge does not write
    are 0)."""
    a_, ap = _u8(a)
    b_, bp = _u8(b)
    m_, mp = _u8(median)
    if not (len(a_) == len(b_) == len(m_)):
        raise ValueError("aligned sequences of different lengths")
    cap = ua.length + ub.length + 2
    cs = np.zeros(cap, np.uint8)
    co = np.zeros(cap, np.int32)
    c1 = np.zeros(cap, np.int32)
    c2 = np.zeros(cap, np.int32)
    n = C.c_int(0)
    P8 = lambda x: x.ctypes.data_as(C.POINTER(C.c_uint8))  # noqa: E731
    P32 = lambda x: x.ctypes.data_as(C.POINTER(C.c_int32))  # noqa: E731
    uas, uao = np.ascontiguousarray(ua.seq), np.ascontiguousarray(ua.offset)
    ubs, ubo = np.ascontiguousarray(ub.seq), np.ascontiguousarray(ub.offset)
    r = lib().orc_union_merge(ap, bp, mp, len(a_), P8(uas), ua.length, ua.cap, P32(uao),
                              P8(ubs), ub.length, ub.cap, P32(ubo), P8(cs), cap, P32(co),
                              P32(c1), P32(c2), C.byref(n))
    if r != 0:
        raise RuntimeError("union_merge fails on this input")
    return Union(cs, n.value, co, c1, c2)


# ---- Sequence.readjust / DOS.readjust (src/sequence.ml:2097-2146,
# src/seqCS.ml:776-952): OCaml glue over the functions above, restated here
# with the same calls and the same order of comparisons ----

def get_closest(cm, a, b):
    """Cost_matrix.Two_D.get_closest (src/cost_matrix.ml:1387): -1 where the
    reference fails."""
    return lib().orc_get_closest(C.byref(cm), int(a), int(b))


def closest_map(s1p, s2p, cm, equal=False):
    """The combine branch of Align.closest (src/sequence.ml:1199-1228): the
    column map of get_closest over (s1', s2'), then remove_gaps2."""
    a, ap = _u8(s1p)
    b, bp = _u8(s2p)
    if len(a) != len(b):
        raise ValueError("aligned sequences of different lengths")
    out = np.zeros(len(a) + 1, np.uint8)
    n = lib().orc_closest_map(ap, bp, len(a), 1 if equal else 0, C.byref(cm),
                              out.ctypes.data_as(C.POINTER(C.c_uint8)))
    if n < 0:
        raise ValueError("No bits on?")
    return out[:n].copy()


def is_empty(s):
    """Sequence.is_empty (src/sequence.ml:241): every symbol is the gap."""
    s = np.asarray(s, np.uint8)
    return len(s) == 0 or bool(np.all(s == GAP))


def select_one(s):
    """Sequence.select_one (src/sequence.ml:2069) with combinations: the
    lowest state of each symbol ("Nothing?" for a 0)."""
    s = np.asarray(s, np.int64)
    if np.any(s == 0):
        raise ValueError("Nothing?")
    return (s & -s).astype(np.uint8)


def align_2_any(s1, s2, cm):
    """Align.align_2 ~first_gap:true (src/sequence.ml:1061): align_affine_3's
    aligned pair under the affine model, cost_2 + create_edited_2 under the
    linear one.  Returns (s1', s2', cost)."""
    s1 = np.asarray(s1, np.uint8)
    s2 = np.asarray(s2, np.uint8)
    if cm.cost_model == 1:
        sw = len(s1) > len(s2)
        x, y = (s2, s1) if sw else (s1, s2)
        r = align_affine_3(x, y, cm, int(sw))
        if sw:
            return r["resj"], r["resi"], r["cost"]
        return r["resi"], r["resj"], r["cost"]
    p1, p2, c, _ = align_2(s1, s2, cm)
    return p1, p2, c


def cost_2_any(s1, s2, cm, deltaw=None):
    """Align.cost_2 (src/sequence.ml:868): the full affine DP under the
    affine model, the banded linear cost otherwise."""
    if cm.cost_model == 1:
        return cost_affine_3(s1, s2, cm)
    return cost_2(s1, s2, cm, deltaw)[0]


def readjust_algn(s1, s2, cm, use_ukk):
    """The `algn` of Sequence.readjust (src/sequence.ml:2099-2118): (cost,
    median) of two sequences."""
    if cm.cost_model == 1:
        if use_ukk:
            p1, p2, c = newkk_align_2(s1, s2, cm)
            return c, median_2_no_gaps(p1, p2, cm)
        sw = len(s1) > len(s2)
        x, y = (s2, s1) if sw else (s1, s2)
        r = align_affine_3(x, y, cm, int(sw))
        return r["cost"], r["median"]
    if use_ukk:
        raise ValueError("the diagonal-store aligner's linear path is not restated")
    p1, p2, c = align_2_any(s1, s2, cm)
    return c, median_2_no_gaps(p1, p2, cm)


def closest(s1, s2, cm):
    """Align.closest (src/sequence.ml:1180-1236), tables with combinations:
    (sequence, cost)."""
    s1 = np.asarray(s1, np.uint8)
    s2 = np.asarray(s2, np.uint8)
    if is_empty(s2):
        return s2.copy(), 0
    if len(s1) == len(s2) and np.array_equal(s1, s2):
        return closest_map(s1, s2, cm, equal=True), 0
    p1, p2, _ = align_2_any(s1, s2, cm)
    r = closest_map(p1, p2, cm)
    return r, cost_2_any(s1, r, cm)


def newkk_closest(s1, s2, cm):
    """NewkkAlign.closest (src/sequence.ml:1997-2060): closest with
    NewkkAlign.align_2 and NewkkAlign.cost_2."""
    s1 = np.asarray(s1, np.uint8)
    s2 = np.asarray(s2, np.uint8)
    if is_empty(s2):
        return s2.copy(), 0
    if len(s1) == len(s2) and np.array_equal(s1, s2):
        return closest_map(s1, s2, cm, equal=True), 0
    p1, p2, _ = newkk_align_2(s1, s2, cm)
    r = closest_map(p1, p2, cm)
    return r, newkk_cost_2(s1, r, cm)


def make_center(a, b, c, cm, use_ukk):
    """make_center of Sequence.readjust (src/sequence.ml:2120-2136)."""
    cab, ab = readjust_algn(a, b, cm, use_ukk)
    cbc, bc = readjust_algn(b, c, cm, use_ukk)
    cac, ac = readjust_algn(a, c, cm, use_ukk)
    cabc, _ = readjust_algn(ab, c, cm, use_ukk)
    cbca, _ = readjust_algn(bc, a, cm, use_ukk)
    cacb, _ = readjust_algn(ac, b, cm, use_ukk)
    cabc += cab
    cbca += cbc
    cacb += cac
    if cabc <= cbca:
        if cabc <= cacb:
            return closest(c, ab, cm)
        return closest(b, ac, cm)
    if cbca < cacb:
        return closest(a, bc, cm)
    return closest(b, ac, cm)


def seq_readjust(a, b, m, parent, cm, use_ukk=False):
    """Sequence.readjust (src/sequence.ml:2097-2146): (cost3, cost2, new
    median, aligned median vs parent x3 as the reference returns it)."""
    newseqm, _ = make_center(a, b, parent, cm, use_ukk)
    _, _, c1 = align_2_any(a, newseqm, cm)
    _, _, c2 = align_2_any(b, newseqm, cm)
    _, aliedmp, cp = align_2_any(parent, newseqm, cm)
    cost2 = c1 + c2
    return cost2 + cp, cost2, newseqm, aliedmp, aliedmp, aliedmp


def max_cost_2_any(a, b, cm):
    """Align.max_cost_2 (src/sequence.ml:652)."""
    if is_empty(a) or is_empty(b):
        return 0
    return worst_2(a, b, cm)


def dos_distance(a, b, cm, use_ukk, missing=0):
    """DOS.distance (src/seqCS.ml:701-775, without USE_VERIFY_COSTS): band
    deltaw = max(|la - lb|, 8)."""
    if is_empty(a) or is_empty(b):
        return missing
    deltaw = max(abs(len(a) - len(b)), 8)
    if use_ukk:
        return newkk_cost_2(a, b, cm)
    return cost_2_any(a, b, cm, deltaw)


def readjust_two(s1, s2, cm, use_ukk):
    """readjust_algn_two_child (src/seqCS.ml:783-813): (median, cost)."""
    if cm.cost_model == 1:
        if use_ukk:
            p1, p2, c = newkk_align_2(s1, s2, cm)
            med = median_2_no_gaps(p1, p2, cm)
        else:
            c, med = readjust_algn(s1, s2, cm, False)
        return select_one(med), c
    if use_ukk:
        raise ValueError("the diagonal-store aligner's linear path is not restated")
    p1, p2, c = align_2_any(s1, s2, cm)
    return select_one(median_2_no_gaps(p1, p2, cm)), c


def dos_readjust(ch1, ch2, parent, mine, cm, use_ukk=False):
    """DOS.readjust (src/seqCS.ml:820-952), ApproxD mode, DNA.  A node is a
    dict with "seq", "cost2", "cost2_max", "cost3", "sum_cost" (and
    "aligned" for the result).  Returns (changed, node, cost3, cost2,
    sum_cost)."""
    e1, e2, ep = is_empty(ch1["seq"]), is_empty(ch2["seq"]), is_empty(parent["seq"])
    mc2, msum, mc3 = int(mine["cost2"]), int(mine["sum_cost"]), int(mine["cost3"])
    sum12 = int(ch1["sum_cost"] + ch2["sum_cost"])

    def node(seq, aligned, c2, c2max, c3, s):
        return {"seq": seq, "aligned": aligned, "cost2": float(c2), "cost2_max": float(c2max),
                "cost3": float(c3), "sum_cost": float(s)}

    def differs(seq):
        return not (len(seq) == len(mine["seq"]) and np.array_equal(seq, mine["seq"]))

    if not e1 and not e2 and not ep:
        c3, c2, m, al1, al2, mwg = seq_readjust(ch1["seq"], ch2["seq"], mine["seq"],
                                                parent["seq"], cm, use_ukk)
        maxc2 = max_cost_2_any(al1, al2, cm)
        s = c2 + sum12
        changed = s != msum or c3 != mc3 or c2 != mc2 or differs(m)
        return changed, node(m, (al1, al2, mwg), c2, maxc2, c3, s), c3, c2, s
    if (e1 and e2) or (e1 and ep) or (e2 and ep):
        r = parent if (e1 and e2) else (ch2 if (e1 and ep) else ch1)
        changed = sum12 != msum or 0 != mc3 or 0 != mc2 or differs(r["seq"])
        out = dict(r)
        out["sum_cost"] = float(sum12)
        return changed, out, 0, 0, sum12
    if ep:
        med, c2 = readjust_two(ch1["seq"], ch2["seq"], cm, use_ukk)
        s = c2 + sum12
        c3 = c2
    else:
        ch = ch1 if e2 else ch2
        med, _ = readjust_two(ch["seq"], parent["seq"], cm, use_ukk)
        c2 = dos_distance(ch["seq"], med, cm, use_ukk)
        s = c2 + sum12
        c3 = c2 + dos_distance(med, parent["seq"], cm, use_ukk)
    changed = s != msum or c3 != mc3 or c2 != mc2 or differs(med)
    return changed, node(med, (med, med, med), c2, 0, c3, s), c3, c2, s


# ---- Sequence.Clip (src/sequence.ml:3441-3575) and Unions.get_positions
# (src/sequence.ml:2869-2904): OCaml bookkeeping around Align.align_2,
# restated with the OCaml's own control flow ----

def _get(s, pos):
    if pos < 0 or pos >= len(s):
        raise IndexError("Sequence.get out of range")
    return int(s[pos])


def clip_count_tip_gaps(seqa, seqb, gap=GAP):
    """Clip.count_tip_gaps (src/sequence.ml:3443)."""
    pos, acc = 0, 0
    while True:
        a, b = _get(seqa, pos), _get(seqb, pos)
        if gap == a or gap == b:
            if gap == a and gap == b:
                pos += 1
            else:
                pos, acc = pos + 1, acc + 1
        else:
            return acc


def clip_correct_distance(seqa, seqb, cost, cm, gap=GAP):
    """Clip.correct_distance (src/sequence.ml:3455)."""
    pos, acc = 0, 0
    while gap == _get(seqa, pos) or gap == _get(seqb, pos):
        acc += cm.cost[(_get(seqa, pos) << 5) + _get(seqb, pos)]
        pos += 1
    return cost if acc == 0 else cost - (acc + cm.gap_open)


def clip_corrected_distance(a, b, cm, missing=0, use_ukk=False):
    """Clip.corrected_distance (src/sequence.ml:3469)."""
    if is_empty(a) or is_empty(b):
        return missing
    if use_ukk:
        sa, sb, c = newkk_align_2(a, b, cm)
    else:
        sa, sb, c = align_2_any(a, b, cm)
    return clip_correct_distance(sa, sb, c, cm)


def _seq_init(f, n):
    """Sequence.init (src/sequence.ml:159): f called from index n - 1 down to
    0, each result prepended."""
    out = []
    for i in range(n - 1, -1, -1):
        out.insert(0, f(i))
    return np.array(out, np.uint8)


def clip_n_fix(s1, s2):
    """Clip.clip_n_fix (src/sequence.ml:3489-3539), fraction 1.00."""
    def aux(s1, s2):
        s1len, s2len = float(len(s1)), float(len(s2))
        if s1len > s2len * 1.00:
            maxlen = int(s2len * 1.00)
            dif = int(s1len) - maxlen
            cnt = [int(s1len)]

            def f(x):
                cnt[0] -= 1
                return _get(s1, 0) if x == 0 else _get(s1, cnt[0])
            s1p = _seq_init(f, maxlen)

            def add_bottom(x):
                return [int(v) for v in x[1:]]

            def fixs1(x):
                newx = add_bottom(x)
                for i in range(dif, -1, -1):
                    newx.insert(0, _get(s1, i))
                return np.array(newx, np.uint8)

            def fixs2(x):
                newx = add_bottom(x)
                g = _get(s1, 0)
                for _ in range(dif, -1, -1):
                    newx.insert(0, g)
                return np.array(newx, np.uint8)
            return s1p, np.asarray(s2, np.uint8), fixs1, fixs2, dif

        def ident(x):
            return np.asarray(x, np.uint8)
        return np.asarray(s1, np.uint8), np.asarray(s2, np.uint8), ident, ident, 0

    if len(s1) >= len(s2):
        return aux(s1, s2)
    a, b, c, d, dif = aux(s2, s1)
    return b, a, d, c, dif


def seq_safe_reverse(x):
    """Sequence.safe_reverse (src/sequence.ml:186)."""
    g = _get(x, 0)
    y = []
    for i in range(1, len(x)):
        y.insert(0, int(x[i]))
    return np.array([g] + y, np.uint8)


def clip_align_2(s1, s2, cm, first_gap=True):
    """Clip.Align.align_2 (src/sequence.ml:3541-3573) on (kind, sequence)."""
    (k1, a), (k2, b) = s1, s2

    def al2(x, y):
        if first_gap:
            return align_2_any(x, y, cm)
        xx = np.concatenate([[GAP], x]).astype(np.uint8)
        yy = np.concatenate([[GAP], y]).astype(np.uint8)
        p, q, c = align_2_any(xx, yy, cm)
        return p[1:], q[1:], c

    def algn_and_fix(x, y):
        x, y, f1, f2, dif = clip_n_fix(x, y)
        x, y, cost = al2(x, y)
        cost = clip_correct_distance(x, y, cost, cm)
        tip = clip_count_tip_gaps(x, y)
        return f1(x), f2(y), cost, dif + tip, x, x

    if (k1, k2) == ("DO", "DO"):
        x, y, cost = al2(a, b)
        return ("DO", x), ("DO", y), cost, 0, ("DO", x), ("DO", y)
    if (k1, k2) in (("Last", "DO"), ("DO", "Last"), ("Last", "Last")):
        x, y, cost, dif, xo, yo = algn_and_fix(seq_safe_reverse(a), seq_safe_reverse(b))
        r = seq_safe_reverse
        return ("Last", r(x)), ("Last", r(y)), cost, dif, ("Last", r(xo)), ("Last", r(yo))
    if (k1, k2) in (("First", "DO"), ("DO", "First"), ("First", "First")):
        x, y, cost, dif, xo, yo = algn_and_fix(a, b)
        return ("First", x), ("First", y), cost, dif, ("First", xo), ("First", yo)
    raise AssertionError("assert false")


def union_get_positions(u, positions):
    """Unions.get_positions (src/sequence.ml:2869-2904)."""
    ln = len(u.offset)
    seq_len = u.length

    def pos_finder(modifier, arr, x, guard=0):
        if guard > 4 * ln + 8:
            raise IndexError("no end")
        if x == 0:
            return int(arr[0])
        if x == ln:
            return pos_finder(-1, arr, x - 1, guard + 1)
        if x < ln:
            if x < 0:
                raise IndexError("bigarray index")
            if arr[x] != 0:
                return int(arr[x])
        return pos_finder(modifier, arr, x + modifier, guard + 1)

    positions = list(positions)
    poslen = len(positions) - 1

    def get_position(arr, pos, xy):
        x, y = xy
        xmod = 1 if pos == poslen else -1
        ymod = -1 if pos == 0 else 1
        return (pos_finder(xmod, arr, ln - seq_len + x), pos_finder(ymod, arr, ln - seq_len + y))

    a = [get_position(u.c1, k, p) for k, p in enumerate(positions)]
    b = [get_position(u.c2, k, p) for k, p in enumerate(positions)]
    return a, b

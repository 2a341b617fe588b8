"""Host-side mirror of POY's ``Sequence.Align`` on the MI355X library.

Names, argument meaning and error behaviour follow ``src/sequence.ml``
(module ``Align``, lines 596-1146, and ``median_2`` / ``median_2_with_gaps``,
lines 473-492).  Sequences are numpy uint8 arrays of 5-bit codes that start
with POY's leading gap (16).  Every computation runs on the GPU through
``libpoyhip``; nothing here computes an alignment on the CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .cost_matrix import GAP, ThreeD, TwoD


def _u8(a):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    return a, a.ctypes.data_as(C.POINTER(C.c_uint8))


class Context:
    """One device context (POY's ``Matrix.default`` scratch plus a stream)."""

    def __init__(self, device: int = 0):
        self._lib = _lib.lib()
        h = C.c_void_p()
        _lib.check(self._lib.poyhip_create(device, C.byref(h)))
        self.h = h
        self.cm: TwoD | None = None
        self.cm3: ThreeD | None = None

    def close(self):
        if self.h:
            self._lib.poyhip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_cm(self, cm: TwoD):
        self._cm_c = cm.to_c()
        _lib.check(self._lib.poyhip_set_cm(self.h, C.byref(self._cm_c)), self.h)
        self.cm = cm

    def set_cm3(self, cm3: ThreeD):
        self._cm3_c = cm3.to_c()
        _lib.check(self._lib.poyhip_set_cm3(self.h, C.byref(self._cm3_c)), self.h)
        self.cm3 = cm3

    def chk(self, rc):
        return _lib.check(rc, self.h)


def _closest_map(ctx: Context, a, b, equal: bool):
    """The column map of Align.closest / NewkkAlign.closest on the device
    (poyhip_closest_2)."""
    x, xp = _u8(a)
    y, yp = _u8(b)
    cap = len(x) + 1
    out = np.zeros(cap, np.uint8)
    n = C.c_int()
    rc = ctx._lib.poyhip_closest_2(ctx.h, xp, yp, len(x), 1 if equal else 0,
                                   out.ctypes.data_as(C.POINTER(C.c_uint8)), cap, C.byref(n))
    if rc == -1:
        raise ValueError("No bits on?")
    ctx.chk(rc)
    return out[:n.value].copy()


class Align:
    """``Sequence.Align`` over a device context."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.L = ctx._lib

    # -- helpers of Align.cost_2 (sequence.ml:852-936) ---------------------
    @staticmethod
    def count_gaps(s) -> int:
        """seq_CAML_count with startNO = 0 (src/seq.c:644-669)."""
        s = np.asarray(s, dtype=np.int64)
        return int(np.count_nonzero((s >= GAP) & ((s & GAP) != 0)))

    @staticmethod
    def deltaw_calc(s1len: int, s2len: int, deltaw: int | None) -> int:
        dif = s1len - s2len
        lower = int(float(s1len) * 0.10)
        if deltaw is None:
            return lower // 2 if dif < lower else 2
        return lower if dif < lower else deltaw

    # -- externals ---------------------------------------------------------
    def c_cost_2(self, s1, s2, deltawh: int) -> int:
        """algn_CAML_simple_2: s1 must not be longer than s2."""
        a, ap = _u8(s1)
        b, bp = _u8(s2)
        out = C.c_int()
        self.ctx.chk(self.L.poyhip_cost_2(self.ctx.h, ap, len(a), bp, len(b), deltawh, C.byref(out)))
        return out.value

    def extract_edited_2(self, s1, s2, swaped: int):
        """algn_CAML_backtrace_2d after c_cost_2 on the same sequences."""
        a, ap = _u8(s1)
        b, bp = _u8(s2)
        cap = len(a) + len(b)
        r1 = np.zeros(cap, np.uint8)
        r2 = np.zeros(cap, np.uint8)
        n = C.c_int()
        self.ctx.chk(self.L.poyhip_backtrace_2d(
            self.ctx.h, ap, len(a), bp, len(b), r1.ctypes.data_as(C.POINTER(C.c_uint8)),
            r2.ctypes.data_as(C.POINTER(C.c_uint8)), cap, C.byref(n), int(swaped)))
        return r1[:n.value].copy(), r2[:n.value].copy()

    def c_cost_2_limit(self, s1, s2, w: int, h: int, s1_st: int, s2_st: int, s1_len: int,
                       s2_len: int) -> int:
        """algn_CAML_limit_2: s1 must not be longer than s2; the limits only
        move the full-plane-vs-band decision (h, s1_st, s2_st are unused)."""
        a, ap = _u8(s1)
        b, bp = _u8(s2)
        out = C.c_int()
        self.ctx.chk(self.L.poyhip_limit_2(self.ctx.h, ap, len(a), bp, len(b), w, h, s1_st,
                                           s2_st, s1_len, s2_len, C.byref(out)))
        return out.value

    def extract_edited_2_limit(self, s1, s2, st_s1: int, st_s2: int, algn_s1: int,
                               algn_s2: int, swaped: int):
        """algn_CAML_backtrace_2d_limit after c_cost_2 / c_cost_2_limit on
        the same sequences: the walk over the window (see poyhip.h)."""
        a, ap = _u8(s1)
        b, bp = _u8(s2)
        cap = len(a) + len(b)
        r1 = np.zeros(cap, np.uint8)
        r2 = np.zeros(cap, np.uint8)
        n = C.c_int()
        self.ctx.chk(self.L.poyhip_backtrace_2d_limit(
            self.ctx.h, ap, len(a), bp, len(b), r1.ctypes.data_as(C.POINTER(C.c_uint8)),
            r2.ctypes.data_as(C.POINTER(C.c_uint8)), cap, C.byref(n), st_s1, st_s2, algn_s1,
            algn_s2, int(swaped)))
        return r1[:n.value].copy(), r2[:n.value].copy()

    def cost_2_limit(self, s1, s2, a: int, b: int, c: int, d: int, w: int | None = None,
                     h: int | None = None) -> int:
        """Align.cost_2_limit (sequence.ml:748-783): default width and height
        min(d, 2 |d - c|) clamped to c and d, positions a of s1 and b of s2
        set to the gap during the call, and the sequence with the longer
        limit passed first, as the OCaml does."""
        default = min(d, 2 * abs(d - c))
        w = default if w is None else w
        h = default if h is None else h
        w = c if w > c else w
        h = d if h > d else h
        s1 = np.array(s1, np.uint8)
        s2 = np.array(s2, np.uint8)
        s1[a] = GAP
        s2[b] = GAP
        if c >= d:
            return self.c_cost_2_limit(s1, s2, w, h, a, b, c, d)
        return self.c_cost_2_limit(s2, s1, h, w, b, a, d, c)

    def cost_2_affine(self, s1, s2) -> int:
        """algn_CAML_cost_affine_3: full affine DP, either order."""
        a, ap = _u8(s1)
        b, bp = _u8(s2)
        out = C.c_int()
        self.ctx.chk(self.L.poyhip_cost_affine_3(self.ctx.h, ap, len(a), bp, len(b), C.byref(out)))
        return out.value

    def c_align_affine_3(self, si, sj, swaped: int):
        """algn_CAML_align_affine_3: si must not be longer than sj."""
        a, ap = _u8(si)
        b, bp = _u8(sj)
        cap = len(a) + len(b) + 2
        bufs = [np.zeros(cap, np.uint8) for _ in range(4)]
        P = [x.ctypes.data_as(C.POINTER(C.c_uint8)) for x in bufs]
        ml, mwl, rl, cost = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        rc = self.L.poyhip_align_affine_3(self.ctx.h, ap, len(a), bp, len(b), int(swaped),
                                          P[0], C.byref(ml), P[1], C.byref(mwl), P[2], P[3],
                                          C.byref(rl), cap, C.byref(cost))
        if rc == -1 and len(b) < len(a):
            raise ValueError("pass the shorter one as first")
        self.ctx.chk(rc)
        return (bufs[0][:ml.value].copy(), bufs[2][:rl.value].copy(), bufs[3][:rl.value].copy(),
                cost.value, bufs[1][:mwl.value].copy())

    # -- OCaml-level functions ---------------------------------------------
    def cost_2(self, s1, s2, deltaw: int | None = None) -> int:
        """Align.cost_2 (sequence.ml:868-936): affine models use the full
        affine DP; linear models put the shorter sequence first and pass
        deltawh = max gap count + deltaw_calc."""
        if self.ctx.cm is not None and self.ctx.cm.cost_model == 1:
            return self.cost_2_affine(s1, s2)
        s1 = np.asarray(s1, np.uint8)
        s2 = np.asarray(s2, np.uint8)
        if len(s1) == 0:
            s1 = np.array([GAP], np.uint8)
        if len(s2) == 0:
            s2 = np.array([GAP], np.uint8)
        gaps = max(self.count_gaps(s1), self.count_gaps(s2))
        if len(s1) <= len(s2):
            return self.c_cost_2(s1, s2, gaps + self.deltaw_calc(len(s1), len(s2), deltaw))
        return self.c_cost_2(s2, s1, gaps + self.deltaw_calc(len(s2), len(s1), deltaw))

    def create_edited_2(self, s1, s2):
        """Align.create_edited_2 (sequence.ml:1019)."""
        if len(s1) <= len(s2):
            return self.extract_edited_2(s1, s2, 0)
        r2, r1 = self.extract_edited_2(s2, s1, 1)
        return r1, r2

    def align_affine_3(self, si, sj):
        """Align.align_affine_3 (sequence.ml:633-649):
        (median, s1', s2', cost, median_with_gaps)."""
        if len(si) <= len(sj):
            m, ri, rj, c, mwg = self.c_align_affine_3(si, sj, 0)
            return m, ri, rj, c, mwg
        m, ri, rj, c, mwg = self.c_align_affine_3(sj, si, 1)
        return m, rj, ri, c, mwg

    def align_2(self, s1, s2, first_gap: bool = True):
        """Align.align_2 (sequence.ml:1061-1082): (s1', s2', cost)."""
        def cmp(a, b):
            if self.ctx.cm is not None and self.ctx.cm.cost_model == 1:
                _, ap, bp, tc, _ = self.align_affine_3(a, b)
                return ap, bp, tc
            tc = self.cost_2(a, b)
            ap, bp = self.create_edited_2(a, b)
            return ap, bp, tc
        if first_gap:
            return cmp(s1, s2)
        a = np.concatenate([[GAP], np.asarray(s1, np.uint8)]).astype(np.uint8)
        b = np.concatenate([[GAP], np.asarray(s2, np.uint8)]).astype(np.uint8)
        ap, bp, tc = cmp(a, b)
        return ap[1:], bp[1:], tc

    def _median_op(self, fn, a, b, extra=1):
        if len(a) != len(b):
            raise ValueError("The size of the sequences is not the same.")
        x, xp = _u8(a)
        y, yp = _u8(b)
        cap = len(x) + extra
        out = np.zeros(cap, np.uint8)
        n = C.c_int()
        rc = fn(self.ctx.h, xp, yp, len(x), out.ctypes.data_as(C.POINTER(C.c_uint8)), cap, C.byref(n))
        if rc == -1:
            raise ValueError("median should not be 0")
        self.ctx.chk(rc)
        return out[:n.value].copy()

    def ancestor_2(self, a, b):
        """Align.ancestor_2 (sequence.ml:1114) / algn_CAML_ancestor_2."""
        return self._median_op(self.L.poyhip_ancestor_2, a, b)

    def median_2(self, a, b):
        """Sequence.median_2 (sequence.ml:483) / seq_CAML_median_2_no_gaps."""
        return self._median_op(self.L.poyhip_median_2, a, b)

    def median_2_with_gaps(self, a, b):
        """Sequence.median_2_with_gaps (sequence.ml:473)."""
        return self._median_op(self.L.poyhip_median_2_with_gaps, a, b, extra=0)

    def max_cost_2(self, a, b) -> int:
        """Align.max_cost_2 (sequence.ml:652): 0 when a side is all gaps."""
        a = np.asarray(a, np.uint8)
        b = np.asarray(b, np.uint8)
        if len(a) == 0 or len(b) == 0 or np.all(a == GAP) or np.all(b == GAP):
            return 0
        if len(a) != len(b):
            raise AssertionError("aligned sequences differ in length")
        x, xp = _u8(a)
        y, yp = _u8(b)
        out = C.c_int()
        self.ctx.chk(self.L.poyhip_worst_2(self.ctx.h, xp, yp, len(x), C.byref(out)))
        return out.value

    def closest(self, s1, s2):
        """Align.closest (sequence.ml:1180-1236) for the DNA tables (with
        combinations, as every device table is): the states of s2 closest to s1 column by column along
        their alignment, gaps removed; (sequence, cost_2 s1 sequence), cost
        0 when s1 = s2 (no alignment) or s2 is all gaps (s2 returned)."""
        s1 = np.asarray(s1, np.uint8)
        s2 = np.asarray(s2, np.uint8)
        if len(s2) == 0 or np.all(s2 == GAP):
            return s2.copy(), 0
        equal = len(s1) == len(s2) and np.array_equal(s1, s2)
        a, b = (s1, s2) if equal else self.align_2(s1, s2)[:2]
        res = _closest_map(self.ctx, a, b, equal)
        return res, (0 if equal else self.cost_2(s1, res))

    def full_median_2(self, a, b):
        """Align.full_median_2 (sequence.ml:1136)."""
        if self.ctx.cm is not None and self.ctx.cm.cost_model == 1:
            return self.align_affine_3(a, b)[0]
        ap, bp, _ = self.align_2(a, b)
        return self.median_2(ap, bp)

    # -- three sequences (sequence.ml:940-968) ------------------------------
    def cost_3(self, s1, s2, s3, uk: int = 0) -> int:
        """Align.cost_3 / algn_CAML_simple_3."""
        a, ap = _u8(s1)
        b, bp = _u8(s2)
        c, cp = _u8(s3)
        out = C.c_int()
        self.ctx.chk(self.L.poyhip_cost_3(self.ctx.h, ap, len(a), bp, len(b), cp, len(c), uk,
                                          C.byref(out)))
        return out.value

    def _outs3(self, cap):
        bufs = [np.zeros(cap, np.uint8) for _ in range(3)]
        return bufs, [x.ctypes.data_as(C.POINTER(C.c_uint8)) for x in bufs]

    def extract_edited_3(self, s1, s2, s3):
        """Align.extract_edited_3 / algn_CAML_backtrack_3d after cost_3."""
        a, ap = _u8(s1)
        b, bp = _u8(s2)
        c, cp = _u8(s3)
        cap = len(a) + len(b) + len(c)
        bufs, P = self._outs3(cap)
        n = C.c_int()
        self.ctx.chk(self.L.poyhip_backtrack_3d(self.ctx.h, ap, len(a), bp, len(b), cp, len(c),
                                                P[0], P[1], P[2], cap, C.byref(n)))
        return tuple(x[:n.value].copy() for x in bufs)

    def align_3(self, s1, s2, s3, uk: int = 0):
        """Align.c_align_3 / algn_CAML_align_3d: ((s1', s2', s3'), cost)."""
        a, ap = _u8(s1)
        b, bp = _u8(s2)
        c, cp = _u8(s3)
        cap = len(a) + len(b) + len(c)
        bufs, P = self._outs3(cap)
        n, cost = C.c_int(), C.c_int()
        self.ctx.chk(self.L.poyhip_align_3d(self.ctx.h, ap, len(a), bp, len(b), cp, len(c), uk,
                                            P[0], P[1], P[2], cap, C.byref(n), C.byref(cost)))
        return tuple(x[:n.value].copy() for x in bufs), cost.value

    def median_3(self, a, b, c):
        """Align.c_median_3 / algn_CAML_median_3 on aligned sequences."""
        x, xp = _u8(a)
        y, yp = _u8(b)
        z, zp = _u8(c)
        out = np.zeros(max(len(x), 1), np.uint8)
        n = C.c_int()
        self.ctx.chk(self.L.poyhip_median_3(self.ctx.h, xp, yp, zp, len(x),
                                            out.ctypes.data_as(C.POINTER(C.c_uint8)), len(out),
                                            C.byref(n)))
        return out[:n.value].copy()


class NewkkAlign:
    """``Sequence.NewkkAlign`` (src/sequence.ml:1831-2062), affine model:
    the diagonal-store Ukkonen aligner of src/newkkonen.c on the device."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.L = ctx._lib

    # -- externals ---------------------------------------------------------
    def newkk_cost2_affine(self, s1, s2, swaped: int = 0) -> int:
        """newkkonen_CAML_algn_affine: s1 must not be longer than s2."""
        a, ap = _u8(s1)
        b, bp = _u8(s2)
        out = C.c_int()
        self.ctx.chk(self.L.poyhip_newkk_cost_2_affine(self.ctx.h, ap, len(a), bp, len(b),
                                                       int(swaped), C.byref(out)))
        return out.value

    def get_alignment(self, s1, s2, swaped: int):
        """newkkonen_CAML_algn_affine + newkkonen_CAML_backtrace_affine on
        the same sequences (get_alignment, src/sequence.ml:1864-1881):
        (s1', s2', cost)."""
        a, ap = _u8(s1)
        b, bp = _u8(s2)
        cap = len(a) + len(b)
        r1 = np.zeros(cap, np.uint8)
        r2 = np.zeros(cap, np.uint8)
        n, cost = C.c_int(), C.c_int()
        self.ctx.chk(self.L.poyhip_newkk_align_2_affine(
            self.ctx.h, ap, len(a), bp, len(b), int(swaped), C.byref(cost),
            r1.ctypes.data_as(C.POINTER(C.c_uint8)), r2.ctypes.data_as(C.POINTER(C.c_uint8)),
            cap, C.byref(n)))
        return r1[:n.value].copy(), r2[:n.value].copy(), cost.value

    # -- OCaml wrappers ----------------------------------------------------
    def align_2(self, s1, s2, first_gap: bool = True):
        """NewkkAlign.align_2 (src/sequence.ml:1883-1913): shorter first with
        swaped = 1 when exchanged, results swapped back; first_gap False
        prepends a gap to both and strips the first column."""
        def cmp(x, y):
            if len(x) <= len(y):
                p1, p2, tc = self.get_alignment(x, y, 0)
                return p1, p2, tc
            p1, p2, tc = self.get_alignment(y, x, 1)
            return p2, p1, tc
        if first_gap:
            return cmp(np.asarray(s1, np.uint8), np.asarray(s2, np.uint8))
        a = np.concatenate([[GAP], np.asarray(s1, np.uint8)]).astype(np.uint8)
        b = np.concatenate([[GAP], np.asarray(s2, np.uint8)]).astype(np.uint8)
        p1, p2, tc = cmp(a, b)
        return p1[1:], p2[1:], tc

    def cost_2(self, s1, s2) -> int:
        """NewkkAlign.cost_2 (src/sequence.ml:1915-1963), affine model:
        shorter first, an empty sequence replaced by a lone gap."""
        a = np.asarray(s1, np.uint8)
        b = np.asarray(s2, np.uint8)
        if len(a) > len(b):
            a, b = b, a
        if len(a) == 0:
            a = np.array([GAP], np.uint8)
        if len(b) == 0:
            b = np.array([GAP], np.uint8)
        return self.newkk_cost2_affine(a, b, 0)

    def closest(self, s1, s2):
        """NewkkAlign.closest (sequence.ml:1997-2060): Align.closest with
        this module's align_2 and cost_2."""
        s1 = np.asarray(s1, np.uint8)
        s2 = np.asarray(s2, np.uint8)
        if len(s2) == 0 or np.all(s2 == GAP):
            return s2.copy(), 0
        equal = len(s1) == len(s2) and np.array_equal(s1, s2)
        a, b = (s1, s2) if equal else self.align_2(s1, s2)[:2]
        res = _closest_map(self.ctx, a, b, equal)
        return res, (0 if equal else self.cost_2(s1, res))

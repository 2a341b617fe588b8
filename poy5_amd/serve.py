"""Continuous-batching alignment server (``poyhip_server_*``), no reference
counterpart.

Sequences are put into a device-resident pool once and named by id; requests
(``align_2d`` / ``cost_2d`` / ``align_affine`` / ``cost_affine``) name two
pool ids.  An ``align_2d`` request may keep its ancestor median in the pool,
so medians computed on the GPU feed later requests without a host round
trip.  The server's worker thread batches whatever has arrived into one of
its slots as soon as a slot is free.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .batch import OPS
from .sequence import Context


class Server:
    def __init__(self, ctx: Context, n_slots=3, max_batch_pairs=4096, max_len=16384,
                 slot_dm_bytes=4 << 30, pool_bytes=256 << 20):
        self.ctx = ctx
        self.L = ctx._lib
        o = _lib.PoyhipServerOpts(n_slots, max_batch_pairs, max_len, slot_dm_bytes, pool_bytes)
        h = C.c_void_p()
        ctx.chk(self.L.poyhip_server_create(ctx.h, C.byref(o), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.L.poyhip_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def put(self, seq) -> int:
        a = np.ascontiguousarray(seq, np.uint8)
        i = C.c_int64()
        self.ctx.chk(self.L.poyhip_server_put_seq(self.h, a.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                  len(a), C.byref(i)))
        return i.value

    def get(self, sid: int, cap: int = 1 << 16) -> np.ndarray:
        out = np.zeros(cap, np.uint8)
        n = C.c_int()
        self.ctx.chk(self.L.poyhip_server_get_seq(self.h, sid, out.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                  cap, C.byref(n)))
        return out[:n.value].copy()

    def submit(self, op: str, a: int, b: int, deltaw: int | None = None, keep_median=False):
        """Returns (request id, median pool id or None)."""
        r, m = C.c_int64(), C.c_int64(-1)
        flags = _lib.REQ_KEEP_MEDIAN if keep_median else 0
        self.ctx.chk(self.L.poyhip_server_submit(self.h, OPS[op], a, b,
                                                 -1 if deltaw is None else deltaw, flags,
                                                 C.byref(r), C.byref(m)))
        return r.value, (m.value if keep_median else None)

    def wait(self, req: int, timeout_ms: int = -1):
        self.ctx.chk(self.L.poyhip_server_wait(self.h, req, timeout_ms))

    def result(self, req: int, cap: int = 1 << 16, release=True):
        """(cost, r1, r2, latency_us); r1/r2 are empty for cost ops."""
        cost, n, lat = C.c_int(), C.c_int(), C.c_double()
        r1 = np.zeros(cap, np.uint8)
        r2 = np.zeros(cap, np.uint8)
        P = lambda x: x.ctypes.data_as(C.POINTER(C.c_uint8))  # noqa: E731
        self.ctx.chk(self.L.poyhip_server_result(self.h, req, C.byref(cost), P(r1), P(r2), cap,
                                                 C.byref(n), C.byref(lat), 1 if release else 0))
        return cost.value, r1[:n.value].copy(), r2[:n.value].copy(), lat.value

    def stats(self):
        b, p, cells = C.c_int64(), C.c_int64(), C.c_double()
        mb, mi = C.c_int(), C.c_int()
        self.ctx.chk(self.L.poyhip_server_stats(self.h, C.byref(b), C.byref(p), C.byref(cells),
                                                C.byref(mb), C.byref(mi)))
        return dict(batches=b.value, pairs=p.value, cells_full=cells.value, max_batch=mb.value,
                    max_inflight=mi.value)

    def replay(self, arrival_us, ops, a, b, deltaw=None, keep=None):
        """Open-loop arrivals from a native thread: (costs, latency_us,
        median ids, wall_us)."""
        n = len(arrival_us)
        arr = np.ascontiguousarray(arrival_us, np.float64)
        opv = np.ascontiguousarray([OPS[o] if isinstance(o, str) else o for o in ops], np.int32)
        av = np.ascontiguousarray(a, np.int64)
        bv = np.ascontiguousarray(b, np.int64)
        dw = np.ascontiguousarray(np.full(n, -1) if deltaw is None else deltaw, np.int32)
        fl = np.ascontiguousarray(np.zeros(n) if keep is None else
                                  np.where(np.asarray(keep), _lib.REQ_KEEP_MEDIAN, 0), np.int32)
        costs = np.zeros(n, np.int32)
        lat = np.zeros(n, np.float64)
        mids = np.zeros(n, np.int64)
        wall = C.c_double()
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        lp = C.POINTER(C.c_int64)
        self.ctx.chk(self.L.poyhip_server_replay(
            self.h, n, arr.ctypes.data_as(dp), opv.ctypes.data_as(ip), av.ctypes.data_as(lp),
            bv.ctypes.data_as(lp), dw.ctypes.data_as(ip), fl.ctypes.data_as(ip),
            costs.ctypes.data_as(ip), lat.ctypes.data_as(dp), mids.ctypes.data_as(lp),
            C.byref(wall)))
        return costs, lat, mids, wall.value

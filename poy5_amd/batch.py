"""Batched alignment on device-resident inputs (no reference counterpart).

A :class:`Batch` plans one set of pairs (largest first, per-pair regions,
chunks that fit the direction-store budget) and runs it with the persistent
kernels of libpoyhip, optionally replayed through a HIP graph.  PyTorch only
provides the device buffers.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .sequence import Context

OPS = {
    "align_affine": _lib.OP_ALIGN_AFFINE,
    "cost_affine": _lib.OP_COST_AFFINE,
    "align_2d": _lib.OP_ALIGN_2D,
    "cost_2d": _lib.OP_COST_2D,
    "newkk_align": _lib.OP_NEWKK_ALIGN,
    "newkk_cost": _lib.OP_NEWKK_COST,
}


def pack_pairs(pairs):
    """[(s1, s2), ...] -> (blob uint8, off1 int64, off2 int64, len1, len2)."""
    lens = []
    for a, b in pairs:
        lens += [len(a), len(b)]
    total = int(sum(lens))
    blob = np.empty(total, np.uint8)
    off1 = np.empty(len(pairs), np.int64)
    off2 = np.empty(len(pairs), np.int64)
    o = 0
    for k, (a, b) in enumerate(pairs):
        off1[k] = o
        blob[o:o + len(a)] = a
        o += len(a)
        off2[k] = o
        blob[o:o + len(b)] = b
        o += len(b)
    len1 = np.array([len(a) for a, _ in pairs], np.int32)
    len2 = np.array([len(b) for _, b in pairs], np.int32)
    return blob, off1, off2, len1, len2


class Batch:
    """A planned batch on one context."""

    def __init__(self, ctx: Context, op: str, len1, len2, deltaw=None, dm_budget: int = 0):
        self.ctx = ctx
        self.L = ctx._lib
        self.op = OPS[op]
        self.len1 = np.ascontiguousarray(len1, np.int32)
        self.len2 = np.ascontiguousarray(len2, np.int32)
        self.n = len(self.len1)
        dw = None
        if deltaw is not None:
            self._dw = np.ascontiguousarray(deltaw, np.int32)
            dw = self._dw.ctypes.data_as(C.POINTER(C.c_int))
        h = C.c_void_p()
        ctx.chk(self.L.poyhip_batch_create(
            ctx.h, self.op, self.n, self.len1.ctypes.data_as(C.POINTER(C.c_int)),
            self.len2.ctypes.data_as(C.POINTER(C.c_int)), dw, dm_budget, C.byref(h)))
        self.h = h
        self.out_bytes = int(self.L.poyhip_batch_out_bytes(h))

    def close(self):
        if self.h:
            self.L.poyhip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def use_graph(self, on: bool = True):
        self.ctx.chk(self.L.poyhip_batch_use_graph(self.h, 1 if on else 0))

    def alloc_outputs(self, device="cuda"):
        cost = torch.empty(self.n, dtype=torch.int32, device=device)
        out = torch.empty(max(self.out_bytes, 1), dtype=torch.uint8, device=device)
        outlen = torch.empty(4 * self.n, dtype=torch.int32, device=device)
        return cost, out, outlen

    def run(self, d_blob, d_off1, d_off2, cost, out, outlen, sync=True):
        """Inputs and outputs are torch tensors on the context's device."""
        self.ctx.chk(self.L.poyhip_batch_run(
            self.h, d_blob.data_ptr(), d_off1.data_ptr(), d_off2.data_ptr(), cost.data_ptr(),
            out.data_ptr(), outlen.data_ptr(), 1 if sync else 0))

    def fill_ms(self) -> float:
        return float(self.L.poyhip_batch_last_fill_ms(self.h))

    def total_ms(self) -> float:
        return float(self.L.poyhip_batch_last_total_ms(self.h))

    def chunks(self) -> int:
        return int(self.L.poyhip_batch_chunks(self.h))

    def stats(self):
        cells = np.zeros(self.n, np.int64)
        slots = np.zeros(self.n, np.int64)
        rounds = np.zeros(self.n, np.int32)
        newk = np.zeros(self.n, np.int32)
        self.ctx.chk(self.L.poyhip_batch_stats(
            self.h, cells.ctypes.data_as(C.POINTER(C.c_int64)),
            slots.ctypes.data_as(C.POINTER(C.c_int64)),
            rounds.ctypes.data_as(C.POINTER(C.c_int)), newk.ctypes.data_as(C.POINTER(C.c_int))))
        return cells, slots, rounds, newk

    def unpack(self, out_host: np.ndarray, outlen_host: np.ndarray, k: int):
        """Four output sequences of pair k (see poyhip_batch_out_offset)."""
        off = int(self.L.poyhip_batch_out_offset(self.h, k))
        cap = int(self.L.poyhip_batch_out_cap(self.h, k))
        res = []
        for r in range(4):
            n = int(outlen_host[4 * k + r])
            base = off + r * cap
            res.append(out_host[base + cap - n: base + cap].copy())
        return res


def run_pairs(ctx: Context, op: str, pairs, deltaw=None, device="cuda", dm_budget: int = 0):
    """Convenience wrapper: host pairs in, host results out.

    Returns (costs, outputs) where outputs[k] is the list of four sequences
    of pair k for align ops (caller order: s1', s2', median, median with
    gaps) and None for cost ops."""
    blob, off1, off2, len1, len2 = pack_pairs(pairs)
    b = Batch(ctx, op, len1, len2, deltaw=deltaw, dm_budget=dm_budget)
    d_blob = torch.from_numpy(blob).to(device)
    d_off1 = torch.from_numpy(off1).to(device)
    d_off2 = torch.from_numpy(off2).to(device)
    cost, out, outlen = b.alloc_outputs(device)
    torch.cuda.synchronize()
    b.run(d_blob, d_off1, d_off2, cost, out, outlen, sync=True)
    costs = cost.cpu().numpy()
    if b.op in (_lib.OP_COST_AFFINE, _lib.OP_COST_2D, _lib.OP_NEWKK_COST):
        b.close()
        return costs, None
    oh = out.cpu().numpy()
    lh = outlen.cpu().numpy()
    outs = [b.unpack(oh, lh, k) for k in range(len(pairs))]
    b.close()
    return costs, outs


class Batch3:
    """A planned batch of triples (``poyhip_batch3_*``)."""

    OPS = {"align_3d": _lib.OP_ALIGN_3D, "cost_3d": _lib.OP_COST_3D}

    def __init__(self, ctx: Context, op: str, len1, len2, len3, dm_budget: int = 0):
        self.ctx = ctx
        self.L = ctx._lib
        self.op = self.OPS[op]
        self.lens = [np.ascontiguousarray(x, np.int32) for x in (len1, len2, len3)]
        self.n = len(self.lens[0])
        h = C.c_void_p()
        ip = C.POINTER(C.c_int)
        ctx.chk(self.L.poyhip_batch3_create(ctx.h, self.op, self.n,
                                            *[x.ctypes.data_as(ip) for x in self.lens],
                                            dm_budget, C.byref(h)))
        self.h = h
        self.out_bytes = int(self.L.poyhip_batch3_out_bytes(h))

    def close(self):
        if self.h:
            self.L.poyhip_batch3_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def alloc_outputs(self, device="cuda"):
        cost = torch.empty(self.n, dtype=torch.int32, device=device)
        out = torch.empty(max(self.out_bytes, 1), dtype=torch.uint8, device=device)
        outlen = torch.empty(4 * self.n, dtype=torch.int32, device=device)
        return cost, out, outlen

    def run(self, d_blob, d_off1, d_off2, d_off3, cost, out, outlen, sync=True):
        self.ctx.chk(self.L.poyhip_batch3_run(
            self.h, d_blob.data_ptr(), d_off1.data_ptr(), d_off2.data_ptr(), d_off3.data_ptr(),
            cost.data_ptr(), out.data_ptr(), outlen.data_ptr(), 1 if sync else 0))

    def fill_ms(self) -> float:
        return float(self.L.poyhip_batch3_last_fill_ms(self.h))

    def chunks(self) -> int:
        return int(self.L.poyhip_batch3_chunks(self.h))

    def unpack(self, out_host: np.ndarray, outlen_host: np.ndarray, k: int):
        """(r1, r2, r3, median_3) of triple k."""
        off = int(self.L.poyhip_batch3_out_offset(self.h, k))
        cap = int(self.L.poyhip_batch3_out_cap(self.h, k))
        res = []
        for r in range(4):
            n = int(outlen_host[4 * k + r])
            if n < 0:
                raise RuntimeError(f"triple {k}: backtrack left the cube")
            base = off + r * cap
            res.append(out_host[base + cap - n: base + cap].copy())
        return res


def run_triples(ctx: Context, op: str, triples, device="cuda"):
    """Host triples in, host results out: (costs, outputs or None)."""
    lens = [np.array([len(t[u]) for t in triples], np.int32) for u in range(3)]
    total = int(sum(int(x.sum()) for x in lens))
    blob = np.empty(max(total, 1), np.uint8)
    offs = [np.empty(len(triples), np.int64) for _ in range(3)]
    o = 0
    for k, t in enumerate(triples):
        for u in range(3):
            offs[u][k] = o
            blob[o:o + len(t[u])] = t[u]
            o += len(t[u])
    b = Batch3(ctx, op, *lens)
    d_blob = torch.from_numpy(blob).to(device)
    d_offs = [torch.from_numpy(x).to(device) for x in offs]
    cost, out, outlen = b.alloc_outputs(device)
    torch.cuda.synchronize()
    b.run(d_blob, *d_offs, cost, out, outlen, sync=True)
    costs = cost.cpu().numpy()
    if b.op == _lib.OP_COST_3D:
        b.close()
        return costs, None
    oh = out.cpu().numpy()
    lh = outlen.cpu().numpy()
    outs = [b.unpack(oh, lh, k) for k in range(len(triples))]
    b.close()
    return costs, outs

"""``Sequence.Unions`` (src/sequence.ml:2759-2912) over the device union
merge (``poyhip_union_merge``, restating src/union.c).

A union is a sequence right-aligned in a buffer of ``cap`` symbols with
three int32 arrays of ``cap`` entries: ``offset`` (counts of the
non-homologous items before each position), ``c1`` / ``c2`` (the children's
positions of each homologous column).  ``leaf`` builds a tip's union,
``union`` merges two children's unions along their aligned sequences and
aligned median (DOS's ``aligned_children``, src/seqCS.ml:1835-1852), and
``union_level`` runs many merges in one launch.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .sequence import Context

JOB_DTYPE = np.dtype([
    ("a", "<i8"), ("b", "<i8"), ("median", "<i8"),
    ("ua_s", "<i8"), ("ub_s", "<i8"), ("c_s", "<i8"),
    ("ua_off", "<i8"), ("ub_off", "<i8"),
    ("c_off", "<i8"), ("c_c1", "<i8"), ("c_c2", "<i8"),
    ("lena", "<i4"), ("ua_len", "<i4"), ("ua_cap", "<i4"), ("ub_len", "<i4"),
    ("ub_cap", "<i4"), ("c_cap", "<i4"),
])


@dataclass
class Union:
    seq: np.ndarray      # uint8, cap symbols, the sequence at the right end
    length: int
    offset: np.ndarray   # int32, cap
    c1: np.ndarray
    c2: np.ndarray

    @property
    def cap(self) -> int:
        return len(self.seq)

    def sequence(self) -> np.ndarray:
        return self.seq[self.cap - self.length:].copy()


def leaf(s) -> Union:
    """Unions.leaf (src/sequence.ml:2833): offsets 0, c1 = c2 = 0..len-1."""
    s = np.asarray(s, np.uint8).copy()
    n = len(s)
    return Union(s, n, np.zeros(n, np.int32), np.arange(n, dtype=np.int32),
                 np.arange(n, dtype=np.int32))


def _p8(x):
    return x.ctypes.data_as(C.POINTER(C.c_uint8))


def _p32(x):
    return x.ctypes.data_as(C.POINTER(C.c_int32))


def union(ctx: Context, a, b, median, ua: Union, ub: Union) -> Union:
    """Unions.union (src/sequence.ml:2849): capacity len(ua) + len(ub) + 2."""
    a = np.ascontiguousarray(a, np.uint8)
    b = np.ascontiguousarray(b, np.uint8)
    m = np.ascontiguousarray(median, np.uint8)
    if not (len(a) == len(b) == len(m)):
        raise ValueError("aligned sequences of different lengths")
    cap = ua.length + ub.length + 2
    cs = np.zeros(cap, np.uint8)
    co, c1, c2 = (np.zeros(cap, np.int32) for _ in range(3))
    n = C.c_int()
    uas, uao = np.ascontiguousarray(ua.seq), np.ascontiguousarray(ua.offset, np.int32)
    ubs, ubo = np.ascontiguousarray(ub.seq), np.ascontiguousarray(ub.offset, np.int32)
    ctx.chk(ctx._lib.poyhip_union_merge(ctx.h, _p8(a), _p8(b), _p8(m), len(a), _p8(uas),
                                        ua.length, ua.cap, _p32(uao), _p8(ubs), ub.length,
                                        ub.cap, _p32(ubo), _p8(cs), cap, _p32(co), _p32(c1),
                                        _p32(c2), C.byref(n)))
    return Union(cs, n.value, co, c1, c2)


def pack_level(merges):
    """Job descriptors and blobs of many merges: merges = [(a, b, median,
    ua, ub), ...].  Returns (jobs, seq blob, int blob, new capacities)."""
    seq_parts, int_parts = [], []
    so = io = 0
    jobs = np.zeros(len(merges), JOB_DTYPE)

    def put_seq(x):
        nonlocal so
        at = so
        seq_parts.append(np.asarray(x, np.uint8))
        so += len(x)
        return at

    def put_int(x):
        nonlocal io
        at = io
        int_parts.append(np.asarray(x, np.int32))
        io += len(x)
        return at

    caps = []
    for k, (a, b, m, ua, ub) in enumerate(merges):
        cap = ua.length + ub.length + 2
        caps.append(cap)
        j = jobs[k]
        j["a"], j["b"], j["median"] = put_seq(a), put_seq(b), put_seq(m)
        j["lena"] = len(a)
        j["ua_s"], j["ua_len"], j["ua_cap"] = put_seq(ua.seq), ua.length, ua.cap
        j["ub_s"], j["ub_len"], j["ub_cap"] = put_seq(ub.seq), ub.length, ub.cap
        j["ua_off"], j["ub_off"] = put_int(ua.offset), put_int(ub.offset)
        j["c_s"], j["c_cap"] = put_seq(np.zeros(cap, np.uint8)), cap
        j["c_off"] = put_int(np.zeros(cap, np.int32))
        j["c_c1"] = put_int(np.zeros(cap, np.int32))
        j["c_c2"] = put_int(np.zeros(cap, np.int32))
    return jobs, np.concatenate(seq_parts), np.concatenate(int_parts), caps


class Level:
    """A packed level resident on the device, run as one launch.  The jobs
    are laid out longest walk first: a wavefront runs as long as its longest
    merge, so merges of similar length share wavefronts."""

    def __init__(self, ctx: Context, merges, device="cuda"):
        self.ctx = ctx
        self.order = sorted(range(len(merges)), key=lambda k: -len(merges[k][0]))
        self.jobs, seq, ints, self.caps = pack_level([merges[k] for k in self.order])
        dev = torch.device(device)
        self.d_seq = torch.from_numpy(seq).to(dev)
        self.d_int = torch.from_numpy(ints).to(dev)
        self.d_jobs = torch.from_numpy(self.jobs.view(np.uint8).copy()).to(dev)
        self.d_len = torch.zeros(len(merges), dtype=torch.int32, device=dev)
        self.d_st = torch.zeros(len(merges), dtype=torch.int32, device=dev)
        torch.cuda.synchronize()

    def run(self, sync=True):
        self.ctx.chk(self.ctx._lib.poyhip_union_merge_batch(
            self.ctx.h, len(self.caps), C.c_void_p(self.d_jobs.data_ptr()),
            C.c_void_p(self.d_seq.data_ptr()), self.d_seq.numel(),
            C.c_void_p(self.d_int.data_ptr()), self.d_int.numel(),
            C.c_void_p(self.d_len.data_ptr()), C.c_void_p(self.d_st.data_ptr()),
            1 if sync else 0))

    def results(self):
        hs, hi = self.d_seq.cpu().numpy(), self.d_int.cpu().numpy()
        hl, hst = self.d_len.cpu().numpy(), self.d_st.cpu().numpy()
        out = [None] * len(self.caps)
        for k, cap in enumerate(self.caps):
            if hst[k] != _lib.POYHIP_OK:
                continue
            j = self.jobs[k]
            out[self.order[k]] = (Union(hs[j["c_s"]: j["c_s"] + cap].copy(), int(hl[k]),
                             hi[j["c_off"]: j["c_off"] + cap].copy(),
                             hi[j["c_c1"]: j["c_c1"] + cap].copy(),
                             hi[j["c_c2"]: j["c_c2"] + cap].copy()))
        return out


def union_level(ctx: Context, merges, device="cuda"):
    """Many merges in one launch: merges = [(a, b, median, ua, ub), ...].
    Returns the new unions (None where the reference fails)."""
    lv = Level(ctx, merges, device)
    lv.run(sync=True)
    return lv.results()

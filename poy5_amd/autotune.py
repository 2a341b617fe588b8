"""Autotuner of the batch kernels (no reference counterpart).

A context has four tuning knobs (``poyhip_set_tuning``): resident
wavefronts per CU of a pair batch's persistent fill, the rotating stripe
schedule of gap-free affine pairs, the workgroup-per-pair linear fill and
the workgroup-per-triple 3D fill.  Results never depend on them; the fill
time does, and by how much depends on the lengths.

:func:`tune_pairs` cuts a pair batch into length buckets (by the longer
sequence of each pair), times every candidate setting on a sample of each
bucket with the batch kernels themselves (fill stage, HIP events, best of a
few runs) and keeps the fastest per bucket.  One persistent launch serves
the whole batch, so the setting applied is the one with the smallest
predicted total: each bucket's sample time scaled by the bucket's share of
the cells.  A sample smaller than the persistent grid would run every
pair at once whatever the occupancy knob, so a bucket's sample is large
(8,192 pairs or the whole bucket), and the default setting is kept unless
another is predicted at least 2% faster.
"""
from __future__ import annotations

import itertools
import time

import numpy as np
import torch

from .batch import Batch
from .sequence import Context

KNOBS = {"waves_per_cu": 1, "affine_rotating": 2, "linear_workgroup": 3, "cube_workgroup": 4}

# bucket edges on the longer sequence of a pair
EDGES = (0, 256, 1024, 4096, 1 << 62)


def get_knobs(ctx: Context) -> dict:
    return {k: int(ctx._lib.poyhip_get_tuning(ctx.h, v)) for k, v in KNOBS.items()}


def set_knobs(ctx: Context, settings: dict) -> None:
    for k, v in settings.items():
        ctx.chk(ctx._lib.poyhip_set_tuning(ctx.h, KNOBS[k], int(v)))


def candidates(op: str) -> list[dict]:
    """Settings tried for a pair op."""
    if op in ("align_affine", "cost_affine"):
        rot = (1, 0) if op == "align_affine" else (1,)
        return [{"waves_per_cu": w, "affine_rotating": r}
                for w, r in itertools.product((8, 12, 16), rot)]
    return [{"waves_per_cu": w, "linear_workgroup": g}
            for w, g in itertools.product((8, 16), (2, 1, 0))]


def _time_fill(ctx, op, d_blob, off1, off2, len1, len2, idx, repeats, device):
    """Best fill time (ms) of the pairs idx as one batch."""
    sub_len1, sub_len2 = len1[idx], len2[idx]
    b = Batch(ctx, op, sub_len1, sub_len2)
    try:
        d1 = torch.from_numpy(np.ascontiguousarray(off1[idx])).to(device)
        d2 = torch.from_numpy(np.ascontiguousarray(off2[idx])).to(device)
        cost, out, outlen = b.alloc_outputs(device)
        best = float("inf")
        for _ in range(repeats + 1):  # the first run warms up
            b.run(d_blob, d1, d2, cost, out, outlen, sync=True)
            best = min(best, b.fill_ms())
        return best
    finally:
        b.close()


def tune_pairs(ctx: Context, op: str, blob, off1, off2, len1, len2, sample=8192, repeats=2,
               seed=1, device="cuda", apply=True, margin=0.02) -> dict:
    """Tunes the knobs for a pair batch (host arrays of the batch layout).
    Returns {"choice", "default", "predicted_ms", "buckets": [...]} and, with
    apply, leaves the chosen setting on the context."""
    t0 = time.perf_counter()
    len1 = np.asarray(len1, np.int32)
    len2 = np.asarray(len2, np.int32)
    longer = np.maximum(len1, len2)
    cells = len1.astype(np.float64) * len2
    rng = np.random.default_rng(seed)
    default = get_knobs(ctx)
    cands = candidates(op)
    keys = sorted({k for c in cands for k in c})
    total = {i: 0.0 for i in range(len(cands))}
    buckets = []
    d_blob = torch.from_numpy(np.asarray(blob)).to(device)
    try:
        for lo, hi in zip(EDGES[:-1], EDGES[1:]):
            members = np.nonzero((longer >= lo) & (longer < hi))[0]
            if len(members) == 0:
                continue
            idx = np.sort(rng.choice(members, size=min(sample, len(members)), replace=False))
            scale = cells[members].sum() / cells[idx].sum()
            ms = []
            for cand in cands:
                set_knobs(ctx, cand)
                ms.append(_time_fill(ctx, op, d_blob, off1, off2, len1, len2, idx, repeats, device))
            for i, t in enumerate(ms):
                total[i] += t * scale
            best = int(np.argmin(ms))
            buckets.append({"longer": [int(lo), int(min(hi, longer.max() + 1))],
                            "pairs": int(len(members)), "sample": int(len(idx)),
                            "best": cands[best],
                            "sample_ms": {str(i): round(t, 3) for i, t in enumerate(ms)}})
    finally:
        set_knobs(ctx, {k: default[k] for k in keys})
    pick = min(total, key=total.get)
    dflt = [i for i, c in enumerate(cands) if all(default[k] == v for k, v in c.items())]
    if dflt and total[pick] > (1.0 - margin) * total[dflt[0]]:
        pick = dflt[0]
    choice = {k: cands[pick][k] for k in keys}
    if apply:
        set_knobs(ctx, choice)
    return {"choice": choice, "default": {k: default[k] for k in keys},
            "candidates": cands, "predicted_ms": {str(i): round(t, 3) for i, t in total.items()},
            "buckets": buckets, "tune_s": round(time.perf_counter() - t0, 3)}

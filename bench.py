#!/usr/bin/env python3
"""Benchmark: batched uneven affine-gap DO alignment on one MI355X.

Workload (config 2): 65,536 DNA pairs, root lengths log-uniform 50-5,000,
per-pair divergence U[2%,25%] at the 10:3 substitution:indel ratio, 5% IUPAC
codes; affine costs substitution 1, gap extension 1, gap opening 3.  One
step = POY's align_affine_3 (banded fill with Ukkonen doubling, backtrace,
median and median-with-gaps) over every pair, inputs already resident in HBM.

Printed JSON (one line): GCUPS on full cells (sum of L1*L2, leading gaps
included) per second of the timed steps, the fill kernel's VALU roofline,
the plain per-call GPU path (one launch per pair through the drop-in call)
on a sample, the CPU oracle on a bounded sample over the host cores, and
the other configurations (1, 4, 5 at full size, each with its roofline and
CPU baseline) plus the serving run.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GCUPS on uneven affine-gap pair batches vs plain per-call; p50/p99 latency"
OPS_PER_CELL = 35          # int32 adds+compares per affine cell with backtrace state
VALU_PEAK_TOPS = 39.3      # 256 CU x 64 lanes x 2.4 GHz int32 VALU ops/s (1e12)
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=65536)
    ap.add_argument("--plain-sample", type=int, default=512)
    ap.add_argument("--cpu-stride", type=int, default=4,
                    help="CPU baseline of the headline runs every N-th pair")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving run (128-taxon SPR rejoin traffic)")
    ap.add_argument("--serve-requests", type=int, default=20000)
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the secondary configurations (1, 4, 5)")
    ap.add_argument("--c1-pairs", type=int, default=1024)
    ap.add_argument("--c4-pairs", type=int, default=2048)
    ap.add_argument("--c5-triples", type=int, default=16384)
    ap.add_argument("--nk-pairs", type=int, default=65536)
    ap.add_argument("--rj-nodes", type=int, default=4096)
    ap.add_argument("--only", choices=["c1", "c4", "c5", "nk", "un", "rj", "serve"], default=None,
                    help="run one secondary configuration or the serving run alone (for "
                         "profiling) and print its record")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default tuning knobs instead of autotuning config 2")
    ap.add_argument("--graph", type=int, default=0,
                    help="replay each step as one HIP graph (no per-kernel events then)")
    ap.add_argument("--dm-budget-gb", type=float, default=0.0)
    ap.add_argument("--copy-calib-gb", type=float, default=0.0,
                    help="also run one device-to-device copy of this size (counter calibration)")
    return ap.parse_args()


def host_cores():
    """Cores this process may use, capped at 16 (the host cores that go with
    one GPU on the measurement nodes)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_job(job):
    """One host process of a CPU baseline: the oracle over its share of a
    sample.  job = (kind, tcm, blob, offsets..., lengths...)."""
    import numpy as np
    import oracle
    L = oracle.lib()
    kind = job[0]
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
    if kind == "affine":
        _, tcm, blob, o1, l1, o2, l2 = job
        cm = oracle.cm_dna(*tcm)
        costs = np.zeros(len(l1), np.int32)
        cells = C.c_int64()
        t0 = time.perf_counter()
        L.orc_batch_align_affine(len(l1), P(blob, C.c_uint8), P(o1, C.c_int64), P(l1, C.c_int),
                                 P(o2, C.c_int64), P(l2, C.c_int), C.byref(cm),
                                 P(costs, C.c_int), C.byref(cells))
    elif kind == "newkk":
        _, tcm, blob, o1, l1, o2, l2 = job
        cm = oracle.cm_dna(*tcm)
        costs = np.zeros(len(l1), np.int32)
        cells = C.c_int64()
        t0 = time.perf_counter()
        L.orc_batch_newkk(len(l1), P(blob, C.c_uint8), P(o1, C.c_int64), P(l1, C.c_int),
                          P(o2, C.c_int64), P(l2, C.c_int), C.byref(cm), P(costs, C.c_int),
                          C.byref(cells))
    elif kind == "linear":
        _, tcm, blob, o1, l1, o2, l2 = job
        cm = oracle.cm_dna(*tcm)
        costs = np.zeros(len(l1), np.int32)
        cells = C.c_int64()
        t0 = time.perf_counter()
        L.orc_batch_align_2(len(l1), P(blob, C.c_uint8), P(o1, C.c_int64), P(l1, C.c_int),
                            P(o2, C.c_int64), P(l2, C.c_int), C.byref(cm), P(costs, C.c_int),
                            C.byref(cells))
    elif kind == "readjust":
        _, tcm, blob, o1, l1, o2, l2, o3, l3 = job
        cm = oracle.cm_dna(*tcm)
        costs = np.zeros(len(l1), np.int32)
        t0 = time.perf_counter()
        for k in range(len(l1)):
            a = blob[o1[k]: o1[k] + l1[k]]
            b = blob[o2[k]: o2[k] + l2[k]]
            p = blob[o3[k]: o3[k] + l3[k]]
            costs[k] = oracle.seq_readjust(a, b, a, p, cm)[0]
    else:
        _, tcm, blob, o1, l1, o2, l2, o3, l3 = job
        c3 = oracle.cm3_from_2d(oracle.cm_dna(*tcm))
        costs = np.zeros(len(l1), np.int32)
        t0 = time.perf_counter()
        L.orc_batch_align_3d(len(l1), P(blob, C.c_uint8), P(o1, C.c_int64), P(l1, C.c_int),
                             P(o2, C.c_int64), P(l2, C.c_int), P(o3, C.c_int64), P(l3, C.c_int),
                             C.byref(c3), P(costs, C.c_int))
    return time.perf_counter() - t0, costs


def _warm_worker(_):
    import oracle
    oracle.lib()
    return 0


def cpu_baseline(pool, procs, kind, tcm, blob, offs, lens, sample, gpu_costs, what):
    """The oracle (a C restatement of POY's algn.c path) on the pairs or
    triples `sample`, LPT-partitioned over `procs` worker processes (forked
    before any GPU work).  GCUPS on full cells."""
    import numpy as np
    nseq = len(offs)
    work = np.ones(len(sample))
    for u in range(nseq):
        work = work * lens[u][sample].astype(np.float64)
    parts = [[] for _ in range(procs)]
    load = [0.0] * procs
    for k in np.argsort(-work):
        q = int(np.argmin(load))
        parts[q].append(k)
        load[q] += work[k]
    jobs, idx = [], []
    for part in parts:
        if not part:
            continue
        sel = sample[np.array(part, np.int64)]
        # pack the share's sequences into one blob
        pieces, newoffs, o = [], [[] for _ in range(nseq)], 0
        for k in sel:
            for u in range(nseq):
                seq = blob[offs[u][k]: offs[u][k] + lens[u][k]]
                newoffs[u].append(o)
                pieces.append(seq)
                o += len(seq)
        sub = np.concatenate(pieces).astype(np.uint8)
        args = [kind, tcm, sub]
        for u in range(nseq):
            args += [np.array(newoffs[u], np.int64), lens[u][sel].astype(np.int32)]
        jobs.append(tuple(args))
        idx.append(sel)
    # the workers load the oracle before the clock starts
    pool.map(_warm_worker, range(len(jobs)))
    t0 = time.perf_counter()
    res = pool.map(cpu_job, jobs)
    wall = time.perf_counter() - t0
    match = all(np.array_equal(c, gpu_costs[i]) for (_, c), i in zip(res, idx))
    cells = float(work.sum())
    return {"value": round(cells / wall / 1e9, 3), "unit": "GCUPS", "cores": len(jobs),
            "kind": "port", "wall_s": round(wall, 3),
            "sample": f"{what} ({len(sample)} of {len(lens[0])}, {cells:.3e} full cells), "
                      f"oracle/oracle.c compiled -O2, {len(jobs)} processes, wall {wall:.2f} s",
            "costs_match_gpu": bool(match)}


def plain_per_call(ctx, al, blob, off, ln, sample_n, seed=1):
    """One drop-in call (algn_CAML_align_affine_3) per pair: upload, fill
    launch, backtrace launch, download, synchronise."""
    import numpy as np
    n = len(ln) // 2
    rng = np.random.default_rng(seed)
    idx = rng.choice(n, size=min(sample_n, n), replace=False)
    lat = []
    cells = 0.0
    seqs = []
    for k in idx:
        a = blob[off[2 * k]: off[2 * k] + ln[2 * k]]
        b = blob[off[2 * k + 1]: off[2 * k + 1] + ln[2 * k + 1]]
        seqs.append((a, b))
    al.align_affine_3(*seqs[0])  # first-call setup outside the timing
    t0 = time.perf_counter()
    for a, b in seqs:
        t1 = time.perf_counter()
        al.align_affine_3(a, b)
        lat.append((time.perf_counter() - t1) * 1e3)
        cells += float(len(a)) * len(b)
    wall = time.perf_counter() - t0
    lat = np.array(lat)
    return {"gcups": cells / wall / 1e9, "pairs": int(len(idx)), "cells": cells,
            "wall_s": wall, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99))}


def _timed(run, steps=2):
    run()
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


# int32 adds and compares per cell of the reference recurrence
OPS_LINEAR = 6    # algn_fill_row: 3 candidates x (add + compare)
OPS_CUBE = 14     # algn_fill_cube: 7 candidates x (add + compare)
OPS_NEWKK = 35    # newkkonen.c's affine cell: four states' adds, minima and
                  # tie compares, counted as for algn.c's affine cell


def unions_workload(args, pool, procs):
    """The union merge (src/union.c) of a tree's first level: the unions of
    16,384 tip pairs of the config-2 workload, their aligned children and
    median with gaps from the batched align_affine_3 (DOS.median's affine
    branch), merged in one launch.  Inputs resident; merges per second and
    the HBM bytes the merges read and write against the HBM roofline; the
    oracle's restatement on a sample."""
    import numpy as np
    import torch
    from poy5_amd import cost_matrix, unions, workloads
    from poy5_amd.batch import Batch
    from poy5_amd.sequence import Context
    n = 16384
    print(f"[bench] union merge: {n} tip pairs", file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    ctx = Context(0)
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1, 3))
    blob, off, ln = workloads.gen_pairs(2, n, seed=7)
    l1, l2 = ln[0::2].astype(np.int32), ln[1::2].astype(np.int32)
    b = Batch(ctx, "align_affine", l1, l2)
    d_blob = torch.from_numpy(blob).to(dev)
    d1 = torch.from_numpy(np.ascontiguousarray(off[0::2])).to(dev)
    d2 = torch.from_numpy(np.ascontiguousarray(off[1::2])).to(dev)
    cost, o, ol = b.alloc_outputs(dev)
    b.run(d_blob, d1, d2, cost, o, ol, sync=True)
    oh, lh = o.cpu().numpy(), ol.cpu().numpy()
    merges = []
    for k in range(n):
        ra, rb, _, mwg = b.unpack(oh, lh, k)
        sa = blob[off[2 * k]: off[2 * k] + ln[2 * k]]
        sb = blob[off[2 * k + 1]: off[2 * k + 1] + ln[2 * k + 1]]
        merges.append((ra, rb, mwg, unions.leaf(sa), unions.leaf(sb)))
    b.close()
    del cost, o, ol, d_blob
    lv = unions.Level(ctx, merges, dev)
    lv.run(sync=True)  # warm-up
    steps = 5
    t0 = time.perf_counter()
    for _ in range(steps):
        lv.run(sync=True)
    dt = (time.perf_counter() - t0) / steps
    res = lv.results()
    ok = sum(r is not None for r in res)
    lena = np.array([len(m[0]) for m in merges], np.float64)
    caps = np.array(lv.caps, np.float64)
    ulen = np.array([m[3].length + m[4].length for m in merges], np.float64)
    # bytes a merge must move: the aligned triple, the children's sequences
    # and offsets, the new union's sequence and three arrays (cleared, then
    # written)
    moved = float((3 * lena + 5 * ulen + 13 * caps).sum())
    rec = {"merges": n, "ok": ok, "ms": round(dt * 1e3, 3),
           "merges_per_s": round(n / dt, 1),
           "columns_per_s": round(float(lena.sum()) / dt, 1),
           "roofline": {"bound": "hbm", "achieved": round(moved / dt / 1e9, 2),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(moved / dt / 1e9 / HBM_PEAK_GBS, 5),
                        "kernel": "k_union_merge (one lane per merge: a serial walk)"}}
    if pool is not None:
        import oracle
        leaves = [(oracle.union_leaf(m[3].seq), oracle.union_leaf(m[4].seq)) for m in merges]
        t0 = time.perf_counter()
        ours = [oracle.union_merge(m[0], m[1], m[2], la, lb) for m, (la, lb) in zip(merges, leaves)]
        cdt = time.perf_counter() - t0
        same = all(g is not None and g.length == o.length and np.array_equal(g.offset, o.offset)
                   and np.array_equal(g.c1, o.c1) and np.array_equal(g.c2, o.c2)
                   for g, o in zip(res, ours))
        rec["cpu_baseline"] = {"value": round(n / cdt, 1), "unit": "merges/s",
                               "cores": 1, "kind": "port",
                               "sample": f"all {n} merges, oracle/union.c through ctypes, one "
                                         f"process, {cdt:.2f} s",
                               "results_match_gpu": bool(same)}
    ctx.close()
    torch.cuda.empty_cache()
    return rec


def readjust_workload(args, pool, procs):
    """DOS.readjust's pairwise centre search (Sequence.readjust,
    src/sequence.ml:2097) for a level of nodes: config 5's triples as
    (child, child, parent), its table (substitution 1, indel 1), ten
    alignments and the closest map per node in four batched rounds.  The
    level takes host sequences and returns host results (uploads, the
    host's choices between rounds and downloads inside the timing); `ms` is
    the mean over three levels after an untimed one."""
    import numpy as np
    import torch
    from poy5_amd import cost_matrix, readjust, workloads
    from poy5_amd.sequence import Context
    n = args.rj_nodes
    print(f"[bench] readjust: {n} nodes", file=sys.stderr, flush=True)
    ctx = Context(0)
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1))
    blob, off, ln = workloads.gen_triples(n)
    tri = [tuple(blob[off[3 * k + u]: off[3 * k + u] + ln[3 * k + u]] for u in range(3))
           for k in range(n)]
    # an untimed level of the same size first (library set-up and the
    # allocator's first reservation), then the mean of three timed levels
    readjust.readjust_level(ctx, tri)
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        res = readjust.readjust_level(ctx, tri)
    dt = (time.perf_counter() - t0) / reps
    # full cells of four of the ten alignments per node: the three input
    # pairs and the new median against the three (the median-against-third
    # and closest alignments are left out, so this bounds the work below)
    l = [ln[u::3].astype(np.float64) for u in range(3)]
    m = np.array([len(r[2]) for r in res], np.float64)
    cells = float(np.sum(l[0] * l[1] + l[1] * l[2] + l[0] * l[2] + m * (l[0] + l[1] + l[2])))
    rec = {"nodes": n, "ms": round(dt * 1e3, 3), "nodes_per_s": round(n / dt, 1),
           "cells_lower_bound": cells, "gcups_lower_bound": round(cells / dt / 1e9, 2),
           "rounds": 4, "alignments_per_node": 10, "levels_timed": reps}
    gpu_cost3 = np.array([r[0] for r in res], np.int32)
    ctx.close()
    torch.cuda.empty_cache()
    if pool is not None:
        sample = np.arange(n)
        lens = [ln[u::3].astype(np.int32) for u in range(3)]
        bl = cpu_baseline(pool, procs, "readjust", (1, 1), blob, [off[0::3], off[1::3], off[2::3]],
                          lens, sample, gpu_cost3, "every node")
        rec["cpu_baseline"] = {"value": round(len(sample) / bl["wall_s"], 2),
                               "unit": "nodes/s", "cores": bl["cores"], "kind": "port",
                               "sample": bl["sample"].replace("oracle/oracle.c",
                                                              "oracle.seq_readjust over oracle/*.c"),
                               "costs_match_gpu": bl["costs_match_gpu"]}
    return rec


def newkk_workload(args, pool, procs):
    """The diagonal-store affine aligner (src/newkkonen.c) on the config-2
    pairs: NewkkAlign.align_2 + ancestor_2 + median_2_with_gaps per pair
    (DOS.median's use_ukk branch), batched, inputs resident.  GCUPS on full
    cells; roofline of the fill stage on the band cells of every round; the
    oracle's restatement (the reference's incremental refill) on a sample."""
    import numpy as np
    import torch
    from poy5_amd import cost_matrix, workloads
    from poy5_amd.batch import Batch
    from poy5_amd.sequence import Context
    dev = torch.device("cuda:0")
    n = args.nk_pairs
    tcm = (1, 1, 3)
    print(f"[bench] diagonal-store aligner: {n} config-2 pairs", file=sys.stderr, flush=True)
    ctx = Context(0)
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(*tcm))
    blob, off, ln = workloads.gen_pairs(2, n)
    l1, l2 = ln[0::2].astype(np.int32), ln[1::2].astype(np.int32)
    b = Batch(ctx, "newkk_align", l1, l2)
    d_blob = torch.from_numpy(blob).to(dev)
    d1 = torch.from_numpy(np.ascontiguousarray(off[0::2])).to(dev)
    d2 = torch.from_numpy(np.ascontiguousarray(off[1::2])).to(dev)
    cost, o, ol = b.alloc_outputs(dev)
    fills = []

    def run():
        b.run(d_blob, d1, d2, cost, o, ol, sync=True)
        fills.append(b.fill_ms())
    dt = _timed(run, steps=2)
    fill_ms = float(np.mean(fills[1:]))
    cells = float(np.sum(l1.astype(np.float64) * l2))
    cb, sl, rounds, _ = b.stats()
    computed = float(cb.sum())
    achieved = computed * OPS_NEWKK / (fill_ms / 1e3) / 1e12
    rec = {"pairs": n, "op": "newkk_align", "tcm": list(tcm), "ms": round(dt * 1e3, 3),
           "cells_full": cells, "cells_computed": computed, "lane_steps": float(sl.sum()),
           "gcups": round(cells / dt / 1e9, 2), "chunks": b.chunks(),
           "mean_rounds": float(rounds.mean()),
           "roofline": {"bound": "valu", "achieved": round(achieved, 3),
                        "peak": VALU_PEAK_TOPS, "unit": "TOP/s",
                        "frac": round(achieved / VALU_PEAK_TOPS, 4),
                        "ops_per_cell": OPS_NEWKK, "fill_ms": round(fill_ms, 3),
                        "kernel": "k_newkk_fill<true,false> (fill stage, every doubling round)"}}
    gpu_costs = cost.cpu().numpy()
    b.close()
    del cost, o, ol, d_blob
    ctx.close()
    torch.cuda.empty_cache()
    if pool is not None:
        stride = 16
        sample = np.arange(0, n, stride)
        rec["cpu_baseline"] = cpu_baseline(
            pool, procs, "newkk", tcm, blob, [off[0::2], off[1::2]], [l1, l2], sample,
            gpu_costs, f"every {stride}th pair")
        rec["cpu_baseline"]["sample"] = rec["cpu_baseline"]["sample"].replace(
            "oracle/oracle.c", "oracle/newkk.c")
    return rec


def secondary(args, pool, procs, which=("c1", "c4", "c5")):
    """The other workloads at their full sizes, batched, inputs resident:
    config 1 (1,024 linear pairs, align_2 + ancestor + median with gaps),
    config 4 (2,048 pairs of 16k-100k, linear sub 1 / indel 1, band
    doubling, align_2 + medians), config 5 (16,384 triples, 3D cube +
    backtrack + median_3).  GCUPS on full cells; roofline of the fill stage
    on the cells the reference's loops evaluate; the oracle on a sample."""
    import numpy as np
    import torch
    from poy5_amd import cost_matrix, workloads
    from poy5_amd.batch import Batch, Batch3
    from poy5_amd.sequence import Context
    dev = torch.device("cuda:0")
    out = {}
    for cfg, n, tcm, steps, cpu_stride, cpu_procs in (
            (1, args.c1_pairs, (1, 2), 3, 1, procs),
            (4, args.c4_pairs, (1, 1), 1, 32, max(1, procs // 2))):
        if f"c{cfg}" not in which:
            continue
        print(f"[bench] config {cfg}: {n} pairs", file=sys.stderr, flush=True)
        ctx = Context(0)
        ctx.set_cm(cost_matrix.of_transformations_and_gaps(*tcm))
        blob, off, ln = workloads.gen_pairs(cfg, n)
        l1, l2 = ln[0::2].astype(np.int32), ln[1::2].astype(np.int32)
        b = Batch(ctx, "align_2d", l1, l2)
        d_blob = torch.from_numpy(blob).to(dev)
        d1 = torch.from_numpy(np.ascontiguousarray(off[0::2])).to(dev)
        d2 = torch.from_numpy(np.ascontiguousarray(off[1::2])).to(dev)
        cost, o, ol = b.alloc_outputs(dev)
        fills = []

        def run():
            b.run(d_blob, d1, d2, cost, o, ol, sync=True)
            fills.append(b.fill_ms())
        dt = _timed(run, steps=steps)
        fill_ms = float(np.mean(fills[1:]))
        cells = float(np.sum(l1.astype(np.float64) * l2))
        cb, sl, rounds, _ = b.stats()
        computed = float(cb.sum())
        achieved = computed * OPS_LINEAR / (fill_ms / 1e3) / 1e12
        rec = {"pairs": n, "op": "align_2d", "tcm": list(tcm), "ms": round(dt * 1e3, 3),
               "cells_full": cells, "cells_computed": computed,
               "gcups": round(cells / dt / 1e9, 2), "chunks": b.chunks(),
               "mean_rounds": float(rounds.mean()),
               "roofline": {"bound": "valu", "achieved": round(achieved, 3),
                            "peak": VALU_PEAK_TOPS, "unit": "TOP/s",
                            "frac": round(achieved / VALU_PEAK_TOPS, 4),
                            "ops_per_cell": OPS_LINEAR, "fill_ms": round(fill_ms, 3),
                            "kernel": "k_linear_mw (fill stage, every doubling round)"}}
        gpu_costs = cost.cpu().numpy()
        b.close()
        del cost, o, ol, d_blob
        ctx.close()
        torch.cuda.empty_cache()
        if pool is not None:
            sample = np.arange(0, n, cpu_stride)
            rec["cpu_baseline"] = cpu_baseline(
                pool, cpu_procs, "linear", tcm, blob, [off[0::2], off[1::2]], [l1, l2],
                sample, gpu_costs, f"every {cpu_stride}th pair" if cpu_stride > 1 else "all pairs")
        out[f"config{cfg}"] = rec
    if "c5" not in which:
        return out
    print(f"[bench] config 5: {args.c5_triples} triples", file=sys.stderr, flush=True)
    ctx = Context(0)
    two = cost_matrix.of_transformations_and_gaps(1, 1)
    ctx.set_cm(two)
    ctx.set_cm3(cost_matrix.three_d_of_two_dim_comb(two))
    n3 = args.c5_triples
    blob, off, ln = workloads.gen_triples(n3)
    lens = [ln[u::3].astype(np.int32) for u in range(3)]
    b3 = Batch3(ctx, "align_3d", *lens)
    d_blob = torch.from_numpy(blob).to(dev)
    ds = [torch.from_numpy(np.ascontiguousarray(off[u::3])).to(dev) for u in range(3)]
    cost, o, ol = b3.alloc_outputs(dev)
    dt = _timed(lambda: b3.run(d_blob, *ds, cost, o, ol, sync=True), steps=1)
    cells = float(np.sum(lens[0].astype(np.float64) * lens[1] * lens[2]))
    fill_ms = b3.fill_ms()
    achieved = cells * OPS_CUBE / (fill_ms / 1e3) / 1e12
    rec = {"triples": n3, "op": "align_3d", "ms": round(dt * 1e3, 3), "cells_full": cells,
           "gcups": round(cells / dt / 1e9, 2), "chunks": b3.chunks(),
           "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": VALU_PEAK_TOPS,
                        "unit": "TOP/s", "frac": round(achieved / VALU_PEAK_TOPS, 4),
                        "ops_per_cell": OPS_CUBE, "fill_ms": round(fill_ms, 3),
                        "kernel": "k_cube_fill (fill stage)"}}
    gpu_costs = cost.cpu().numpy()
    b3.close()
    del cost, o, ol, d_blob
    ctx.close()
    torch.cuda.empty_cache()
    if pool is not None:
        stride = 1024
        sample = np.arange(0, n3, stride)
        rec["cpu_baseline"] = cpu_baseline(
            pool, max(1, procs // 2), "cube", (1, 1), blob, [off[0::3], off[1::3], off[2::3]],
            lens, sample, gpu_costs, f"every {stride}th triple")
    out["config5"] = rec
    return out


def serving(n_req):
    """Continuous batching on the server: down pass of a 128-taxon tree
    (align_2 + ancestor medians kept on the GPU), then cost_2 rejoin
    requests from Poisson arrivals at 0.25/0.5/0.9 of the measured
    capacity (linear tcm sub 1 / indel 2)."""
    from poy5_amd import cost_matrix, spr
    from poy5_amd.sequence import Context
    from poy5_amd.serve import Server
    ctx = Context(0)
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 2))
    srv = Server(ctx, n_slots=3, max_batch_pairs=2048, max_len=4096, slot_dm_bytes=2 << 30,
                 pool_bytes=64 << 20)
    try:
        parent, children, leaves = spr.simulate()
        t0 = time.perf_counter()
        mid = spr.downpass(srv, parent, children, leaves)
        down_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        emid = spr.edge_medians(srv, parent, mid)
        edge_s = time.perf_counter() - t0
        res = spr.run_loads(srv, parent, children, mid, emid, n_req=n_req)
        res["workload"] = ("128-taxon Yule tree, root 1500, JC + 2% indels; down pass "
                           "(127 align_2 + ancestor, medians kept in the device pool), edge "
                           "medians (254 align_2 + ancestor), then cost_2(clade median, edge "
                           "median) rejoin requests with deltaw = max(|la-lb|, 8)")
        res["downpass_ms"] = round(down_s * 1e3, 3)
        res["edge_medians_ms"] = round(edge_s * 1e3, 3)
        res["server"] = srv.stats()
        return res
    finally:
        srv.close()
        ctx.close()


def load_traffic():
    p = os.path.join(ROOT, "profiles", "affine_fill_config2_traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X and "
              f"has no multi-GPU path (asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)
    # the CPU baseline's worker processes are forked first, from a process
    # that has touched neither the GPU nor the HIP runtime
    pool, procs = None, host_cores()
    if not args.no_cpu:
        import multiprocessing as mp
        pool = mp.get_context("fork").Pool(procs)
    try:
        if args.only == "serve":
            print(json.dumps(serving(args.serve_requests)), flush=True)
        elif args.only == "rj":
            print(json.dumps(readjust_workload(args, pool, procs)), flush=True)
        elif args.only == "un":
            print(json.dumps(unions_workload(args, pool, procs)), flush=True)
        elif args.only == "nk":
            rec = newkk_workload(args, pool, procs)
            if args.copy_calib_gb > 0:
                rec["copy_calibration_bytes"] = copy_calibration(args.copy_calib_gb)
            print(json.dumps(rec), flush=True)
        elif args.only:
            rec = secondary(args, pool, procs, which=(args.only,))
            if args.copy_calib_gb > 0:
                rec["copy_calibration_bytes"] = copy_calibration(args.copy_calib_gb)
            print(json.dumps(rec), flush=True)
        else:
            run_bench(args, pool, procs)
    finally:
        if pool is not None:
            pool.close()
            pool.join()


def copy_calibration(gb):
    """A device-to-device copy of gb GiB: the scale of the FETCH_SIZE and
    WRITE_SIZE counters against known bytes."""
    import torch
    nb = int(gb * (1 << 30))
    x = torch.ones(nb, dtype=torch.uint8, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    del x, y
    return nb


def run_bench(args, pool, procs):
    import numpy as np
    import torch

    from poy5_amd import cost_matrix, workloads
    from poy5_amd.batch import Batch
    from poy5_amd.sequence import Align, Context

    print(f"[bench] generating config-2 workload ({args.pairs} pairs)", file=sys.stderr, flush=True)
    blob, off, ln = workloads.gen_pairs(2, args.pairs)
    len1 = ln[0::2].astype(np.int32)
    len2 = ln[1::2].astype(np.int32)
    cells_full = float(np.sum(len1.astype(np.float64) * len2))

    ctx = Context(0)
    ctx.set_cm(cost_matrix.of_transformations_and_gaps(1, 1, 3))
    tune = None
    if not args.no_tune:
        # outside the timed region: knobs per length bucket on samples of
        # the batch, the best predicted single setting applied
        from poy5_amd import autotune
        print("[bench] autotuning", file=sys.stderr, flush=True)
        tune = autotune.tune_pairs(ctx, "align_affine", blob, off[0::2], off[1::2], len1, len2)
        print(f"[bench] autotune choice {tune['choice']}", file=sys.stderr, flush=True)
    budget = int(args.dm_budget_gb * (1 << 30))
    b = Batch(ctx, "align_affine", len1, len2, dm_budget=budget)
    dev = torch.device("cuda:0")
    d_blob = torch.from_numpy(blob).to(dev)
    d_off1 = torch.from_numpy(np.ascontiguousarray(off[0::2])).to(dev)
    d_off2 = torch.from_numpy(np.ascontiguousarray(off[1::2])).to(dev)
    cost, out, outlen = b.alloc_outputs(dev)
    torch.cuda.synchronize()

    # first run: per-pair band statistics and the costs every later step
    # must reproduce
    print("[bench] first run", file=sys.stderr, flush=True)
    b.run(d_blob, d_off1, d_off2, cost, out, outlen, sync=True)
    cells_band, slots, rounds, _ = b.stats()
    cells_computed = float(cells_band.sum())
    lane_slots = float(slots.sum())
    gpu_costs = cost.cpu().numpy()
    nchunks = b.chunks()

    if args.graph:
        b.use_graph(True)
    for _ in range(args.warmup):
        b.run(d_blob, d_off1, d_off2, cost, out, outlen, sync=True)
    torch.cuda.synchronize()
    ctx.chk(ctx._lib.poyhip_sync(ctx.h))
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    # one host synchronisation per step (a step lasts hundreds of ms): the
    # fill stage's HIP events of each timed step are read back after it
    fill_ms_steps = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.run(d_blob, d_off1, d_off2, cost, out, outlen, sync=True)
        fill_ms_steps.append(b.fill_ms())
    ctx.chk(ctx._lib.poyhip_sync(ctx.h))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt * 1e3 / args.steps
    gcups = cells_full / (ms_per_step / 1e3) / 1e9
    assert np.array_equal(cost.cpu().numpy(), gpu_costs)

    # dominant kernel: the fill stage (gap-free and general fill kernels,
    # one launch of each per chunk; the gap-free kernel also runs each pair's
    # backtrace after its fill); ops are algorithmic: OPS_PER_CELL per band
    # cell the reference's loops evaluate, backtrace included
    roofline = None
    if not args.graph:
        fill_ms = float(np.mean(fill_ms_steps))
        launch_ms = fill_ms / nchunks
        cells_launch = cells_computed / nchunks
        achieved = cells_launch * OPS_PER_CELL / (launch_ms / 1e3) / 1e12
        traffic = load_traffic()
        roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": VALU_PEAK_TOPS,
                    "unit": "TOP/s", "frac": round(achieved / VALU_PEAK_TOPS, 4),
                    "traffic": traffic.get("bytes_per_launch") if traffic else None,
                    "kernel": "k_affine_fill<true,true> (fill + backtrace) + k_affine_fill<true,false> "
                              "(fill stage)",
                    "ops_per_cell": OPS_PER_CELL, "launches_per_step": nchunks,
                    "cells_per_launch": cells_launch, "avg_launch_ms": round(launch_ms, 3),
                    "timing": "HIP events on the library stream, every timed step"}
        # one direction byte per evaluated cell, plus the sequences
        dm_bytes = cells_launch
        seq_bytes = float(blob.size) / nchunks
        roofline["hbm_secondary"] = {
            "algorithmic_bytes": dm_bytes + seq_bytes,
            "achieved_GBs": round((dm_bytes + seq_bytes) / (launch_ms / 1e3) / 1e9, 1),
            "peak_GBs": HBM_PEAK_GBS}
        if traffic:
            roofline["traffic_source"] = traffic.get("source")

    result = {
        "metric": METRIC, "value": round(gcups, 2), "unit": "GCUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "config2: 65,536 uneven affine DNA pairs (log-uniform 50-5000, "
                               "divergence 2-25%, sub 1 / ext 1 / open 3), align_affine_3 per pair",
                   "pairs": int(args.pairs), "cells_full": cells_full,
                   "cells_computed": cells_computed, "lane_steps": lane_slots,
                   "lane_occupancy": round(cells_computed / max(lane_slots, 1.0), 4),
                   "mean_rounds": float(rounds.mean()),
                   "graph": bool(args.graph), "dm_budget_gb": args.dm_budget_gb,
                   "autotune": tune},
        "roofline": roofline,
    }
    if roofline is None:
        result["roofline_note"] = "graph replay: per-kernel events unavailable, run with --graph 0"
    # the headline batch's direction store is sized to most of the device:
    # release it before the other sections plan their own
    b.close()
    del cost, out, outlen
    torch.cuda.empty_cache()
    if not args.no_plain:
        print("[bench] plain per-call sample", file=sys.stderr, flush=True)
        al = Align(ctx)
        pl = plain_per_call(ctx, al, blob, off, ln, args.plain_sample)
        result["plain_per_call"] = pl
        result["batched_vs_plain"] = round(gcups / pl["gcups"], 2)
        result["p50_ms"] = round(pl["p50_ms"], 3)
        result["p99_ms"] = round(pl["p99_ms"], 3)
    if pool is not None:
        print("[bench] CPU baseline (config 2 sample)", file=sys.stderr, flush=True)
        sample = np.arange(0, args.pairs, args.cpu_stride)
        result["cpu_baseline"] = cpu_baseline(
            pool, procs, "affine", (1, 1, 3), blob, [off[0::2], off[1::2]], [len1, len2], sample,
            gpu_costs, f"every {args.cpu_stride}th pair of the workload")
        result["cpu_baseline"]["host_cpus"] = os.cpu_count()
    if not args.no_extra:
        print("[bench] secondary configurations", file=sys.stderr, flush=True)
        result["secondary"] = secondary(args, pool, procs)
        result["secondary"]["newkk_config2"] = newkk_workload(args, pool, procs)
        result["secondary"]["union_merge_level"] = unions_workload(args, pool, procs)
        result["secondary"]["readjust_level_config5"] = readjust_workload(args, pool, procs)
    if not args.no_serve:
        print("[bench] serving: 128-taxon SPR rejoin traffic", file=sys.stderr, flush=True)
        result["serving"] = serving(args.serve_requests)
    if args.copy_calib_gb > 0:
        result["copy_calibration_bytes"] = copy_calibration(args.copy_calib_gb)
    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

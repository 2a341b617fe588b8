"""Serving workload of a tree search (SPR rejoin traffic) on the server.

A seeded 128-taxon Yule tree (birth rate 1, branch lengths exponential with
mean 0.05 substitutions/site) carries sequences of root length 1,500 evolved
under Jukes-Cantor substitutions plus 2% indel events per branch.  The
client mirrors what POY's DO search asks of ``Sequence.Align``:

* the down pass computes every internal node's median as the ancestor of
  the aligned child medians (``align_2`` + ``ancestor_2``,
  src/seqCS.ml:985-1084), keeping each median in the server's device pool;
* every edge (v, parent v) carries its own median, the ancestor of the
  two node medians (the edge data ``cost_fn`` prices a rejoin against,
  src/allDirChar.ml:2132-2178);
* an SPR round clips a subtree and prices its rejoin at every edge of the
  remaining tree with ``cost_2(clade median, edge median)`` and
  ``DOS.distance``'s band ``deltaw = max(|la - lb|, 8)``
  (src/seqCS.ml:701-744), all on pool-resident medians.

Rejoin requests arrive as Poisson traffic from a native load generator
(``poyhip_server_replay``), at fractions of the measured batched capacity.
"""
from __future__ import annotations

import numpy as np

BASES = np.array([1, 2, 4, 8], np.uint8)


def yule_tree(n_taxa: int, rng):
    """Returns (parent, children, branch) with nodes 0..2n-2, root = 0.
    Pure-birth: repeatedly split a uniformly chosen current leaf."""
    parent = [-1]
    children = [[]]
    leaves = [0]
    while len(leaves) < n_taxa:
        k = int(rng.integers(len(leaves)))
        v = leaves.pop(k)
        for _ in range(2):
            parent.append(v)
            children.append([])
            children[v].append(len(parent) - 1)
            leaves.append(len(parent) - 1)
    branch = rng.exponential(0.05, size=len(parent))
    branch[0] = 0.0
    return np.array(parent), children, branch


def evolve(seq, t, rng, indel=0.02):
    """Jukes-Cantor substitutions for branch length t plus indel events at
    2% of sites (half insertions, half deletions, geometric length mean 3)."""
    p_sub = 0.75 * (1.0 - np.exp(-4.0 * t / 3.0))
    out = []
    k = 0
    n = len(seq)
    while k < n:
        if rng.random() < indel:
            ln = int(rng.geometric(1.0 / 3.0))
            if rng.random() < 0.5:
                k += ln
                continue
            out.extend(int(x) for x in rng.choice(BASES, ln))
        c = int(seq[k])
        if rng.random() < p_sub:
            c = int(rng.choice(BASES))  # JC: uniform over the four states
        out.append(c)
        k += 1
    return np.array(out, np.uint8)


def simulate(n_taxa=128, root_len=1500, seed=0x504F59003):
    """Leaf sequences (leading gap included) of a seeded tree."""
    rng = np.random.default_rng(seed)
    parent, children, branch = yule_tree(n_taxa, rng)
    seqs = [None] * len(parent)
    seqs[0] = rng.choice(BASES, root_len).astype(np.uint8)
    order = [0]
    for v in order:
        for c in children[v]:
            seqs[c] = evolve(seqs[v], branch[c], rng)
            order.append(c)
    leaf = [v for v in range(len(parent)) if not children[v]]
    with_gap = {v: np.concatenate([[16], seqs[v]]).astype(np.uint8) for v in leaf}
    return parent, children, with_gap


def post_order(children, root=0):
    out, stack = [], [(root, False)]
    while stack:
        v, done = stack.pop()
        if done or not children[v]:
            out.append(v)
            continue
        stack.append((v, True))
        for c in children[v]:
            stack.append((c, False))
    return out


def downpass(server, parent, children, leaves):
    """Median of every node in the device pool: {node: pool id}.  Internal
    nodes are submitted level by level from the leaves so each level is
    batched together; a request naming a median still being computed waits
    for it inside the server."""
    mid = {v: server.put(s) for v, s in leaves.items()}
    depth = {}
    for v in post_order(children):
        depth[v] = 0 if not children[v] else 1 + max(depth[c] for c in children[v])
    reqs = []
    for d in range(1, max(depth.values()) + 1):
        for v in [u for u in depth if depth[u] == d]:
            a, b = children[v]
            r, m = server.submit("align_2d", mid[a], mid[b], keep_median=True)
            mid[v] = m
            reqs.append(r)
    for r in reqs:
        server.wait(r)
        server.result(r)
    return mid


def edge_medians(server, parent, mid):
    """Median of every edge (v, parent v), v not the root: {v: pool id}.
    An edge's data in the all-directions tree is the median of its two end
    nodes, computed like a down-pass median (align_2 + ancestor)."""
    emid, reqs = {}, []
    for v in mid:
        if parent[v] < 0:
            continue
        r, m = server.submit("align_2d", mid[v], mid[parent[v]], keep_median=True)
        emid[v] = m
        reqs.append(r)
    for r in reqs:
        server.wait(r)
        server.result(r)
    return emid


def dos_deltaw(la: int, lb: int) -> int:
    """``DOS.distance``'s band for ``cost_2`` (src/seqCS.ml:738-744): the
    length difference, at least 8."""
    d = abs(int(la) - int(lb))
    return d if d > 8 else 8


def subtree(children, v):
    out, st = [], [v]
    while st:
        u = st.pop()
        out.append(u)
        st.extend(children[u])
    return set(out)


def rejoin_requests(parent, children, mid, emid, n_req, rng):
    """(a ids, b ids, clipped nodes, edge nodes): cost_2(clade median, edge
    median) for the edges of the tree left after a clip, clip after clip,
    until n_req requests.  The clipped subtree's own edges, and the edge
    above the clip's parent (which the clip merges with the sibling's edge,
    still priced through the sibling's edge median), are not junctions."""
    nodes = [v for v in mid if parent[v] >= 0]
    a, b, cv, ev = [], [], [], []
    while len(a) < n_req:
        c = int(rng.choice(nodes))
        inside = subtree(children, c)
        for v in nodes:
            if v in inside or v == parent[c]:
                continue
            a.append(mid[c])
            b.append(emid[v])
            cv.append(c)
            ev.append(v)
            if len(a) == n_req:
                break
    return (np.array(a, np.int64), np.array(b, np.int64), np.array(cv, np.int64),
            np.array(ev, np.int64))


def poisson_arrivals(n, rate_per_s, rng):
    """Arrival offsets (us) of a Poisson process."""
    gaps = rng.exponential(1e6 / rate_per_s, size=n)
    return np.cumsum(gaps) - gaps[0]


def rejoin_deltaw(server, a, b):
    """Per-request ``DOS.distance`` band of the pool medians a[k], b[k];
    also returns the pool lengths read."""
    lens = {}
    for x in np.concatenate([a, b]):
        if int(x) not in lens:
            lens[int(x)] = len(server.get(int(x)))
    dw = np.array([dos_deltaw(lens[int(x)], lens[int(y)]) for x, y in zip(a, b)], np.int32)
    return dw, lens


def run_loads(server, parent, children, mid, emid, n_req=4000, loads=(0.25, 0.5, 0.9), seed=7):
    """Capacity (all requests at once) and Poisson runs at fractions of it.
    Returns a dict of capacity, and per load p50/p99 latency and GCUPS."""
    rng = np.random.default_rng(seed)
    a, b, _, _ = rejoin_requests(parent, children, mid, emid, n_req, rng)
    dw, lens = rejoin_deltaw(server, a, b)
    cells_all = float(sum(float(lens[int(x)]) * lens[int(y)] for x, y in zip(a, b)))
    ops = ["cost_2d"] * n_req
    costs0, lat0, _, wall0 = server.replay(np.zeros(n_req), ops, a, b, deltaw=dw)
    cap_rps = n_req / (wall0 / 1e6)
    res = {"requests": n_req, "cells_full_per_request": cells_all / n_req,
           "capacity_rps": cap_rps, "capacity_gcups": cells_all / (wall0 / 1e6) / 1e9,
           "saturated_p50_ms": float(np.percentile(lat0, 50)) / 1e3,
           "saturated_p99_ms": float(np.percentile(lat0, 99)) / 1e3, "loads": []}
    for rho in loads:
        arr = poisson_arrivals(n_req, rho * cap_rps, rng)
        costs, lat, _, wall = server.replay(arr, ops, a, b, deltaw=dw)
        # wall = arrivals + the drain of the last batches (about one latency):
        # the rate served while requests arrive is n over the arrival span
        # when the drain stays near the latency, i.e. the server kept up
        span = float(arr[-1])
        res["loads"].append({
            "load": rho, "offered_rps": rho * cap_rps, "achieved_rps": n_req / (wall / 1e6),
            "arrival_span_ms": span / 1e3, "drain_ms": (wall - span) / 1e3,
            "kept_up": bool((wall - span) / 1e3 < 2.0 * float(np.percentile(lat, 99)) / 1e3),
            "gcups": cells_all / (wall / 1e6) / 1e9,
            "p50_ms": float(np.percentile(lat, 50)) / 1e3,
            "p99_ms": float(np.percentile(lat, 99)) / 1e3,
            "costs_equal_saturated": bool(np.array_equal(costs, costs0))})
    return res

"""``Sequence.readjust`` and ``DOS.readjust`` (src/sequence.ml:2097-2146,
src/seqCS.ml:776-952) on the MI355X library: the pairwise centre search
that re-optimises a node's median from its two children and its parent.

``readjust`` / ``dos_readjust`` mirror the OCaml one node at a time through
the drop-ins (``Align``, ``NewkkAlign``, ``poyhip_closest_2``).
``readjust_level`` runs the same search for many nodes at once: four rounds
of batched alignments (the three pairwise medians, the three
median-against-third costs, the closest sequence of the chosen pair, the
new median against children and parent) and batched column maps, with
every sequence in one device arena; the host only plans the rounds and
makes the comparisons the OCaml makes.  Nodes are dicts with "seq",
"cost2", "cost2_max", "cost3" and "sum_cost", as DOS keeps them.

One table serves as both ``c2_full`` and ``c2_original``: the two differ
only for combination levels above 1, which the device tables do not take.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .batch import Batch
from .cost_matrix import GAP
from .sequence import Align, Context, NewkkAlign


def is_empty(s) -> bool:
    """Sequence.is_empty (sequence.ml:241): every symbol is the gap."""
    s = np.asarray(s, np.uint8)
    return len(s) == 0 or bool(np.all(s == GAP))


def select_one(s):
    """Sequence.select_one (sequence.ml:2069) for sets with combinations:
    the lowest state of each symbol."""
    s = np.asarray(s, np.int64)
    if np.any(s == 0):
        raise ValueError("Nothing?")
    return (s & -s).astype(np.uint8)


def _affine(ctx: Context) -> bool:
    return ctx.cm is not None and ctx.cm.cost_model == 1


def _algn(ctx: Context, s1, s2, use_ukk: bool):
    """The `algn` of Sequence.readjust (sequence.ml:2099-2118): (cost,
    median)."""
    al = Align(ctx)
    if _affine(ctx):
        if use_ukk:
            p1, p2, c = NewkkAlign(ctx).align_2(s1, s2)
            return c, al.median_2(p1, p2)
        m, _, _, c, _ = al.align_affine_3(s1, s2)
        return c, m
    if use_ukk:
        raise ValueError("NewkkAlign under the linear model is not built (its fill is defective "
                         "in the reference)")
    p1, p2, c = al.align_2(s1, s2)
    return c, al.median_2(p1, p2)


def make_center(ctx: Context, a, b, c, use_ukk: bool = False):
    """make_center (sequence.ml:2120-2136): the closest sequence to one of
    the three from the median of the other two, the pair whose median is
    cheapest to reach the third."""
    cab, ab = _algn(ctx, a, b, use_ukk)
    cbc, bc = _algn(ctx, b, c, use_ukk)
    cac, ac = _algn(ctx, a, c, use_ukk)
    cabc = _algn(ctx, ab, c, use_ukk)[0] + cab
    cbca = _algn(ctx, bc, a, use_ukk)[0] + cbc
    cacb = _algn(ctx, ac, b, use_ukk)[0] + cac
    al = Align(ctx)
    if cabc <= cbca:
        if cabc <= cacb:
            return al.closest(c, ab)
        return al.closest(b, ac)
    if cbca < cacb:
        return al.closest(a, bc)
    return al.closest(b, ac)


def readjust(ctx: Context, a, b, m, parent, use_ukk: bool = False):
    """Sequence.readjust (sequence.ml:2097-2146): (cost3, cost2, new median,
    aliedmp, aliedmp, aliedmp) -- the aligned new median against the parent
    three times, as the reference returns it."""
    newseqm, _ = make_center(ctx, a, b, parent, use_ukk)
    al = Align(ctx)
    c1 = al.align_2(a, newseqm)[2]
    c2 = al.align_2(b, newseqm)[2]
    _, aliedmp, cp = al.align_2(parent, newseqm)
    cost2 = c1 + c2
    return cost2 + cp, cost2, newseqm, aliedmp, aliedmp, aliedmp


def dos_distance(ctx: Context, a, b, use_ukk: bool, missing: int = 0) -> int:
    """DOS.distance (seqCS.ml:701-775): band deltaw = max(|la - lb|, 8)."""
    if is_empty(a) or is_empty(b):
        return missing
    deltaw = max(abs(len(a) - len(b)), 8)
    if use_ukk:
        return NewkkAlign(ctx).cost_2(a, b)
    return Align(ctx).cost_2(a, b, deltaw)


def _two_child(ctx: Context, s1, s2, use_ukk: bool):
    """readjust_algn_two_child (seqCS.ml:783-813): (median, cost)."""
    c, med = _algn(ctx, s1, s2, use_ukk)
    return select_one(med), c


def _node(seq, aligned, c2, c2max, c3, s):
    return {"seq": seq, "aligned": aligned, "cost2": float(c2), "cost2_max": float(c2max),
            "cost3": float(c3), "sum_cost": float(s)}


def dos_readjust(ctx: Context, ch1, ch2, parent, mine, use_ukk: bool = False, _res=None):
    """DOS.readjust (seqCS.ml:820-952), ApproxD mode: (changed, node, cost3,
    cost2, sum_cost)."""
    e1, e2, ep = is_empty(ch1["seq"]), is_empty(ch2["seq"]), is_empty(parent["seq"])
    mc2, msum, mc3 = int(mine["cost2"]), int(mine["sum_cost"]), int(mine["cost3"])
    sum12 = int(ch1["sum_cost"] + ch2["sum_cost"])

    def differs(seq):
        return not (len(seq) == len(mine["seq"]) and np.array_equal(seq, mine["seq"]))

    if not e1 and not e2 and not ep:
        c3, c2, m, al1, al2, mwg = (_res if _res is not None else
                                    readjust(ctx, ch1["seq"], ch2["seq"], mine["seq"],
                                             parent["seq"], use_ukk))
        maxc2 = 0 if is_empty(al1) or is_empty(al2) else Align(ctx).max_cost_2(al1, al2)
        s = c2 + sum12
        changed = s != msum or c3 != mc3 or c2 != mc2 or differs(m)
        return changed, _node(m, (al1, al2, mwg), c2, maxc2, c3, s), c3, c2, s
    if (e1 and e2) or (e1 and ep) or (e2 and ep):
        # the or-pattern binds the first alternative that matches
        r = parent if (e1 and e2) else (ch2 if (e1 and ep) else ch1)
        changed = sum12 != msum or 0 != mc3 or 0 != mc2 or differs(r["seq"])
        out = dict(r)
        out["sum_cost"] = float(sum12)
        return changed, out, 0, 0, sum12
    if ep:
        med, c2 = _two_child(ctx, ch1["seq"], ch2["seq"], use_ukk)
        c3 = c2
    else:
        ch = ch1 if e2 else ch2
        med, _ = _two_child(ctx, ch["seq"], parent["seq"], use_ukk)
        c2 = dos_distance(ctx, ch["seq"], med, use_ukk)
        c3 = c2 + dos_distance(ctx, med, parent["seq"], use_ukk)
    s = c2 + sum12
    changed = s != msum or c3 != mc3 or c2 != mc2 or differs(med)
    return changed, _node(med, (med, med, med), c2, 0, c3, s), c3, c2, s


# ---- many nodes at once ---------------------------------------------------

POST_JOB = np.dtype([("a", "<i8"), ("b", "<i8"), ("out", "<i8"), ("len", "<i4"),
                     ("cap", "<i4")])


class _Arena:
    """One device byte blob that holds a level's inputs and every round's
    outputs (bump allocation, 16-byte aligned), so each batch reads and
    writes it through offsets."""

    def __init__(self, nbytes: int, device):
        self.t = torch.zeros(max(nbytes, 16), dtype=torch.uint8, device=device)
        self.top = 0

    def take(self, n: int) -> int:
        at = (self.top + 15) & ~15
        if at + n > self.t.numel():
            raise MemoryError(f"readjust arena: {at + n} > {self.t.numel()} bytes")
        self.top = at + n
        return at

    def put_all(self, hosts):
        """Places host sequences, each at a 16-byte boundary: one transfer of
        the bytes back to back, spread to their places on the device;
        returns [(at, len), ...]."""
        base = self.take(0)
        ln = np.fromiter((len(h) for h in hosts), np.int64, len(hosts))
        size = (ln + 15) & ~15
        rel = np.zeros(len(hosts), np.int64)
        starts = np.zeros(len(hosts), np.int64)
        if len(hosts) > 1:
            np.cumsum(size[:-1], out=rel[1:])
            np.cumsum(ln[:-1], out=starts[1:])
        o = int(size.sum())
        self.take(o)
        total = int(ln.sum())
        if total:
            dev = self.t.device
            flat = torch.from_numpy(np.concatenate(hosts).astype(np.uint8, copy=False)).to(dev)
            shift = torch.from_numpy(base + rel - starts).to(dev)
            idx = torch.repeat_interleave(shift, torch.from_numpy(ln).to(dev), output_size=total)
            idx += torch.arange(total, dtype=torch.int64, device=dev)
            self.t[idx] = flat
        return list(zip((base + rel).tolist(), ln.tolist()))

    def fetch(self, at, ln):
        """Host copies of the regions (offsets at, lengths ln): the bytes are
        gathered on the device into one buffer and copied over at once;
        returns (flat bytes, start of each region in them)."""
        at = np.asarray(at, np.int64)
        ln = np.asarray(ln, np.int64)
        starts = np.zeros(len(ln), np.int64)
        if len(ln) > 1:
            np.cumsum(ln[:-1], out=starts[1:])
        total = int(ln.sum())
        if total == 0:
            return np.zeros(0, np.uint8), starts
        dev = self.t.device
        base = torch.from_numpy(at - starts).to(dev)
        idx = torch.repeat_interleave(base, torch.from_numpy(ln).to(dev), output_size=total)
        idx += torch.arange(total, dtype=torch.int64, device=dev)
        return self.t[idx].cpu().numpy(), starts


class _Round:
    """One batch of pairs over arena sequences: the two sides as (offset,
    length) arrays in; costs and the output regions out."""

    def __init__(self, ctx: Context, arena: _Arena, op: str, at1, l1, at2, l2):
        dev = arena.t.device
        self.b = Batch(ctx, op, np.asarray(l1, np.int32), np.asarray(l2, np.int32))
        self.n = len(l1)
        self.out_at = arena.take(self.b.out_bytes)
        d1 = torch.from_numpy(np.ascontiguousarray(at1, np.int64)).to(dev)
        d2 = torch.from_numpy(np.ascontiguousarray(at2, np.int64)).to(dev)
        cost = torch.empty(self.n, dtype=torch.int32, device=dev)
        outlen = torch.empty(4 * self.n, dtype=torch.int32, device=dev)
        out = arena.t[self.out_at:self.out_at + max(self.b.out_bytes, 1)]
        self.b.run(arena.t, d1, d2, cost, out, outlen, sync=True)
        self.costs = cost.cpu().numpy().astype(np.int64)
        self.lens = outlen.cpu().numpy().reshape(-1, 4).astype(np.int64)
        self.off = np.zeros(self.n, np.int64)
        cap = np.zeros(self.n, np.int32)
        ctx.chk(self.b.L.poyhip_batch_out_layout(self.b.h,
                                                 self.off.ctypes.data_as(C.POINTER(C.c_int64)),
                                                 cap.ctypes.data_as(C.POINTER(C.c_int))))
        self.cap = cap.astype(np.int64)
        self.b.close()

    def region(self, r: int):
        """(offsets, lengths) of output region r of every pair."""
        n = self.lens[:, r]
        return self.out_at + self.off + r * self.cap + self.cap - n, n


def _post(ctx: Context, arena: _Arena, op: int, at1, at2, ln):
    """Column maps of aligned pairs (two offset arrays, one length array) into
    new arena regions; returns (offsets, lengths)."""
    n = len(ln)
    if n == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    ln = np.asarray(ln, np.int64)
    cap = ln + 1
    size = (cap + 15) & ~15
    out = arena.take(int(size.sum())) + np.concatenate([[0], np.cumsum(size)[:-1]])
    jobs = np.zeros(n, POST_JOB)
    jobs["a"], jobs["b"], jobs["out"], jobs["len"], jobs["cap"] = at1, at2, out, ln, cap
    dev = arena.t.device
    dj = torch.from_numpy(jobs.view(np.uint8).copy()).to(dev)
    dl = torch.zeros(n, dtype=torch.int32, device=dev)
    ds = torch.zeros(n, dtype=torch.int32, device=dev)
    ctx.chk(ctx._lib.poyhip_pairs_post_batch(
        ctx.h, op, n, C.c_void_p(dj.data_ptr()), C.c_void_p(arena.t.data_ptr()),
        arena.t.numel(), C.c_void_p(dl.data_ptr()), C.c_void_p(ds.data_ptr()), 1))
    got, st = dl.cpu().numpy().astype(np.int64), ds.cpu().numpy()
    if np.any(st != _lib.POYHIP_OK):
        raise ValueError("No bits on?" if op in (_lib.POST_CLOSEST, _lib.POST_CLOSEST_EQ)
                         else "column map failed")
    return out + cap - got, got


def readjust_level(ctx: Context, triples, use_ukk: bool = False, device="cuda"):
    """Sequence.readjust for many (a, b, parent) triples: a list of (cost3,
    cost2, new median, aliedmp) as ``readjust`` gives them."""
    n = len(triples)
    if n == 0:
        return []
    affine = _affine(ctx)
    if use_ukk and not affine:
        raise ValueError("NewkkAlign under the linear model is not built (its fill is defective "
                         "in the reference)")
    align_op = "align_affine" if affine else "align_2d"
    algn_op = "newkk_align" if use_ukk else align_op
    seqs = [[np.asarray(x, np.uint8) for x in t] for t in triples]
    tot = sum(len(a) + len(b) + len(c) for a, b, c in seqs)
    # every round's outputs are bounded by a multiple of the node's three
    # lengths (aligned pairs of at most twice the input, four regions each)
    arena = _Arena(64 * tot + 1024 * n + 4096, device)
    flat = arena.put_all([x for t in seqs for x in t])
    at = np.array([o for o, _ in flat], np.int64).reshape(n, 3)   # a, b, parent
    ln = np.array([q for _, q in flat], np.int64).reshape(n, 3)

    def algn_round(at1, l1, at2, l2, medians=True):
        """Costs and median regions of the pairs, as `algn` computes them."""
        r = _Round(ctx, arena, algn_op, at1, l1, at2, l2)
        if not medians:
            return r.costs, None
        if affine and not use_ukk:
            return r.costs, r.region(2)
        (x, n1), (y, _) = r.region(0), r.region(1)
        return r.costs, _post(ctx, arena, _lib.POST_MEDIAN_2, x, y, n1)

    def inter(*cols):
        """Interleaves per-node columns into per-pair order (node-major)."""
        return np.stack(cols, axis=1).reshape(-1)

    # first batch: the pairwise medians (a, b), (b, c), (a, c) of every node
    c1, (m_at, m_ln) = algn_round(inter(at[:, 0], at[:, 1], at[:, 0]),
                                  inter(ln[:, 0], ln[:, 1], ln[:, 0]),
                                  inter(at[:, 1], at[:, 2], at[:, 2]),
                                  inter(ln[:, 1], ln[:, 2], ln[:, 2]))
    m_at, m_ln, c1 = m_at.reshape(n, 3), m_ln.reshape(n, 3), c1.reshape(n, 3)
    # second batch: each median against the third sequence: ab-c, bc-a, ac-b
    third = np.array([2, 0, 1])
    c2, _ = algn_round(m_at.reshape(-1), m_ln.reshape(-1), at[:, third].reshape(-1),
                       ln[:, third].reshape(-1), medians=False)
    tot3 = c2.reshape(n, 3) + c1          # cabc, cbca, cacb
    cabc, cbca, cacb = tot3[:, 0], tot3[:, 1], tot3[:, 2]
    # make_center's comparisons: which median (0 ab, 1 bc, 2 ac) against
    # which sequence (2 c, 0 a, 1 b)
    pick = np.where(cabc <= cbca, np.where(cabc <= cacb, 0, 2), np.where(cbca < cacb, 1, 2))
    s1_at, s1_ln = at[np.arange(n), third[pick]], ln[np.arange(n), third[pick]]
    s2_at, s2_ln = m_at[np.arange(n), pick], m_ln[np.arange(n), pick]
    # Align.closest: an all-gap median is kept, equal sequences take the
    # masked map, the others their alignment's map
    flat, st = arena.fetch(s2_at, s2_ln)
    # symbols other than the gap per median
    run = np.concatenate([[0], np.cumsum(flat != GAP)])
    empty = run[st + s2_ln] - run[st] == 0
    new_at, new_ln = s2_at.copy(), s2_ln.copy()
    cand = np.nonzero(~empty)[0]
    # only a chosen third sequence of the median's length can equal it
    same = ln[cand, third[pick[cand]]] == s2_ln[cand]
    eq = np.array([k for k in cand[same]
                   if np.array_equal(seqs[k][third[pick[k]]], flat[st[k]:st[k] + s2_ln[k]])],
                  np.int64)
    al = np.setdiff1d(cand, eq)
    if len(eq):
        new_at[eq], new_ln[eq] = _post(ctx, arena, _lib.POST_CLOSEST_EQ, s1_at[eq], s2_at[eq],
                                       s1_ln[eq])
    if len(al):
        r3 = _Round(ctx, arena, align_op, s1_at[al], s1_ln[al], s2_at[al], s2_ln[al])
        (x, n1), (y, _) = r3.region(0), r3.region(1)
        new_at[al], new_ln[al] = _post(ctx, arena, _lib.POST_CLOSEST, x, y, n1)
    # last batch: a, b and the parent against the new median
    r4 = _Round(ctx, arena, align_op, at.reshape(-1), ln.reshape(-1), np.repeat(new_at, 3),
                np.repeat(new_ln, 3))
    c4 = r4.costs.reshape(n, 3)
    p_at, p_ln = r4.region(1)
    fm, sm = arena.fetch(new_at, new_ln)
    pl = p_ln[2::3]
    fp, sp = arena.fetch(p_at[2::3], pl)
    out = []
    for k in range(n):
        cost2 = int(c4[k, 0] + c4[k, 1])
        out.append((cost2 + int(c4[k, 2]), cost2, fm[sm[k]:sm[k] + new_ln[k]].copy(),
                    fp[sp[k]:sp[k] + pl[k]].copy()))
    return out


def dos_readjust_level(ctx: Context, nodes, use_ukk: bool = False, device="cuda"):
    """DOS.readjust for many (ch1, ch2, parent, mine) node tuples: nodes whose
    three neighbours are all non-empty share one readjust_level; the others
    take their branches one by one."""
    full = [k for k, (c1, c2, p, _) in enumerate(nodes)
            if not (is_empty(c1["seq"]) or is_empty(c2["seq"]) or is_empty(p["seq"]))]
    res = readjust_level(ctx, [(nodes[k][0]["seq"], nodes[k][1]["seq"], nodes[k][2]["seq"])
                               for k in full], use_ukk, device)
    by_node = dict(zip(full, res))
    out = []
    for k, (c1, c2, p, m) in enumerate(nodes):
        r = by_node.get(k)
        pre = None if r is None else (r[0], r[1], r[2], r[3], r[3], r[3])
        out.append(dos_readjust(ctx, c1, c2, p, m, use_ukk, _res=pre))
    return out

"""Brute-force model of POY's four-state affine DP (tiny inputs only).

The value of each state at each cell -- H (horizontal gap run), V (vertical
gap run), E (gap-gap block) and C (aligned column) -- is computed as the
minimum over every state-labelled path from the boundary, each path priced
move by move.  No row buffers, no recurrence loop: a depth-first enumeration
of all paths.  The move costs are read from the reference's cell functions:

  H <- H  shx[j]                     H <- C  gop[j] + ge_j[j]
        (FILL_EXTEND_HORIZONTAL, src/algn.c:1283-1304)
  V <- V  vext[i]                    V <- C  go_i[i] + ge_i[i]
        (FILL_EXTEND_VERTICAL, src/algn.c:1327-1348)
  E <- E  0 if both symbols carry the gap bit, else HIGH
  E <- C  0 / 2*go / HIGH by the gap bits of both symbols and of their
          predecessors (FILL_EXTEND_BLOCK_DIAGONAL, src/algn.c:1376-1399)
  C <- C  d     C <- V  d (+gop[j] when s_i carries the gap bit)
  C <- H  d (+go_i[i] when s_j carries it)     C <- E  d + max(gop[j], go_i[i])
        (FILL_CLOSE_BLOCK_DIAGONAL, src/algn.c:1436-1492)

with d = cost(s_i & 15, s_j & 15), the gap opening rule HAS_GAP_OPENING
(src/algn.c:1240-1253) and the extension vectors of algn_newkk_test_aff
(src/algn.c:2186-2260) / algn_fill_plane_3_aff_nobt (src/algn.c:1987-2031).
HIGH (src/algn.c:37) is an ordinary number here, as in the reference, so a
min over paths equals the min-plus recurrence exactly.

Two boundaries:
  * "full": initialize_matrices_affine_nobt + algn_fill_plane_3_aff_nobt
    (src/algn.c:1822-1864, 1987-2110), the cost-only entry point, with the
    effect of its row buffers overlapping: algn_CAML_cost_affine_3 carves
    the four state arrays 2*largest ints apart (src/algn.c:2488-2491) but
    fills them with a row offset of lenj+2 = largest+1, so the second row's
    last entry of each array is the first entry of the next array.  Stated
    as rules (derived from that layout, not emulated):
      - in every even row i >= 2, the start-of-row writes of H[0] and E[0]
        land on V(i-1, lenj) and C(i-1, lenj), so V(i, lenj) has only
        HIGH sources: V(i, lenj) = HIGH + min(vext, go_i + ge_i);
      - when lenj = 1, in every odd row the cell's own V(i, 1) and E(i, 1)
        land on H(i-1, 0) and V(i-1, 0) before C(i, 1) reads them, so C's
        "from H" and "from V" candidates come from V(i, 1) and E(i, 1);
  * "band": the first round of algn_newkk_test_aff at half-width newk
    (row 0 rewritten at src/algn.c:2226-2247, band rows 2251-2256, forced
    HIGH at the left border and at the right border, src/algn.c:2128-2150).

From the state values the direction flags follow declaratively (strict '<'
keeps BEGIN_*, ties give END_*; ALIGN_TO_* and DO_* are every tied argmin;
HORIZONTAL_EQ_VERTICAL when the cell minimum equals both H and V,
src/algn.c:1936-1981), and `walk` restates backtrace_aff's mode machine
(src/algn.c:1531-1819) over them.
"""
from __future__ import annotations

HIGH = 1_000_000
GAP = 16

# direction flags (src/algn.c:1185-1199)
ALIGN_TO_ALIGN, ALIGN_TO_VERTICAL, ALIGN_TO_HORIZONTAL, ALIGN_TO_DIAGONAL = 1, 2, 4, 8
BEGIN_BLOCK, END_BLOCK, BEGIN_VERTICAL, END_VERTICAL = 16, 32, 64, 128
BEGIN_HORIZONTAL, END_HORIZONTAL = 256, 512
DO_ALIGN, DO_VERTICAL, DO_HORIZONTAL, HORIZONTAL_EQ_VERTICAL, DO_DIAGONAL = (
    1024, 2048, 4096, 8192, 16384)

H, V, E, C = "H", "V", "E", "C"


def has_go(idx, prev, curr, go):
    if idx == 1 and (curr & GAP):
        return 0
    if idx > 1 and not (prev & GAP) and (curr & GAP):
        return 0
    return go


class Model:
    """si (rows) must not be longer than sj; cost/prepend/median are the
    flat 5-bit tables of a cost model (index (a << 5) + b)."""

    def __init__(self, si, sj, cost, prepend, median, go, mode, newk=None, overlap=True):
        self.si = [int(x) for x in si]
        self.sj = [int(x) for x in sj]
        self.cost, self.prepend, self.med, self.go = cost, prepend, median, go
        self.mode = mode
        self.overlap = overlap and mode == "full"
        self.leni, self.lenj = len(self.si) - 1, len(self.sj) - 1
        li, lj = self.leni, self.lenj
        si_, sj_ = self.si, self.sj
        self.ge_j = [prepend[sj_[j]] for j in range(lj + 1)]
        self.gop = [0] + [has_go(j, sj_[j - 1], sj_[j], go) for j in range(1, lj + 1)]

        def gap_close(prev, curr):
            return (prev & GAP) and not (curr & GAP)

        # row-0 horizontal increments of the band round (including j = 1)
        self.shx0 = [0] + [self.gop[j] + self.ge_j[j] if gap_close(sj_[j - 1], sj_[j])
                           else self.ge_j[j] for j in range(1, lj + 1)]
        # interior increments: position 1 is reset to the plain extension
        self.shx = list(self.shx0)
        if lj >= 1:
            self.shx[1] = self.ge_j[1]
        self.ge_i = [cost[(a << 5) + GAP] for a in si_]
        self.go_i = [0] + [has_go(i, si_[i - 1], si_[i], go) for i in range(1, li + 1)]
        self.vext = [0] + [self.go_i[i] + self.ge_i[i] if i > 1 and gap_close(si_[i - 1], si_[i])
                           else self.ge_i[i] for i in range(1, li + 1)]
        if mode == "band":
            self.newk = newk
        self._paths()

    # ---- geometry -----------------------------------------------------------
    def startj(self, i):
        return 0 if self.mode == "full" else max(i - self.newk, 0)

    def endj(self, i):
        if self.mode == "full":
            return self.lenj
        e = i + self.lenj - self.leni + self.newk
        return e if e <= self.lenj - 1 else self.lenj

    def in_band(self, i, j):
        return i == 0 or self.startj(i) <= j <= self.endj(i)

    def left(self, i, j):
        return self.mode == "band" and j == self.startj(i)

    def right(self, i, j):
        return self.mode == "band" and (j - i) - (self.lenj - self.leni + 1) + 1 == self.newk

    # ---- boundary values and moves -------------------------------------------
    def given(self):
        """(cell, state) -> value for states that no move reaches."""
        g = {}
        go, lj = self.go, self.lenj
        if self.mode == "full":
            g[(0, 0, C)], g[(0, 0, E)], g[(0, 0, H)], g[(0, 0, V)] = 0, HIGH, go, go
            run = go
            for j in range(1, lj + 1):
                run += self.ge_j[j]
                g[(0, j, H)] = run
                g[(0, j, C)] = g[(0, j, E)] = g[(0, j, V)] = HIGH
            for i in range(1, self.leni + 1):
                g[(i, 0, H)] = g[(i, 0, C)] = g[(i, 0, E)] = HIGH
                if self.overlap and i % 2 == 0 and lj >= 1:
                    g[(i, lj, V)] = HIGH + min(self.vext[i], self.go_i[i] + self.ge_i[i])
        else:
            g[(0, 0, C)], g[(0, 0, V)], g[(0, 0, H)], g[(0, 0, E)] = 0, HIGH, go, HIGH
            run = 0
            for j in range(1, lj + 1):
                run += self.shx0[j]
                g[(0, j, C)] = g[(0, j, H)] = run
                g[(0, j, V)] = g[(0, j, E)] = HIGH
            for i in range(1, self.leni + 1):
                for j in range(self.startj(i), self.endj(i) + 1):
                    if self.left(i, j):
                        g[(i, j, H)] = HIGH
                    if self.right(i, j):
                        g[(i, j, V)] = HIGH
                    if j == 0:
                        g[(i, j, C)] = g[(i, j, E)] = HIGH
        return g

    def moves_into(self, i, j):
        """[(target state, source (i, j, state), cost)] for cell (i, j), i >= 1."""
        out = []
        si, sj, go = self.si, self.sj, self.go
        ic, jc = si[i], sj[j]
        if self.mode == "full" and j == 0:
            return [(V, (i - 1, 0, V), self.vext[i])]
        if not self.left(i, j):
            out.append((H, (i, j - 1, H), self.shx[j]))
            out.append((H, (i, j - 1, C), self.gop[j] + self.ge_j[j]))
        if not self.right(i, j):
            out.append((V, (i - 1, j, V), self.vext[i]))
            out.append((V, (i - 1, j, C), self.go_i[i] + self.ge_i[i]))
        if j >= 1:
            ip = si[i - 1]
            jp = sj[j] if (self.mode == "band" and j == self.startj(i)) else sj[j - 1]
            both = bool(ic & GAP) and bool(jc & GAP)
            clean = not (ip & GAP) and not (jp & GAP)
            out.append((E, (i - 1, j - 1, E), 0 if both else HIGH))
            out.append((E, (i - 1, j - 1, C), (0 if clean else 2 * go) if both else HIGH))
            d = self.cost[((ic & 15) << 5) + (jc & 15)]
            out.append((C, (i - 1, j - 1, C), d))
            fv_src, fh_src = (i - 1, j - 1, V), (i - 1, j - 1, H)
            if self.overlap and self.lenj == 1 and i % 2 == 1:
                fv_src, fh_src = (i, j, E), (i, j, V)
            out.append((C, fv_src, d + (self.gop[j] if ic & GAP else 0)))
            out.append((C, fh_src, d + (self.go_i[i] if jc & GAP else 0)))
            out.append((C, (i - 1, j - 1, E), d + max(self.gop[j], self.go_i[i])))
        return out

    # ---- enumeration -----------------------------------------------------------
    def _paths(self):
        given = self.given()
        succ = {}
        for i in range(1, self.leni + 1):
            for j in range(self.startj(i), self.endj(i) + 1):
                for st, src, w in self.moves_into(i, j):
                    if (i, j, st) in given:
                        continue  # forced value: no move reaches it
                    if not self.in_band(src[0], src[1]):
                        raise AssertionError(f"move from outside the band {src} -> {(i, j)}")
                    succ.setdefault(src, []).append(((i, j, st), w))
        best = dict(given)

        def dfs(node, acc):
            for nxt, w in succ.get(node, ()):
                v = acc + w
                if nxt not in best or v < best[nxt]:
                    best[nxt] = v
                dfs(nxt, v)
        for node, v in given.items():
            dfs(node, v)
        self.val = best

    # ---- results -----------------------------------------------------------------
    def v(self, i, j, s):
        return self.val[(i, j, s)]

    def fc(self, i, j):
        return min(self.v(i, j, s) for s in (H, V, E, C))

    def cost_at_end(self):
        if self.leni == 0:
            # no DP row: the initialisation's cell minimum (H of row 0)
            return self.go + sum(self.ge_j[1:]) if self.lenj > 0 else 0
        return self.fc(self.leni, self.lenj)

    def flags(self, i, j):
        assert i >= 1 and self.in_band(i, j), (i, j)
        f = 0
        if not self.left(i, j):
            ext = self.v(i, j - 1, H) + self.shx[j]
            opn = self.v(i, j - 1, C) + self.gop[j] + self.ge_j[j]
            f |= BEGIN_HORIZONTAL if ext < opn else END_HORIZONTAL
        if not self.right(i, j):
            ext = self.v(i - 1, j, V) + self.vext[i]
            opn = self.v(i - 1, j, C) + self.go_i[i] + self.ge_i[i]
            f |= BEGIN_VERTICAL if ext < opn else END_VERTICAL
        if j >= 1:
            src = {s: self.v(i - 1, j - 1, s) for s in (H, V, E, C)}
            cands = {}
            for st, (a, b, s), w in self.moves_into(i, j):
                if (a, b) == (i - 1, j - 1):
                    cands.setdefault(st, []).append((s, src[s] + w))
            e_ext = dict(cands[E])[E]
            e_opn = dict(cands[E])[C]
            f |= BEGIN_BLOCK if e_ext < e_opn else END_BLOCK
            cb = dict(cands[C])
            m = min(cb.values())
            for s, bit in ((C, ALIGN_TO_ALIGN), (V, ALIGN_TO_VERTICAL), (H, ALIGN_TO_HORIZONTAL),
                           (E, ALIGN_TO_DIAGONAL)):
                if cb[s] == m:
                    f |= bit
        vals = {s: self.v(i, j, s) for s in (H, V, E, C)}
        m = min(vals.values())
        for s, bit in ((H, DO_HORIZONTAL), (V, DO_VERTICAL), (E, DO_DIAGONAL), (C, DO_ALIGN)):
            if vals[s] == m:
                f |= bit
        if m == vals[H] == vals[V]:
            f |= HORIZONTAL_EQ_VERTICAL
        return f

    def walk(self, swaped):
        """backtrace_aff over the brute-force flags: (median, medianwg, resi,
        resj), each left to right."""
        si, sj = self.si, self.sj
        med, mwg, ri, rj = [], [], [], []

        def choose(diag, algn, vert, hrzn):
            order = (("V", vert), ("H", hrzn)) if not swaped else (("H", hrzn), ("V", vert))
            for name, flag in order + (("D", diag),):
                if flag:
                    return name
            return "A"
        i, j = self.leni, self.lenj
        mode = "T"
        while i != 0 and j != 0:
            d = self.flags(i, j)
            ic, jc = si[i], sj[j]
            if mode == "T":
                mode = choose(d & DO_DIAGONAL, d & DO_ALIGN, d & DO_VERTICAL, d & DO_HORIZONTAL)
            elif mode == "V":
                if d & (END_VERTICAL | HORIZONTAL_EQ_VERTICAL):
                    mode = "T"
                if not ic & GAP:
                    med.append(ic | GAP)
                    mwg.append(ic | GAP)
                else:
                    mwg.append(GAP)
                ri.append(ic)
                rj.append(GAP)
                i -= 1
            elif mode == "H":
                if d & (END_HORIZONTAL | HORIZONTAL_EQ_VERTICAL):
                    mode = "T"
                if not jc & GAP:
                    med.append(jc | GAP)
                    mwg.append(jc | GAP)
                else:
                    mwg.append(GAP)
                ri.append(GAP)
                rj.append(jc)
                j -= 1
            elif mode == "D":
                if d & END_BLOCK:
                    mode = "T"
                ri.append(ic)
                rj.append(jc)
                mwg.append(GAP)
                i -= 1
                j -= 1
            else:
                mode = choose(d & ALIGN_TO_DIAGONAL, d & ALIGN_TO_ALIGN, d & ALIGN_TO_VERTICAL,
                              d & ALIGN_TO_HORIZONTAL)
                p = self.med[((ic & 15) << 5) + (jc & 15)]
                med.append(p)
                mwg.append(p)
                ri.append(ic)
                rj.append(jc)
                i -= 1
                j -= 1
        while i != 0:
            ic = si[i]
            if not ic & GAP:
                med.append(ic | GAP)
                mwg.append(ic | GAP)
            else:
                mwg.append(GAP)
            ri.append(ic)
            rj.append(GAP)
            i -= 1
        while j != 0:
            jc = sj[j]
            if not jc & GAP:
                med.append(jc | GAP)
                mwg.append(jc | GAP)
            else:
                mwg.append(GAP)
            ri.append(GAP)
            rj.append(jc)
            j -= 1
        ri.append(GAP)
        rj.append(GAP)
        mwg.append(GAP)
        med = med[::-1]
        if not med or med[0] != GAP:
            med = [GAP] + med
        return med, mwg[::-1], ri[::-1], rj[::-1]

"""Brute-force model of POY's three-way cube fill (tiny inputs only).

algn_fill_cube (src/algn.c:2593-2763) walks the cube plane by plane (s1),
row by row (s2), cell by cell (s3), reading each cell's predecessors through
three row pointers.  At plane i >= 1 the per-plane row loop advances those
pointers L2 - 1 times while the plane itself advances L2 rows, so they drift:
in linear row numbering rho = i * L2 + r the predecessor rows of row rho are

    upper = (i-1)(L2-1) + r       (geometrically (i-1, r))
    prev  = i (L2-1) + r - 1      (geometrically (i, r-1))
    diag  = (i-1)(L2-1) + r - 1   (geometrically (i-1, r-1))

and row 0 of plane i reads row (i-1)(L2-1).  This module states that as a
graph of cells with one priced edge per move and takes, for every cell, the
minimum over every path from the origin (an exhaustive depth-first
enumeration, no recurrence loop).  Move costs use the 3D table with the
leading-gap entries s[0] standing for gaps, as the reference does:

    P3 upper[k]   + c(s1_i, -, s3_0)    S3 upper[k-1] + c(s1_i, -, s3_k)
    P1 prev[k]    + c(-, s2_r, s3_0)    S1 prev[k-1]  + c(-, s2_r, s3_k)
    P2 diag[k]    + c(s1_i, s2_r, s3_0) S2 diag[k-1]  + c(s1_i, s2_r, s3_k)
    SS cur[k-1]   + c(-, -, s3_k)

(fill_parallel 2552, fill_moved 2524, the SS pass 2739-2755; plane 0 and
row 0 of each plane use the subsets at 2622-2704).  A later candidate
replaces the running minimum only when strictly smaller, so a cell's
direction is the first minimising move in the order P3, P1, P2, S3, S1, S2,
SS.  `walk` restates backtrack_3d (src/algn.c:3345-3421), which moves
geometrically through the direction cube.
"""
from __future__ import annotations

GAP = 16
P1, P2, P3, S1, S2, S3, SS = 1, 2, 4, 8, 16, 32, 64  # src/matrices.h:34-40


class Cube:
    def __init__(self, s1, s2, s3, cost3):
        self.s1 = [int(x) for x in s1]
        self.s2 = [int(x) for x in s2]
        self.s3 = [int(x) for x in s3]
        self.c = cost3
        self.L1, self.L2, self.L3 = len(self.s1), len(self.s2), len(self.s3)
        self._enumerate()

    def cost3(self, a, b, c):
        return self.c[(((a << 5) + b) << 5) + c]

    def preds(self, i, r, k):
        """[(move, (row rho, k'), cost)] in the reference's candidate order."""
        L2, s1, s2, s3 = self.L2, self.s1, self.s2, self.s3
        g = GAP
        out = []
        if i == 0:
            if r == 0:
                if k >= 1:
                    out.append((SS, (0, k - 1), self.cost3(g, g, s3[k])))
                return out
            prv = r - 1
            out.append((P1, (prv, k), self.cost3(g, s2[r], s3[0])))
            if k >= 1:
                out.append((S1, (prv, k - 1), self.cost3(g, s2[r], s3[k])))
                out.append((SS, (r, k - 1), self.cost3(g, g, s3[k])))
            return out
        rho = i * L2 + r
        a = s1[i]
        if r == 0:
            dg = (i - 1) * (L2 - 1)
            out.append((P3, (dg, k), self.cost3(a, g, s3[0])))
            if k >= 1:
                out.append((S3, (dg, k - 1), self.cost3(a, g, s3[k])))
                out.append((SS, (rho, k - 1), self.cost3(g, g, s3[k])))
            return out
        up = (i - 1) * (L2 - 1) + r
        pv = i * (L2 - 1) + r - 1
        dg = (i - 1) * (L2 - 1) + r - 1
        b = s2[r]
        out.append((P3, (up, k), self.cost3(a, g, s3[0])))
        out.append((P1, (pv, k), self.cost3(g, b, s3[0])))
        out.append((P2, (dg, k), self.cost3(a, b, s3[0])))
        if k >= 1:
            out.append((S3, (up, k - 1), self.cost3(a, g, s3[k])))
            out.append((S1, (pv, k - 1), self.cost3(g, b, s3[k])))
            out.append((S2, (dg, k - 1), self.cost3(a, b, s3[k])))
            out.append((SS, (rho, k - 1), self.cost3(g, g, s3[k])))
        return out

    def _enumerate(self):
        L1, L2, L3 = self.L1, self.L2, self.L3
        succ = {}
        for i in range(L1):
            for r in range(L2):
                for k in range(L3):
                    node = (i * L2 + r, k)
                    for mv, src, w in self.preds(i, r, k):
                        succ.setdefault(src, []).append((node, w))
        best = {(0, 0): 0}

        def dfs(node, acc):
            for nxt, w in succ.get(node, ()):
                v = acc + w
                if nxt not in best or v < best[nxt]:
                    best[nxt] = v
                dfs(nxt, v)
        dfs((0, 0), 0)
        self.val = best
        self.dirs = {}
        for i in range(L1):
            for r in range(L2):
                for k in range(L3):
                    cands = self.preds(i, r, k)
                    if not cands:
                        continue
                    m = best[(i * L2 + r, k)]
                    for mv, src, w in cands:
                        if best[src] + w == m:
                            self.dirs[(i * L2 + r, k)] = mv
                            break

    def cost(self):
        return self.val[((self.L1 - 1) * self.L2 + self.L2 - 1, self.L3 - 1)]

    def dm_code(self, i, r, k):
        """Direction code of cell (i, r, k); the origin holds S2 as the
        reference's first-row initialisation writes (src/algn.c:2622)."""
        if (i, r, k) == (0, 0, 0):
            return S2
        return self.dirs[(i * self.L2 + r, k)]

    def walk(self):
        """backtrack_3d: (r1, r2, r3) left to right, or None when the walk
        takes a symbol before the start of a sequence.  As in the reference,
        the position in the direction cube (a linear index moved by plane,
        line and cell strides) and the three sequence indices are tracked
        separately; the walk ends when the linear index reaches the origin."""
        L1, L2, L3 = self.L1, self.L2, self.L3
        plane, line = L2 * L3, L3
        end = L1 * L2 * L3 - 1
        idx = [L1 - 1, L2 - 1, L3 - 1]
        seqs = (self.s1, self.s2, self.s3)
        out = ([], [], [])
        # each code: (takes s1, takes s2, takes s3), checked in the
        # reference's order S2, S3, S1, P3, SS, P1, P2
        moves = ((S2, (1, 1, 1)), (S3, (1, 0, 1)), (S1, (0, 1, 1)), (P3, (1, 0, 0)),
                 (SS, (0, 0, 1)), (P1, (0, 1, 0)), (P2, (1, 1, 0)))
        cap = L1 + L2 + L3
        while end > 0 and len(out[0]) < cap:
            i, rem = divmod(end, plane)
            r, k = divmod(rem, line)
            d = self.dm_code(i, r, k)
            take = next(t for code, t in moves if d & code)
            for u in range(3):
                if take[u]:
                    if idx[u] < 0:
                        return None
                    out[u].append(seqs[u][idx[u]])
                    idx[u] -= 1
                else:
                    out[u].append(GAP)
            end -= take[0] * plane + take[1] * line + take[2]
        return tuple(x[::-1] for x in out)

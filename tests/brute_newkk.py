"""Path enumeration for the diagonal-store aligner's affine model
(src/newkkonen.c:125-160 and update_internal_cell, 680-870).

A path starts at (0, 0) in the align state (cost 0) or in either gap state
(cost gap_open), as newkk_algn seeds it (src/newkkonen.c:1390), and moves one
cell at a time:
  align    (i-1, j-1) -> (i, j): from align + sub(a&15, b&15); from the
           vertical state + cost(a&15, b) (+ go when a carries the gap bit);
           from the horizontal state + cost(a, b&15) (+ go when b does); from
           the gap-gap state + sub + go(a) + go(b);
  gap-gap  (i-1, j-1) -> (i, j): from align + go(a) + go(b) (whatever the
           symbols), or from gap-gap at no cost when a and b both carry the
           gap bit;
  vertical (i-1, j) -> (i, j): from vertical + cost(a, gap) + extgo(a), or
           from align + cost(a, gap) + go(a);
  horizontal (i, j-1) -> (i, j): the same with b.
go(x) is 0 when x opens a run of gap-bit symbols (or is the first symbol
and carries the bit), else gap_open; extgo(x) is gap_open when x leaves such
a run.  Every path is enumerated explicitly (tiny inputs only); the cheapest
path into a cell in each state is that state's value there.
"""
GAP = 16
ALIGN, GG, VERT, HORZ = 0, 1, 2, 3


def go_of(s, idx, go):
    x, p = s[idx], (s[idx - 1] if idx > 0 else 0)
    if idx == 1 and x & GAP:
        return 0
    if idx > 1 and not (p & GAP) and (x & GAP):
        return 0
    return go


def extgo_of(s, idx, go):
    x, p = s[idx], (s[idx - 1] if idx > 0 else 0)
    return go if (p & GAP) and not (x & GAP) else 0


def _steps(s1, s2, cost, go, i, j, st):
    out = []
    n1, n2 = len(s1), len(s2)
    if i + 1 < n1 and j + 1 < n2:
        a, b = s1[i + 1], s2[j + 1]
        g1, g2 = go_of(s1, i + 1, go), go_of(s2, j + 1, go)
        sub = cost(a & 15, b & 15)
        both_gap = bool(a & GAP) and bool(b & GAP)
        if st == ALIGN:
            out.append((i + 1, j + 1, ALIGN, sub))
            out.append((i + 1, j + 1, GG, g1 + g2))
        elif st == VERT:
            out.append((i + 1, j + 1, ALIGN, cost(a & 15, b) + (go if a & GAP else 0)))
        elif st == HORZ:
            out.append((i + 1, j + 1, ALIGN, cost(a, b & 15) + (go if b & GAP else 0)))
        else:
            out.append((i + 1, j + 1, ALIGN, sub + g1 + g2))
            if both_gap:
                out.append((i + 1, j + 1, GG, 0))
    if i + 1 < n1:
        ge = cost(s1[i + 1], GAP)
        if st == VERT:
            out.append((i + 1, j, VERT, ge + extgo_of(s1, i + 1, go)))
        elif st == ALIGN:
            out.append((i + 1, j, VERT, ge + go_of(s1, i + 1, go)))
    if j + 1 < n2:
        ge = cost(s2[j + 1], GAP)
        if st == HORZ:
            out.append((i, j + 1, HORZ, ge + extgo_of(s2, j + 1, go)))
        elif st == ALIGN:
            out.append((i, j + 1, HORZ, ge + go_of(s2, j + 1, go)))
    return out


def state_minima(s1, s2, cost, go):
    """{(i, j, state): cheapest path cost} over the paths from (0, 0).
    cost(a, b) is the raw 5-bit table lookup."""
    s1 = [int(x) for x in s1]
    s2 = [int(x) for x in s2]
    best = {}
    stack = [(0, 0, ALIGN, 0), (0, 0, VERT, go), (0, 0, HORZ, go)]
    while stack:
        i, j, st, c = stack.pop()
        key = (i, j, st)
        # a dearer way into the same cell and state has no cheaper future
        if key in best and c >= best[key]:
            continue
        best[key] = c
        for ni, nj, nst, dc in _steps(s1, s2, cost, go, i, j, st):
            stack.append((ni, nj, nst, c + dc))
    return best

"""Host side of the serving workload (no GPU): the tree simulator, the
rejoin junctions and DOS.distance's band."""
import numpy as np

from poy5_amd import spr


def test_dos_deltaw_rule():
    # src/seqCS.ml:738-744: the length difference, at least 8
    assert spr.dos_deltaw(100, 100) == 8
    assert spr.dos_deltaw(100, 108) == 8
    assert spr.dos_deltaw(100, 109) == 9
    assert spr.dos_deltaw(1500, 1320) == 180


def test_yule_tree_shape_and_leaf_sequences():
    parent, children, leaves = spr.simulate(n_taxa=32, root_len=200, seed=5)
    assert len(parent) == 2 * 32 - 1
    assert all(len(c) in (0, 2) for c in children)
    assert len(leaves) == 32
    for s in leaves.values():
        assert s[0] == 16 and np.all(np.isin(s[1:], [1, 2, 4, 8]))


def test_rejoin_junctions_exclude_the_clipped_subtree():
    parent, children, leaves = spr.simulate(n_taxa=24, root_len=50, seed=6)
    mid = {v: 1000 + v for v in range(len(parent))}
    emid = {v: 5000 + v for v in range(len(parent)) if parent[v] >= 0}
    a, b, cv, ev = spr.rejoin_requests(parent, children, mid, emid, 500, np.random.default_rng(1))
    assert len(a) == 500
    for x, y, c, e in zip(a, b, cv, ev):
        assert x == mid[c] and y == emid[e]
        assert parent[c] >= 0 and e not in spr.subtree(children, c) and e != parent[c]

"""Pins the oracle's converter restatement (oracle/xo_convert.c) with
hand-derived trees and invariants of src/model/OctreeConstruction.h and
src/model/Octree.cpp.
"""
import numpy as np
import pytest

from oracle import xo
from tests import helpers as H

LEAF = 0x80000000


def _leaf_boxes(nodes, side):
    """Walk the tree (children x*4+y*2+z, OctreeConstruction.h:183-190) and
    return (index, pos, extent, depth) for each reachable leaf."""
    out = []

    def walk(i, pos, ext, depth):
        n = nodes[i]
        if n[9] & LEAF:
            out.append((i, pos, ext, depth))
            return
        h = ext // 2
        c = 0
        for xo_ in (0, h):
            for yo in (0, h):
                for zo in (0, h):
                    walk(int(n[c]), (pos[0] + xo_, pos[1] + yo, pos[2] + zo), h, depth + 1)
                    c += 1

    walk(0, (0, 0, 0), side, 0)
    return out


def test_single_voxel_hand_derived():
    """One non-zero voxel at (0,0,0) of 8^3 with --chan-diff 0: root, three
    internal levels with 7 empty leaves each, and 1+7 leaves at the bottom =
    25 nodes.  --dag dedupes equal leaves only within a depth (is_leaf_depth
    is part of the node, Octree.cpp:35-40): 3 internal + 4 unique leaves = 7."""
    g = np.zeros((8, 8, 8, 4), np.uint8)
    g[0, 0, 0] = (10, 20, 30, 255)
    nodes, st = xo.build_octree(g)
    assert nodes.shape[0] == 25 and st["total_nodes"] == 25 and st["depth"] == 3
    assert st["total_leaves"] == 22 and st["unique_leaves"] == 22
    # root is node 0, internal, depth 0; reverse post-order => children > parent
    assert nodes[0, 9] == 0
    for i in range(nodes.shape[0]):
        if not nodes[i, 9] & LEAF:
            assert (nodes[i, :8] > i).all()
        else:
            assert (nodes[i, :8] == 0).all()
    # The voxel sits in child 0 at every level; its leaf holds the colour.
    i = 0
    for _ in range(3):
        i = int(nodes[i, 0])
    assert nodes[i, 9] == LEAF | 3
    assert nodes[i, 8] == (10 | 20 << 8 | 30 << 16 | 255 << 24)
    # internal nodes store the integer mean over their region (Grid.cpp:86-131)
    assert nodes[0, 8] == (0 | 0 << 8 | 0 << 16 | 0 << 24)  # 255/512 -> 0
    dag, st2 = xo.build_octree(g, tree=xo.TREE_DAG)
    assert dag.shape[0] == 7 and st2["unique_leaves"] == 4 and st2["total_nodes"] == 25


def test_reverse_post_order_numbering():
    """OctreeBuilder::build (OctreeConstruction.h:92-112): final index = N-1 -
    post-order insertion index.  Recompute the post-order by walking the tree
    and check it."""
    g = H.random_grid((8, 8, 8), density=0.2, seed=4)
    nodes, st = xo.build_octree(g)
    N = nodes.shape[0]
    order = []

    def post(i):
        if not nodes[i, 9] & LEAF:
            for c in range(8):
                post(int(nodes[i, c]))
        order.append(i)

    post(0)
    assert len(order) == N
    assert [N - 1 - k for k in range(N)] == order


@pytest.mark.parametrize("shape", [(8, 8, 8), (5, 7, 3), (16, 12, 9)])
def test_lossless_tree_reproduces_grid(shape):
    """--chan-diff 0: every voxel's leaf carries that voxel's colour; leaves
    outside the grid are {0, colour 0, LEAF|depth} (OctreeConstruction.h:141-154)."""
    g = H.blocky_grid(16, 2, seed=2)[: shape[0], : shape[1], : shape[2]].copy()
    nodes, st = xo.build_octree(g)
    side = st["side"]
    Z, Y, X = g.shape[:3]
    for i, pos, ext, depth in _leaf_boxes(nodes, side):
        assert nodes[i, 9] == LEAF | depth
        x0, y0, z0 = pos
        region = g[z0:z0 + ext, y0:y0 + ext, x0:x0 + ext]
        if x0 >= X or y0 >= Y or z0 >= Z:
            assert nodes[i, 8] == 0
            continue
        assert region.shape[:3] == (ext, ext, ext), "a partly-outside node must split"
        col = region.reshape(-1, 4)
        assert (col == col[0]).all()
        packed = int(col[0, 0]) | int(col[0, 1]) << 8 | int(col[0, 2]) << 16 | int(col[0, 3]) << 24
        assert nodes[i, 8] == packed


def test_dag_expands_to_sparse_tree():
    g = H.blocky_grid(16, 2, seed=5)
    sparse, _ = xo.build_octree(g)
    dag, _ = xo.build_octree(g, tree=xo.TREE_DAG)
    assert dag.shape[0] < sparse.shape[0]

    def expand(nodes, i):
        n = nodes[i]
        if n[9] & LEAF:
            return (int(n[8]), int(n[9]))
        return (int(n[8]), int(n[9]), tuple(expand(nodes, int(c)) for c in n[:8]))

    assert expand(sparse, 0) == expand(dag, 0)
    # first-occurrence dedup: DAG nodes are pairwise distinct
    assert len({bytes(r) for r in dag}) == dag.shape[0]


def test_rope_neighbours_brute_force():
    """generate_ropes (Octree.cpp:181-201): rope k of a leaf is the node found
    at depth <= leaf depth that contains the face-neighbour point; 0 outside."""
    g = H.blocky_grid(16, 2, seed=6)
    nodes, st = xo.build_octree(g, tree=xo.TREE_ROPE)
    sparse, _ = xo.build_octree(g)
    side = st["side"]
    # ropes only change leaf children
    inner = (sparse[:, 9] & LEAF) == 0
    assert (nodes[inner] == sparse[inner]).all()

    def find(pos, max_depth):
        if any(p < 0 or p >= side for p in pos):
            return 0
        i, off, ext, d = 0, [0, 0, 0], side, 0
        while True:
            ext //= 2
            if nodes[i, 9] & LEAF or ext == 0 or d == max_depth:
                return i
            c = 0
            for a in range(3):
                if pos[a] >= off[a] + ext:
                    c |= 4 >> a
                    off[a] += ext
            i = int(nodes[i, c])
            d += 1

    for i, pos, ext, depth in _leaf_boxes(nodes, side):
        for a in range(3):
            p = list(pos)
            p[a] += ext
            assert nodes[i, 2 * a] == find(p, depth)
            p = list(pos)
            p[a] -= ext
            assert nodes[i, 2 * a + 1] == find(p, depth)
        assert nodes[i, 6] == 0 and nodes[i, 7] == 0


def test_vol_scan_and_stddev_scan_brute_force():
    g = H.random_grid((9, 6, 7), density=0.6, seed=8)
    rng = np.random.default_rng(0)
    for _ in range(20):
        lo = rng.integers(0, 8, 3)
        hi = lo + rng.integers(1, 6, 3)
        a, md = xo.vol_scan(g, lo, hi)
        sub = g[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]].reshape(-1, 4).astype(np.int64)
        if sub.shape[0] == 0:
            assert (a, md) == (0, 0)
            continue
        avg = sub.sum(0) // sub.shape[0]
        assert a == int(avg[0] | avg[1] << 8 | avg[2] << 16 | avg[3] << 24)
        assert md == int((sub.max(0) - sub.min(0)).max())
        a2, sd = xo.stddev_scan(g, lo, hi)
        mu = sub.sum(0) / sub.shape[0]
        ref = np.sqrt(((sub - mu) ** 2).sum() / sub.shape[0])
        assert abs(sd - ref) <= 1e-9 * (1 + ref)
        assert a2 == int(np.floor(mu).astype(np.int64) @ np.array([1, 256, 65536, 16777216]))


def test_chan_diff_and_std_dev_thresholds():
    g = np.zeros((4, 4, 4, 4), np.uint8)
    g[..., 3] = 255
    g[0, 0, 0, 0] = 9  # one channel differs by 9
    n_lo, _ = xo.build_octree(g, chan_diff=8)
    n_hi, _ = xo.build_octree(g, chan_diff=9)
    assert n_lo.shape[0] > 1 and n_hi.shape[0] == 1
    # std-dev over all channels / n (Grid.cpp:175-213): 9 in one of 64 voxels
    mu = 9 / 64
    sd = np.sqrt(((9 - mu) ** 2 + 63 * mu ** 2) / 64)
    n_a, _ = xo.build_octree(g, heuristic=xo.HEUR_STD_DEV, std_dev=sd * 0.999)
    n_b, _ = xo.build_octree(g, heuristic=xo.HEUR_STD_DEV, std_dev=sd * 1.001)
    assert n_a.shape[0] > 1 and n_b.shape[0] == 1

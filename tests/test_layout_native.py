"""Host check of the derived octree device layout (csrc/xd_layout.cpp), built
with its own harness (tests/native/layout_check.cpp): renumbering, tags,
{color, depth}, rope depths, the parallel numbering and depth check of large
trees, and the start table -- the node find() reaches from the table for
seeded points (cell boundaries included) equals the node the shader's walk
from the root reaches over the .svo words, with the same offset and side."""
import os
import subprocess

import numpy as np
import pytest

from oracle import xo
from tests import helpers as H
from xenodon_amd import abi, synth
from xenodon_amd import build as native_build

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("layout") / "layout_check")
    cmd = [native_build.hipcc(), "-O2", "-std=c++17", "-pthread", "-ffp-contract=off", "-I", native_build.INCLUDE,
           "-I", native_build.CSRC, os.path.join(HERE, "native", "layout_check.cpp"),
           os.path.join(native_build.CSRC, "xd_layout.cpp"), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    return out


def _check(harness, tmp_path, nodes, wide=False, points=20000):
    path = tmp_path / "tree.words"
    np.ascontiguousarray(nodes, dtype=np.uint32).tofile(path)
    r = subprocess.run([harness, str(path), "1" if wide else "0", str(points)], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-1000:]
    return r.stdout


def test_layout_of_small_trees(harness, tmp_path):
    g = synth.shells_128()
    for tree in (xo.TREE_SPARSE, xo.TREE_ROPE, xo.TREE_DAG):
        nodes, _ = xo.build_octree(g, tree=tree)
        out = _check(harness, tmp_path, nodes)
        assert "ok:" in out
    rope, _ = xo.build_octree(g, tree=xo.TREE_ROPE)
    assert "ropes packed 1" in _check(harness, tmp_path, rope)
    assert "wide 1" in _check(harness, tmp_path, rope, wide=True)
    # a single leaf, and a ragged blocky grid
    one, _ = xo.build_octree(np.full((1, 1, 1, 4), 9, np.uint8))
    assert "ok:" in _check(harness, tmp_path, one, points=500)
    blocky, _ = xo.build_octree(H.blocky_grid(48, 4, seed=3), tree=xo.TREE_ROPE, chan_diff=12)
    assert "ok:" in _check(harness, tmp_path, blocky)


def test_layout_of_a_large_tree_numbered_in_parallel(harness, tmp_path):
    """The 512x361x512 CT volume's lossless rope tree (20 M nodes) takes the
    subtree-parallel numbering and depth check."""
    nodes, st = abi.build_octree(synth.ct_body(), tree=abi.TREE_ROPE)
    assert st["node_count"] >= 1 << 22
    out = _check(harness, tmp_path, nodes, points=50000)
    assert "tree 1" in out and "ropes packed 1" in out

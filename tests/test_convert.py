"""The product converter (xd_build_octree, CPU) against the oracle's
restatement of build_octree: identical node buffers and statistics."""
import numpy as np
import pytest

from oracle import xo
from tests import helpers as H
from xenodon_amd import abi, synth

CASES = [
    ("rand16", lambda: H.random_grid((16, 16, 16), density=0.3, seed=1)),
    ("odd", lambda: H.random_grid((5, 13, 9), density=0.5, seed=2)),
    ("blocky64", lambda: H.blocky_grid(64, 4, seed=3)),
    ("shells64", lambda: synth.shells((64, 48, 40), seed=4, n_shells=8, radii=(5.0, 20.0))),
]
OPTIONS = [
    dict(heuristic=0, chan_diff=0, tree=0),
    dict(heuristic=0, chan_diff=0, tree=1),
    dict(heuristic=0, chan_diff=0, tree=2),
    dict(heuristic=0, chan_diff=40, tree=0),
    dict(heuristic=0, chan_diff=40, tree=1),
    dict(heuristic=1, std_dev=20.0, tree=0),
    dict(heuristic=1, std_dev=5.5, tree=2),
]


@pytest.mark.parametrize("name,make", CASES)
@pytest.mark.parametrize("opt", OPTIONS)
def test_product_converter_matches_oracle(name, make, opt):
    g = make()
    o_nodes, o_st = xo.build_octree(g, heuristic=opt["heuristic"], chan_diff=opt.get("chan_diff", 0),
                                    std_dev=opt.get("std_dev", 0.0), tree=opt["tree"])
    p_nodes, p_st = abi.build_octree(g, heuristic=opt["heuristic"], chan_diff=opt.get("chan_diff", 0),
                                     std_dev=opt.get("std_dev", 0.0), tree=opt["tree"])
    assert p_st == o_st
    assert np.array_equal(p_nodes, o_nodes)


@pytest.mark.slow
def test_product_converter_matches_oracle_at_ct_size():
    """The 512x361x512 CT-like workload volume, lossless rope tree (the tree
    the octree traversals are benchmarked on): parallel top levels included."""
    g = synth.ct_body()
    o_nodes, o_st = xo.build_octree(g, tree=xo.TREE_ROPE)
    p_nodes, p_st = abi.build_octree(g, tree=abi.TREE_ROPE)
    assert p_st == o_st
    assert np.array_equal(p_nodes, o_nodes)

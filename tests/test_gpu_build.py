"""GPU octree builder (xd_build_octree_device): the node buffer and stats are
bit-identical to the host converter's (itself pinned against the oracle and
hand-derived trees in test_convert.py / test_oracle_convert.py), for sparse
and rope trees, lossless and lossy channel-difference builds, power-of-two
and ragged grids."""
import numpy as np
import pytest

from tests import helpers as H
from xenodon_amd import abi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    if abi.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    c = abi.Context(0)
    yield c
    c.close()


def _grids():
    yield "one", np.full((1, 1, 1, 4), 77, np.uint8)
    yield "tiny", H.random_grid((2, 3, 1), density=0.6, seed=1)
    yield "odd", H.random_grid((5, 7, 9), density=0.4, seed=2)
    yield "rand16", H.random_grid((16, 16, 16), density=0.3, seed=3)
    yield "ragged", H.random_grid((33, 17, 40), density=0.2, seed=4)
    yield "blocky64", H.blocky_grid(64, 4, seed=5)
    yield "uniform", np.full((24, 24, 24, 4), 130, np.uint8)
    yield "empty", np.zeros((20, 9, 31, 4), np.uint8)
    yield "shells128", synth.shells_128()
    yield "lognormal128", synth.lognormal(side=128, base=64)


@pytest.mark.parametrize("tree", [abi.TREE_SPARSE, abi.TREE_ROPE])
@pytest.mark.parametrize("chan_diff", [0, 24])
def test_device_build_matches_host_converter(ctx, tree, chan_diff):
    for name, g in _grids():
        ref, ref_st = abi.build_octree(g, chan_diff=chan_diff, tree=tree)
        vol = ctx.grid_volume(g)
        got, st = ctx.build_octree(vol, chan_diff=chan_diff, tree=tree)
        vol.free()
        assert st == ref_st, (name, st, ref_st)
        assert got.shape == ref.shape and np.array_equal(got, ref), name


def test_device_build_refuses_dags_and_svo_input(ctx):
    g = H.random_grid((8, 8, 8), density=0.3, seed=9)
    vol = ctx.grid_volume(g)
    with pytest.raises(abi.XdError):
        ctx.build_octree(vol, tree=abi.TREE_DAG)
    nodes, st = abi.build_octree(g)
    sv = ctx.svo_volume(nodes, st["side"])
    with pytest.raises(abi.XdError):
        ctx.build_octree(sv)

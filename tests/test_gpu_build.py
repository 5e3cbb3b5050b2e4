"""GPU octree builder (xd_build_octree_device): the node buffer and stats are
bit-identical to the host converter's (itself pinned against the oracle and
hand-derived trees in test_convert.py / test_oracle_convert.py), for sparse,
DAG and rope trees, lossless and lossy channel-difference builds, std-dev
builds at several limits, power-of-two and ragged grids."""
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


@pytest.mark.parametrize("chan_diff", [0, 24])
def test_device_dag_matches_host_converter(ctx, chan_diff):
    """DAG builds: equal nodes of a level merged, the first in post-order kept
    (OctreeConstruction.h:19-30, 72-112)."""
    for name, g in _grids():
        ref, ref_st = abi.build_octree(g, chan_diff=chan_diff, tree=abi.TREE_DAG)
        vol = ctx.grid_volume(g)
        got, st = ctx.build_octree(vol, chan_diff=chan_diff, tree=abi.TREE_DAG)
        vol.free()
        assert st == ref_st, (name, st, ref_st)
        assert got.shape == ref.shape and np.array_equal(got, ref), name


def _smooth(side, seed):
    """A grid whose regions have std-devs spread over a range, from a
    smoothed random field (values 0..255 in all four channels)."""
    rng = np.random.default_rng(seed)
    f = rng.random((side // 4 + 2,) * 3)
    f = np.repeat(np.repeat(np.repeat(f, 4, 0), 4, 1), 4, 2)[:side, :side, :side]
    g = np.empty((side, side, side, 4), np.uint8)
    for c in range(4):
        g[..., c] = np.clip(f * 200 + rng.integers(0, 40 + 10 * c, f.shape), 0, 255)
    return g


@pytest.mark.parametrize("tree", [abi.TREE_SPARSE, abi.TREE_DAG, abi.TREE_ROPE])
def test_device_std_dev_matches_host_converter(ctx, tree):
    """std-dev builds (Grid.cpp:139-214): the split of a node of up to 16^3
    voxels from the reference's sequential FP64 sum, of a larger one from a
    double-double sum with the sequential sum's error bound, and the host's
    comparison where that bound straddles the limit."""
    grids = [(n, g) for n, g in _grids()] + [("smooth64", _smooth(64, 7)), ("smooth40", _smooth(40, 8)[:, :33])]
    for limit in (0.0, 3.5, 17.25, 60.0):
        for name, g in grids:
            ref, ref_st = abi.build_octree(g, heuristic=abi.HEURISTIC_STD_DEV, std_dev=limit, tree=tree)
            vol = ctx.grid_volume(g)
            got, st = ctx.build_octree(vol, heuristic=abi.HEURISTIC_STD_DEV, std_dev=limit, tree=tree)
            vol.free()
            assert st == ref_st, (name, limit, st, ref_st)
            assert got.shape == ref.shape and np.array_equal(got, ref), (name, limit)


def test_device_build_refuses_svo_input_and_bad_arguments(ctx):
    g = H.random_grid((8, 8, 8), density=0.3, seed=9)
    vol = ctx.grid_volume(g)
    with pytest.raises(abi.XdError, match="std_dev must be >= 0"):
        ctx.build_octree(vol, heuristic=abi.HEURISTIC_STD_DEV, std_dev=-1.0)
    with pytest.raises(abi.XdError, match="unknown tree type"):
        ctx.build_octree(vol, tree=7)
    nodes, st = abi.build_octree(g)
    sv = ctx.svo_volume(nodes, st["side"])
    with pytest.raises(abi.XdError):
        ctx.build_octree(sv)

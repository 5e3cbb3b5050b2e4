"""Batched form (one persistent ray-tile-queue launch over many views): every
frame is bit-identical to the plain one-launch render of the same view, for
every launch shape the autotuner can pick."""
import numpy as np
import pytest

from oracle import xo
from tests import helpers as H
from xenodon_amd import abi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    if abi.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    ctx = abi.Context(0)
    shells = synth.shells_128()
    nodes, st = abi.build_octree(shells, tree=abi.TREE_ROPE)
    blocky = H.blocky_grid(64, 4, seed=41)
    bnodes, bst = abi.build_octree(blocky, tree=abi.TREE_ROPE)
    vols = {
        "shells_grid": (ctx.grid_volume(shells), [128, 128, 128]),
        "shells_svo": (ctx.svo_volume(nodes, st["side"]), [st["side"]] * 3),
        "blocky_grid": (ctx.grid_volume(blocky), [64, 64, 64]),
        "blocky_svo": (ctx.svo_volume(bnodes, bst["side"]), [bst["side"]] * 3),
    }
    yield ctx, vols
    ctx.close()


def _views(ctx, vols, seed):
    rng = np.random.default_rng(seed)
    plan = [("shells_grid", abi.DDA), ("shells_svo", abi.ESVO), ("shells_svo", abi.SVO_DF),
            ("shells_svo", abi.SVO_NAIVE), ("shells_svo", abi.SVO_ROPE), ("blocky_grid", abi.DDA),
            ("blocky_svo", abi.SVO_DF), ("blocky_svo", abi.ESVO), ("blocky_svo", abi.SVO_ROPE),
            ("shells_svo", abi.SVO_DF)]
    sizes = [(256, 256), (300, 200), (64, 80), (1, 1), (320, 240), (77, 131), (200, 150), (128, 96), (33, 65),
             (160, 120)]
    out = []
    cams = H.orbit_cameras(len(plan), 1.8, seed=seed)
    for (vname, t), (W, Hh), (f, up, eye) in zip(plan, sizes, cams):
        vol, md = vols[vname]
        e = float(rng.uniform(1.0, 4.0))
        out.append((vol, t, abi.make_camera(f, up, eye), abi.make_uniforms(W, Hh, md, emission=e), ctx.frame(W, Hh)))
    return out


def _check_against_plain(ctx, views):
    for vol, t, cam, uni, fr in views:
        got = fr.download()
        ref, _ = ctx.render_image(vol, t, cam, uni)
        assert np.array_equal(got, ref), abi.TRAVERSAL_NAMES[t]


def test_batch_matches_plain_renders(env):
    ctx, vols = env
    b = ctx.batch()
    views = _views(ctx, vols, seed=1)
    b.set_views(views)
    ms, rays = b.render()
    assert rays == sum(v[3].output_region.extent[0] * v[3].output_region.extent[1] for v in views)
    assert ms > 0
    _check_against_plain(ctx, views)


@pytest.mark.parametrize("tile_w,grab,waves,classes,inter,mix,rpl", [(16, 1, 0, 0, 1, 1, 1), (32, 2, 8, 1, 0, 0, 2),
                                                                     (8, 4, 8, 0, 2, 1, 4), (8, 1, 0, 1, 1, 0, 2),
                                                                     (16, 2, 0, 1, 2, 1, 1), (8, 1, 8, 0, 2, 2, 2),
                                                                     (32, 1, 8, 0, 5, 1, 4)])
def test_batch_launch_shapes_do_not_change_pixels(env, tile_w, grab, waves, classes, inter, mix, rpl):
    """Every launch shape -- tile shape, grab size, register budget, kernels,
    tile order, class mixing or XCD affinity, rays per lane (mixed per
    traversal) -- renders the plain form's pixels."""
    ctx, vols = env
    b = ctx.batch()
    b.set_tuning([tile_w, 8, tile_w, 32, 16], grab, waves, classes, inter, mix, [rpl, 1, 4, rpl, 2])
    views = _views(ctx, vols, seed=2)
    b.set_views(views)
    b.render()
    _check_against_plain(ctx, views)


def test_batch_reuses_graph_with_new_cameras_and_autotunes(env):
    ctx, vols = env
    b = ctx.batch()
    views = _views(ctx, vols, seed=3)
    b.set_views(views)
    b.render()
    chosen = b.autotune(trials=1)
    assert all(w in (8, 16, 32) for w in chosen["tile_w"]) and 1 <= chosen["tiles_per_grab"] <= 8
    views2 = _views(ctx, vols, seed=4)  # new cameras, sizes kept
    b.set_views(views2)
    for _ in range(3):  # graph replays, each checked against plain renders
        b.render()
        _check_against_plain(ctx, views2)


@pytest.mark.parametrize("with_octree_view", [False, True])
def test_batch_matches_oracle_on_c1(env, with_octree_view):
    """Four C1 dda views in one launch.  Alone they run the DDA class kernel
    (a one-traversal batch always does); with one svo-naive view added the
    batch runs the all-traversal kernel on the same dda views."""
    ctx, vols = env
    grid = synth.shells_128()
    ov = xo.OracleVolume(grid=grid)
    vol, md = vols["shells_grid"]
    b = ctx.batch()
    b.set_tuning([8] * 5, 1, 0, 0, 1)  # the four views form one interleaved group
    cams = H.orbit_cameras(4, 2 * np.sqrt(3), seed=7)
    views = [(vol, abi.DDA, abi.make_camera(*c), abi.make_uniforms(256, 256, md, emission=4.0), ctx.frame(256, 256))
             for c in cams]
    extra = []
    if with_octree_view:
        svol, smd = vols["shells_svo"]
        extra = [(svol, abi.SVO_NAIVE, abi.make_camera(*cams[0]), abi.make_uniforms(64, 48, smd, emission=4.0),
                  ctx.frame(64, 48))]
    b.set_views(views + extra)
    b.render()
    for (v, t, cam, uni, fr), c in zip(views, cams):
        ref = xo.render(xo.DDA, ov, xo.make_camera(*c), xo.make_uniforms(256, 256, md, emission=4.0))["rgba"]
        assert np.array_equal(fr.download(), ref)
    _check_against_plain(ctx, extra)


def test_batch_rejects_bad_views(env):
    ctx, vols = env
    b = ctx.batch()
    vol, md = vols["shells_svo"]
    with pytest.raises(abi.XdError) as e:
        b.set_views([(vol, abi.DDA, abi.make_camera((0, 0, 1), (0, 1, 0), (0.5, 0.5, -2)),
                      abi.make_uniforms(8, 8, md), ctx.frame(8, 8))])
    assert e.value.status == 4
    with pytest.raises(abi.XdError):
        b.set_views([(vol, abi.ESVO, abi.make_camera((0, 0, 1), (0, 1, 0), (0.5, 0.5, -2)),
                      abi.make_uniforms(8, 8, md), ctx.frame(9, 8))])


@pytest.mark.parametrize("classes,waves", [(1, 8), (1, 0), (0, 8)])
def test_dda_batch_edge_views(env, classes, waves):
    """DDA in the batch kernels on a ragged grid (29 x 50 x 37: bricks cut by
    the grid's faces): eyes inside the grid, axis-aligned views (rays with
    zero direction components), and a grazing oblique view; every frame
    equals the plain render."""
    ctx, _ = env
    grid = np.ascontiguousarray(H.blocky_grid(64, 4, seed=43)[:37, :50, :29])
    vol = ctx.grid_volume(grid)
    md = [29, 50, 37]
    cams = [H.look_at((0.5, 0.5, 0.5), (1.5, 0.5, 0.5)),               # inside, along +x
            H.look_at((0.3, 0.6, 0.45), (0.3, 0.6, -1.0)),             # inside, along -z
            H.look_at((0.5, 0.5, 2.5), (0.5, 0.5, 0.0)),               # outside, along -z
            H.look_at((-0.4, 0.02, 0.5), (1.2, 0.03, 0.52)),           # grazing the y = 0 face
            H.look_at((1.7, 1.3, -0.6), (0.45, 0.55, 0.5))]            # oblique
    views = [(vol, abi.DDA, abi.make_camera(f, up, eye), abi.make_uniforms(97, 61, md, emission=2.5), ctx.frame(97, 61))
             for f, up, eye in cams]
    if classes == 0:
        # one octree view, so that the batch holds two traversals and the dda
        # views run in the all-traversal kernel (a batch of one traversal
        # always runs that traversal's own kernel)
        svol, smd = env[1]["blocky_svo"]
        views.append((svol, abi.SVO_NAIVE, abi.make_camera(*cams[4]), abi.make_uniforms(40, 30, smd, emission=2.5),
                      ctx.frame(40, 30)))
    b = ctx.batch()
    tw = b.tuning()
    b.set_tuning(tw["tile_w"], tw["tiles_per_grab"], waves, classes, tw["interleave"], tw["mix_classes"])
    b.set_views(views)
    b.render()
    lit = 0
    for v, t, cam, uni, fr in views:
        got = fr.download()
        ref, _ = ctx.render_image(v, t, cam, uni)
        assert np.array_equal(got, ref)
        lit += int(got[..., :3].any())
        fr.free()
    assert lit >= 4
    b.free()
    vol.free()


def _macro_words(dims_xyz):
    """Words of a grid's macro-cell bitmask, from the footprint the library
    states: bricks of 256 B + 1 occupancy byte, then the bitmask."""
    nb = int(np.prod([(d + 3) // 4 for d in dims_xyz]))
    return (abi.volume_footprint(abi.VOLUME_GRID, tuple(dims_xyz)) - nb * 257) // 4


@pytest.mark.parametrize("classes,waves", [(1, 8), (1, 0), (0, 8), (0, 0)])
def test_dda_macro_mask_in_lds_and_in_memory(env, classes, waves):
    """A batch wave keeps its own LDS copy of a DDA view's macro-cell bitmask
    when the mask fits its LDS region, and reads it from device memory when it
    does not.  Views of a grid whose mask exceeds every region (1189 words)
    alternate with views of two small-mask grids, so waves copy, keep, drop
    and copy masks again; with one kernel for all traversals, esvo and svo-df
    tiles overwrite the region between DDA tiles.  Every frame equals the
    plain render."""
    ctx, vols = env
    rng = np.random.default_rng(7)
    dims = (520, 72, 520)  # x, y, z
    nb = (dims[2] // 8, dims[1] // 8, dims[0] // 8)
    cols = rng.integers(0, 256, nb + (4,)).astype(np.uint8)
    cols[rng.random(nb) < 0.7] = 0
    cols[..., 3] = 255
    big = np.ascontiguousarray(np.repeat(np.repeat(np.repeat(cols, 8, 0), 8, 1), 8, 2))
    assert _macro_words(dims) == 1189 and _macro_words((128, 128, 128)) == 128
    bvol = ctx.grid_volume(big)
    md = list(dims)
    cams = H.orbit_cameras(4, 1.9, seed=11)
    views = []
    for k, (f, up, eye) in enumerate(cams):
        views.append((bvol, abi.DDA, abi.make_camera(f, up, eye), abi.make_uniforms(160, 90, md, emission=2.0),
                      ctx.frame(160, 90)))
        small, smd = vols["shells_grid" if k % 2 == 0 else "blocky_grid"]
        views.append((small, abi.DDA, abi.make_camera(f, up, eye), abi.make_uniforms(96, 72, smd, emission=3.0),
                      ctx.frame(96, 72)))
    if classes == 0:
        svol, smd = vols["shells_svo"]
        for t in (abi.ESVO, abi.SVO_DF):
            f, up, eye = cams[1]
            views.append((svol, t, abi.make_camera(f, up, eye), abi.make_uniforms(128, 96, smd, emission=2.0),
                          ctx.frame(128, 96)))
    b = ctx.batch()
    b.set_tuning([8] * 5, 1, waves, classes, 1, 1)  # views of one size interleaved tile by tile
    b.set_views(views)
    for _ in range(2):
        b.render()
        _check_against_plain(ctx, views)
    assert sum(int(v[4].download()[..., :3].any()) for v in views) >= len(views) - 1
    for v in views:
        v[4].free()
    b.free()
    bvol.free()

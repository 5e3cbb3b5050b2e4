"""xd_batch_render_static, the allocation-free batched entry point: its view
and tile tables and queue counters live in a static device array.  Its frames
must equal those of xd_batch_render (the batch object's own device tables) on
the same views, bit for bit, on two calls in a row -- the second with fewer
and smaller views, so that anything the first call left in the array would
show -- and it must refuse inputs beyond its stated limits before launching."""
import numpy as np
import pytest

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


PLAN = [("shells_grid", abi.DDA, 256, 256), ("shells_svo", abi.ESVO, 300, 200), ("shells_svo", abi.SVO_DF, 64, 80),
        ("shells_svo", abi.SVO_NAIVE, 1, 1), ("shells_svo", abi.SVO_ROPE, 320, 240),
        ("blocky_grid", abi.DDA, 77, 131), ("blocky_svo", abi.SVO_DF, 200, 150), ("blocky_svo", abi.ESVO, 128, 96),
        ("blocky_svo", abi.SVO_ROPE, 33, 65), ("shells_svo", abi.SVO_DF, 160, 120)]


def _views(ctx, vols, plan, seed):
    rng = np.random.default_rng(seed)
    cams = H.orbit_cameras(len(plan), 1.8, seed=seed)
    out = []
    for (vname, t, W, Hh), (f, up, eye) in zip(plan, cams):
        vol, md = vols[vname]
        e = float(rng.uniform(1.0, 4.0))
        out.append((vol, t, abi.make_camera(f, up, eye), abi.make_uniforms(W, Hh, md, emission=e)))
    return out


def _both(ctx, views, tuning=None):
    """Frames of the static entry point and of a batch object on the same views."""
    fs = [ctx.frame(u.output_region.extent[0], u.output_region.extent[1]) for (_, _, _, u) in views]
    fb = [ctx.frame(u.output_region.extent[0], u.output_region.extent[1]) for (_, _, _, u) in views]
    ms, rays = abi.render_static(ctx, [v + (f,) for v, f in zip(views, fs)], tuning)
    assert ms > 0 and rays == sum(u.output_region.extent[0] * u.output_region.extent[1] for (_, _, _, u) in views)
    b = ctx.batch()
    if tuning is not None:
        b.set_tuning(**tuning)
    b.set_views([v + (f,) for v, f in zip(views, fb)])
    b.render()
    got = [f.download() for f in fs]
    ref = [f.download() for f in fb]
    b.free()
    for f in fs + fb:
        f.free()
    return got, ref


def _assert_equal(got, ref, views):
    for g, r, (_, t, _, u) in zip(got, ref, views):
        assert g.shape == r.shape
        if not np.array_equal(g, r):
            d = np.abs(g.astype(int) - r.astype(int))
            raise AssertionError(f"{abi.TRAVERSAL_NAMES[t]} {u.output_region.extent[0]}x{u.output_region.extent[1]}: "
                                 f"{int((d > 0).any(-1).sum())} pixels differ, max {int(d.max())} LSB")


def test_static_matches_batch_on_two_calls(env):
    ctx, vols = env
    first = _views(ctx, vols, PLAN, seed=3)
    got, ref = _both(ctx, first)
    _assert_equal(got, ref, first)
    assert sum(int(g[..., :3].any()) for g in got) >= 6  # most views see their volume
    # second call: fewer, smaller views (other cameras) in the same array
    small = [(v, t, max(1, W // 3), max(1, Hh // 4)) for (v, t, W, Hh) in PLAN[1:7]]
    second = _views(ctx, vols, small, seed=4)
    got, ref = _both(ctx, second)
    _assert_equal(got, ref, second)


def test_static_per_class_kernels(env):
    """The same entry point under a tuning with one kernel per traversal."""
    ctx, vols = env
    views = _views(ctx, vols, PLAN[:6], seed=5)
    tun = {"tile_w": [16, 8, 32, 8, 16], "tiles_per_grab": 2, "waves_per_simd": 0, "class_kernels": 1,
           "interleave": 0, "mix_classes": 0}
    got, ref = _both(ctx, views, tun)
    _assert_equal(got, ref, views)


def test_static_refuses_beyond_its_limits(env):
    """Over 256 views, or over 2^21 wave tiles, is refused before any launch."""
    ctx, vols = env
    vol, md = vols["shells_svo"]
    cam = abi.make_camera(*H.orbit_cameras(1, 1.8, seed=6)[0])
    small = ctx.frame(8, 8)
    views = [(vol, abi.ESVO, cam, abi.make_uniforms(8, 8, md), small)] * 257
    with pytest.raises(abi.XdError, match="257 views exceed the static limit of 256"):
        abi.render_static(ctx, views)
    big = ctx.frame(1920, 1080)  # 32400 wave tiles each: 65 views hold 2106000 > 2^21
    views = [(vol, abi.ESVO, cam, abi.make_uniforms(1920, 1080, md), big)] * 65
    with pytest.raises(abi.XdError, match="wave tiles exceed the static limit"):
        abi.render_static(ctx, views)
    small.free()
    big.free()

"""The 1024^3 log-normal workload at full size: DDA on the 4 GiB grid and ESVO
on its rope SVO (one allocation above 4 GB, built on the GPU) at 3840x2160
over frames of the reference camera path.  Frames are rendered whole on the
GPU and compared bit for bit with the oracle on bands of rows (the oracle
renders the same output regions of the same display).  The many-views batch
(64 views over the 128^3, 512x361x512 and 1024^3 volumes in one launch) is
checked here too, since it needs the 1024^3 volumes resident."""
import os

import numpy as np
import pytest

from oracle import xo
from xenodon_amd import abi, synth, workloads

pytestmark = pytest.mark.gpu

W, H = 3840, 2160
CAMERA_FILE = os.path.join(os.path.dirname(__file__), "golden", "camera.txt")
BANDS = [(0, 540, 3840, 2), (0, 1080, 3840, 2), (1200, 1620, 1200, 4)]  # x, y, w, h of checked regions
EMISSION = 1.0


def _cameras():
    v = np.array(open(CAMERA_FILE).read().split(), np.float32).reshape(-1, 9)
    return [(r[0:3], r[3:6], r[6:9]) for r in v[[0, 75, 149]]]


@pytest.fixture(scope="module")
def c3():
    if abi.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    grid = synth.lognormal(side=1024)
    ctx = abi.Context(0)
    gv = ctx.grid_volume(grid)
    nodes, st = ctx.build_octree(gv, tree=abi.TREE_ROPE)
    sv = ctx.svo_volume(nodes, st["side"])
    yield ctx, grid, gv, nodes, st, sv
    ctx.close()


def _check(ctx, vol, ov, t, md):
    threads = min(16, os.cpu_count() or 1)
    for k, (f, up, eye) in enumerate(_cameras()):
        img, _ = ctx.render_image(vol, t, abi.make_camera(f, up, eye), abi.make_uniforms(W, H, md, emission=EMISSION))
        assert (img[..., 3] == 255).all()
        for (x, y, w, h) in BANDS:
            u = xo.make_uniforms(w, h, md, emission=EMISSION, out_offset=(x, y), disp_offset=(0, 0),
                                 disp_extent=(W, H))
            ref = xo.render(t, ov, xo.make_camera(f, up, eye), u, threads=threads)["rgba"]
            assert np.array_equal(img[y:y + h, x:x + w], ref), (abi.TRAVERSAL_NAMES[t], k, x, y)
        assert img[..., :3].any()  # the volume is in view
        # the workload's emission coefficient leaves at most 1 % of the pixels clamped
        assert (img[..., :3] == 255).any(-1).mean() <= 0.01


@pytest.mark.timeout(600)
def test_c3_dda_4k_bit_exact(c3):
    ctx, grid, gv, nodes, st, sv = c3
    assert gv.info().device_bytes >= 1 << 32
    _check(ctx, gv, xo.OracleVolume(grid=grid), abi.DDA, [1024] * 3)


@pytest.mark.timeout(600)
def test_c3_esvo_on_svo_above_4gb_bit_exact(c3):
    ctx, grid, gv, nodes, st, sv = c3
    assert st["side"] == 1024 and st["depth"] == 10
    assert nodes.nbytes > 1 << 32 and sv.info().device_bytes > 1 << 32
    _check(ctx, sv, xo.OracleVolume(nodes=nodes, side=1024), abi.ESVO, [1024] * 3)


@pytest.mark.timeout(600)
def test_c3_raw_octree_words_with_64bit_offsets(c3):
    """The .svo words kept as they are (the layout above 2^31 nodes): one
    6.7 GB buffer addressed with 64-bit offsets; frames equal the tagged
    layout's."""
    ctx, grid, gv, nodes, st, sv = c3
    with abi.Context(0) as raw_ctx:
        raw_ctx.set_option("octree_layout", 1)
        rv = raw_ctx.svo_volume(nodes, 1024)
        assert rv.info().device_bytes == nodes.nbytes
        f, up, eye = _cameras()[1]
        for t in (abi.ESVO, abi.SVO_DF):
            u = abi.make_uniforms(960, 540, [1024] * 3, emission=EMISSION)
            a, _ = raw_ctx.render_image(rv, t, abi.make_camera(f, up, eye), u)
            b, _ = ctx.render_image(sv, t, abi.make_camera(f, up, eye), u)
            assert np.array_equal(a, b), abi.TRAVERSAL_NAMES[t]


@pytest.mark.timeout(600)
@pytest.mark.parametrize("tuning", [None, {"interleave": 2, "waves_per_simd": 8, "mix_classes": 1},
                                    {"interleave": 2, "waves_per_simd": 8, "mix_classes": 1, "class_kernels": 1}])
def test_mixed_views_batch_bit_exact(c3, tuning):
    """64 views (dda / svo-df / esvo; W x H log-uniform in 256..1920 x
    256..1080; orbit cameras at distance 0.5..3) over three resident volumes
    in one batched launch: every frame equals the plain form's frame of the
    same view, and a band of rows of one view per traversal equals the
    oracle."""
    ctx, grid, gv, nodes, st, sv = c3
    scenes = {"lognormal1024": (grid, gv, [1024] * 3, nodes, sv, [1024] * 3, 1.0)}
    owned = []
    for name, g, em in (("shells128", synth.shells_128(), 4.0), ("ct512", synth.ct_body(), 1.6)):
        v = ctx.grid_volume(g)
        n, s2 = ctx.build_octree(v, tree=abi.TREE_ROPE)
        t2 = ctx.svo_volume(n, s2["side"])
        owned += [v, t2]
        scenes[name] = (g, v, list(g.shape[2::-1]), n, t2, [s2["side"]] * 3, em)
    spec = workloads.mixed_views(sorted(scenes))
    views = []
    for name, t, cam, W, H in spec:
        g, v, gmd, n, t2, smd, em = scenes[name]
        vol, md = (v, gmd) if t == abi.DDA else (t2, smd)
        views.append((vol, t, abi.make_camera(*cam), abi.make_uniforms(W, H, md, emission=em), ctx.frame(W, H)))
    b = ctx.batch()
    if tuning:
        tw = b.tuning()
        b.set_tuning(tw["tile_w"], tw["tiles_per_grab"], tuning["waves_per_simd"],
                     tuning.get("class_kernels", tw["class_kernels"]),
                     tuning["interleave"], tuning["mix_classes"])
    b.set_views(views)
    b.render()
    checked = set()
    for (name, t, cam, W, H), (vol, _, gcam, u, f) in zip(spec, views):
        got = f.download()
        ref, _ = ctx.render_image(vol, t, gcam, u)
        assert np.array_equal(got, ref), (name, abi.TRAVERSAL_NAMES[t], W, H)
        if t not in checked and got[..., :3].any():
            checked.add(t)
            g, v, gmd, n, t2, smd, em = scenes[name]
            ov = xo.OracleVolume(grid=g) if t == abi.DDA else xo.OracleVolume(nodes=n, side=smd[0])
            y = H // 2
            ou = xo.make_uniforms(W, 2, gmd if t == abi.DDA else smd, emission=em, out_offset=(0, y),
                                  disp_offset=(0, 0), disp_extent=(W, H))
            oref = xo.render(t, ov, xo.make_camera(*cam), ou, threads=min(16, os.cpu_count() or 1))["rgba"]
            assert np.array_equal(got[y:y + 2], oref), (name, abi.TRAVERSAL_NAMES[t], W, H)
    assert checked == {abi.DDA, abi.SVO_DF, abi.ESVO}
    b.free()
    for v in views:
        v[4].free()
    for v in owned:
        v.free()

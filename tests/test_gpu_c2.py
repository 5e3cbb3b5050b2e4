"""The headline configuration at full size: the 512x361x512 CT-like grid and
its lossless rope octree, all five traversals at 1920x1080 on frames of the
reference camera path (resources: camera.txt; orbit, the zoom, the
axis-aligned frames 111-112 and frames from inside the volume).

* every plain frame equals the oracle on bands of rows spread over the image
  and on a block at its centre (the oracle renders the same output regions of
  the same 1920x1080 display);
* the bench's default batched launch -- ten frames x five traversals in one
  persistent all-traversal kernel, the frames the default 20-step run samples
  over the path -- equals the plain frames bit for bit.
"""
import os

import numpy as np
import pytest

from oracle import xo
from xenodon_amd import abi, synth, workloads

pytestmark = pytest.mark.gpu

W, H = 1920, 1080
EMISSION = 1.6
CAMERA_FILE = os.path.join(os.path.dirname(__file__), "golden", "camera.txt")
FRAMES = [0, 60, 74, 90, 111, 112, 149]
# (x, y, w, h) regions checked against the oracle: 2-row bands down the
# image, and a 96 x 64 block at the centre
REGIONS = [(0, y, W, 2) for y in (0, 216, 432, 538, 648, 864, 1078)] + [(912, 508, 96, 64)]
ALL = [abi.DDA, abi.SVO_NAIVE, abi.ESVO, abi.SVO_DF, abi.SVO_ROPE]


def _cameras():
    v = np.array(open(CAMERA_FILE).read().split(), np.float32).reshape(-1, 9)
    return [(r[0:3], r[3:6], r[6:9]) for r in v]


@pytest.fixture(scope="module")
def c2():
    if abi.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    grid = synth.ct_body()
    assert grid.shape == (512, 361, 512, 4)
    nodes, st = abi.build_octree(grid, tree=abi.TREE_ROPE)
    assert st["side"] == 512 and st["depth"] == 9
    ctx = abi.Context(0)
    gv = ctx.grid_volume(grid)
    sv = ctx.svo_volume(nodes, st["side"])
    ovs = {"grid": xo.OracleVolume(grid=grid), "svo": xo.OracleVolume(nodes=nodes, side=st["side"])}
    yield ctx, gv, sv, ovs
    ctx.close()


def _vol(t, gv, sv, ovs):
    if t == abi.DDA:
        return gv, [512, 361, 512], ovs["grid"]
    return sv, [512] * 3, ovs["svo"]


@pytest.mark.timeout(600)
@pytest.mark.parametrize("t", ALL, ids=[abi.TRAVERSAL_NAMES[t] for t in ALL])
def test_c2_plain_frames_match_oracle(c2, t):
    ctx, gv, sv, ovs = c2
    vol, md, ov = _vol(t, gv, sv, ovs)
    cams = _cameras()
    threads = min(16, os.cpu_count() or 1)
    lit = 0
    for k in FRAMES:
        f, up, eye = cams[k]
        img, _ = ctx.render_image(vol, t, abi.make_camera(f, up, eye), abi.make_uniforms(W, H, md, emission=EMISSION))
        assert (img[..., 3] == 255).all()
        for (x, y, w, h) in REGIONS:
            u = xo.make_uniforms(w, h, md, emission=EMISSION, out_offset=(x, y), disp_extent=(W, H))
            ref = xo.render(t, ov, xo.make_camera(f, up, eye), u, threads=threads)["rgba"]
            got = img[y:y + h, x:x + w]
            if not np.array_equal(got, ref):
                d = np.abs(got.astype(int) - ref.astype(int))
                raise AssertionError(f"{abi.TRAVERSAL_NAMES[t]} frame {k} region {(x, y, w, h)}: "
                                     f"{int((d > 0).any(-1).sum())} pixels differ, max {int(d.max())} LSB")
            lit += int(got[..., :3].any())
    assert lit >= len(FRAMES) * 3  # the volume is in view on most checked regions


@pytest.mark.timeout(600)
def test_c2_default_batched_launch_matches_plain(c2):
    """The launch `bench.py` times by default: groups of ten frames (the
    frames a 20-step run spreads over the path) x five traversals in one
    persistent all-traversal kernel at the library's default launch shape."""
    ctx, gv, sv, ovs = c2
    cams = _cameras()
    steps, group = 20, 10
    b = ctx.batch()
    tun = b.tuning()
    assert tun["class_kernels"] == 0 and tun["waves_per_simd"] == 8  # the all-traversal kernel, k_batch_w8<5>
    frames = [ctx.frame(W, H) for _ in range(group * len(ALL))]
    for k0 in range(0, steps, group):
        views = []
        for j in range(group):
            f, up, eye = cams[workloads.path_frame(k0 + j, steps, len(cams))]
            cam = abi.make_camera(f, up, eye)
            for i, t in enumerate(ALL):
                vol, md, _ = _vol(t, gv, sv, ovs)
                views.append((vol, t, cam, abi.make_uniforms(W, H, md, emission=EMISSION), frames[j * len(ALL) + i]))
        b.set_views(views)
        ms, rays = b.render()
        assert rays == len(views) * W * H and ms > 0
        for j, (vol, t, cam, uni, fr) in enumerate(views):
            got = fr.download()
            ref, _ = ctx.render_image(vol, t, cam, uni)
            if not np.array_equal(got, ref):
                d = np.abs(got.astype(int) - ref.astype(int))
                raise AssertionError(f"{abi.TRAVERSAL_NAMES[t]} step {k0 + j // len(ALL)}: "
                                     f"{int((d > 0).any(-1).sum())} pixels differ, max {int(d.max())} LSB")
    for fr in frames:
        fr.free()
    b.free()

"""GPU parity: the HIP kernels (through the C-ABI) against the oracle's FP32
restatement of the reference shaders, on the same seeded inputs.

The kernels restate the shaders operation by operation with IEEE single
precision, so frames are expected to be bit-identical to the oracle.  The
looser tolerance stated for the shaders on a Vulkan CPU driver (<= 1 LSB per
8-bit channel on >= 99.9 % of pixels, never more than 2 LSB) is asserted as
well, and the exact-match fraction is reported.
"""
import numpy as np
import pytest

from oracle import xo
from tests import helpers as H
from xenodon_amd import abi, synth

pytestmark = pytest.mark.gpu

TRAVERSALS = [abi.DDA, abi.SVO_NAIVE, abi.ESVO, abi.SVO_DF, abi.SVO_ROPE]
C1_EMISSION = 4.0


@pytest.fixture(scope="module")
def ctx():
    if abi.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    c = abi.Context(0)
    yield c
    c.close()


def _compare(gpu, ref):
    d = np.abs(gpu.astype(np.int32) - ref.astype(np.int32))[..., :3]
    return {"exact": float((d == 0).all(-1).mean()), "le1": float((d <= 1).mean()), "max": int(d.max())}


def _assert_parity(gpu, ref, label, exact=True):
    s = _compare(gpu, ref)
    print(f"{label}: exact {s['exact']:.6f} <=1LSB {s['le1']:.6f} max {s['max']}")
    assert (gpu[..., 3] == 255).all()
    assert s["max"] <= 2 and s["le1"] >= 0.999, (label, s)
    if exact:
        assert s["exact"] == 1.0, (label, s)


class Scene:
    """A grid, its lossless rope octree, and matching oracle/GPU volumes."""

    def __init__(self, ctx, grid):
        self.grid = grid
        nodes, st = xo.build_octree(grid, tree=xo.TREE_ROPE)
        self.nodes, self.side = nodes, st["side"]
        self.o_grid = xo.OracleVolume(grid=grid)
        self.o_svo = xo.OracleVolume(nodes=nodes, side=self.side)
        self.g_grid = ctx.grid_volume(grid)
        self.g_svo = ctx.svo_volume(nodes, self.side)

    def pick(self, t):
        if t == abi.DDA:
            return self.o_grid, self.g_grid, list(self.grid.shape[2::-1][:3])
        return self.o_svo, self.g_svo, [self.side] * 3


def _render_both(ctx, scene, t, f, up, eye, W, Hh, emission, **kw):
    ov, gv, md = scene.pick(t)
    ocam = xo.make_camera(f, up, eye)
    gcam = abi.make_camera(f, up, eye)
    ou = xo.make_uniforms(W, Hh, md, emission=emission, **kw)
    gu = abi.make_uniforms(W, Hh, md, emission=emission, **kw)
    ref = xo.render(t, ov, ocam, ou)["rgba"]
    img, ms = ctx.render_image(gv, t, gcam, gu)
    return img, ref


@pytest.fixture(scope="module")
def small_scenes(ctx):
    return {
        "rand16": Scene(ctx, H.random_grid((16, 16, 16), density=0.35, seed=21)),
        "rand_odd": Scene(ctx, H.random_grid((7, 9, 13), density=0.4, seed=22)),
        "blocky32": Scene(ctx, H.blocky_grid(32, 4, seed=23)),
    }


@pytest.mark.parametrize("t", TRAVERSALS)
@pytest.mark.parametrize("scene_name", ["rand16", "rand_odd", "blocky32"])
def test_small_volumes_bit_exact(ctx, small_scenes, scene_name, t):
    scene = small_scenes[scene_name]
    cams = H.orbit_cameras(3, 1.8, seed=5) + [H.look_at((0.45, 0.55, 0.4), (0.9, 0.2, 0.7))]
    for i, (f, up, eye) in enumerate(cams):
        for (W, Hh) in ((64, 48), (37, 29)):
            img, ref = _render_both(ctx, scene, t, f, up, eye, W, Hh, emission=0.9)
            _assert_parity(img, ref, f"{scene_name}/{abi.TRAVERSAL_NAMES[t]}/cam{i}/{W}x{Hh}")


@pytest.fixture(scope="module")
def c1_scene(ctx):
    return Scene(ctx, synth.shells_128())


def c1_orbit(n=10):
    """10 frames on a circular orbit at twice the unit-cube diagonal."""
    cams = []
    for k in range(n):
        th = 2 * np.pi * k / n
        eye = np.array([0.5, 0.5, 0.5]) + 2 * np.sqrt(3) * np.array([np.sin(th), 0.0, -np.cos(th)])
        cams.append(H.look_at(eye, (0.5, 0.5, 0.5)))
    return cams


@pytest.mark.parametrize("t", TRAVERSALS)
def test_c1_shells_orbit_bit_exact(ctx, c1_scene, t):
    """128^3 shells volume, 256x256, the 10-frame orbit (plain DDA workload;
    the octree traversals run on its lossless rope SVO)."""
    for k, (f, up, eye) in enumerate(c1_orbit()):
        img, ref = _render_both(ctx, c1_scene, t, f, up, eye, 256, 256, emission=C1_EMISSION)
        _assert_parity(img, ref, f"c1/{abi.TRAVERSAL_NAMES[t]}/frame{k}")


def _saturated_fraction(img):
    return float((img[..., :3] == 255).any(-1).mean())


def test_c1_emission_saturates_at_most_one_percent(ctx, c1_scene):
    """The 128^3 workload's emission coefficient (4.0) leaves at most 1 % of
    the pixels of the orbit with a clamped channel, for every traversal."""
    for t in TRAVERSALS:
        _, gv, md = c1_scene.pick(t)
        sat = [_saturated_fraction(ctx.render_image(gv, t, abi.make_camera(f, up, eye),
                                                    abi.make_uniforms(256, 256, md, emission=C1_EMISSION))[0])
               for f, up, eye in c1_orbit()]
        print(f"{abi.TRAVERSAL_NAMES[t]}: saturated pixels {max(sat):.4%} (worst frame)")
        assert max(sat) <= 0.01, (abi.TRAVERSAL_NAMES[t], sat)


@pytest.mark.parametrize("t", TRAVERSALS)
def test_saturating_emission_bit_exact(ctx, c1_scene, t):
    """A saturating run (emission 40): the store's clamp to [0, 1] before the
    unorm8 conversion matches the oracle when most of the volume's pixels
    clamp."""
    f, up, eye = c1_orbit()[1]
    img, ref = _render_both(ctx, c1_scene, t, f, up, eye, 128, 128, emission=40.0)
    _assert_parity(img, ref, f"saturating/{abi.TRAVERSAL_NAMES[t]}")
    covered = float((img[..., :3] > 0).any(-1).mean())
    assert _saturated_fraction(img) > 0.5 * covered > 0


@pytest.mark.parametrize("t", TRAVERSALS)
def test_camera_inside_volume(ctx, c1_scene, t):
    f, up, eye = H.look_at((0.52, 0.47, 0.45), (0.1, 0.9, 0.95))
    img, ref = _render_both(ctx, c1_scene, t, f, up, eye, 160, 120, emission=C1_EMISSION)
    _assert_parity(img, ref, f"inside/{abi.TRAVERSAL_NAMES[t]}")


@pytest.mark.parametrize("t", TRAVERSALS)
def test_output_regions_stitch_to_display(ctx, c1_scene, t):
    """Two outputs splitting one display (multi-device sort-first layout of
    the headless config) render exactly the pixels of the full display."""
    f, up, eye = c1_orbit()[3]
    W, Hh = 96, 64
    full, _ = _render_both(ctx, c1_scene, t, f, up, eye, W, Hh, emission=C1_EMISSION)
    _, gv, md = c1_scene.pick(t)
    cam = abi.make_camera(f, up, eye)
    left_u = abi.make_uniforms(40, Hh, md, emission=C1_EMISSION, out_offset=(0, 0), disp_extent=(W, Hh))
    right_u = abi.make_uniforms(W - 40, Hh, md, emission=C1_EMISSION, out_offset=(40, 0), disp_extent=(W, Hh))
    left, _ = ctx.render_image(gv, t, cam, left_u)
    right, _ = ctx.render_image(gv, t, cam, right_u)
    assert np.array_equal(np.concatenate([left, right], axis=1), full)


@pytest.mark.parametrize("t", TRAVERSALS)
def test_block_order_and_tile_shape_keep_frames_bit_exact(ctx, c1_scene, t):
    """Plain launches in every XCD block order (dispatch order, one band per
    XCD, interleaved runs) and every wave tile shape give the oracle's frame
    (200x120 leaves partial blocks and runs at the image's end)."""
    f, up, eye = c1_orbit()[6]
    ov, gv, md = c1_scene.pick(t)
    ref = xo.render(t, ov, xo.make_camera(f, up, eye), xo.make_uniforms(200, 120, md, emission=C1_EMISSION))["rgba"]
    try:
        for chunks in (0, 1, 3, 64):
            for tile in (8, 16, 32):
                ctx.set_option("xcd_chunks", chunks)
                ctx.set_tile(tile)
                img, _ = ctx.render_image(gv, t, abi.make_camera(f, up, eye),
                                          abi.make_uniforms(200, 120, md, emission=C1_EMISSION))
                _assert_parity(img, ref, f"order{chunks}/tile{tile}/{abi.TRAVERSAL_NAMES[t]}")
    finally:
        ctx.set_option("xcd_chunks", 0)
        ctx.set_tile(8)
    with pytest.raises(abi.XdError):
        ctx.set_option("xcd_chunks", 65)


@pytest.mark.parametrize("t", TRAVERSALS)
def test_word_counts_match_oracle(ctx, small_scenes, c1_scene, t):
    """The instrumented kernel counts exactly the 32-bit volume words the
    reference algorithm reads (the roofline's algorithmic bytes)."""
    for scene in (small_scenes["blocky32"], c1_scene):
        ov, gv, md = scene.pick(t)
        f, up, eye = c1_orbit()[2]
        ocam, gcam = xo.make_camera(f, up, eye), abi.make_camera(f, up, eye)
        ou = xo.make_uniforms(80, 60, md, emission=1.0)
        gu = abi.make_uniforms(80, 60, md, emission=1.0)
        ref = int(xo.render(t, ov, ocam, ou, words=True)["words"].sum())
        got = ctx.count_words(gv, t, gcam, gu)
        assert got == ref, (abi.TRAVERSAL_NAMES[t], got, ref)


def test_incompatible_traversal_is_rejected(ctx, small_scenes):
    scene = small_scenes["rand16"]
    cam = abi.make_camera((0, 0, 1), (0, 1, 0), (0.5, 0.5, -1.5))
    u = abi.make_uniforms(8, 8, [16, 16, 16])
    with pytest.raises(abi.XdError) as e:
        ctx.render_image(scene.g_svo, abi.DDA, cam, u)
    assert e.value.status == 4
    assert str(e.value) == "Shader 'dda' is incompatible with model type 'svo' (requires 'tiff')"
    with pytest.raises(abi.XdError) as e:
        ctx.render_image(scene.g_grid, abi.ESVO, cam, u)
    assert str(e.value) == "Shader 'esvo' is incompatible with model type 'tiff' (requires 'svo')"


def test_rope_traversal_on_plain_tree_matches_reference_behaviour(ctx, small_scenes):
    """svo-rope on a tree without ropes integrates only the first leaf (leaf
    children are 0 = the root, which ends svo_rope.comp's loop)."""
    g = small_scenes["blocky32"].grid
    nodes, st = xo.build_octree(g, tree=xo.TREE_SPARSE)
    ov = xo.OracleVolume(nodes=nodes, side=st["side"])
    gv = ctx.svo_volume(nodes, st["side"])
    assert gv.info().is_rope == 0
    f, up, eye = H.look_at((1.4, 0.8, -1.0), (0.5, 0.5, 0.5))
    ref = xo.render(xo.SVO_ROPE, ov, xo.make_camera(f, up, eye), xo.make_uniforms(48, 48, [32] * 3, emission=2.0))
    img, _ = ctx.render_image(gv, abi.SVO_ROPE, abi.make_camera(f, up, eye),
                              abi.make_uniforms(48, 48, [32] * 3, emission=2.0))
    assert np.array_equal(img, ref["rgba"])


def test_malformed_octree_is_rejected(ctx):
    nodes = np.zeros((2, 10), np.uint32)
    nodes[0, :8] = 7  # child out of range
    with pytest.raises(abi.XdError) as e:
        ctx.svo_volume(nodes, 2)
    assert e.value.status == 5
    nodes[0, :8] = 0  # internal root pointing to itself: a cycle
    with pytest.raises(abi.XdError) as e:
        ctx.svo_volume(nodes, 2)
    assert "cycle" in str(e.value)


@pytest.mark.parametrize("layout", [1, 2])
@pytest.mark.parametrize("t", [abi.SVO_NAIVE, abi.ESVO, abi.SVO_DF, abi.SVO_ROPE])
def test_raw_and_wide_octree_layouts_bit_exact(t, layout):
    """Trees above 2^31 nodes keep the .svo node words on the device (layout
    1), trees of 2^30..2^31 nodes get 31-bit child indices without the empty
    tag (layout 2); the context option forces either layout on a small tree
    so their kernels are checked here."""
    with abi.Context(0) as ctx:
        ctx.set_option("octree_layout", layout)
        grid = synth.shells_128()
        nodes, st = xo.build_octree(grid, tree=xo.TREE_ROPE)
        gv = ctx.svo_volume(nodes, st["side"])
        ov = xo.OracleVolume(nodes=nodes, side=st["side"])
        for k, (f, up, eye) in enumerate(c1_orbit()[:4]):
            u_args = (200, 150, [st["side"]] * 3)
            img, _ = ctx.render_image(gv, t, abi.make_camera(f, up, eye), abi.make_uniforms(*u_args, emission=4.0))
            ref = xo.render(t, ov, xo.make_camera(f, up, eye), xo.make_uniforms(*u_args, emission=4.0))
            _assert_parity(img, ref["rgba"], f"layout{layout}/{abi.TRAVERSAL_NAMES[t]}/frame{k}")
            words = ctx.count_words(gv, t, abi.make_camera(f, up, eye), abi.make_uniforms(*u_args, emission=4.0))
            ref_w = xo.render(t, ov, xo.make_camera(f, up, eye), xo.make_uniforms(*u_args, emission=4.0),
                              words=True)["words"].sum()
            assert words == int(ref_w)


@pytest.mark.parametrize("t", [abi.SVO_NAIVE, abi.ESVO, abi.SVO_DF, abi.SVO_ROPE])
def test_tree_with_inexact_depth_fields_bit_exact(ctx, t):
    """svo-df sizes a popped node from its depth field (svo_df.comp:58) and
    svo-rope sizes a rope target the same way (svo_rope.comp:30); a tree whose
    internal depth fields are off renders exactly as the shaders would, and
    the device flags it so no depth is taken from the tree's structure."""
    g = H.blocky_grid(32, 2, seed=17)
    nodes, st = xo.build_octree(g, tree=xo.TREE_ROPE)
    bad = nodes.copy()
    internal = np.nonzero((bad[:, 9] & 0x80000000) == 0)[0][1:]  # all but the root
    assert internal.size > 4
    bad[internal[::2], 9] += 1
    good_v = ctx.svo_volume(nodes, st["side"])
    bad_v = ctx.svo_volume(bad, st["side"])
    assert good_v.info().depth_fields_exact == 1 and bad_v.info().depth_fields_exact == 0
    ov = xo.OracleVolume(nodes=bad, side=st["side"])
    for k, (f, up, eye) in enumerate(H.orbit_cameras(3, 1.6, seed=12)):
        args = (64, 48, [st["side"]] * 3)
        img, _ = ctx.render_image(bad_v, t, abi.make_camera(f, up, eye), abi.make_uniforms(*args, emission=2.0))
        ref = xo.render(t, ov, xo.make_camera(f, up, eye), xo.make_uniforms(*args, emission=2.0))["rgba"]
        _assert_parity(img, ref, f"inexact-depth/{abi.TRAVERSAL_NAMES[t]}/cam{k}")


@pytest.mark.parametrize("t", [abi.SVO_NAIVE, abi.ESVO, abi.SVO_DF, abi.SVO_ROPE])
def test_dag_and_single_leaf_trees_bit_exact(ctx, t):
    """--dag trees share subtrees between parents (the device renumbering
    keeps a shared node's first index); a uniform volume is one leaf root."""
    grids = {"dag": H.blocky_grid(64, 4, seed=31), "uniform": np.full((16, 16, 16, 4), 90, np.uint8)}
    for name, g in grids.items():
        nodes, st = xo.build_octree(g, tree=xo.TREE_DAG if name == "dag" else xo.TREE_ROPE)
        if name == "dag":
            sparse, _ = xo.build_octree(g)
            assert nodes.shape[0] < sparse.shape[0]
        else:
            assert nodes.shape[0] == 1
        gv = ctx.svo_volume(nodes, st["side"])
        ov = xo.OracleVolume(nodes=nodes, side=st["side"])
        for k, (f, up, eye) in enumerate(H.orbit_cameras(2, 1.7, seed=9) + [H.look_at((0.4, 0.6, 0.5), (1, 0, 1))]):
            args = (72, 56, [st["side"]] * 3)
            img, _ = ctx.render_image(gv, t, abi.make_camera(f, up, eye), abi.make_uniforms(*args, emission=1.5))
            ref = xo.render(t, ov, xo.make_camera(f, up, eye), xo.make_uniforms(*args, emission=1.5))["rgba"]
            _assert_parity(img, ref, f"{name}/{abi.TRAVERSAL_NAMES[t]}/cam{k}")


def _chain_tree(depth, target, seed):
    """A tree whose every level splits once more towards `target`: at each
    depth the node holds seven leaves (some with zero RGB) and one child that
    continues, down to a coloured leaf at `depth`.  Node words as the .svo
    format has them: 8 children, RGBA colour, leaf bit | depth."""
    rng = np.random.default_rng(seed)
    rows = [[0] * 10]
    cur = 0
    for d in range(depth):
        k = sum((int(target[a] * (1 << (d + 1))) & 1) << (2 - a) for a in range(3))
        for c in range(8):
            rows.append([0] * 10)
            idx = len(rows) - 1
            rows[cur][c] = idx
            if c == k and d + 1 < depth:
                rows[idx][9] = d + 1  # internal, exact depth field
                nxt = idx
            else:
                rgb = 0 if rng.random() < 0.3 else int(rng.integers(1, 1 << 24))
                rows[idx][8] = 0xFF000000 | rgb
                rows[idx][9] = 0x80000000 | (d + 1)
        if d + 1 < depth:
            cur = nxt
    return np.array(rows, np.uint32)


@pytest.mark.parametrize("t", [abi.SVO_NAIVE, abi.ESVO, abi.SVO_DF, abi.SVO_ROPE])
def test_deep_tree_bit_exact(ctx, t):
    """A tree 22 levels deep (side 2^22; the device accepts up to 23), seen
    from an eye inside its deepest cells: every descent, pop and rope step
    down to the finest level, where the corners and sides are the smallest
    dyadic values of the walk, renders as the oracle does."""
    depth, target = 22, (0.3712345, 0.6123457, 0.5317891)
    nodes = _chain_tree(depth, target, seed=5)
    side = 1 << depth
    if t == abi.SVO_ROPE:
        xo.lib().xo_generate_ropes(nodes.ctypes.data, nodes.shape[0], side)
    gv = ctx.svo_volume(nodes, side)
    assert gv.info().depth_fields_exact == 1
    ov = xo.OracleVolume(nodes=nodes, side=side)
    eyes = [np.array(target) + np.array(o) for o in [(3e-7, 2e-7, -4e-7), (-2e-6, 1e-6, 3e-6), (0.4, 0.3, -0.5)]]
    for k, eye in enumerate(eyes):
        f, up, e = H.look_at(tuple(eye), target)
        args = (64, 48, [side] * 3)
        img, _ = ctx.render_image(gv, t, abi.make_camera(f, up, e), abi.make_uniforms(*args, emission=2.0))
        ref = xo.render(t, ov, xo.make_camera(f, up, e), xo.make_uniforms(*args, emission=2.0))["rgba"]
        _assert_parity(img, ref, f"deep/{abi.TRAVERSAL_NAMES[t]}/eye{k}")
        assert ref[..., :3].any()

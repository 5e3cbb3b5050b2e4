"""Independent pins of the oracle's image geometry, voxel-ratio branch, store
rounding and word counters.

Each check raises AssertionError when the oracle disagrees with a statement
derived without the oracle's own ray() (the pinhole model in tests/helpers.py,
hand-counted reads, the Vulkan unorm8 rule).  test_oracle_projection.py runs
them on the oracle; test_oracle_mutants.py runs them on oracle builds with one
deliberate mistake planted (oracle/xo_trace_body.h, XO_MUTANT) and shows that
each mistake fails the checks named for it in MUTANT_CATCHES.

Camera conventions under test: resources/common.glsl:45-56 (ray), :68-72
(voxel_emission_coeff), the per-shader main()s (dda.comp:55-73,
svo_naive.comp:73-89, svo_df.comp:69-85, esvo.comp:138-157,
svo_rope.comp:137-153), and the host's translation / voxel_ratio
(src/render/Renderer.cpp:62).
"""
import numpy as np
from scipy.spatial import ConvexHull

from oracle import xo
from tests import helpers as H

ALL = [xo.DDA, xo.SVO_NAIVE, xo.ESVO, xo.SVO_DF, xo.SVO_ROPE]
COLOUR = np.array([200, 120, 40])


def _uniform_volumes(side=8, colour=COLOUR):
    g = np.zeros((side, side, side, 4), np.uint8)
    g[..., :3] = colour
    g[..., 3] = 255
    nodes, st = xo.build_octree(g, tree=xo.TREE_ROPE)
    return xo.OracleVolume(grid=g), xo.OracleVolume(nodes=nodes, side=st["side"])


def _vol(t, gv, sv):
    return gv if t == xo.DDA else sv


def check_chord_field(traversals=ALL):
    """A uniform cube seen from an eye inside it: every pixel must equal
    c * e * (distance from the eye to the cube's surface along the pinhole
    direction of that pixel).  The eye is off-centre and the camera tilted,
    and the display is not square, so a mirrored or flipped image, a shift of
    half a pixel, the aspect applied to the wrong axis or another field of
    view all change the chord at most pixels by far more than FP32 error."""
    gv, sv = _uniform_volumes()
    eye = np.array([0.31, 0.62, 0.44])
    f, up, _ = H.look_at(eye, (0.93, 0.35, 0.81))
    W, Hh, e = 40, 24, 0.5
    cam = xo.make_camera(f, up, eye)
    for t in traversals:
        vol = _vol(t, gv, sv)
        u = xo.make_uniforms(W, Hh, vol.model_dim, emission=e)
        got = xo.render(t, vol, cam, u, linear=True)["linear"].astype(np.float64)
        exp = np.zeros_like(got)
        for j in range(Hh):
            for i in range(W):
                d = H.pinhole_dir(f, up, W, Hh, i, j)
                exp[j, i] = COLOUR / 255.0 * e * H.exit_distance(eye, d, (1, 1, 1))
        err = np.abs(got - exp) / exp
        assert err.max() < 2e-5, (t, float(err.max()), np.unravel_index(err.argmax(), err.shape))


def check_centre_pixel_is_forward(traversals=ALL):
    """On an even W x H display, pixel (W/2, H/2) samples uv = (1/2, 1/2), the
    plane centre, so its ray is `forward` itself: its chord is the distance to
    the surface along `forward`.  A pixel-centre (+0.5) convention moves it."""
    gv, sv = _uniform_volumes()
    eye = np.array([0.27, 0.58, 0.36])
    f, up, _ = H.look_at(eye, (0.9, 0.7, 0.95))
    W, Hh, e = 32, 20, 1.0
    cam = xo.make_camera(f, up, eye)
    expect = COLOUR / 255.0 * e * H.exit_distance(eye, f, (1, 1, 1))
    for t in traversals:
        vol = _vol(t, gv, sv)
        u = xo.make_uniforms(1, 1, vol.model_dim, emission=e, out_offset=(W // 2, Hh // 2), disp_extent=(W, Hh))
        got = xo.render(t, vol, cam, u, linear=True)["linear"][0, 0].astype(np.float64)
        assert np.allclose(got, expect, rtol=2e-6), (t, got, expect)


def check_edge_pixel_angle(traversals=ALL):
    """The camera looks along +z from the centre of a uniform cube.  Pixel
    (0, H/2) sits on the left edge of the image plane, half a unit from its
    centre, so its ray makes the angle atan(1/2) with forward and leaves the
    cube through z = 1 after 0.5 / cos(atan(1/2)) = 0.5 * sqrt(5/4).  Pixel
    (W/2, 0) is (H/2)/W units above the centre: chord 0.5 * sqrt(1 + (H/2W)^2).
    A 60 degree field of view, or the aspect applied to x, changes them."""
    gv, sv = _uniform_volumes()
    eye = np.array([0.5, 0.5, 0.5])
    f, up = (0.0, 0.0, 1.0), (0.0, 1.0, 0.0)
    W, Hh, e = 40, 24, 1.0
    cam = xo.make_camera(f, up, eye)
    cases = {(0, Hh // 2): 0.5 * np.sqrt(1.25), (W // 2, 0): 0.5 * np.sqrt(1 + (Hh / (2.0 * W)) ** 2)}
    for t in traversals:
        vol = _vol(t, gv, sv)
        for (i, j), chord in cases.items():
            u = xo.make_uniforms(1, 1, vol.model_dim, emission=e, out_offset=(i, j), disp_extent=(W, Hh))
            got = xo.render(t, vol, cam, u, linear=True)["linear"][0, 0].astype(np.float64)
            assert np.allclose(got, COLOUR / 255.0 * e * chord, rtol=2e-6), (t, (i, j), got)


def check_single_voxel_projection(traversals=ALL):
    """One lit voxel of a 16^3 grid, seen from inside the grid: the lit pixels
    must be exactly those whose sample point (the pixel corner) falls inside
    the forward projection of the voxel's box onto the image plane (pixels
    within 0.02 px of the projected outline are not judged).  This pins the
    image orientation -- row 0 on the side `up` points to, columns along
    up x forward -- and the pixel-corner convention on a small target."""
    g = np.zeros((16, 16, 16, 4), np.uint8)
    vx, vy, vz = 11, 4, 12
    g[vz, vy, vx] = (255, 255, 255, 255)
    nodes, st = xo.build_octree(g, tree=xo.TREE_ROPE)
    gv, sv = xo.OracleVolume(grid=g), xo.OracleVolume(nodes=nodes, side=st["side"])
    eye = np.array([0.15, 0.8, 0.1])
    centre = (np.array([vx, vy, vz]) + 0.5) / 16.0
    # look past the voxel so that it lands left of and below the image centre
    # (near pixel (26.7, 27.9) of 64 x 40), where a mirror or a flip moves it
    f, up, _ = H.look_at(eye, centre + (0.12, 0.1, 0.0))
    W, Hh = 64, 40
    corners = [(np.array([vx + a, vy + b, vz + c]) / 16.0) for a in (0, 1) for b in (0, 1) for c in (0, 1)]
    pts = np.array([H.project_point(f, up, eye, W, Hh, p) for p in corners])
    hull = ConvexHull(pts)
    jj, ii = np.mgrid[0:Hh, 0:W]
    q = np.stack([ii.ravel(), jj.ravel(), np.ones(ii.size)], 1).astype(np.float64)
    dist = (q @ hull.equations.T).max(1).reshape(Hh, W)  # < 0 inside
    inside, outside = dist < -0.02, dist > 0.02
    assert inside.sum() >= 20, "the voxel must cover a few pixels"
    cam = xo.make_camera(f, up, eye)
    for t in traversals:
        vol = _vol(t, gv, sv)
        u = xo.make_uniforms(W, Hh, vol.model_dim, emission=1.0)
        lit = xo.render(t, vol, cam, u, linear=True)["linear"].max(-1) > 0
        assert lit[inside].all() and not lit[outside].any(), (t, int((~lit[inside]).sum()),
                                                               int(lit[outside].sum()))


def check_voxel_ratio_world_chord(traversals=ALL):
    """A uniform volume with voxel_ratio r spans [0, r] in camera space
    (render.txt:33-41); the host passes translation / r (Renderer.cpp:62).
    The pixel must be c * e * (chord of the camera-space ray in [0, r]): the
    ray is normalised in model space and the emission coefficient restores
    the camera-space length (common.glsl:54-55,68-72)."""
    gv, sv = _uniform_volumes()
    r = np.array([1.0, 0.5, 2.0])
    eye = np.array([0.3, 0.31, 0.7])  # inside [0, r]
    f, up, _ = H.look_at(eye, (0.95, 0.05, 1.9))
    W, Hh, e = 24, 16, 0.7
    cam = xo.make_camera(f, up, eye / r)
    for t in traversals:
        vol = _vol(t, gv, sv)
        u = xo.make_uniforms(W, Hh, vol.model_dim, emission=e, voxel_ratio=r)
        got = xo.render(t, vol, cam, u, linear=True)["linear"].astype(np.float64)
        for j in range(Hh):
            for i in range(W):
                d = H.pinhole_dir(f, up, W, Hh, i, j)
                exp = COLOUR / 255.0 * e * H.exit_distance(eye, d, r)
                assert np.allclose(got[j, i], exp, rtol=2e-5), (t, (i, j), got[j, i], exp)


def check_store_rounding(traversals=ALL):
    """imageStore into an rgba8 image converts with the Vulkan unorm rule:
    clamp to [0, 1], times 255, round to nearest.  Checked on every pixel of
    a random volume whose values cover the whole range; exact halves (where
    the rule allows either neighbour) are not judged."""
    g = H.random_grid((12, 12, 12), density=0.5, seed=21)
    nodes, st = xo.build_octree(g, tree=xo.TREE_ROPE)
    gv, sv = xo.OracleVolume(grid=g), xo.OracleVolume(nodes=nodes, side=st["side"])
    f, up, eye = H.look_at((1.4, 1.1, -0.7), (0.5, 0.45, 0.5))
    cam = xo.make_camera(f, up, eye)
    for t in traversals:
        vol = _vol(t, gv, sv)
        u = xo.make_uniforms(48, 40, vol.model_dim, emission=1.3)
        out = xo.render(t, vol, cam, u, linear=True)
        lin = out["linear"].astype(np.float64)
        scaled = np.clip(lin, 0.0, 1.0) * 255.0
        tie = np.abs(scaled - np.floor(scaled) - 0.5) < 1e-3
        want = H.unorm8(lin)
        got = out["rgba"][..., :3].astype(np.int32)
        ok = (got == want) | tie
        assert ok.all(), (t, int((~ok).sum()))
        assert (out["rgba"][..., 3] == 255).all()
        frac = scaled - np.floor(scaled)
        assert ((frac > 0.6) & (frac < 0.95)).sum() > 100, "needs values that truncation would get wrong"


def _two_level_tree():
    """Hand-built tree: root (node 0) with an internal child 0 (node 1, eight
    leaves 9..16 at depth 2) and leaf children 1..7 (nodes 2..8, depth 1).
    Ropes along +z for the leaves a +z ray through x = y = 1/8 meets:
    9 -> 10 -> 2 -> none (children[4] is the +z rope, Octree.cpp:192-199)."""
    LEAF = 1 << 31
    nodes = np.zeros((17, 10), np.uint32)
    nodes[0, :8] = np.arange(1, 9)
    nodes[0, 9] = 0
    nodes[1, :8] = np.arange(9, 17)
    nodes[1, 9] = 1
    for n in range(2, 9):
        nodes[n, 8] = 0x00A0B0C0
        nodes[n, 9] = LEAF | 1
    for n in range(9, 17):
        nodes[n, 8] = 0x00304050
        nodes[n, 9] = LEAF | 2
    nodes[9, 4] = 10
    nodes[10, 4] = 2
    return nodes


def check_word_counts_hand_counted():
    """Words each octree traversal reads for one +z ray through x = y = 1/8
    of the two-level tree above, counted by hand from the shaders:
      svo-naive (svo_naive.comp:15,25,66): 3 steps; leaf 9 and leaf 10 are
        found with root ild + child + node 1 ild + child + leaf ild, then the
        colour (6 each); leaf 2 with root ild + child + leaf ild + colour (4):
        16.
      svo-df (svo_df.comp:25,33,34,58): root child 0 = child word + ild (2);
        in node 1, children 0 and 1 are hit leaves (child + ild + colour = 3
        each) and children 2..7 miss (1 each): 12; the pop back to the root
        reads the root's ild (1); root child 1 is a hit leaf (3) and children
        2..7 miss (6): 2 + 12 + 1 + 3 + 6 = 24.
      esvo (esvo.comp:73,75,76): root child 0 = child + ild (2), then leaves
        9, 10 and 2 at child + ild + colour (3 each): 11.
      svo-rope (svo_rope.comp:15,25,34,44,30,104,109,124,130): find to leaf
        9 (5) + colour + rope (7); leaf 10 via find_relative = ild twice +
        colour + rope (4); leaf 2 the same (4): 15.
    The ray is the exact forward direction (pixel (1, 1) of a 2 x 2 display),
    which adjust_ray turns into (2^-23, 2^-23, 1)."""
    nodes = _two_level_tree()
    vol = xo.OracleVolume(nodes=nodes, side=4)
    cam = xo.make_camera((0.0, 0.0, 1.0), (0.0, 1.0, 0.0), (0.125, 0.125, -1.0))
    u = xo.make_uniforms(1, 1, [4, 4, 4], out_offset=(1, 1), disp_extent=(2, 2))
    expect = {xo.SVO_NAIVE: 16, xo.SVO_DF: 24, xo.ESVO: 11, xo.SVO_ROPE: 15}
    for t, n in expect.items():
        out = xo.render(t, vol, cam, u, linear=True, words=True)
        assert int(out["words"][0, 0]) == n, (t, int(out["words"][0, 0]), n)
        # the ray crosses leaves 9, 10 (0.25 each) and 2 (0.5)
        want = 0.25 * (np.array([0x50, 0x40, 0x30]) / 255.0) * 2 + 0.5 * np.array([0xC0, 0xB0, 0xA0]) / 255.0
        assert np.allclose(out["linear"][0, 0], want, rtol=1e-5), (t, out["linear"][0, 0])


CHECKS = {
    "chord_field": check_chord_field,
    "centre_pixel": check_centre_pixel_is_forward,
    "edge_angle": check_edge_pixel_angle,
    "single_voxel": check_single_voxel_projection,
    "voxel_ratio": check_voxel_ratio_world_chord,
    "store_rounding": check_store_rounding,
    "word_counts": check_word_counts_hand_counted,
}

# mutant id (oracle/xo_trace_body.h) -> the checks that must fail under it
MUTANT_CATCHES = {
    1: ["chord_field", "single_voxel"],                 # right = cross(dir, up)
    2: ["chord_field", "centre_pixel", "single_voxel"],  # pixel centre
    3: ["chord_field", "edge_angle", "single_voxel"],   # aspect on x
    4: ["chord_field", "single_voxel"],                 # upside down
    5: ["voxel_ratio"],                                 # ratio, not ratio^2
    6: ["store_rounding"],                              # truncating store
    7: ["voxel_ratio"],                                 # no division by ratio
    8: ["word_counts"],                                 # df pop word
    9: ["word_counts"],                                 # esvo colour word
    10: ["chord_field", "edge_angle", "single_voxel"],  # 60 degree FOV
}

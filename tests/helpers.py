"""Shared test helpers: tiny seeded volumes, cameras and an FP64 brute-force
emission integral that is independent of every traversal algorithm.

The brute-force integral is the textbook slab clip: for each non-empty voxel
box, the ray's chord length inside the box times the voxel's RGB, summed in
FP64.  It pins what all five traversals compute (the line integral of the
piecewise-constant emission field along the ray, resources/dda.comp:45,
svo_naive.comp:66-67, svo_df.comp:34-35, esvo.comp:76-77,
svo_rope.comp:103-105,124-126).
"""
import numpy as np


def random_grid(shape_zyx, density=0.3, seed=0, alpha=255):
    rng = np.random.default_rng(seed)
    z, y, x = shape_zyx
    occ = rng.random((z, y, x)) < density
    g = rng.integers(1, 256, (z, y, x, 4)).astype(np.uint8)
    g[~occ] = 0
    g[..., 3] = np.where(occ, alpha, 0)
    return g


def blocky_grid(side, block, seed=0):
    """A grid made of uniform cubic blocks, so that a lossless octree has
    leaves at several depths."""
    rng = np.random.default_rng(seed)
    nb = side // block
    cols = rng.integers(0, 256, (nb, nb, nb, 4)).astype(np.uint8)
    cols[rng.random((nb, nb, nb)) < 0.4] = 0
    cols[..., 3] = 255
    g = np.repeat(np.repeat(np.repeat(cols, block, 0), block, 1), block, 2)
    return np.ascontiguousarray(g)


def look_at(eye, target, world_up=(0.0, 1.0, 0.0)):
    """Camera (forward, up, translation) with up perpendicular to forward, as
    the camera files require (resources/help/render.txt: 'the viewing direction
    and the up vector must be perpendicular')."""
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    r = np.cross(np.asarray(world_up, np.float64), f)
    if np.linalg.norm(r) < 1e-6:
        r = np.cross((1.0, 0.0, 0.0), f)
    r /= np.linalg.norm(r)
    up = np.cross(f, r)
    return f, up, eye


def orbit_cameras(n, distance, seed=0, center=(0.5, 0.5, 0.5), jitter=0.3):
    rng = np.random.default_rng(seed)
    cams = []
    for i in range(n):
        th = 2 * np.pi * (i + rng.random() * 0.5) / n
        ph = (rng.random() - 0.5) * jitter * np.pi
        eye = np.array(center) + distance * np.array([np.cos(ph) * np.sin(th), np.sin(ph), -np.cos(ph) * np.cos(th)])
        tgt = np.array(center) + (rng.random(3) - 0.5) * 0.2
        cams.append(look_at(eye, tgt))
    return cams


def exact_integral(grid, ro, rd, scale=1.0, offset=0.0):
    """FP64 sum over voxels of rgb * |ray ∩ box| for t >= 0.

    grid: [Z, Y, X, 4] uint8; voxel (x, y, z) spans [x, x+1]*scale + offset.
    ro, rd: ray origin and direction in the same space.
    """
    nz = np.argwhere(grid[..., :3].any(-1))
    if nz.shape[0] == 0:
        return np.zeros(3)
    lo = nz[:, ::-1].astype(np.float64) * scale + offset
    hi = lo + scale
    rd = np.asarray(rd, np.float64)
    ro = np.asarray(ro, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / rd
        t1 = (lo - ro) * inv
        t2 = (hi - ro) * inv
    tmin = np.maximum(np.minimum(t1, t2).max(1), 0.0)
    tmax = np.maximum(t1, t2).min(1)
    chord = np.clip(tmax - tmin, 0.0, None)
    col = grid[nz[:, 0], nz[:, 1], nz[:, 2], :3].astype(np.float64) / 255.0
    return (col * chord[:, None]).sum(0)


def dda_shifted_integral(grid, ro, rd):
    """The integral the reference DDA computes when its entry point rounds to
    just below the lower face of the grid: ivec3() truncates toward zero
    (resources/dda.comp:31) while side_dist uses floor() (dda.comp:36), so
    every texel fetch along the ray is offset by +1 on the entry axis.  This
    returns the exact FP64 integral of the grid shifted that way (voxel v takes
    the colour of v + e_axis, zero past the end)."""
    Z, Y, X = grid.shape[:3]
    rd = np.asarray(rd, np.float64)
    ro = np.asarray(ro, np.float64)
    inv = 1.0 / rd
    dims = np.array([X, Y, Z], np.float64)
    t1 = (0 - ro) * inv
    t2 = (dims - ro) * inv
    axis = int(np.argmax(np.minimum(t1, t2)))  # x, y or z
    np_axis = 2 - axis                           # grid is [Z, Y, X]
    shifted = np.zeros_like(grid)
    src = [slice(None)] * 4
    dst = [slice(None)] * 4
    src[np_axis] = slice(1, None)
    dst[np_axis] = slice(0, -1)
    shifted[tuple(dst)] = grid[tuple(src)]
    return exact_integral(shifted, ro, rd)


def unorm8(x):
    x = np.nan_to_num(np.asarray(x, np.float64), nan=0.0)
    return np.rint(np.clip(x, 0.0, 1.0) * 255.0).astype(np.int32)


# ---------------------------------------------------------------------------
# The reference camera stated geometrically (resources/common.glsl:45-56 and
# resources/help/render.txt:33-41), written independently of the oracle's
# ray() so that it can pin it:
#   * a pinhole at the eye, the image plane one unit ahead along `forward`
#     and one unit wide, so the horizontal field of view is 2*atan(1/2);
#   * square pixels, so the plane is H/W units tall;
#   * columns run along up x forward, rows run away from `up` (row 0 is the
#     top of the image, the side `up` points to);
#   * pixel (i, j) is sampled at its corner, not its centre:
#     plane point (i/W - 1/2, (H/2 - j)/W);
#   * the volume spans [0, voxel_ratio] in camera space.
# ---------------------------------------------------------------------------


def camera_basis(forward, up):
    f = np.asarray(forward, np.float64)
    f = f / np.linalg.norm(f)
    u = np.asarray(up, np.float64)
    u = u / np.linalg.norm(u)
    r = np.cross(u, f)
    return f, r / np.linalg.norm(r), u


def pinhole_dir(forward, up, W, H, i, j):
    """Unit direction of pixel (i, j) of a W x H display (FP64)."""
    f, r, u = camera_basis(forward, up)
    d = f + (i / W - 0.5) * r + ((H / 2.0 - j) / W) * u
    return d / np.linalg.norm(d)


def project_point(forward, up, eye, W, H, p):
    """Continuous pixel coordinates (i, j) of world point p: the inverse of
    pinhole_dir, so pixel (i, j) looks exactly at it."""
    f, r, u = camera_basis(forward, up)
    v = np.asarray(p, np.float64) - np.asarray(eye, np.float64)
    depth = v @ f
    x = (v @ r) / depth
    y = (v @ u) / depth
    return (x + 0.5) * W, H / 2.0 - y * W


def exit_distance(eye, d, box_hi):
    """Distance from `eye` inside the box [0, box_hi] to the box boundary
    along unit direction d (FP64)."""
    eye = np.asarray(eye, np.float64)
    hi = np.asarray(box_hi, np.float64)
    t = [((hi[k] if d[k] > 0 else 0.0) - eye[k]) / d[k] for k in range(3) if d[k] != 0]
    return min(t)

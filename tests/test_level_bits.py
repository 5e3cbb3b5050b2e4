"""The octree kernels' start table rests on one arithmetic fact: the child
choice that find() makes at level l (svo_naive.comp:21-25: pos >= offset +
extent, with offset and extent built by FP32 halvings and additions) equals
bit (22 - l) of level_bits(pos) = floor(clamp(pos) * 2^23) for every FP32
position, including values on and next to the cell boundaries, negatives,
values at or above 1, infinities and NaN.  This checks that equivalence in
IEEE single precision (numpy float32 rounds every operation) on the walk's
first 23 levels -- the levels any tree of side <= 2^23 has."""
import numpy as np

LEVELS = 23


def _walk_choices(p):
    """The comparisons of find()'s descent, level by level, in FP32."""
    p = p.astype(np.float32)
    offset = np.zeros_like(p)
    extent = np.float32(1.0)
    out = []
    for _ in range(LEVELS):
        extent = np.float32(extent * np.float32(0.5))
        c = p >= (offset + extent).astype(np.float32)
        offset = (offset + np.where(c, np.float32(1.0), np.float32(0.0)) * extent).astype(np.float32)
        out.append(c)
    return np.stack(out, axis=1)


def _level_bits(p):
    """The device function's arithmetic: 0 unless p > 0, all ones unless
    p < 1, else the truncation of p * 2^23 (exact: a power-of-two scaling)."""
    p = p.astype(np.float32)
    with np.errstate(invalid="ignore"):
        scaled = (p * np.float32(8388608.0)).astype(np.float32)
        q = np.where(p > 0, np.where(p < 1, np.floor(np.nan_to_num(scaled)), 8388607.0), 0.0).astype(np.uint32)
    return np.stack([((q >> np.uint32(22 - l)) & np.uint32(1)).astype(bool) for l in range(LEVELS)], axis=1)


def _edge_values():
    vals = [0.0, -0.0, 1.0, -1.0, 2.0, 1e-45, -1e-45, 1e-30, np.inf, -np.inf, np.nan]
    one = np.float32(1.0)
    vals.append(np.nextafter(one, np.float32(0)))
    rng = np.random.default_rng(7)
    for m in range(1, LEVELS + 1):
        ks = rng.integers(0, 1 << m, size=64)
        for k in ks:
            d = np.float32(k) / np.float32(1 << m)
            vals += [d, np.nextafter(d, np.float32(2)), np.nextafter(d, np.float32(-1))]
    return np.array(vals, np.float32)


def test_level_bits_equal_the_walks_comparisons():
    rng = np.random.default_rng(11)
    p = np.concatenate([_edge_values(), rng.uniform(-0.25, 1.25, 200_000).astype(np.float32),
                        rng.uniform(0, 1e-6, 20_000).astype(np.float32),
                        (1 - rng.uniform(0, 1e-6, 20_000)).astype(np.float32)])
    walk = _walk_choices(p)
    bits = _level_bits(p)
    bad = np.nonzero((walk != bits).any(axis=1))[0]
    assert bad.size == 0, [(float(p[i]), np.nonzero(walk[i] != bits[i])[0].tolist()) for i in bad[:5]]


def test_offset_after_l_levels_is_the_bits_times_the_extent():
    """The walk's offset after l levels equals (level_bits >> (23 - l)) * 2^-l
    exactly, the value the start table's walk resumes from."""
    rng = np.random.default_rng(12)
    p = np.concatenate([_edge_values(), rng.uniform(-0.25, 1.25, 50_000).astype(np.float32)])
    bits = _level_bits(p)
    q = np.zeros(p.shape, np.uint64)
    for l in range(LEVELS):
        q = (q << np.uint64(1)) | bits[:, l].astype(np.uint64)
    offset = np.zeros_like(p)
    extent = np.float32(1.0)
    for l in range(1, 7):
        extent = np.float32(extent * np.float32(0.5))
        c = p >= (offset + extent).astype(np.float32)
        offset = (offset + np.where(c, np.float32(1.0), np.float32(0.0)) * extent).astype(np.float32)
        want = ((q >> np.uint64(LEVELS - l)).astype(np.float32) * np.float32(2.0 ** -l)).astype(np.float32)
        assert np.array_equal(offset.view(np.uint32), want.view(np.uint32)), l

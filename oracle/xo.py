"""ctypes binding of the oracle library (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product
(xenodon_amd) never imports this module.
"""
import contextlib
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

DDA, SVO_NAIVE, ESVO, SVO_DF, SVO_ROPE = range(5)
TRAVERSAL_NAMES = {"dda": DDA, "svo-naive": SVO_NAIVE, "esvo": ESVO, "svo-df": SVO_DF, "svo-rope": SVO_ROPE}
HEUR_CHAN_DIFF, HEUR_STD_DEV = 0, 1
TREE_SPARSE, TREE_DAG, TREE_ROPE = 0, 1, 2


class Camera(ctypes.Structure):
    _fields_ = [("forward", ctypes.c_float * 4), ("up", ctypes.c_float * 4), ("translation", ctypes.c_float * 4)]


class Uniforms(ctypes.Structure):
    _fields_ = [
        ("out_offset", ctypes.c_int32 * 2),
        ("out_extent", ctypes.c_uint32 * 2),
        ("disp_offset", ctypes.c_int32 * 2),
        ("disp_extent", ctypes.c_uint32 * 2),
        ("voxel_ratio", ctypes.c_float * 4),
        ("model_dim", ctypes.c_uint32 * 4),
        ("emission_coeff", ctypes.c_float),
    ]


class Volume(ctypes.Structure):
    _fields_ = [
        ("grid", ctypes.c_void_p),
        ("dims", ctypes.c_uint64 * 3),
        ("nodes", ctypes.c_void_p),
        ("node_count", ctypes.c_uint64),
        ("side", ctypes.c_uint64),
        ("max_depth", ctypes.c_uint32),
    ]


class BuildStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in
                ("total_leaves", "unique_leaves", "total_nodes", "depth", "side", "node_count")]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load(path):
    L = ctypes.CDLL(path)
    P = ctypes.POINTER
    L.xo_render.argtypes = [ctypes.c_int, P(Volume), P(Camera), P(Uniforms), ctypes.c_int, ctypes.c_int,
                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.xo_trace_ray.argtypes = [ctypes.c_int, P(Volume), P(ctypes.c_double), P(ctypes.c_double), ctypes.c_int,
                               P(ctypes.c_double), P(ctypes.c_uint64)]
    L.xo_ray_dir.argtypes = [P(Camera), P(Uniforms), ctypes.c_uint32, ctypes.c_uint32, P(ctypes.c_float)]
    L.xo_iteration_cap.argtypes = [ctypes.c_int, P(Volume), P(ctypes.c_uint32)]
    L.xo_iteration_cap.restype = ctypes.c_uint64
    L.xo_build_octree.argtypes = [ctypes.c_void_p, P(ctypes.c_uint64), ctypes.c_int, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_int, P(ctypes.c_void_p), P(BuildStats)]
    L.xo_generate_ropes.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64]
    L.xo_octree_find.argtypes = [ctypes.c_void_p, ctypes.c_uint64, P(ctypes.c_uint64), ctypes.c_uint64]
    L.xo_octree_find.restype = ctypes.c_uint64
    L.xo_vol_scan.argtypes = [ctypes.c_void_p, P(ctypes.c_uint64), P(ctypes.c_uint64), P(ctypes.c_uint64),
                              P(ctypes.c_uint32), P(ctypes.c_uint32)]
    L.xo_stddev_scan.argtypes = [ctypes.c_void_p, P(ctypes.c_uint64), P(ctypes.c_uint64), P(ctypes.c_uint64),
                                 P(ctypes.c_uint32), P(ctypes.c_double)]
    L.xo_free.argtypes = [ctypes.c_void_p]
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _load(LIB_PATH)
    return _lib


def build_mutant(mutant, out_dir):
    """Builds the oracle with one deliberate mistake planted (-DXO_MUTANT=k,
    the list is in xo_trace_body.h) into out_dir and returns the library path.
    Used only by the tests that show each pin of the oracle fails under the
    mistake it is meant to catch."""
    path = os.path.join(out_dir, f"liboracle_m{int(mutant)}.so")
    subprocess.run(["make", "-s", "-C", HERE, f"MUTANT={int(mutant)}", f"OUT={path}", path], check=True)
    return path


@contextlib.contextmanager
def use_library(path):
    """Routes every call of this module to the library at `path` (a mutant
    build) for the duration of the block."""
    global _lib
    saved = lib()
    _lib = _load(path)
    try:
        yield
    finally:
        _lib = saved


def make_camera(forward, up, translation):
    c = Camera()
    c.forward[:] = [float(v) for v in forward] + [0.0]
    c.up[:] = [float(v) for v in up] + [0.0]
    c.translation[:] = [float(v) for v in translation] + [0.0]
    return c


def make_uniforms(width, height, model_dim, emission=1.0, voxel_ratio=(1.0, 1.0, 1.0), out_offset=(0, 0),
                  disp_offset=(0, 0), disp_extent=None):
    u = Uniforms()
    u.out_offset[:] = list(out_offset)
    u.out_extent[:] = [width, height]
    u.disp_offset[:] = list(disp_offset)
    u.disp_extent[:] = list(disp_extent) if disp_extent is not None else [width, height]
    u.voxel_ratio[:] = [float(v) for v in voxel_ratio] + [0.0]
    u.model_dim[:] = [int(v) for v in model_dim] + [0]
    u.emission_coeff = float(emission)
    return u


class OracleVolume:
    """Keeps numpy buffers alive next to the ctypes descriptor."""

    def __init__(self, grid=None, nodes=None, side=None):
        self.v = Volume()
        self.grid = None
        self.nodes = None
        if grid is not None:
            g = np.ascontiguousarray(grid, dtype=np.uint8)
            assert g.ndim == 4 and g.shape[3] == 4, "grid must be [Z, Y, X, 4]"
            self.grid = g
            self.v.grid = g.ctypes.data
            self.v.dims[:] = [g.shape[2], g.shape[1], g.shape[0]]
        if nodes is not None:
            n = np.ascontiguousarray(nodes, dtype=np.uint32).reshape(-1, 10)
            self.nodes = n
            self.v.nodes = n.ctypes.data
            self.v.node_count = n.shape[0]
            self.v.side = int(side)
            self.v.max_depth = int((n[:, 9] & 0x7FFFFFFF).max()) if n.shape[0] else 0

    @property
    def model_dim(self):
        if self.grid is not None:
            return list(self.v.dims)
        return [self.v.side] * 3


def render(traversal, vol, cam, uni, precision=32, threads=0, linear=False, words=False):
    L = lib()
    W, H = uni.out_extent[0], uni.out_extent[1]
    rgba = np.zeros((H, W, 4), np.uint8)
    lin = np.zeros((H, W, 3), np.float32) if linear else None
    wc = np.zeros((H, W), np.uint64) if words else None
    rc = L.xo_render(traversal, ctypes.byref(vol.v), ctypes.byref(cam), ctypes.byref(uni), precision, threads,
                     rgba.ctypes.data, lin.ctypes.data if linear else None, wc.ctypes.data if words else None)
    if rc != 0:
        raise ValueError(f"xo_render failed: {rc}")
    out = {"rgba": rgba}
    if linear:
        out["linear"] = lin
    if words:
        out["words"] = wc
    return out


def trace_ray(traversal, vol, ro, rd, precision=32):
    L = lib()
    a = (ctypes.c_double * 3)(*ro)
    b = (ctypes.c_double * 3)(*rd)
    o = (ctypes.c_double * 3)()
    w = ctypes.c_uint64()
    rc = L.xo_trace_ray(traversal, ctypes.byref(vol.v), a, b, precision, o, ctypes.byref(w))
    if rc != 0:
        raise ValueError(f"xo_trace_ray failed: {rc}")
    return np.array(list(o)), int(w.value)


def ray_dir(cam, uni, px, py):
    o = (ctypes.c_float * 3)()
    lib().xo_ray_dir(ctypes.byref(cam), ctypes.byref(uni), px, py, o)
    return np.array(list(o), np.float32)


def iteration_cap(traversal, vol, model_dim):
    md = (ctypes.c_uint32 * 3)(*[int(v) for v in model_dim])
    return int(lib().xo_iteration_cap(traversal, ctypes.byref(vol.v), md))


def build_octree(grid, heuristic=HEUR_CHAN_DIFF, chan_diff=0, std_dev=0.0, tree=TREE_SPARSE):
    """Returns (nodes [N,10] uint32, stats dict)."""
    L = lib()
    g = np.ascontiguousarray(grid, dtype=np.uint8)
    dims = (ctypes.c_uint64 * 3)(g.shape[2], g.shape[1], g.shape[0])
    out = ctypes.c_void_p()
    st = BuildStats()
    rc = L.xo_build_octree(g.ctypes.data, dims, heuristic, chan_diff, std_dev, tree, ctypes.byref(out),
                           ctypes.byref(st))
    if rc != 0:
        raise RuntimeError(f"xo_build_octree failed: {rc}")
    n = st.node_count
    buf = (ctypes.c_uint32 * (n * 10)).from_address(out.value)
    nodes = np.frombuffer(buf, dtype=np.uint32).reshape(n, 10).copy()
    L.xo_free(out)
    stats = {f: getattr(st, f) for f, _ in BuildStats._fields_}
    return nodes, stats


def octree_find(nodes, side, pos, max_depth):
    n = np.ascontiguousarray(nodes, dtype=np.uint32)
    p = (ctypes.c_uint64 * 3)(*[int(v) % (1 << 64) for v in pos])
    r = lib().xo_octree_find(n.ctypes.data, side, p, max_depth)
    return None if r == (1 << 64) - 1 else int(r)


def vol_scan(grid, bmin, bmax):
    g = np.ascontiguousarray(grid, dtype=np.uint8)
    dims = (ctypes.c_uint64 * 3)(g.shape[2], g.shape[1], g.shape[0])
    a = ctypes.c_uint32()
    m = ctypes.c_uint32()
    lib().xo_vol_scan(g.ctypes.data, dims, (ctypes.c_uint64 * 3)(*bmin), (ctypes.c_uint64 * 3)(*bmax),
                      ctypes.byref(a), ctypes.byref(m))
    return a.value, m.value


def stddev_scan(grid, bmin, bmax):
    g = np.ascontiguousarray(grid, dtype=np.uint8)
    dims = (ctypes.c_uint64 * 3)(g.shape[2], g.shape[1], g.shape[0])
    a = ctypes.c_uint32()
    s = ctypes.c_double()
    lib().xo_stddev_scan(g.ctypes.data, dims, (ctypes.c_uint64 * 3)(*bmin), (ctypes.c_uint64 * 3)(*bmax),
                         ctypes.byref(a), ctypes.byref(s))
    return a.value, s.value

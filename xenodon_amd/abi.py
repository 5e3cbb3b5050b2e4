"""ctypes binding of the C-ABI in include/xenodon.h (libxenodon_amd.so).

This is plumbing only: every call goes straight to the native library, and
importing fails loudly when the library has not been built -- there is no
CPU fallback anywhere in the package.
"""
import ctypes
import os
import weakref

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libxenodon_amd.so")

XD_OK = 0
(XD_ERR_INVALID_ARGUMENT, XD_ERR_DEVICE, XD_ERR_OUT_OF_MEMORY, XD_ERR_INCOMPATIBLE, XD_ERR_FORMAT, XD_ERR_IO,
 XD_ERR_REJECTED, XD_ERR_INTERNAL) = range(1, 9)
DDA, SVO_NAIVE, ESVO, SVO_DF, SVO_ROPE = range(5)
# src/main_loop.cpp:37-43 option names
TRAVERSALS = {"dda": DDA, "svo-naive": SVO_NAIVE, "esvo": ESVO, "svo-df": SVO_DF, "svo-rope": SVO_ROPE}
TRAVERSAL_NAMES = {v: k for k, v in TRAVERSALS.items()}
VOLUME_GRID, VOLUME_SVO = 0, 1


class XdError(RuntimeError):
    """Error raised by a C-ABI call; .status is the xd_status code."""

    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


class Camera(ctypes.Structure):
    _fields_ = [("forward", ctypes.c_float * 4), ("up", ctypes.c_float * 4), ("translation", ctypes.c_float * 4)]


class Rect(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_int32 * 2), ("extent", ctypes.c_uint32 * 2)]


class Uniforms(ctypes.Structure):
    _fields_ = [
        ("output_region", Rect),
        ("display_region", Rect),
        ("voxel_ratio", ctypes.c_float * 4),
        ("model_dim", ctypes.c_uint32 * 4),
        ("emission_coeff", ctypes.c_float),
    ]


class VolumeInfo(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("max_depth", ctypes.c_uint32),
        ("dims", ctypes.c_uint64 * 3),
        ("side", ctypes.c_uint64),
        ("node_count", ctypes.c_uint64),
        ("device_bytes", ctypes.c_uint64),
        ("is_rope", ctypes.c_uint32),
        ("depth_fields_exact", ctypes.c_uint32),
    ]


class View(ctypes.Structure):
    _fields_ = [("volume", ctypes.c_void_p), ("traversal", ctypes.c_int32), ("camera", Camera),
                ("uniforms", Uniforms), ("frame", ctypes.c_void_p)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("tile_w", ctypes.c_uint32 * 5), ("tiles_per_grab", ctypes.c_uint32),
                ("waves_per_simd", ctypes.c_uint32), ("class_kernels", ctypes.c_uint32),
                ("interleave", ctypes.c_uint32), ("mix_classes", ctypes.c_uint32),
                ("rays_per_lane", ctypes.c_uint32 * 5)]


class BuildStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in
                ("total_leaves", "unique_leaves", "total_nodes", "depth", "side", "node_count")]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("frame_arena_bytes", ctypes.c_uint64),
                ("max_active", ctypes.c_uint32), ("tiles_per_launch", ctypes.c_uint32),
                ("launches_in_flight", ctypes.c_uint32), ("readback_threads", ctypes.c_uint32),
                ("tuning", BatchTuning), ("latency_target_ms", ctypes.c_double)]


class Request(ctypes.Structure):
    _fields_ = [("volume_id", ctypes.c_uint32), ("traversal", ctypes.c_int32), ("camera", Camera),
                ("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("emission_coeff", ctypes.c_float),
                ("voxel_ratio", ctypes.c_float * 3), ("dst", ctypes.c_void_p), ("release_ms", ctypes.c_double),
                ("user_tag", ctypes.c_uint64)]


class Completion(ctypes.Structure):
    _fields_ = [("request_id", ctypes.c_uint64), ("user_tag", ctypes.c_uint64), ("status", ctypes.c_int32),
                ("launches", ctypes.c_uint32), ("arrival_ms", ctypes.c_double), ("start_ms", ctypes.c_double),
                ("device_done_ms", ctypes.c_double), ("done_ms", ctypes.c_double), ("rays", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("resident_bytes", ctypes.c_uint64), ("budget_bytes", ctypes.c_uint64),
                ("volumes", ctypes.c_uint32), ("pending", ctypes.c_uint32), ("active", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("completed", ctypes.c_uint64), ("launches", ctypes.c_uint64),
                ("tiles", ctypes.c_uint64), ("rays", ctypes.c_uint64), ("max_requests_per_launch", ctypes.c_uint32),
                ("max_launches_overlapped", ctypes.c_uint32), ("busy_ms", ctypes.c_double),
                ("shed", ctypes.c_uint64), ("rays_per_ms", ctypes.c_double)]


HEURISTIC_CHAN_DIFF, HEURISTIC_STD_DEV = 0, 1
TREE_SPARSE, TREE_DAG, TREE_ROPE = 0, 1, 2

_lib = None


def _declare(L):
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    sigs = {
        "xd_last_error": ([], ctypes.c_char_p),
        "xd_device_count": ([P(ctypes.c_int)], ctypes.c_int),
        "xd_device_describe": ([ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t, P(ctypes.c_uint32),
                                P(ctypes.c_uint32), P(ctypes.c_uint32), P(ctypes.c_uint32), P(ctypes.c_uint64)],
                               ctypes.c_int),
        "xd_open": ([ctypes.c_int, P(vp)], ctypes.c_int),
        "xd_close": ([vp], ctypes.c_int),
        "xd_context_set_tile": ([vp, ctypes.c_uint32], ctypes.c_int),
        "xd_context_set_option": ([vp, ctypes.c_char_p, ctypes.c_int64], ctypes.c_int),
        "xd_volume_from_grid": ([vp, vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, P(vp)], ctypes.c_int),
        "xd_volume_from_svo": ([vp, vp, ctypes.c_uint64, ctypes.c_uint64, P(vp)], ctypes.c_int),
        "xd_volume_info_get": ([vp, P(VolumeInfo)], ctypes.c_int),
        "xd_volume_free": ([vp], ctypes.c_int),
        "xd_frame_create": ([vp, ctypes.c_uint32, ctypes.c_uint32, P(vp)], ctypes.c_int),
        "xd_frame_free": ([vp], ctypes.c_int),
        "xd_frame_device_ptr": ([vp, P(vp)], ctypes.c_int),
        "xd_render": ([vp, vp, ctypes.c_int, P(Camera), P(Uniforms), vp, P(ctypes.c_double)], ctypes.c_int),
        "xd_frame_download": ([vp, vp, ctypes.c_uint64], ctypes.c_int),
        "xd_count_words": ([vp, vp, ctypes.c_int, P(Camera), P(Uniforms), P(ctypes.c_uint64)], ctypes.c_int),
        "xd_build_octree": ([vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                             ctypes.c_double, ctypes.c_int, P(vp), P(BuildStats)], ctypes.c_int),
        "xd_free": ([vp], None),
        "xd_build_octree_device": ([vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, P(vp),
                                    P(BuildStats)], ctypes.c_int),
        "xd_batch_create": ([vp, P(vp)], ctypes.c_int),
        "xd_batch_set_views": ([vp, P(View), ctypes.c_uint32], ctypes.c_int),
        "xd_batch_set_tuning": ([vp, P(BatchTuning)], ctypes.c_int),
        "xd_batch_get_tuning": ([vp, P(BatchTuning)], ctypes.c_int),
        "xd_batch_autotune": ([vp, ctypes.c_uint32, P(BatchTuning)], ctypes.c_int),
        "xd_batch_render": ([vp, P(ctypes.c_double), P(ctypes.c_uint64)], ctypes.c_int),
        "xd_batch_free": ([vp], ctypes.c_int),
        "xd_batch_render_static": ([vp, P(View), ctypes.c_uint32, P(BatchTuning), P(ctypes.c_double),
                                    P(ctypes.c_uint64)], ctypes.c_int),
        "xd_volume_footprint": ([vp, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                 ctypes.c_uint64, P(ctypes.c_uint64)], ctypes.c_int),
        "xd_server_create": ([vp, P(ServerConfig), P(vp)], ctypes.c_int),
        "xd_server_set_latency_target": ([vp, ctypes.c_double], ctypes.c_int),
        "xd_server_load_grid": ([vp, vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, P(ctypes.c_uint32)],
                                ctypes.c_int),
        "xd_server_load_svo": ([vp, vp, ctypes.c_uint64, ctypes.c_uint64, P(ctypes.c_uint32)], ctypes.c_int),
        "xd_server_unload": ([vp, ctypes.c_uint32], ctypes.c_int),
        "xd_server_submit": ([vp, P(Request), P(ctypes.c_uint64)], ctypes.c_int),
        "xd_server_poll": ([vp, P(Completion), ctypes.c_uint32, P(ctypes.c_uint32), ctypes.c_int32], ctypes.c_int),
        "xd_server_stats_get": ([vp, P(ServerStats)], ctypes.c_int),
        "xd_server_now_ms": ([vp], ctypes.c_double),
        "xd_server_free": ([vp], ctypes.c_int),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    return L


def lib():
    """The loaded native library.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m xenodon_amd.build` "
                              "(the HIP path has no fallback)")
        _lib = _declare(ctypes.CDLL(LIB_PATH))
    return _lib


def check(status):
    if status != XD_OK:
        msg = lib().xd_last_error()
        raise XdError(status, msg.decode() if msg else f"xd status {status}")


def device_count():
    n = ctypes.c_int()
    st = lib().xd_device_count(ctypes.byref(n))
    return n.value if st == XD_OK else 0


def device_describe(index):
    """(name, (pci domain, bus, device, function), total bytes) of a device."""
    name = ctypes.create_string_buffer(256)
    f = [ctypes.c_uint32() for _ in range(4)]
    total = ctypes.c_uint64()
    check(lib().xd_device_describe(index, name, 256, *[ctypes.byref(v) for v in f], ctypes.byref(total)))
    return name.value.decode(), tuple(v.value for v in f), total.value


def make_camera(forward, up, translation):
    c = Camera()
    c.forward[:] = [float(v) for v in forward[:3]] + [0.0]
    c.up[:] = [float(v) for v in up[:3]] + [0.0]
    c.translation[:] = [float(v) for v in translation[:3]] + [0.0]
    return c


def make_uniforms(width, height, model_dim, emission=1.0, voxel_ratio=(1.0, 1.0, 1.0), out_offset=(0, 0),
                  disp_offset=(0, 0), disp_extent=None):
    u = Uniforms()
    u.output_region.offset[:] = list(out_offset)
    u.output_region.extent[:] = [width, height]
    u.display_region.offset[:] = list(disp_offset)
    u.display_region.extent[:] = list(disp_extent) if disp_extent is not None else [width, height]
    u.voxel_ratio[:] = [float(v) for v in voxel_ratio] + [0.0]
    u.model_dim[:] = [int(v) for v in model_dim] + [0]
    u.emission_coeff = float(emission)
    return u


class Context:
    """One device + one HIP stream (xd_open)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib().xd_open(device, ctypes.byref(h)))
        self.h = h
        self._children = weakref.WeakSet()

    def close(self):
        if self.h:
            # xd_close frees the context's remaining volumes and frames
            for c in list(self._children):
                c.h = None
            lib().xd_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def grid_volume(self, grid):
        """grid: uint8 array [Z, Y, X, 4] (x fastest)."""
        g = np.ascontiguousarray(grid, dtype=np.uint8)
        if g.ndim != 4 or g.shape[3] != 4:
            raise ValueError("grid must be [Z, Y, X, 4] uint8")
        h = ctypes.c_void_p()
        check(lib().xd_volume_from_grid(self.h, g.ctypes.data, g.shape[2], g.shape[1], g.shape[0], ctypes.byref(h)))
        return self._adopt(Volume(self, h))

    def svo_volume(self, nodes, side):
        """nodes: uint32 [N, 10] (children[8], color, is_leaf_depth)."""
        n = np.ascontiguousarray(nodes, dtype=np.uint32).reshape(-1, 10)
        h = ctypes.c_void_p()
        check(lib().xd_volume_from_svo(self.h, n.ctypes.data, n.shape[0], int(side), ctypes.byref(h)))
        return self._adopt(Volume(self, h))

    def frame(self, width, height):
        h = ctypes.c_void_p()
        check(lib().xd_frame_create(self.h, width, height, ctypes.byref(h)))
        return self._adopt(Frame(self, h, width, height))

    def batch(self):
        h = ctypes.c_void_p()
        check(lib().xd_batch_create(self.h, ctypes.byref(h)))
        return self._adopt(Batch(self, h))

    def server(self, hbm_budget_bytes=0, frame_arena_bytes=0, max_active=0, tiles_per_launch=0,
               launches_in_flight=0, readback_threads=0, tuning=None, latency_target_ms=0.0):
        c = ServerConfig()
        c.latency_target_ms = float(latency_target_ms)
        c.hbm_budget_bytes = hbm_budget_bytes
        c.frame_arena_bytes = frame_arena_bytes
        c.max_active = max_active
        c.tiles_per_launch = tiles_per_launch
        c.launches_in_flight = launches_in_flight
        c.readback_threads = readback_threads
        if tuning:
            c.tuning.tile_w[:] = list(tuning.get("tile_w", [8] * 5))
            c.tuning.tiles_per_grab = tuning.get("tiles_per_grab", 1)
            c.tuning.waves_per_simd = tuning.get("waves_per_simd", 0)
            c.tuning.class_kernels = tuning.get("class_kernels", 0)
            c.tuning.interleave = tuning.get("interleave", 0)
            c.tuning.mix_classes = tuning.get("mix_classes", 0)
            c.tuning.rays_per_lane[:] = list(tuning.get("rays_per_lane", [1] * 5))
        h = ctypes.c_void_p()
        check(lib().xd_server_create(self.h, ctypes.byref(c), ctypes.byref(h)))
        return self._adopt(Server(self, h))

    def build_octree(self, grid_volume, chan_diff=0, tree=None, heuristic=None, std_dev=0.0):
        """GPU octree build from a grid volume in HBM (both heuristics; sparse,
        DAG or rope trees); same (nodes, stats) as build_octree."""
        out = ctypes.c_void_p()
        st = BuildStats()
        check(lib().xd_build_octree_device(self.h, grid_volume.h, HEURISTIC_CHAN_DIFF if heuristic is None else heuristic,
                                           chan_diff, float(std_dev), TREE_SPARSE if tree is None else tree,
                                           ctypes.byref(out), ctypes.byref(st)))
        return _take_nodes(out, st)

    def set_tile(self, tile_w):
        check(lib().xd_context_set_tile(self.h, tile_w))

    def set_option(self, name, value):
        check(lib().xd_context_set_option(self.h, name.encode(), int(value)))

    def _adopt(self, child):
        self._children.add(child)
        return child

    def render(self, volume, traversal, cam, uni, frame):
        """Plain one-launch render; returns the launch time in ms."""
        ms = ctypes.c_double()
        check(lib().xd_render(self.h, volume.h, traversal, ctypes.byref(cam), ctypes.byref(uni), frame.h,
                              ctypes.byref(ms)))
        return ms.value

    def count_words(self, volume, traversal, cam, uni):
        w = ctypes.c_uint64()
        check(lib().xd_count_words(self.h, volume.h, traversal, ctypes.byref(cam), ctypes.byref(uni),
                                   ctypes.byref(w)))
        return w.value

    def render_image(self, volume, traversal, cam, uni):
        """Render into a fresh frame and return ([H, W, 4] uint8, ms)."""
        W, H = uni.output_region.extent[0], uni.output_region.extent[1]
        f = self.frame(W, H)
        try:
            ms = self.render(volume, traversal, cam, uni, f)
            return f.download(), ms
        finally:
            f.free()


def view_array(views):
    """ctypes xd_view array of (volume, traversal, camera, uniforms, frame) tuples."""
    arr = (View * max(1, len(views)))()
    for i, (vol, t, cam, uni, fr) in enumerate(views):
        arr[i].volume = vol.h
        arr[i].traversal = t
        arr[i].camera = cam
        arr[i].uniforms = uni
        arr[i].frame = fr.h
    return arr


def batch_tuning(tile_w, tiles_per_grab=1, waves_per_simd=8, class_kernels=0, interleave=2, mix_classes=1,
                 rays_per_lane=None):
    t = BatchTuning()
    t.rays_per_lane[:] = list(rays_per_lane or [1] * 5)
    t.tile_w[:] = list(tile_w)
    t.tiles_per_grab = tiles_per_grab
    t.waves_per_simd = waves_per_simd
    t.class_kernels = class_kernels
    t.interleave = interleave
    t.mix_classes = mix_classes
    return t


def render_static(ctx, views, tuning=None):
    """xd_batch_render_static: one batched launch of `views` whose tables live
    in the library's static device array (no device allocation).  tuning: a
    dict of xd_batch_tuning fields, or None for the defaults.  Returns (ms, rays)."""
    ms = ctypes.c_double()
    rays = ctypes.c_uint64()
    t = None if tuning is None else ctypes.byref(batch_tuning(**tuning))
    check(lib().xd_batch_render_static(ctx.h, view_array(views), len(views), t, ctypes.byref(ms),
                                       ctypes.byref(rays)))
    return ms.value, rays.value


def tuning_dict(t):
    return {"tile_w": list(t.tile_w), "tiles_per_grab": t.tiles_per_grab, "waves_per_simd": t.waves_per_simd,
            "class_kernels": t.class_kernels, "interleave": t.interleave, "mix_classes": t.mix_classes,
            "rays_per_lane": list(t.rays_per_lane)}


class Batch:
    """Persistent ray-tile-queue launch over many views (xd_batch_*)."""

    def __init__(self, ctx, h):
        self.ctx = ctx
        self.h = h
        self._keep = []

    def set_views(self, views):
        """views: list of (volume, traversal, camera, uniforms, frame)."""
        arr = view_array(views)
        self._keep = list(views)
        check(lib().xd_batch_set_views(self.h, arr, len(views)))

    def tuning(self):
        t = BatchTuning()
        check(lib().xd_batch_get_tuning(self.h, ctypes.byref(t)))
        return tuning_dict(t)

    def set_tuning(self, tile_w, tiles_per_grab=1, waves_per_simd=8, class_kernels=0, interleave=2, mix_classes=1,
                   rays_per_lane=None):
        """Defaults are the library's starting launch shape."""
        t = batch_tuning(tile_w, tiles_per_grab, waves_per_simd, class_kernels, interleave, mix_classes, rays_per_lane)
        check(lib().xd_batch_set_tuning(self.h, ctypes.byref(t)))

    def autotune(self, trials=3):
        t = BatchTuning()
        check(lib().xd_batch_autotune(self.h, trials, ctypes.byref(t)))
        return tuning_dict(t)

    def render(self):
        """Returns (ms, rays)."""
        ms = ctypes.c_double()
        rays = ctypes.c_uint64()
        check(lib().xd_batch_render(self.h, ctypes.byref(ms), ctypes.byref(rays)))
        return ms.value, rays.value

    def free(self):
        if self.h:
            lib().xd_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """Served form (xd_server_*): resident volumes admitted by HBM footprint
    and a request loop with continuous batching."""

    def __init__(self, ctx, h):
        self.ctx = ctx
        self.h = h
        self._dst = {}  # request id -> host array the server writes into

    def load_grid(self, grid):
        g = np.ascontiguousarray(grid, dtype=np.uint8)
        vid = ctypes.c_uint32()
        check(lib().xd_server_load_grid(self.h, g.ctypes.data, g.shape[2], g.shape[1], g.shape[0],
                                        ctypes.byref(vid)))
        return vid.value

    def load_svo(self, nodes, side):
        n = np.ascontiguousarray(nodes, dtype=np.uint32).reshape(-1, 10)
        vid = ctypes.c_uint32()
        check(lib().xd_server_load_svo(self.h, n.ctypes.data, n.shape[0], int(side), ctypes.byref(vid)))
        return vid.value

    def unload(self, volume_id):
        check(lib().xd_server_unload(self.h, volume_id))

    def submit(self, volume_id, traversal, cam, width, height, emission=1.0, voxel_ratio=(1.0, 1.0, 1.0),
               keep=True, release_ms=0.0, tag=0):
        """Queues a request; returns (request id, host array it fills, or None
        when keep=False)."""
        r = Request()
        r.volume_id = volume_id
        r.traversal = traversal
        r.camera = cam
        r.width, r.height = width, height
        r.emission_coeff = emission
        r.voxel_ratio[:] = list(voxel_ratio)
        dst = np.zeros((height, width, 4), np.uint8) if keep else None
        r.dst = dst.ctypes.data if keep else None
        r.release_ms = release_ms
        r.user_tag = tag
        rid = ctypes.c_uint64()
        check(lib().xd_server_submit(self.h, ctypes.byref(r), ctypes.byref(rid)))
        if keep:
            self._dst[rid.value] = dst
        return rid.value, dst

    def poll(self, max_n=256, timeout_ms=-1):
        """Finished requests as dicts (status, times in ms, rays, image)."""
        arr = (Completion * max_n)()
        n = ctypes.c_uint32()
        check(lib().xd_server_poll(self.h, arr, max_n, ctypes.byref(n), timeout_ms))
        out = []
        for i in range(n.value):
            c = arr[i]
            out.append({"id": c.request_id, "tag": c.user_tag, "status": c.status, "launches": c.launches,
                        "arrival_ms": c.arrival_ms, "start_ms": c.start_ms, "device_done_ms": c.device_done_ms,
                        "done_ms": c.done_ms, "rays": c.rays, "image": self._dst.pop(c.request_id, None)})
        return out

    def set_latency_target(self, ms):
        """Latency target (ms) for admissions from now on; 0 = none."""
        check(lib().xd_server_set_latency_target(self.h, float(ms)))

    def stats(self):
        st = ServerStats()
        check(lib().xd_server_stats_get(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_ if k != "reserved"}

    def now_ms(self):
        return lib().xd_server_now_ms(self.h)

    def free(self):
        if self.h:
            lib().xd_server_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def volume_footprint(kind, dims=(0, 0, 0), node_count=0, ctx=None):
    """HBM bytes a volume will hold (xd_volume_footprint); needs no device."""
    b = ctypes.c_uint64()
    check(lib().xd_volume_footprint(ctx.h if ctx else None, kind, dims[0], dims[1], dims[2], node_count,
                                    ctypes.byref(b)))
    return b.value


class Volume:
    def __init__(self, ctx, h):
        self.ctx = ctx
        self.h = h

    def info(self):
        i = VolumeInfo()
        check(lib().xd_volume_info_get(self.h, ctypes.byref(i)))
        return i

    def free(self):
        if self.h:
            lib().xd_volume_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Frame:
    def __init__(self, ctx, h, w, hgt):
        self.ctx = ctx
        self.h = h
        self.width = w
        self.height = hgt

    def download(self):
        out = np.zeros((self.height, self.width, 4), np.uint8)
        check(lib().xd_frame_download(self.h, out.ctypes.data, 0))
        return out

    def device_ptr(self):
        p = ctypes.c_void_p()
        check(lib().xd_frame_device_ptr(self.h, ctypes.byref(p)))
        return p.value

    def free(self):
        if self.h:
            lib().xd_frame_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def build_octree(grid, heuristic=HEURISTIC_CHAN_DIFF, chan_diff=0, std_dev=0.0, tree=TREE_SPARSE):
    """`xenodon convert` on a [Z, Y, X, 4] uint8 grid.  Returns (nodes uint32
    [N, 10], stats dict); the node buffer is the reference converter's."""
    g = np.ascontiguousarray(grid, dtype=np.uint8)
    out = ctypes.c_void_p()
    st = BuildStats()
    check(lib().xd_build_octree(g.ctypes.data, g.shape[2], g.shape[1], g.shape[0], heuristic, chan_diff,
                                float(std_dev), tree, ctypes.byref(out), ctypes.byref(st)))
    return _take_nodes(out, st)


def _take_nodes(out, st):
    n = st.node_count
    try:
        buf = (ctypes.c_uint32 * (n * 10)).from_address(out.value)
        nodes = np.frombuffer(buf, dtype=np.uint32).reshape(n, 10).copy()
    finally:
        lib().xd_free(out)
    return nodes, {f: getattr(st, f) for f, _ in BuildStats._fields_}

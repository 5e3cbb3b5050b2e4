"""Benchmark: Mrays/s of the five traversals at 1920x1080 on the 512x361x512
CT-like volume over the reference's 150-frame benchmark camera path, in the
plain, batched and served forms.

A step renders one camera frame with each of the five traversals
(dda on the grid; svo-naive, esvo, svo-df, svo-rope on its lossless rope
octree).  Inputs are HBM-resident before timing.
  plain    one launch per traversal per frame -- the reference's form
           (src/render/Renderer.cpp:55-103); per-traversal Mrays/s use the
           launches' own HIP-event time (RenderStats semantics,
           src/render/RenderStats.cpp:46-83).
  batched  `batch_frames` steps (x 5 traversals) per persistent-kernel
           hipGraph launch.  `value` = total rays / wall time of the timed
           batched steps.
  mixed    64 views (dda / svo-df / esvo, sizes log-uniform in 256..1920 x
           256..1080, orbit cameras) over the resident 128^3, 512x361x512
           and 1024^3 volumes in one batched launch, against the plain form
           on the same views.
  served   Poisson request arrivals (seeded cameras, log-uniform sizes from
           128x128 to 1920x1080, traversal uniform over all five) against
           the resident 128^3 and 512x361x512 volumes; p50/p99 latency from
           arrival to the frame in host memory, and aggregate Mrays/s; for the
           requests released at once, also the plain form on the same
           requests.

--workload picks the configuration: ct512 (the default, above), shells128
(the 128^3 shells grid, dda at 256x256 over a 10-frame orbit at twice the
unit-cube diagonal, cycled; 150 frames per batched launch) or lognormal1024
(the 1024^3 log-normal grid and its rope SVO built on the GPU, dda and esvo at
3840x2160 over camera.txt).  The mixed-view and served legs run with ct512
only.

The roofline object covers the dominant kernel: algorithmic bytes (4 B per
volume word the reference algorithm reads, counted by the instrumented kernel
variant on the same frames, + 4 B per pixel) over its mean launch time.
cpu_baseline is the oracle (CPU restatement of the same shaders) on the host
cores over a bounded sample of the same frames.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from xenodon_amd import abi, synth, workloads  # noqa: E402

CAMERA_FILE = os.path.join(ROOT, "tests", "golden", "camera.txt")
WIDTH, HEIGHT = 1920, 1080
SERVER_CONFIG = {}  # xd_server_config fields set from the command line (tiles_per_launch, launches_in_flight)
EMISSION = 1.6  # <= 1 % saturated pixels on this volume and path
TRAV_ORDER = ["dda", "svo-naive", "esvo", "svo-df", "svo-rope"]
HBM_PEAK_GBS = 8000.0
# --workload: image size, traversals, emission coefficient, camera set and
# the committed rocprofv3 traffic summary of each measured configuration
WORKLOADS = {
    "ct512": {"size": (1920, 1080), "traversals": TRAV_ORDER, "emission": EMISSION, "cameras": "path",
              "steps": 150, "batch_frames": 10, "profile": "hbm_traffic_ct512_1080p.json",
              "desc": "512x361x512 RGBA8 grid, all five traversals at 1920x1080, camera.txt path",
              "data": "synthetic (seeded CT-like 512x361x512 RGBA8 grid + its lossless rope SVO; reference "
                      "camera.txt)"},
    "shells128": {"size": (256, 256), "traversals": ["dda"], "emission": 4.0, "cameras": "orbit",
                  "steps": 1500, "batch_frames": 150, "profile": "hbm_traffic_shells128_256.json",
                  "desc": "128^3 RGBA8 shells grid, dda at 256x256, 10-frame circular orbit at twice the "
                          "unit-cube diagonal (cycled)",
                  "data": "synthetic (seeded 128^3 RGBA8 grid: 24 ellipsoidal shells + 5 % noise voxels)"},
    "lognormal1024": {"size": (3840, 2160), "traversals": ["dda", "esvo"], "emission": 1.0, "cameras": "path",
                      "steps": 150, "batch_frames": 10, "profile": "hbm_traffic_lognormal1024_2160p.json",
                      # the autotuner's pick here: one persistent kernel per traversal, run in
                      # turn, so the grid's and the octree's working sets do not share the
                      # caches
                      "tuning": {"class_kernels": 1},
                      "desc": "1024^3 RGBA8 log-normal grid, dda on the grid and esvo on its rope SVO (one "
                              "allocation above 4 GB) at 3840x2160, camera.txt path",
                      "data": "synthetic (seeded 1024^3 log-normal RGBA8 grid, P(k) ~ k^-3, + its lossless rope "
                              "SVO built on the GPU; reference camera.txt)"},
}


def orbit_cameras(n=10):
    """n frames on a circular orbit around the unit cube's centre at twice
    its diagonal, looking at the centre."""
    cams = []
    for k in range(n):
        th = 2 * np.pi * k / n
        eye = np.array([0.5, 0.5, 0.5]) + 2 * np.sqrt(3) * np.array([np.sin(th), 0.0, -np.cos(th)])
        cams.append(workloads.look_at(eye, (0.5, 0.5, 0.5)))
    return cams


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def read_cameras(path):
    """ScriptCameraController format: 9 floats per line (forward, up,
    position), read until EOF (src/camera/ScriptCameraController.cpp:4-41)."""
    vals = open(path).read().split()
    if len(vals) % 9:
        raise ValueError("Syntax error in camera input file")
    v = np.array(vals, np.float32).reshape(-1, 9)
    return [(r[0:3], r[3:6], r[6:9]) for r in v]


def host_cpu():
    """The host's CPU model, its logical CPUs, the CPUs this process may run
    on, and the threads the CPU baseline uses: all of those CPUs, or the
    job's share where OMP_NUM_THREADS sets one (a shared host allots each
    GPU job a share of its cores that way)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(usable, share) if share > 0 else usable
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable,
            "omp_num_threads": share or None, "threads": threads}


def spread_order(n):
    """0..n-1 in van der Corput order: every prefix is spread over the range."""
    bits = max(1, (n - 1).bit_length())
    keys = []
    for i in range(1 << bits):
        r = int(format(i, f"0{bits}b")[::-1], 2)
        if r < n:
            keys.append(r)
    return keys


def cpu_baseline(grid, nodes, side, cams, frame_ids, host, budget_s=30.0, width=WIDTH, height=HEIGHT,
                 emission=EMISSION, travs=TRAV_ORDER):
    """The oracle on the host's cores: full frames of the timed steps' camera
    frames (frame_ids, in an order whose every prefix is spread over them),
    every traversal of the workload per frame, stopping after the frame that
    crosses budget_s seconds."""
    from oracle import xo  # the checker, timed here only as the CPU baseline

    threads = host["threads"]
    og = xo.OracleVolume(grid=grid)
    osv = xo.OracleVolume(nodes=nodes, side=side) if any(n != "dda" for n in travs) else None
    rays = 0
    secs = 0.0
    per_rays = {n: 0 for n in travs}
    per_secs = {n: 0.0 for n in travs}
    done = []
    for k in frame_ids:
        f, up, eye = cams[k]
        for name in travs:
            t = abi.TRAVERSALS[name]
            vol, md = (og, list(grid.shape[2::-1][:3])) if t == abi.DDA else (osv, [side] * 3)
            u = xo.make_uniforms(width, height, md, emission=emission)
            t0 = time.perf_counter()
            xo.render(t, vol, xo.make_camera(f, up, eye), u, threads=threads)
            dt = time.perf_counter() - t0
            rays += width * height
            secs += dt
            per_rays[name] += width * height
            per_secs[name] += dt
        done.append(k)
        if secs > budget_s:
            break
    value = rays / secs / 1e6
    return {
        "value": round(value, 3),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle FP32 restatement on {threads} threads of {host['cpu_model']} (host: "
                  f"{host['logical_cpus']} logical CPUs, {host['usable_cpus']} usable by this process"
                  + (f", job share OMP_NUM_THREADS={host['omp_num_threads']}" if host["omp_num_threads"] else "")
                  + f"): camera frames {done} of the timed steps' frames at {width}x{height}, {len(travs)} "
                  f"traversal{'s' if len(travs) > 1 else ''} each ({rays} rays, {secs:.1f} s)",
        "per_core_mrays": round(value / threads, 4),
        "host": host,
        "per_traversal_mrays": {n: round(per_rays[n] / per_secs[n] / 1e6, 3) for n in travs if per_secs[n]},
    }


def plain_same_requests(ctx, volumes, reqs):
    """The plain form on a served request mix: one launch per request, one
    after another; Mrays/s over the summed launch times (the plain form's
    definition) and over the wall time of the loop."""
    vols = {}
    for name, (kind, d, md) in volumes.items():
        vols[name] = (ctx.grid_volume(d) if kind == "grid" else ctx.svo_volume(*d)), md
    frames = {}
    ms_sum, rays = 0.0, 0
    t0 = time.perf_counter()
    for name, t, cam, W, H in reqs:
        vol, md = vols[name]
        f = frames.get((W, H))
        if f is None:
            f = frames[(W, H)] = ctx.frame(W, H)
        ms_sum += ctx.render(vol, t, abi.make_camera(*cam), abi.make_uniforms(W, H, md, emission=EMISSION), f)
        rays += W * H
    wall = time.perf_counter() - t0
    for f in frames.values():
        f.free()
    for vol, _ in vols.values():
        vol.free()
    return {"requests": len(reqs), "mrays_s": round(rays / ms_sum / 1e3, 1),
            "mrays_s_wall": round(rays / wall / 1e6, 1)}


def mixed_views_leg(ctx, scenes, tuning, reps):
    """The many-views batch: 64 views over the resident volumes (traversals
    dda / svo-df / esvo, W x H log-uniform in 256..1920 x 256..1080, orbit
    cameras at distance 0.5..3), one batched launch per repetition, against
    the plain form on the same views (one launch per view).  Both rates are
    over device time.  scenes: name -> (grid volume, grid model_dim, svo
    volume, svo model_dim, emission)."""
    spec = workloads.mixed_views(sorted(scenes))
    views = []
    for name, t, cam, W, H in spec:
        gv, gmd, sv, smd, em = scenes[name]
        vol, md = (gv, gmd) if t == abi.DDA else (sv, smd)
        views.append((vol, t, abi.make_camera(*cam), abi.make_uniforms(W, H, md, emission=em), ctx.frame(W, H)))
    rays = sum(W * H for *_, W, H in spec)
    for v in views:  # warm
        ctx.render(*v)
    plain_ms = 0.0
    for _ in range(reps):
        for v in views:
            plain_ms += ctx.render(*v)
    b = ctx.batch()
    b.set_tuning(tuning["tile_w"], tuning["tiles_per_grab"], tuning["waves_per_simd"], tuning["class_kernels"],
                 tuning["interleave"], tuning["mix_classes"], tuning.get("rays_per_lane"))
    b.set_views(views)
    b.render()  # warm
    batch_ms = [b.render()[0] for _ in range(reps)]
    b.free()
    for v in views:
        v[4].free()
    per_trav = {abi.TRAVERSAL_NAMES[t]: sum(1 for _, tt, *_ in spec if tt == t) for t in (abi.DDA, abi.SVO_DF, abi.ESVO)}
    batched = rays * reps / (sum(batch_ms) * 1e3)
    plain = rays * reps / (plain_ms * 1e3)
    return {"views": len(spec), "volumes": sorted(scenes), "views_per_traversal": per_trav, "rays_per_launch": rays,
            "batched_mrays_s": round(batched, 1), "plain_mrays_s": round(plain, 1),
            "batched_vs_plain": round(batched / plain, 3), "batched_launch_ms": round(float(np.mean(batch_ms)), 3)}


def load_served(srv, volumes):
    """Admits every volume (name -> (kind, data, model_dim)) into the server;
    returns name -> (volume id, kind)."""
    ids = {}
    for name, (kind, d, md) in volumes.items():
        ids[name] = (srv.load_grid(d) if kind == "grid" else srv.load_svo(*d)), kind
    return ids


def served_leg(ctx, volumes, rate, duration_s, seed, tuning, n_saturated=0, plain=False, server=None,
               latency_target_ms=0.0):
    """Poisson arrivals at `rate` req/s for `duration_s` (or n_saturated
    requests all released at once).  volumes: name -> (kind, data, model_dim).
    Returns latency percentiles (ms) and aggregate Mrays/s of the requests
    served; with plain, also the plain form's rate on the same requests.
    server = (server, ids) from load_served reuses resident volumes; otherwise
    a server is made for the leg.  latency_target_ms: the server's latency target
    for this leg (requests that cannot start within it are shed and counted)."""
    if server is None:
        srv = ctx.server(tuning=tuning, latency_target_ms=latency_target_ms, **SERVER_CONFIG)
        ids = load_served(srv, volumes)
    else:
        srv, ids = server
        srv.set_latency_target(latency_target_ms)
    st0 = srv.stats()
    n = n_saturated if n_saturated else int(rate * duration_s)
    named, rng = workloads.request_mix({k: v[1] for k, v in ids.items()}, n, seed)
    offsets = np.zeros(n) if n_saturated else np.cumsum(rng.exponential(1000.0 / rate, n))
    reqs = [(ids[name][0], t, abi.make_camera(*cam), W, H) for name, t, cam, W, H in named]
    t_base = srv.now_ms() + 20.0
    lead_ms = 10.0

    def feed():
        for i, (vid, t, cam, W, H) in enumerate(reqs):
            release = t_base + offsets[i]
            wait = (release - lead_ms - srv.now_ms()) / 1000.0
            if wait > 0:
                time.sleep(wait)
            srv.submit(vid, t, cam, W, H, emission=EMISSION, release_ms=release, tag=i)

    th = threading.Thread(target=feed)
    th.start()
    done = []
    while len(done) < n:
        done += srv.poll(512, timeout_ms=50)
    th.join()
    st = srv.stats()
    if server is None:
        srv.free()
    if server is not None:
        srv.set_latency_target(0.0)
    shed = [c for c in done if c["status"] == abi.XD_ERR_REJECTED]
    bad = [c for c in done if c["status"] not in (abi.XD_OK, abi.XD_ERR_REJECTED)]
    if bad or (shed and not latency_target_ms):
        c = (bad or shed)[0]
        raise RuntimeError(f"{len(bad) + len(shed)} served requests failed (status {c['status']})")
    done = [c for c in done if c["status"] == abi.XD_OK]
    if not done:
        raise RuntimeError(f"all {n} requests were shed under a {latency_target_ms} ms target")
    lat = np.array([c["done_ms"] - c["arrival_ms"] for c in done])
    dev = np.array([c["device_done_ms"] - c["arrival_ms"] for c in done])
    wall = max(c["done_ms"] for c in done) - min(c["arrival_ms"] for c in done)
    rays = sum(c["rays"] for c in done)
    res = {"rate_req_s": rate if not n_saturated else None, "requests": n, "duration_s": round(wall / 1e3, 3),
           "mrays_s": round(rays / wall / 1e3, 1), "p50_ms": round(float(np.percentile(lat, 50)), 3),
           "p99_ms": round(float(np.percentile(lat, 99)), 3),
           "p50_device_ms": round(float(np.percentile(dev, 50)), 3),
           "p99_device_ms": round(float(np.percentile(dev, 99)), 3),
           "mean_rays_per_request": round(rays / len(done)), "launches": st["launches"] - st0["launches"],
           # summed launch durations over the wall time: launches overlap on
           # the device, so this can exceed 1 (it is not a busy fraction)
           "launch_ms_sum_over_wall": round((st["busy_ms"] - st0["busy_ms"]) / wall, 3)}
    if latency_target_ms:
        res["latency_target_ms"] = latency_target_ms
        res["served"] = len(done)
        res["shed"] = len(shed)
        res["shed_frac"] = round(len(shed) / n, 4)
    if server is None:
        res["max_requests_per_launch"] = st["max_requests_per_launch"]
    if plain:
        res["plain_same_requests"] = plain_same_requests(ctx, volumes, named)
        res["served_vs_plain"] = round(res["mrays_s"] / res["plain_same_requests"]["mrays_s"], 3)
    return res


def copy_calibration(gib):
    """A plain device copy of known size, for scaling the FETCH_SIZE /
    WRITE_SIZE counters of a rocprofv3 --pmc run."""
    import torch

    n = int(gib * (1 << 30)) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": n * 4, "launches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="camera frames timed, each with every traversal of the workload (default: 150; 1500 for "
                         "shells128, its 10-frame orbit cycled)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--frame-order", choices=["spread", "consecutive"], default=None,
                    help="spread (default for camera.txt): step k of n renders frame round(k * 150 / n), so any "
                         "number of steps samples the orbit, the zoom and the frames from inside the volume; "
                         "consecutive: step k renders frame k (the default for the cycled 10-frame orbit)")
    ap.add_argument("--batch-frames", type=int, default=None,
                    help="camera frames per batched launch (default: 10; 150 for shells128's 256x256 frames)")
    ap.add_argument("--per-traversal-groups", type=int, default=3,
                    help="launches timed per traversal for the per-traversal batched rates (0: skip)")
    ap.add_argument("--autotune", action="store_true",
                    help="autotune the batch on frames spread over the path (default: the library's launch shape, "
                         "chosen by sweeps on this workload)")
    ap.add_argument("--no-autotune", action="store_true", help="the default; kept for older command lines")
    ap.add_argument("--tuning", default="", help="batch tuning fields as JSON, applied over the defaults "
                                                 "(without --autotune)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="ct512",
                    help="ct512 (default): 512x361x512 volume, five traversals at 1920x1080; shells128: 128^3, "
                         "dda at 256x256 on a 10-frame orbit; lognormal1024: 1024^3, dda and esvo at 3840x2160. "
                         "The mixed-view and served legs run with ct512 only")
    ap.add_argument("--traversals", default="", help="comma-separated subset (default: the workload's)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=8, help="camera frames in the CPU-baseline sample")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="only run a device copy of this size (HBM counter calibration)")
    ap.add_argument("--tile", type=int, default=0, help="per-wave tile width override (8/16/32)")
    ap.add_argument("--plain-order", choices=["traversal", "frame"], default="traversal",
                    help="plain form: each traversal over the whole path in turn (as the reference's benchmark "
                         "runs), or all traversals per frame")
    ap.add_argument("--xcd-chunks", type=int, default=-1,
                    help="plain launches: image block runs per XCD (0: hardware order; default: the library's)")
    ap.add_argument("--mixed-views-reps", type=int, default=3,
                    help="timed launches of the 64-view mixed batch (0: skip)")
    ap.add_argument("--mixed-views-large", type=int, default=1,
                    help="include the 1024^3 log-normal volume in the mixed batch (0: two volumes only)")
    ap.add_argument("--serve-rates", default="200,1000,2000,3000", help="Poisson request rates (req/s); empty: skip")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--serve-tuning", default="", help="served-form tuning fields as JSON, over the defaults")
    ap.add_argument("--serve-tiles-per-launch", type=int, default=0, help="served form: wave tiles per step "
                                                                             "(0: the library's)")
    ap.add_argument("--serve-in-flight", type=int, default=0, help="served form: steps queued on the device "
                                                                   "(0: the library's)")
    ap.add_argument("--c5", action="store_true",
                    help="also serve requests over the 2048^3 and 1024^3 log-normal volumes next to the others")
    ap.add_argument("--c5-rates", default="50,200,500,1000,2000")
    ap.add_argument("--c5-seconds", type=float, default=10.0)
    ap.add_argument("--c5-latency-target-ms", type=float, default=100.0,
                    help="latency target of the 2048^3 served legs run a second time with shedding (0: skip)")
    ap.add_argument("--c5-target-rates", default="1000,2000")
    ap.add_argument("--serve-saturated", type=int, default=400,
                    help="requests released at once for the served form's peak rate (0: skip)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        return 2
    wl = WORKLOADS[args.workload]
    width, height = wl["size"]
    if args.steps is None:
        args.steps = wl["steps"]
    if args.batch_frames is None:
        args.batch_frames = wl["batch_frames"]
    emission = wl["emission"]
    travs = [t for t in args.traversals.split(",") if t] or list(wl["traversals"])
    SERVER_CONFIG.update(tiles_per_launch=args.serve_tiles_per_launch, launches_in_flight=args.serve_in_flight)
    if args.workload != "ct512":
        # the mixed-view and served legs are workloads of their own, run
        # with the default volume
        args.mixed_views_reps = 0
        args.serve_rates = ""
        args.serve_saturated = 0
        args.c5 = False

    ctx = abi.Context(0)
    if args.tile:
        ctx.set_tile(args.tile)
    if args.xcd_chunks >= 0:
        ctx.set_option("xcd_chunks", args.xcd_chunks)
    t0 = time.time()
    grid = {"ct512": synth.ct_body, "shells128": synth.shells_128,
            "lognormal1024": lambda: synth.lognormal(side=1024)}[args.workload]()
    dims = list(grid.shape[2::-1][:3])
    log(f"volume {'x'.join(map(str, dims))} generated in {time.time() - t0:.1f}s, fnv64 "
        f"{synth.volume_hash(grid):016x}")
    t0 = time.time()
    gvol = ctx.grid_volume(grid)
    if args.workload == "lognormal1024":
        nodes, st = ctx.build_octree(gvol, tree=abi.TREE_ROPE)  # the GPU builder (bit-identical nodes)
    else:
        nodes, st = abi.build_octree(grid, tree=abi.TREE_ROPE)
    side = st["side"]
    log(f"lossless rope SVO: {st['node_count']} nodes, depth {st['depth']}, built in {time.time() - t0:.1f}s")
    cams = read_cameras(CAMERA_FILE) if wl["cameras"] == "path" else orbit_cameras()

    svol = ctx.svo_volume(nodes, side)
    frame = ctx.frame(width, height)
    vols = {"dda": (gvol, dims)}
    for n in TRAV_ORDER[1:]:
        vols[n] = (svol, [side] * 3)
    log(f"resident: grid {gvol.info().device_bytes / 1e6:.0f} MB, svo {svol.info().device_bytes / 1e6:.0f} MB")

    def uni(md):
        return abi.make_uniforms(width, height, md, emission=emission)

    if args.frame_order is None:
        args.frame_order = "spread" if wl["cameras"] == "path" else "consecutive"

    def frame_index(k):
        if args.frame_order == "spread":
            return workloads.path_frame(k, args.steps, len(cams))
        return k % len(cams)

    def frame_of(k):
        return cams[frame_index(k)]

    def step(k, ms_acc):
        f, up, eye = frame_of(k)
        cam = abi.make_camera(f, up, eye)
        for name in travs:
            vol, md = vols[name]
            ms = ctx.render(vol, abi.TRAVERSALS[name], cam, uni(md), frame)
            if ms_acc is not None:
                ms_acc[name].append(ms)

    # ---------------- plain form: one launch per traversal per frame --------
    ms_acc = {n: [] for n in travs}
    if args.plain_order == "traversal":
        # the reference's benchmark: one traversal over the whole camera path
        # per run (`xenodon render ... -s <shader>`), then the next
        plain_wall = 0.0
        for name in travs:
            vol, md = vols[name]
            for k in range(args.warmup):
                f, up, eye = frame_of(k)
                ctx.render(vol, abi.TRAVERSALS[name], abi.make_camera(f, up, eye), uni(md), frame)
            t_start = time.perf_counter()
            for k in range(args.steps):
                f, up, eye = frame_of(k)
                ms_acc[name].append(ctx.render(vol, abi.TRAVERSALS[name], abi.make_camera(f, up, eye), uni(md),
                                               frame))
            plain_wall += time.perf_counter() - t_start
    else:
        for k in range(args.warmup):
            step(k, None)
        t_start = time.perf_counter()
        for k in range(args.steps):
            step(k, ms_acc)
        plain_wall = time.perf_counter() - t_start
    log(f"plain: {args.steps} steps in {plain_wall:.2f}s")
    rays_per_frame = width * height
    total_rays = rays_per_frame * args.steps * len(travs)
    plain_value = total_rays / plain_wall / 1e6
    per_trav = {n: rays_per_frame * len(v) / (sum(v) * 1e3) for n, v in ms_acc.items()}
    mean_ms = {n: float(np.mean(v)) for n, v in ms_acc.items()}

    # clamped pixels at this emission coefficient (untimed): worst of frames
    # spread along the path
    saturated = {}
    for name in travs:
        vol, md = vols[name]
        worst = 0.0
        for k in range(0, args.steps, max(1, args.steps // 10)):
            f, up, eye = frame_of(k)
            ctx.render(vol, abi.TRAVERSALS[name], abi.make_camera(f, up, eye), uni(md), frame)
            img = frame.download()
            worst = max(worst, float((img[..., :3] == 255).any(-1).mean()))
        saturated[name] = round(worst, 5)

    # algorithmic bytes of every traversal for every frame (untimed)
    words_of = {}
    for name in travs:
        vol, md = vols[name]
        for k in range(args.steps):
            f, up, eye = frame_of(k)
            words_of[(name, k)] = ctx.count_words(vol, abi.TRAVERSALS[name], abi.make_camera(f, up, eye), uni(md))
    algo_bytes = {n: sum(4 * words_of[(n, k)] + 4 * rays_per_frame for k in range(args.steps)) / args.steps
                  for n in travs}
    plain_roof = {n: round(algo_bytes[n] / (mean_ms[n] * 1e-3) / 1e9, 1) for n in travs}

    # ---------------- batched form: one persistent launch per group --------
    # a group is `batch_frames` consecutive steps x the traversals = views
    B = max(1, min(args.batch_frames, args.steps))
    frames_pool = [ctx.frame(width, height) for _ in range(B * len(travs))]
    batch = ctx.batch()

    def group_views(k0):
        views = []
        for j in range(B):
            f, up, eye = frame_of(k0 + j)
            cam = abi.make_camera(f, up, eye)
            for i, name in enumerate(travs):
                vol, md = vols[name]
                views.append((vol, abi.TRAVERSALS[name], cam, uni(md), frames_pool[j * len(travs) + i]))
        return views

    # tune on frames spread over the whole path (orbit, zoom and inside)
    tune_views = []
    for j in range(B):
        f, up, eye = frame_of((j * args.steps) // B)
        cam = abi.make_camera(f, up, eye)
        for i, name in enumerate(travs):
            vol, md = vols[name]
            tune_views.append((vol, abi.TRAVERSALS[name], cam, uni(md), frames_pool[j * len(travs) + i]))
    batch.set_views(tune_views)
    if args.tuning or (wl.get("tuning") and not args.autotune):
        t = dict(batch.tuning())
        t.update(wl.get("tuning", {}))
        t.update(json.loads(args.tuning or "{}"))
        batch.set_tuning(t["tile_w"], t["tiles_per_grab"], t["waves_per_simd"], t["class_kernels"], t["interleave"],
                         t["mix_classes"], t.get("rays_per_lane"))
    tuning = batch.autotune(trials=2) if args.autotune else batch.tuning()
    log(f"batched: tuning {tuning}")
    for w in range(max(1, args.warmup // B)):
        batch.set_views(group_views(w * B))
        batch.render()
    groups = [k0 for k0 in range(0, args.steps - B + 1, B)]
    batch_ms = []
    batch_rays = 0
    t_start = time.perf_counter()
    for k0 in groups:
        batch.set_views(group_views(k0))
        ms, rays = batch.render()
        batch_ms.append(ms)
        batch_rays += rays
    batch_wall = time.perf_counter() - t_start
    value = batch_rays / batch_wall / 1e6
    log(f"batched: {len(groups)} launches of {B} steps in {batch_wall:.2f}s")
    batch_algo = [sum(4 * words_of[(n, k0 + j)] + 4 * rays_per_frame for j in range(B) for n in travs)
                  for k0 in groups]
    mean_batch_ms = float(np.mean(batch_ms))
    achieved = float(np.mean(batch_algo)) / (mean_batch_ms * 1e-3) / 1e9

    # per-traversal batched rate: groups of one traversal only
    per_trav_batched = {}
    for name in (travs if args.per_traversal_groups > 0 else []):
        vol, md = vols[name]
        sub = ctx.batch()
        sub.set_tuning(tuning["tile_w"], tuning["tiles_per_grab"], tuning["waves_per_simd"], tuning["class_kernels"],
                       tuning["interleave"], tuning["mix_classes"], tuning.get("rays_per_lane"))
        ms_sum = 0.0
        rays_sum = 0
        n_groups = min(len(groups), args.per_traversal_groups)
        for g in range(n_groups):
            k0 = groups[(g * len(groups)) // n_groups]  # spread over the path
            views = []
            for j in range(B * len(travs)):
                f, up, eye = frame_of(k0 + j % B)
                views.append((vol, abi.TRAVERSALS[name], abi.make_camera(f, up, eye), uni(md),
                              frames_pool[j]))
            views = views[:B]
            sub.set_views(views)
            sub.render()  # warm
            ms, rays = sub.render()
            ms_sum += ms
            rays_sum += rays
        per_trav_batched[name] = rays_sum / (ms_sum * 1e3)
        sub.free()

    # the committed rocprofv3 summary of this same configuration (the same
    # command line): HBM bytes per batched launch (FETCH_SIZE + WRITE_SIZE,
    # scaled by a device-copy calibration) and that kernel's mean duration in
    # the profiled run's kernel trace give the measured DRAM rate; its stall
    # counters name what limits the kernel
    run_config = {"workload": args.workload, "steps": len(groups) * B, "warmup": args.warmup,
                  "frame_order": args.frame_order, "batch_frames": B, "traversals": travs,
                  "tuning": {k: tuning[k] for k in sorted(tuning)}}
    traffic = None
    measured = {"profile": None}
    profile_summary = os.path.join(ROOT, "profiles", wl["profile"])
    if os.path.exists(profile_summary):
        prof = json.load(open(profile_summary))
        if prof.get("config") == run_config:
            traffic = prof["hbm_bytes_per_launch"]["batch"]
            k = prof["per_kernel"]["batch"]
            prof_ms = k["mean_launch_ms"]
            gbs = traffic / (prof_ms * 1e-3) / 1e9
            measured = {"profile": "profiles/" + wl["profile"], "profile_mean_launch_ms": prof_ms,
                        "measured_hbm_GBs": round(gbs, 1), "measured_frac": round(gbs / HBM_PEAK_GBS, 4),
                        "l2_hit": k.get("l2_hit"), "valu_lane_utilisation": k.get("valu_lane_utilisation"),
                        "wave_cycles": k.get("wave_cycles"), "limiter": prof.get("limiter")}
        else:
            measured = {"profile": f"profiles/{wl['profile']} covers another configuration ({prof.get('config')})"}
    # The path has no matrix work, so HBM bandwidth is the roofline that
    # bounds it; `limiter` says from the counters how far below that ceiling
    # the kernel runs and what holds it there
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                # a batch of one traversal always runs that traversal's own kernel
                "kernel": "xd::" + ("k_batch_w8" if tuning["waves_per_simd"] == 8 else "k_batch")
                          + (f"<{','.join(str(abi.TRAVERSALS[t]) for t in travs)}> (persistent, one per traversal)"
                             if tuning["class_kernels"] or len(travs) == 1
                             else "<5> (persistent, all five traversals)"),
                "algorithmic_bytes_per_launch": round(float(np.mean(batch_algo))),
                "mean_launch_ms": round(mean_batch_ms, 4), **measured,
                "plain_per_traversal_algorithmic_GBs": plain_roof}

    shells = synth.shells_128()
    snodes, sst = abi.build_octree(shells, tree=abi.TREE_ROPE)

    # ---------------- many views of three volumes in one batched launch ------
    mixed = None
    c3_host = None
    if args.mixed_views_reps > 0:
        sg, ss = ctx.grid_volume(shells), ctx.svo_volume(snodes, sst["side"])
        scenes = {"shells128": (sg, [128] * 3, ss, [sst["side"]] * 3, 4.0),
                  "ct512": (gvol, [512, 361, 512], svol, [side] * 3, emission)}
        extra = []
        if args.mixed_views_large:
            t0 = time.time()
            big = synth.lognormal(side=1024)
            bg = ctx.grid_volume(big)
            bnodes, bst = ctx.build_octree(bg, tree=abi.TREE_ROPE)
            bs = ctx.svo_volume(bnodes, bst["side"])
            if args.c5:
                c3_host = (big, bnodes)
            del big, bnodes
            scenes["lognormal1024"] = (bg, [1024] * 3, bs, [bst["side"]] * 3, 1.0)
            extra = [bg, bs]
            log(f"1024^3 log-normal grid + rope SVO ({bs.info().device_bytes / 1e9:.1f} GB) resident in "
                f"{time.time() - t0:.1f}s")
        mixed = mixed_views_leg(ctx, scenes, tuning, args.mixed_views_reps)
        log(f"mixed views: batched {mixed['batched_mrays_s']} Mrays/s, plain {mixed['plain_mrays_s']} Mrays/s")
        for v in [sg, ss] + extra:
            v.free()

    # ---------------- served form: Poisson arrivals, continuous batching -----
    served = []
    rates = [float(r) for r in args.serve_rates.split(",") if r]
    svols = {"ct512_grid": ("grid", grid, [512, 361, 512]), "ct512_svo": ("svo", (nodes, side), [side] * 3),
             "shells128_grid": ("grid", shells, [128] * 3),
             "shells128_svo": ("svo", (snodes, sst["side"]), [sst["side"]] * 3)}
    # served steps are small and mixed: 8x8 tiles, one tile per grab, one
    # kernel for every traversal per step
    serve_tuning = {"tile_w": [8] * 5, "tiles_per_grab": 1, "waves_per_simd": 0, "class_kernels": 0}
    if args.serve_tuning:
        serve_tuning.update(json.loads(args.serve_tuning))
    if rates or args.serve_saturated:
        for k, r in enumerate(rates):
            res = served_leg(ctx, svols, r, args.serve_seconds, seed=1000 + k, tuning=serve_tuning)
            log(f"served @ {r:.0f} req/s: {res['mrays_s']} Mrays/s, p50 {res['p50_ms']} ms, p99 {res['p99_ms']} ms")
            served.append(res)
        if args.serve_saturated:
            res = served_leg(ctx, svols, 0, 0, seed=999, tuning=serve_tuning, n_saturated=args.serve_saturated,
                             plain=True)
            log(f"served, {args.serve_saturated} requests at once: {res['mrays_s']} Mrays/s, p99 {res['p99_ms']} ms; "
                f"plain on the same requests: {res['plain_same_requests']['mrays_s']} Mrays/s")
            served.append(res)

    # ---------------- served form over the 2048^3 volumes (opt-in) ----------
    served_c5 = None
    if args.c5:
        t0 = time.time()
        if c3_host is None:
            big = synth.lognormal(side=1024)
            bg = ctx.grid_volume(big)
            c3_host = (big, ctx.build_octree(bg, tree=abi.TREE_ROPE)[0])
            bg.free()
            del big
        huge = synth.lognormal(side=2048)
        hg = ctx.grid_volume(huge)
        hnodes, hst = ctx.build_octree(hg, tree=abi.TREE_ROPE)
        hg.free()
        log(f"2048^3 grid + rope octree ({hnodes.shape[0]} nodes) built in {time.time() - t0:.1f}s")
        c5vols = dict(svols)
        c5vols.update({"lognormal1024_grid": ("grid", c3_host[0], [1024] * 3),
                       "lognormal1024_svo": ("svo", (c3_host[1], 1024), [1024] * 3),
                       "lognormal2048_grid": ("grid", huge, [2048] * 3),
                       "lognormal2048_svo": ("svo", (hnodes, hst["side"]), [hst["side"]] * 3)})
        srv = ctx.server(tuning=serve_tuning, **SERVER_CONFIG)
        t0 = time.time()
        ids = load_served(srv, c5vols)
        log(f"served volumes resident ({srv.stats()['resident_bytes'] / 1e9:.1f} GB) in {time.time() - t0:.1f}s")
        served_c5 = []
        for k, r in enumerate(float(x) for x in args.c5_rates.split(",") if x):
            res = served_leg(ctx, c5vols, r, args.c5_seconds, seed=2000 + k, tuning=serve_tuning, server=(srv, ids))
            log(f"served (2048^3 resident) @ {r:.0f} req/s: {res['mrays_s']} Mrays/s, p99 {res['p99_ms']} ms")
            served_c5.append(res)
        if args.c5_latency_target_ms > 0:
            # the same arrivals with a latency target: what cannot start within
            # it is shed instead of queued
            for k, r in enumerate(float(x) for x in args.c5_target_rates.split(",") if x):
                res = served_leg(ctx, c5vols, r, args.c5_seconds, seed=2100 + k, tuning=serve_tuning,
                                 server=(srv, ids), latency_target_ms=args.c5_latency_target_ms)
                log(f"served (2048^3 resident, target {args.c5_latency_target_ms:.0f} ms) @ {r:.0f} req/s: "
                    f"{res['mrays_s']} Mrays/s, p50 {res['p50_ms']} ms, p99 {res['p99_ms']} ms, "
                    f"shed {res['shed_frac']:.1%}")
                served_c5.append(res)
        res = served_leg(ctx, c5vols, 0, 0, seed=2999, tuning=serve_tuning, n_saturated=args.serve_saturated or 400,
                         plain=True, server=(srv, ids))
        log(f"served (2048^3 resident), requests at once: {res['mrays_s']} Mrays/s; plain "
            f"{res['plain_same_requests']['mrays_s']} Mrays/s")
        served_c5.append(res)
        srv.free()
        del huge, hnodes, c3_host

    cpu = None
    if not args.no_cpu_baseline:
        # the timed steps' own frames, in an order whose every prefix is spread
        # over them, up to --cpu-frames of them and about 30 s of CPU work
        step_frames = [frame_index(k) for k in range(len(groups) * B)]
        order = []
        for i in spread_order(len(step_frames)):
            if step_frames[i] not in order:
                order.append(step_frames[i])
        cpu = cpu_baseline(grid, nodes, side, cams, order[:args.cpu_frames], host_cpu(), width=width,
                           height=height, emission=emission, travs=travs)

    out = {
        "metric": "Mrays/s per traversal @1080p, batched/served vs plain; p50/p99 request latency",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": len(groups) * B,
        "warmup": args.warmup,
        "ms_per_step": round(batch_wall / (len(groups) * B) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": wl["data"],
        "config": {"workload": wl["desc"], "frame_order": args.frame_order,
                   "frames_timed": sorted({frame_index(k) for k in range(len(groups) * B)}),
                   "form": f"batched: persistent ray-tile queue, {B} frames x {len(travs)} traversals per "
                           "hipGraph launch", "emission_coeff": emission,
                   "saturated_pixel_fraction": saturated, "traversals": travs,
                   "frames": len(groups) * B, "batch_tuning": tuning},
        "batched_vs_plain": round(value / plain_value, 3),
        "plain": {"value": round(plain_value, 2), "wall_s": round(plain_wall, 3), "order": args.plain_order,
                  "per_traversal_mrays": {n: round(v, 1) for n, v in per_trav.items()},
                  "per_traversal_mean_ms": {n: round(v, 4) for n, v in mean_ms.items()}},
        # each traversal alone in launches of its own: those run that
        # traversal's class kernel, not the all-traversal kernel `value` times
        "batched_per_traversal_class_kernel_mrays": {n: round(v, 1) for n, v in per_trav_batched.items()},
        "mixed_views": mixed,
        "served": served,
        "served_2048": served_c5,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

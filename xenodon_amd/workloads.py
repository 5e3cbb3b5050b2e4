"""Seeded view and request mixes of the multi-view workloads.

  mixed_views()   the many-views batch: W and H log-uniform in 256..1920 and
                  256..1080, an orbit camera at a random angle and distance
                  U[0.5, 3] looking at the volume's centre, the traversal
                  uniform in {dda, svo-df, esvo} and the volume uniform among
                  the resident ones.
  request_mix()   served requests: W x H log-uniform in 128..1920 x 128..1080,
                  a uniform traversal among all five, a volume of the kind the
                  traversal needs and an orbit camera at distance U[0.5, 3].
  path_frame()    which frame of a camera path step k of an n-step run
                  renders: the steps are spread evenly over the whole path.

Cameras are (forward, up, eye) triples in unit-cube space, the reference's
camera-file convention (render.txt:33-41).
"""
import numpy as np

from xenodon_amd import abi

SEED_MIXED_VIEWS = 0x5EED0004C0FFEE04


def path_frame(k, steps, n_frames):
    """Frame of an n_frames-long camera path rendered by step k of `steps`:
    round(k * n_frames / steps) (halves up), so any number of steps samples
    the whole path -- for camera.txt the orbit, the zoom and the frames from
    inside the volume -- rather than its first frames."""
    return ((2 * k * n_frames + steps) // (2 * steps)) % n_frames


def look_at(eye, target, world_up=(0.0, 1.0, 0.0)):
    eye, target = np.asarray(eye, np.float64), np.asarray(target, np.float64)
    f = target - eye
    f /= np.linalg.norm(f)
    up = np.asarray(world_up, np.float64)
    up = up - f * np.dot(up, f)
    up /= np.linalg.norm(up)
    return f, up, eye


def orbit_camera(rng):
    """Eye at distance U[0.5, 3] from the unit cube's centre, random azimuth,
    elevation U[-0.6, 0.6] rad; looks at the centre."""
    ang = rng.uniform(0, 2 * np.pi)
    elev = rng.uniform(-0.6, 0.6)
    d = rng.uniform(0.5, 3.0)
    eye = 0.5 + d * np.array([np.cos(ang) * np.cos(elev), np.sin(elev), np.sin(ang) * np.cos(elev)])
    return look_at(eye, (0.5, 0.5, 0.5))


def _log_uniform(rng, lo, hi):
    return int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))


def mixed_views(volume_names, n=64, seed=SEED_MIXED_VIEWS):
    """n views as (volume name, traversal, (forward, up, eye), W, H)."""
    rng = np.random.default_rng(seed)
    travs = (abi.DDA, abi.SVO_DF, abi.ESVO)
    out = []
    for _ in range(n):
        W = _log_uniform(rng, 256, 1920)
        H = _log_uniform(rng, 256, 1080)
        t = travs[int(rng.integers(len(travs)))]
        name = volume_names[int(rng.integers(len(volume_names)))]
        out.append((name, t, orbit_camera(rng), W, H))
    return out


def request_mix(kinds, n, seed):
    """n requests as (volume name, traversal, (forward, up, eye), W, H);
    kinds: volume name -> "grid" | "svo".  Also returns the generator, for
    arrival times drawn after the requests."""
    grids = [k for k, kind in kinds.items() if kind == "grid"]
    svos = [k for k, kind in kinds.items() if kind == "svo"]
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n):
        t = int(rng.integers(5))
        pool = grids if t == abi.DDA else svos
        name = pool[int(rng.integers(len(pool)))]
        W = _log_uniform(rng, 128, 1920)
        H = _log_uniform(rng, 128, 1080)
        reqs.append((name, t, orbit_camera(rng), W, H))
    return reqs, rng

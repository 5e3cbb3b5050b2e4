"""bench.py on the host: the one-GPU guard, the workload table and the CPU
baseline leg (the oracle on a small frame)."""
import os
import subprocess
import sys

import numpy as np

import bench
from xenodon_amd import abi, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_workload_table():
    assert set(bench.WORKLOADS) == {"ct512", "shells128", "lognormal1024"}
    for name, wl in bench.WORKLOADS.items():
        assert all(t in abi.TRAVERSALS for t in wl["traversals"]), name
        assert wl["cameras"] in ("path", "orbit")
        assert os.path.exists(os.path.join(ROOT, "profiles", wl["profile"])), name
        assert wl["steps"] % wl["batch_frames"] == 0, name
    assert bench.WORKLOADS["ct512"]["traversals"] == bench.TRAV_ORDER
    assert bench.WORKLOADS["ct512"]["size"] == (bench.WIDTH, bench.HEIGHT)


def test_orbit_cameras():
    cams = bench.orbit_cameras()
    assert len(cams) == 10
    for f, up, eye in cams:
        eye = np.asarray(eye, np.float64)
        assert abs(np.linalg.norm(eye - 0.5) - 2 * np.sqrt(3)) < 1e-5
        to_centre = (0.5 - eye) / np.linalg.norm(0.5 - eye)
        assert np.allclose(np.asarray(f, np.float64), to_centre, atol=1e-5)
    assert len(bench.read_cameras(bench.CAMERA_FILE)) == 150


def test_cpu_baseline_contract():
    grid = synth.shells_128()
    nodes, st = abi.build_octree(grid, tree=abi.TREE_ROPE)
    host = dict(bench.host_cpu(), threads=2)
    res = bench.cpu_baseline(grid, nodes, st["side"], bench.orbit_cameras(), [3, 7], host, budget_s=0.0,
                             width=32, height=24, emission=4.0, travs=["dda", "esvo"])
    assert res["unit"] == "Mrays/s" and res["kind"] == "port" and res["cores"] == 2
    assert res["value"] > 0 and set(res["per_traversal_mrays"]) == {"dda", "esvo"}
    assert "32x24" in res["sample"] and "2 traversals" in res["sample"] and "frames [3]" in res["sample"]
    assert host["cpu_model"] in res["sample"]


def test_host_cpu_threads_follow_the_job_share(monkeypatch):
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    assert bench.host_cpu()["threads"] == 1
    monkeypatch.delenv("OMP_NUM_THREADS")
    h = bench.host_cpu()
    assert h["threads"] == h["usable_cpus"] >= 1


def test_spread_order_prefixes_cover_the_range():
    for n in (1, 2, 7, 20, 150):
        o = bench.spread_order(n)
        assert sorted(o) == list(range(n))
    o = bench.spread_order(20)
    assert max(o[:4]) >= 15 and min(o[:4]) == 0  # four frames already span the path


def test_steps_sample_the_whole_camera_path():
    from xenodon_amd import workloads
    frames = [workloads.path_frame(k, 20, 150) for k in range(20)]
    assert frames[0] == 0 and frames[-1] >= 140 and len(set(frames)) == 20
    assert [workloads.path_frame(k, 150, 150) for k in range(150)] == list(range(150))

"""`xenodon render --headless` end to end on the GPU: camera script, several
output regions stitched onto one frame, PNG frames equal to the oracle's
render of the whole display, and the --stats-output text
(src/render/RenderStats.cpp:127-158)."""
import re
import subprocess

import numpy as np
import pytest
from PIL import Image

from oracle import xo
from tests import helpers as H
from xenodon_amd import abi
from xenodon_amd import io as xio

pytestmark = pytest.mark.gpu

W, HH = 96, 48
# two regions of one frame on GPU 0 (a headless config has no comment syntax)
CONF = """device {
    vkindex = 0
    offset = (0, 0)
    extent = (40, 48)
}
device {
    vkindex = 0
    offset = (40, 0)
    extent = (56, 48)
}
"""


@pytest.fixture(scope="module")
def scene(tmp_path_factory):
    if abi.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    d = tmp_path_factory.mktemp("cli")
    g = H.blocky_grid(32, 4, seed=19)
    xio.save_tiff(d / "v.tif", g)
    nodes, st = xo.build_octree(g, tree=xo.TREE_ROPE)
    xio.save_svo(d / "v.svo", nodes, st["side"])
    cams = H.orbit_cameras(3, 1.7, seed=23)
    with open(d / "cams.txt", "w") as f:
        for fw, up, eye in cams:
            f.write(" ".join(f"{float(np.float32(v)):.9g}" for v in (*fw, *up, *eye)) + "\n")
    (d / "h.conf").write_text(CONF)
    return d, g, nodes, st["side"], cams


def _run(d, *args):
    r = subprocess.run([xio.CLI_PATH, "render", *map(str, args)], capture_output=True, text=True, cwd=d,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Error" not in r.stdout, r.stdout
    return r.stdout


def _oracle(vol, t, cam, e, ratio=(1.0, 1.0, 1.0)):
    fw, up, eye = cam
    tr = np.float32(eye) / np.float32(ratio)
    u = xo.make_uniforms(W, HH, vol.model_dim, emission=e, voxel_ratio=ratio)
    return xo.render(t, vol, xo.make_camera(fw, up, tr), u)["rgba"]


def test_render_grid_frames_and_stats(scene):
    d, g, nodes, side, cams = scene
    out = _run(d, "--headless", "h.conf", "v.tif", "--camera", "cams.txt", "-e", "3",
               "--stats-output", "s.txt", "--output", "f-{:0>3}.png")
    assert "Setup: 2 devices, with 1, 1 outputs" in out
    assert "Using shader 'dda'" in out and "Total resolution: 96x48 pixels" in out
    vol = xo.OracleVolume(grid=g)
    for k, cam in enumerate(cams):
        img = np.asarray(Image.open(d / f"f-{k:03d}.png").convert("RGBA"))
        assert np.array_equal(img, _oracle(vol, xo.DDA, cam, 3.0)), k
    lines = (d / "s.txt").read_text().splitlines()
    num = r"[-+0-9.e]+"
    assert re.fullmatch(r"total rays: 13824", lines[0])
    assert re.fullmatch(rf"total render time: {num}", lines[1])
    assert re.fullmatch(rf"total mray/s: {num}", lines[2])
    assert re.fullmatch(rf"average fps: {num}", lines[3])
    assert lines[4] == "frames: 3"
    assert lines[5] == ("# Frame number: total rays, outputs, total render time, max render time, min render "
                        "time, mray/s")
    for k in range(3):
        m = re.fullmatch(rf"frame {k}: 4608 rays, 2, ({num}) ms, ({num}) ms, ({num}) ms, ({num}) mray/s",
                         lines[6 + k])
        assert m, lines[6 + k]
        tot, mx, mn, mr = map(float, m.groups())
        assert mn <= mx <= tot and mr > 0
    assert len(lines) == 9


@pytest.mark.parametrize("shader", ["svo-naive", "esvo", "svo-df", "svo-rope"])
def test_render_svo_with_repeat_and_voxel_ratio(scene, shader):
    d, g, nodes, side, cams = scene
    tag = shader.replace("-", "")
    _run(d, "--headless", "h.conf", "v.svo", "-s", shader, "--camera", "cams.txt", "-e", "2", "-r", "1:2:1.5",
         "--repeat", "2", "--output", tag + "-{}.png", "-q", "--stats-output", tag + ".txt")
    vol = xo.OracleVolume(nodes=nodes, side=side)
    t = abi.TRAVERSALS[shader]
    for k in range(6):
        img = np.asarray(Image.open(d / f"{tag}-{k}.png").convert("RGBA"))
        assert np.array_equal(img, _oracle(vol, t, cams[k // 2], 2.0, (1.0, 2.0, 1.5))), (shader, k)
    assert "frames: 6" in (d / (tag + ".txt")).read_text()


def test_discard_output_and_errors(scene):
    d, g, nodes, side, cams = scene
    out = _run(d, "--headless", "h.conf", "v.tif", "--camera", "cams.txt", "--discard-output",
               "--stats-output", "disc.txt")
    assert "Saving frame" not in out and "frames: 3" in (d / "disc.txt").read_text()
    r = subprocess.run([xio.CLI_PATH, "render", "--headless", "h.conf", "v.svo", "-s", "dda", "--camera",
                        "cams.txt"], capture_output=True, text=True, cwd=d, timeout=300)
    assert "Error: Shader 'dda' is incompatible with model type 'svo' (requires 'tiff')" in r.stdout
    (d / "bad_cams.txt").write_text("0 0 1 0 1 0 0.5 0.5\n")
    r = subprocess.run([xio.CLI_PATH, "render", "--headless", "h.conf", "v.tif", "--camera", "bad_cams.txt"],
                       capture_output=True, text=True, cwd=d, timeout=300)
    assert "Error: Syntax error in camera input file" in r.stdout


@pytest.mark.parametrize("opts,tree,kw", [([], xo.TREE_SPARSE, {}),
                                          (["--rope", "--chan-diff", "20"], xo.TREE_ROPE, {"chan_diff": 20}),
                                          (["--dag"], xo.TREE_DAG, {}),
                                          (["--dag", "--std-dev", "30.5"], xo.TREE_DAG,
                                           {"heuristic": xo.HEUR_STD_DEV, "std_dev": 30.5})])
def test_convert_on_the_gpu_matches_oracle(scene, tmp_path, opts, tree, kw):
    """`xenodon convert --gpu 0` builds on the device and writes the same
    .svo as the reference converter's restatement."""
    g = H.blocky_grid(48, 4, seed=23)
    g[5:9, 3:7, 20:30, :3] = 0
    src, dst = tmp_path / "g.tif", tmp_path / "g.svo"
    xio.save_tiff(src, g)
    r = subprocess.run([xio.CLI_PATH, "convert", "--gpu", "0", *opts, str(src), str(dst)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0 and "Error" not in r.stdout, r.stdout + r.stderr
    nodes, side = xio.load_svo(dst)
    ref, st = xo.build_octree(g, tree=tree, **kw)
    assert side == st["side"] and np.array_equal(nodes, ref)


def test_sysinfo_prints_each_device_pci_address():
    """`xenodon sysinfo` prints every GPU with its PCI address in the format of
    src/sysinfo.cpp:53-59 ("{:0>8}:{:0>2}:{:0>2}.{}": decimal domain, bus,
    device and function).  The address is checked against the PCI devices the
    kernel lists under /sys/bus/pci/devices (hex "dddd:bb:dd.f"), where the
    box exposes them."""
    import os
    import re

    n = abi.device_count()
    if n < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    r = subprocess.run([xio.CLI_PATH, "sysinfo"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == "System setup information"
    addrs = re.findall(r"^\tpci bus: (\d{8}):(\d{2,}):(\d{2,})\.(\d+)$", r.stdout, re.M)
    assert len(addrs) == n, r.stdout
    assert len(re.findall(r"^GPU \d+:\n\tname: '.+'\n\ttype: DiscreteGpu\n\tpci bus: ", r.stdout, re.M)) == n
    sysfs = "/sys/bus/pci/devices"
    listed = set(os.listdir(sysfs)) if os.path.isdir(sysfs) else set()
    for i, a in enumerate(addrs):
        dom, bus, dev, fn = (int(v) for v in a)
        assert (dom, bus, dev, fn) == abi.device_describe(i)[1]
        assert bus < 256 and dev < 32 and fn < 8
        if listed:
            assert f"{dom:04x}:{bus:02x}:{dev:02x}.{fn:x}" in listed, (a, sorted(listed)[:8])

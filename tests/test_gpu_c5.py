"""The 2048^3 log-normal workload in the served form.  The 32 GiB grid and its
lossless rope octree (built on the GPU; 1.17e9 nodes, 46.6 GB, in the tagged
device layout with 31-bit indices, kept as .svo words only above 2^31) are
admitted by a server next to the 128^3 shells volumes.  Requests of all five
traversals, sizes log-uniform in 128..1920 x 128..1080 and orbit cameras
arrive at Poisson times (500 req/s).  Every served frame equals the plain
render of the same view, and bands of rows of a dda and an esvo request on
the 2048^3 volumes equal the oracle."""
import os
import time

import numpy as np
import pytest

from oracle import xo
from xenodon_amd import abi, synth, workloads

pytestmark = pytest.mark.gpu

EMISSION = 1.0
N_REQUESTS = 30
RATE = 500.0  # req/s


@pytest.fixture(scope="module")
def c5():
    if abi.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    ctx = abi.Context(0)
    t0 = time.time()
    grid = synth.lognormal(side=2048)
    print(f"2048^3 grid generated in {time.time() - t0:.0f}s", flush=True)
    gv = ctx.grid_volume(grid)
    nodes, st = ctx.build_octree(gv, tree=abi.TREE_ROPE)
    print(f"rope octree built on the GPU: {time.time() - t0:.0f}s", flush=True)
    sv = ctx.svo_volume(nodes, st["side"])
    print(f"volumes resident: {time.time() - t0:.0f}s", flush=True)
    yield ctx, grid, gv, nodes, st, sv
    ctx.close()


@pytest.mark.timeout(900)
def test_c5_volumes_resident(c5):
    ctx, grid, gv, nodes, st, sv = c5
    assert gv.info().device_bytes >= 32 << 30
    assert st["side"] == 2048 and st["depth"] == 11
    assert 1 << 30 < nodes.shape[0] < 1 << 31
    # 40 B per node either way; the tagged layout adds the 128 MiB start table
    assert sv.info().device_bytes == nodes.nbytes + 8 * 256 ** 3
    print(f"2048^3 rope octree: {nodes.shape[0]} nodes, {nodes.nbytes / 1e9:.1f} GB")


@pytest.mark.timeout(900)
def test_c5_served_requests_match_plain_and_oracle(c5):
    ctx, grid, gv, nodes, st, sv = c5
    shells = synth.shells_128()
    snodes, sst = abi.build_octree(shells, tree=abi.TREE_ROPE)
    data = {"ln2048_grid": ("grid", grid, [2048] * 3), "ln2048_svo": ("svo", (nodes, 2048), [2048] * 3),
            "shells_grid": ("grid", shells, [128] * 3), "shells_svo": ("svo", (snodes, sst["side"]), [sst["side"]] * 3)}
    plain_vols = {"ln2048_grid": gv, "ln2048_svo": sv,
                  "shells_grid": ctx.grid_volume(shells), "shells_svo": ctx.svo_volume(snodes, sst["side"])}
    reqs, rng = workloads.request_mix({k: v[0] for k, v in data.items()}, N_REQUESTS, seed=55)
    offsets = np.cumsum(rng.exponential(1000.0 / RATE, N_REQUESTS))
    assert {t for _, t, *_ in reqs} == set(range(5))

    srv = ctx.server()
    ids = {name: (srv.load_grid(d) if kind == "grid" else srv.load_svo(*d)) for name, (kind, d, md) in data.items()}
    t0 = srv.now_ms() + 5.0
    sub = {}
    for i, (name, t, cam, W, H) in enumerate(reqs):
        rid, _ = srv.submit(ids[name], t, abi.make_camera(*cam), W, H, emission=EMISSION, release_ms=t0 + offsets[i],
                            tag=i)
        sub[rid] = i
    done = []
    while len(done) < N_REQUESTS:
        done += srv.poll(64, timeout_ms=-1)
    srv.free()
    assert all(c["status"] == 0 for c in done)
    images = {c["tag"]: c["image"] for c in done}

    oracle_checked = set()
    for i, (name, t, cam, W, H) in enumerate(reqs):
        md = data[name][2]
        u = abi.make_uniforms(W, H, md, emission=EMISSION)
        ref, _ = ctx.render_image(plain_vols[name], t, abi.make_camera(*cam), u)
        assert np.array_equal(images[i], ref), (i, name, abi.TRAVERSAL_NAMES[t], W, H)
        if name.startswith("ln2048") and t in (abi.DDA, abi.ESVO) and t not in oracle_checked and ref[..., :3].any():
            oracle_checked.add(t)
            ov = xo.OracleVolume(grid=grid) if t == abi.DDA else xo.OracleVolume(nodes=nodes, side=2048)
            y = H // 2
            ou = xo.make_uniforms(W, 2, md, emission=EMISSION, out_offset=(0, y), disp_offset=(0, 0),
                                  disp_extent=(W, H))
            oref = xo.render(t, ov, xo.make_camera(*cam), ou, threads=min(16, os.cpu_count() or 1))["rgba"]
            assert np.array_equal(ref[y:y + 2], oref), (i, name, abi.TRAVERSAL_NAMES[t])
    assert oracle_checked == {abi.DDA, abi.ESVO}
    plain_vols["shells_grid"].free()
    plain_vols["shells_svo"].free()


@pytest.mark.timeout(900)
@pytest.mark.parametrize("waves", [8, 0])
def test_c5_dda_class_kernel_batch_matches_plain(c5, waves):
    """The DDA class kernel (macro words of the next steps loaded during the
    current gathers) on the 2048^3 grid, whose texel offsets need 64 bits:
    every batched frame equals the plain render."""
    ctx, grid, gv, nodes, st, sv = c5
    rng = np.random.default_rng(77)
    views = []
    for W, H in ((640, 360), (256, 256), (333, 217)):
        cam = abi.make_camera(*workloads.orbit_camera(rng))
        views.append((gv, abi.DDA, cam, abi.make_uniforms(W, H, [2048] * 3, emission=EMISSION), ctx.frame(W, H)))
    b = ctx.batch()
    tw = b.tuning()
    b.set_tuning(tw["tile_w"], tw["tiles_per_grab"], waves, 1, tw["interleave"], tw["mix_classes"])
    b.set_views(views)
    b.render()
    lit = 0
    for vol, t, cam, u, f in views:
        got = f.download()
        ref, _ = ctx.render_image(vol, t, cam, u)
        assert np.array_equal(got, ref)
        lit += int(got[..., :3].any())
        f.free()
    assert lit >= 2
    b.free()

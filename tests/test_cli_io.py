"""The CLI's file formats (CPU): TIFF stacks read with TIFFReadRGBAImage
semantics (checked against libtiff itself where it is installed), .svo and
PNG round trips, the output-path formatter, and `xenodon convert` against the
oracle's converter.  No GPU needed."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
from PIL import Image

from oracle import xo
from tests import helpers as H
from xenodon_amd import io as xio

LIBTIFF = "/opt/conda/lib/libtiff.so"


def _rgba_grid(shape_zyx, seed, alpha=255):
    rng = np.random.default_rng(seed)
    g = rng.integers(0, 256, (*shape_zyx, 4), dtype=np.uint8)
    if alpha is not None:
        g[..., 3] = alpha
    return g


def test_tiff_round_trip(tmp_path):
    g = _rgba_grid((5, 7, 9), seed=1)
    p = tmp_path / "g.tif"
    xio.save_tiff(p, g)
    assert np.array_equal(xio.load_tiff(p), g)


def _libtiff_rgba(path):
    """Every page of `path` through libtiff's TIFFReadRGBAImage, as [Z, Y, X, 4]."""
    L = ctypes.CDLL(LIBTIFF)
    L.TIFFOpen.restype = ctypes.c_void_p
    L.TIFFOpen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    L.TIFFNumberOfDirectories.argtypes = [ctypes.c_void_p]
    L.TIFFSetDirectory.argtypes = [ctypes.c_void_p, ctypes.c_uint16]
    L.TIFFReadRGBAImage.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_int]
    L.TIFFClose.argtypes = [ctypes.c_void_p]
    t = L.TIFFOpen(str(path).encode(), b"r")
    assert t
    pages = []
    try:
        for i in range(L.TIFFNumberOfDirectories(t)):
            assert L.TIFFSetDirectory(t, i) == 1
            w, h = ctypes.c_uint32(), ctypes.c_uint32()
            L.TIFFGetField(ctypes.c_void_p(t), ctypes.c_uint32(256), ctypes.byref(w))
            L.TIFFGetField(ctypes.c_void_p(t), ctypes.c_uint32(257), ctypes.byref(h))
            raster = np.zeros((h.value, w.value), np.uint32)
            assert L.TIFFReadRGBAImage(t, w.value, h.value, raster.ctypes.data, 0) == 1
            pages.append(raster.view(np.uint8).reshape(h.value, w.value, 4))
    finally:
        L.TIFFClose(t)
    return np.stack(pages)


needs_libtiff = pytest.mark.skipif(not os.path.exists(LIBTIFF), reason="libtiff is not installed")


@needs_libtiff
@pytest.mark.parametrize("mode", ["RGB", "RGBA", "L", "P", "LA"])
@pytest.mark.parametrize("compression", [None, "tiff_lzw", "tiff_deflate", "packbits"])
def test_tiff_reader_matches_libtiff(tmp_path, mode, compression):
    rng = np.random.default_rng(hash((mode, compression)) & 0xFFFF)
    pages = []
    for z in range(3):
        a = rng.integers(0, 256, (11, 13, 4), dtype=np.uint8)
        a[2:5, 3:9] = a[2, 3]  # runs, so the compressors have repeats to find
        img = Image.fromarray(a, "RGBA").convert(mode) if mode != "P" else \
            Image.fromarray(a[..., :3], "RGB").quantize(colors=37)
        pages.append(img)
    p = tmp_path / f"s_{mode}_{compression}.tif"
    kw = {} if compression is None else {"compression": compression}
    pages[0].save(p, save_all=True, append_images=pages[1:], **kw)
    ref = _libtiff_rgba(p)
    got = xio.load_tiff(p)
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)


@needs_libtiff
def test_own_tiff_writer_matches_libtiff_and_premultiplies_like_it(tmp_path):
    g = _rgba_grid((4, 6, 10), seed=3, alpha=None)  # random alpha: unassociated, premultiplied on read
    p = tmp_path / "w.tif"
    xio.save_tiff(p, g)
    ref = _libtiff_rgba(p)
    assert np.array_equal(xio.load_tiff(p), ref)
    a = g[..., 3:4].astype(np.uint32)
    expect = g.copy()
    expect[..., :3] = ((g[..., :3].astype(np.uint32) * a + 127) // 255).astype(np.uint8)
    assert np.array_equal(ref, expect)


def test_tiff_errors(tmp_path):
    with pytest.raises(xio.IoError, match="Failed to open"):
        xio.load_tiff(tmp_path / "missing.tif")
    (tmp_path / "junk.tif").write_bytes(b"not a tiff at all")
    with pytest.raises(xio.IoError, match="Failed to open"):
        xio.load_tiff(tmp_path / "junk.tif")
    imgs = [Image.fromarray(np.zeros((4, 5, 3), np.uint8)), Image.fromarray(np.zeros((4, 6, 3), np.uint8))]
    imgs[0].save(tmp_path / "ragged.tif", save_all=True, append_images=imgs[1:])
    with pytest.raises(xio.IoError, match="Dimensions of layer 1 differ from previous dimensions"):
        xio.load_tiff(tmp_path / "ragged.tif")
    Image.fromarray(np.zeros((4, 5), np.uint16)).save(tmp_path / "deep.tif")
    with pytest.raises(xio.IoError, match="bits per sample"):
        xio.load_tiff(tmp_path / "deep.tif")


def test_svo_round_trip_and_errors(tmp_path):
    g = H.blocky_grid(16, 2, seed=5)
    nodes, st = xo.build_octree(g, tree=xo.TREE_ROPE)
    p = tmp_path / "t.svo"
    xio.save_svo(p, nodes, st["side"])
    raw = p.read_bytes()
    assert raw[:8] == b"XNDN-SVO" and len(raw) == 24 + 40 * nodes.shape[0]
    assert int.from_bytes(raw[8:16], "little") == st["side"]
    assert int.from_bytes(raw[16:24], "little") == nodes.shape[0]
    back, side = xio.load_svo(p)
    assert side == st["side"] and np.array_equal(back, nodes)
    (tmp_path / "short.svo").write_bytes(raw[:-4])
    with pytest.raises(xio.IoError, match="File size does not match number of nodes"):
        xio.load_svo(tmp_path / "short.svo")
    (tmp_path / "id.svo").write_bytes(b"XNDN-SVX" + raw[8:])
    with pytest.raises(xio.IoError, match="Invalid format id"):
        xio.load_svo(tmp_path / "id.svo")
    with pytest.raises(xio.IoError, match="Failed to open"):
        xio.load_svo(tmp_path / "none.svo")


def test_png_writer(tmp_path):
    img = _rgba_grid((1, 37, 53), seed=9, alpha=None)[0]
    p = tmp_path / "f.png"
    xio.save_png(p, img)
    back = np.asarray(Image.open(p).convert("RGBA"))
    assert np.array_equal(back, img)


def test_output_path_format():
    assert xio.format_output_path("out-{}.png", 7) == "out-7.png"
    assert xio.format_output_path("out-{:0>3}.png", 7) == "out-007.png"
    assert xio.format_output_path("f{0}_{:>4}|{:<3}|{:^5}", 12) == "f12_  12|12 | 12  "
    assert xio.format_output_path("{{lit}}-{:d}", 3) == "{lit}-3"
    assert xio.format_output_path("{:03}", 5) == "005"
    assert xio.format_output_path("plain.png", 5) == "plain.png"
    for bad in ["{1}", "{:x}", "{", "a}b"]:
        with pytest.raises(xio.IoError, match="Failed to format output filename"):
            xio.format_output_path(bad, 1)


def _cli(*args, cwd=None):
    r = subprocess.run([xio.CLI_PATH, *map(str, args)], capture_output=True, text=True, cwd=cwd, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.parametrize("opts,tree,kw", [
    ([], xo.TREE_SPARSE, {}),
    (["--dag"], xo.TREE_DAG, {}),
    (["--rope", "--chan-diff", "20"], xo.TREE_ROPE, {"chan_diff": 20}),
    (["--std-dev", "12.5"], xo.TREE_SPARSE, {"heuristic": xo.HEUR_STD_DEV, "std_dev": 12.5}),
])
def test_convert_cli_matches_oracle(tmp_path, opts, tree, kw):
    g = H.blocky_grid(32, 4, seed=11)
    g[5:9, 3:7, 20:30, :3] = 0
    src, dst = tmp_path / "v.tif", tmp_path / "v.svo"
    xio.save_tiff(src, g)
    out = _cli("convert", *opts, src, dst)
    nodes, side = xio.load_svo(dst)
    ref, st = xo.build_octree(g, tree=tree, **kw)
    assert side == st["side"] and np.array_equal(nodes, ref)
    assert "32x32x32 = 32768 pixels" in out
    assert f" Unique nodes: {ref.shape[0]:,} (" in out
    assert f" Depth: {st['depth']}" in out


def test_cli_argument_errors(tmp_path):
    cases = [
        (["render"], "Error: Missing required positional argument <volume path>"),
        (["render", "v.tif"], "Error: Missing required backend --xorg, --headless or --direct"),
        (["render", "--headless", "a", "--direct", "b", "v.tif"],
         "Error: --xorg, --headless and --direct are mutually exclusive"),
        (["render", "-e", "x", "v.tif"], "Error: Invalid value for <emission coefficient> of parameter -e"),
        (["render", "-e", "1", "--emission-coeff", "2", "v.tif"],
         "Error: Duplicate specification of parameter --emission-coeff/-e"),
        (["render", "-q", "-q", "v.tif"], "Error: Duplicate specification of flag --quiet/-q"),
        (["render", "--bogus", "v.tif"], "Error: Unrecognized option --bogus"),
        (["render", "-r", "1:0:1", "v.tif"], "Error: Invalid value for <voxel dimension ratio> of parameter -r"),
        (["render", "--headless", "c", "--discard-output", "--output", "x", "v.tif"],
         "Error: --dont-save and --output are mutually exclusive"),
        (["render", "--output", "x", "--xorg", "v.tif"], "Error: --output requires --headless"),
        (["render", "a.tif", "b.tif"], "Error: Unexpected positional argument 'b.tif'"),
        (["convert", "--chan-diff", "256", "a", "b"], "Error: Invalid value for <channel difference> of parameter --chan-diff"),
        (["convert", "--chan-diff", "1", "--std-dev", "2", "a", "b"],
         "Error: --std-dev and --chan-diff are mutually exclusive"),
        (["convert", "missing.tif", "o.svo"], "Error reading 'missing.tif': Failed to open"),
        (["convert", "--gpu", "x", "a", "b"], "Error: Invalid value for <gpu index> of parameter --gpu"),
        (["convert", "--gpu", "0", "--dag", "--rope", "a", "b"], "Error: --dag and --rope are mutually exclusive"),
        (["help", "a", "b"], "Error: Invalid usage of subcommand 'help'"),
        (["frobnicate"], "Error: Invalid subcommand 'frobnicate'"),
    ]
    for args, msg in cases:
        out = _cli(*args, cwd=tmp_path)
        assert msg in out, (args, out)


@pytest.mark.parametrize("text,msg", [
    ("device {\n vkindex = 0\n offset = (0, 0)\n}\n", "Configuration error: Missing key 'extent'"),
    ("device {\n vkindex = 0\n vkindex = 1\n offset = (0, 0)\n extent = (4, 4)\n}\n",
     "Configuration error: Ambiguous key 'vkindex'"),
    ("device {\n vkindex = x\n}\n", "Parse error at line 2, col 13: Expected numeric character, found 'x'"),
    ("device {\n colour = 1\n}\n", "Unexpected key 'colour'"),
    ("\n \n", "Configuration error: At least one device entry is required"),
    # headless configs are parsed multiline, where '#' is not a comment
    # (src/core/Parser.cpp:16-19, src/core/Config.h:140)
    ("# nothing here\n", "Parse error at line 1, col 1: Unexpected character '#'"),
    ("device {\n vkindex = 0 # first GPU\n}\n", "Parse error at line 2, col 14: Unexpected character '#'"),
    ("device { vkindex = 0 offset = (0, 0) extent = (4, 4) } trailing", "Unexpected key 'trailing'"),
])
def test_headless_config_errors(tmp_path, text, msg):
    (tmp_path / "h.conf").write_text(text)
    out = _cli("render", "--headless", "h.conf", "v.tif", cwd=tmp_path)
    assert "Error: Failed to initialize backend: Failed to read config file 'h.conf': " in out
    assert msg in out, out


def test_help_topics():
    assert "xenodon render [options] <volume path>" in _cli("help", "render")
    assert "vkindex" in _cli("help", "headless-config")
    assert _cli().startswith("Usage:")
    assert "Error: No such topic nope" in _cli("help", "nope")

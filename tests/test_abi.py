"""CPU checks of the C-ABI boundary: the library loads and exports every
symbol include/xenodon.h declares.  No compute calls (no GPU here)."""
import ctypes
import os
import re
import subprocess

from xenodon_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    inc = os.path.join(ROOT, "include")
    for name in os.listdir(inc):
        if not name.endswith(".h"):
            continue
        text = open(os.path.join(inc, name)).read()
        syms |= set(re.findall(r"XD_API\s+[\w\s\*]+?\b(xd_\w+)\s*\(", text))
    return syms


def test_header_declares_the_plain_path():
    syms = declared_symbols()
    for s in ("xd_open", "xd_close", "xd_volume_from_grid", "xd_volume_from_svo", "xd_render",
              "xd_frame_download", "xd_last_error", "xd_count_words"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(abi.LIB_PATH)
    missing = [s for s in sorted(declared_symbols()) if not hasattr(lib, s)]
    assert not missing, missing


def test_abi_struct_sizes_match_header():
    # push constant 48 B (Renderer.h:37), Rect2D 16 B, uniforms 68 B (+pad)
    assert ctypes.sizeof(abi.Camera) == 48
    assert ctypes.sizeof(abi.Rect) == 16
    assert ctypes.sizeof(abi.Uniforms) == 68
    assert ctypes.sizeof(abi.VolumeInfo) == 64


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """Every struct of include/xenodon.h has the size and field offsets the
    ctypes mirror assumes, as laid out by the C compiler."""
    structs = {"xd_camera": abi.Camera, "xd_rect": abi.Rect, "xd_uniforms": abi.Uniforms,
               "xd_volume_info": abi.VolumeInfo, "xd_view": abi.View, "xd_batch_tuning": abi.BatchTuning,
               "xd_build_stats": abi.BuildStats, "xd_server_config": abi.ServerConfig, "xd_request": abi.Request,
               "xd_completion": abi.Completion, "xd_server_stats": abi.ServerStats}
    lines = ['#include "xenodon.h"', "#include <stddef.h>", "#include <stdio.h>", "int main(void) {"]
    expect = []
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        expect.append(f"{cname} {ctypes.sizeof(py)}")
        for field, _ in py._fields_:
            lines.append(f'printf("{cname}.{field} %zu\\n", offsetof({cname}, {field}));')
            expect.append(f"{cname}.{field} {getattr(py, field).offset}")
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    assert [g for g in got if g] == expect


def test_volume_footprint_without_device():
    # 512x361x512: 128*91*128 bricks of 256 B + 1 occupancy byte, plus a
    # 32x23x32-cell macro bitmask (736 words)
    assert abi.volume_footprint(abi.VOLUME_GRID, (512, 361, 512)) == 128 * 91 * 128 * 257 + 736 * 4
    # 128^3: 32^3 bricks, 16^3 macro cells of 8 voxels (128 words)
    assert abi.volume_footprint(abi.VOLUME_GRID, (128, 128, 128)) == 32 ** 3 * 257 + 128 * 4
    # octree: 32 B of child words + 8 B of {color, depth} per node, plus the
    # 256^3-cell start table of 8 B entries
    assert abi.volume_footprint(abi.VOLUME_SVO, node_count=1000) == 40 * 1000 + 8 * 256 ** 3
    for bad in [dict(kind=abi.VOLUME_GRID, dims=(0, 4, 4)), dict(kind=abi.VOLUME_SVO, node_count=0),
                dict(kind=7, dims=(4, 4, 4))]:
        try:
            abi.volume_footprint(**bad)
        except abi.XdError:
            continue
        raise AssertionError(bad)


def test_errors_are_reported_not_raised_without_device():
    if abi.device_count() > 0:
        return  # covered by the GPU tests
    h = ctypes.c_void_p()
    st = abi.lib().xd_open(0, ctypes.byref(h))
    assert st != 0 and not h.value
    assert abi.lib().xd_last_error()

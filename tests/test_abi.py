"""CPU checks of the C-ABI boundary: the library loads and exports every
symbol include/xenodon.h declares.  No compute calls (no GPU here)."""
import ctypes
import os
import re

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


def test_errors_are_reported_not_raised_without_device():
    if abi.device_count() > 0:
        return  # covered by the GPU tests
    h = ctypes.c_void_p()
    st = abi.lib().xd_open(0, ctypes.byref(h))
    assert st != 0 and not h.value
    assert abi.lib().xd_last_error()

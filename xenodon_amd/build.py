"""Builds the native parts of xenodon_amd in-tree.

  xenodon_amd/libxenodon_amd.so   HIP kernels + C-ABI (include/xenodon.h), gfx950
  xenodon_amd/bin/xenodon         C++ CLI (render / convert / help / sysinfo)
  xenodon_amd/libxd_io.so         the CLI's file formats (TIFF, .svo, PNG) as a C library

Run `python -m xenodon_amd.build` (or __graft_entry__.build()).  hipcc
cross-compiles for gfx950 without a GPU.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libxenodon_amd.so")
BIN_DIR = os.path.join(PKG, "bin")
ARCH = os.environ.get("XD_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["xd_kernels.hip", "xd_runtime.hip", "xd_batch.hip", "xd_serve.hip", "xd_build.hip"]
HOST_SOURCES = ["xd_convert.cpp", "xd_layout.cpp"]

# IEEE-strict float semantics (bit parity with the CPU restatement), and no
# packed FP32 instructions (v_pk_*_f32 are kept out of generated code).
HIPCC_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def hipcc():
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return p


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... -o {cmd[cmd.index('-o') + 1] if '-o' in cmd else ''}")
    return r


def build_library():
    objs = []
    objdir = os.path.join(ROOT, "_build")
    os.makedirs(objdir, exist_ok=True)
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        run([hipcc(), *HIPCC_FLAGS, "-I", INCLUDE, "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(objdir, src.replace(".cpp", ".o"))
        run([hipcc(), "-O3", "-std=c++17", "-fPIC", "-pthread", "-fvisibility=hidden", "-Wall", "-ffp-contract=off",
             "-I", INCLUDE, "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


SYNTH_LIB = os.path.join(PKG, "libxd_synth.so")


def build_synth():
    tmp = SYNTH_LIB + ".tmp"
    run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off", "-fvisibility=hidden",
         "-Wall", "-o", tmp, os.path.join(CSRC, "xd_synth.cpp")])
    os.replace(tmp, SYNTH_LIB)
    return SYNTH_LIB


CLI_SOURCES = ["main.cpp", "cli_util.cpp", "volume_io.cpp"]
CLI = os.path.join(BIN_DIR, "xenodon")


def build_cli():
    """The xenodon CLI (host C++), linked against the library in-tree."""
    os.makedirs(BIN_DIR, exist_ok=True)
    cli_dir = os.path.join(PKG, "cli")
    tmp = CLI + ".tmp"
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-Wno-unused-parameter", "-I", INCLUDE,
         *[os.path.join(cli_dir, s) for s in CLI_SOURCES], "-o", tmp, "-L", PKG, "-lxenodon_amd", "-lz",
         f"-Wl,-rpath-link,{rocm_lib}", "-Wl,-rpath,$ORIGIN/.."])
    os.replace(tmp, CLI)
    return CLI


IO_LIB = os.path.join(PKG, "libxd_io.so")


def build_io():
    """The CLI's file formats as a C library (xenodon_amd/io.py)."""
    cli_dir = os.path.join(PKG, "cli")
    tmp = IO_LIB + ".tmp"
    run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", "-Wall", "-I", INCLUDE,
         *[os.path.join(cli_dir, s) for s in ("volume_io.cpp", "cli_util.cpp", "volume_io_capi.cpp")],
         "-o", tmp, "-lz"])
    os.replace(tmp, IO_LIB)
    return IO_LIB


def build_all():
    build_library()
    build_synth()
    build_cli()
    build_io()
    return LIB


if __name__ == "__main__":
    print(build_all())

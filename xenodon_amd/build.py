"""Builds the native parts of xenodon_amd in-tree.

  xenodon_amd/libxenodon_amd.so   HIP kernels + C-ABI (include/xenodon.h), gfx950
  xenodon_amd/bin/xenodon         C++ CLI (render / convert / help / sysinfo)

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

HIP_SOURCES = ["xd_kernels.hip", "xd_runtime.hip", "xd_batch.hip", "xd_serve.hip"]
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


def build_all():
    build_library()
    build_synth()
    return LIB


if __name__ == "__main__":
    print(build_all())

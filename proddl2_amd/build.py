"""Build script for libpdfft.so (gfx950) and the host CLI.

Run ``python -m proddl2_amd.build`` (or ``__graft_entry__.build()``).  The
library is built in-tree under ``proddl2_amd/_lib`` so that it travels with the
source tree to the GPU machine.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_DIR = os.path.join(PKG, "_lib")
LIB = os.path.join(LIB_DIR, "libpdfft.so")
CLI = os.path.join(LIB_DIR, "proddl-dock-fft")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PD_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib() -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = [os.path.join(PKG, "csrc", f) for f in ("pdfft.hip", "pdprop.hip", "frame.cpp")]
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *srcs])
    os.replace(tmp, LIB)
    return LIB


def build_cli() -> str:
    src = os.path.join(PKG, "host", "proddl_dock_fft.cpp")
    if not os.path.exists(src):
        return ""
    tmp = CLI + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", tmp, src,
          "-L", LIB_DIR, "-lpdfft", f"-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, CLI)
    return CLI


REFERENCE = os.environ.get("PD_REFERENCE", "/root/reference")


def build_oracle() -> str:
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "-B"])
    # the reference's own gdiam, a checker of the minimum-volume box: only where
    # the reference sources are present
    if os.path.exists(os.path.join(REFERENCE, "src", "PRODDL", "External", "Gdiam", "gdiam.cpp")):
        _run(["make", "-C", odir, "-B", "ref", f"REF={REFERENCE}"])
    return os.path.join(odir, "_build", "libpdoracle.so")


def build_all() -> None:
    build_lib()
    build_cli()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what:
        build_all()
    else:
        if "lib" in what:
            build_lib()
        if "cli" in what:
            build_cli()
        if "oracle" in what:
            build_oracle()

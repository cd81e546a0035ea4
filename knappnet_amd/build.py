"""Builds the native libraries in-tree.

- ``knappnet_amd/libnetbc.so``: the HIP device library (gfx950) with the C ABI
  declared in ``include/netbc.h``.
- ``knappnet_amd/libnetgen.so``: the seeded workload generators (host C).

Run ``python -m knappnet_amd.build`` (``__graft_entry__.build()`` does the
same and also builds the CPU oracle used by the tests).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HIP_SOURCES = ["graph.hip", "batch.hip", "small.hip", "capi.hip"]
ARCH = os.environ.get("NETBC_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # FP64 results must follow the reference's operation order: no fused
    # multiply-add, no fast-math.
    "-ffp-contract=off",
    "-fvisibility=hidden",
    # No packed FP32 instructions in device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or Path(c).exists()):
            return c
    raise RuntimeError("hipcc not found")


def build_device_library(out: Path | None = None, defines: tuple = ()) -> Path:
    out = out or PKG / "libnetbc.so"
    tmp = out.with_suffix(".so.tmp")
    srcs = [str(CSRC / s) for s in HIP_SOURCES]
    _run([hipcc(), *HIPCC_FLAGS, *[f"-D{d}" for d in defines], "-I", str(ROOT / "include"),
          *srcs, "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_generators() -> Path:
    out = PKG / "libnetgen.so"
    tmp = out.with_suffix(".so.tmp")
    _run(["gcc", "-O3", "-std=c11", "-fPIC", "-shared", "-pthread", "-Wall",
          str(CSRC / "generate.c"), "-o", str(tmp), "-lm"])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_generators()
    build_device_library()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "gen" in what:
        build_generators()
    if "hip" in what:
        build_device_library()
    if "all" in what:
        build_all()

"""Builds the HIP engine (gfx950) into paracosm_amd/libcsm_mi355x.so and the
drop-in CLI into paracosm_amd/csm.  Run as `python -m paracosm_amd.build`."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libcsm_mi355x.so")
CLI = os.path.join(HERE, "csm")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CSM_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    # packed FP32 instructions disabled on purpose for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib(out: str = LIB, extra=()) -> str:
    src = os.path.join(HERE, "csrc", "csm_engine.hip")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, *extra, "-shared", src, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_cli() -> str:
    src = os.path.join(HERE, "csrc", "csm_cli.cpp")
    if not os.path.exists(src):
        return ""
    _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), src,
          "-o", CLI, f"-L{HERE}", "-lcsm_mi355x", "-Wl,-rpath,$ORIGIN"])
    return CLI


def build() -> None:
    build_lib()
    build_cli()


if __name__ == "__main__":
    build()
    sys.exit(0)

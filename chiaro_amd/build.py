"""In-tree build of the native libraries.

  libchiaro_depth.so  HIP kernels + C ABI for gfx950 (hipcc)
  libchiaro_scene.so  synthetic capture generator (gcc)

Every target is built without floating-point contraction or fast-math: the
reference (Rust) never fuses a*b+c, and bit-exact parity depends on it.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
CSRC = PACKAGE / "csrc"
LIB = PACKAGE / "lib"
ROOT = PACKAGE.parent

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    # packed FP32 (v_pk_*_f32) stays off on these boxes
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result",
]
C_FLAGS = ["-O2", "-std=gnu11", "-fPIC", "-march=x86-64", "-ffp-contract=off", "-fno-fast-math"]


def _run(cmd, cwd=None):
    proc = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {' '.join(map(str, cmd))}")
    return proc


def build_depth() -> Path:
    LIB.mkdir(exist_ok=True)
    objs = []
    tmp = LIB / "_obj"
    tmp.mkdir(exist_ok=True)
    for src in ("depth_kernels.hip", "depth_zncc.hip", "proposal.hip", "scene_render.hip",
                "align_kernels.hip", "synth_kernels.hip", "depth_api.hip"):
        obj = tmp / (src + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    out = LIB / "libchiaro_depth.so"
    part = LIB / "libchiaro_depth.so.tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-o", str(part), *objs])
    os.replace(part, out)
    shutil.rmtree(tmp, ignore_errors=True)
    return out


def build_scene() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libchiaro_scene.so"
    part = LIB / "libchiaro_scene.so.tmp"
    _run(["gcc", *C_FLAGS, "-shared", "-o", str(part), str(CSRC / "scene.c"), "-lm", "-lpthread"])
    os.replace(part, out)
    return out


def build_all() -> list[Path]:
    return [build_scene(), build_depth()]


if __name__ == "__main__":
    for p in build_all():
        print(p)

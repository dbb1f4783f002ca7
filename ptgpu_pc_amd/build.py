"""Build recipe for the gfx950 library (ptgpu_pc_amd/libptgpu.so) and the CPU oracle.

Every HIP translation unit is compiled for gfx950 only, with FP contraction off and IEEE-rounded
division/sqrt (the path's parity contract depends on every FP32 expression rounding as written),
and with packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libptgpu.so"
ORACLE_DIR = ROOT / "oracle"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]

SOURCES = sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stdout}")


def build_library(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = BUILD / (src.name + ".o")
        objs.append(obj)
        lang = ["-x", "hip"]
        cmds.append([HIPCC, *HIP_FLAGS, *lang, "-I", str(ROOT / "include"), "-c", str(src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ORACLE_DIR), "liboracle.so"])
    return ORACLE_DIR / "liboracle.so"


def build_reference_twin() -> Path | None:
    """Compiles the reference's own host path into oracle/_ref/ when the reference tree is present."""
    ref = Path("/root/reference/include/geomfunc.h")
    if not ref.exists():
        return None
    _run(["make", "-s", "-C", str(ORACLE_DIR), "ref"])
    return ORACLE_DIR / "_ref" / "libref.so"


def build_all() -> None:
    build_library()
    build_oracle()
    build_reference_twin()


if __name__ == "__main__":
    build_all()
    print(f"built {LIB}", file=sys.stderr)

"""Build the in-tree shared libraries.

libmialign.so  — HIP kernels + C ABI for gfx950 (hipcc)
libmasynth.so  — seeded synthetic workload generator (host C, OpenMP)

Both land next to this file so the built objects travel with the tree.
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
PKG_ROOT = HERE.parent.parent  # mojo-align_amd/
CSRC = PKG_ROOT / "csrc"
INCLUDE = PKG_ROOT.parent / "include"
BUILD_DIR = PKG_ROOT / "_build"

HIP_SOURCES = ["runtime.hip", "index.hip", "sort.hip", "map.hip", "serve.hip", "bam.hip", "bgzf.hip", "fastq.hip",
               "rescue.hip"]
HIP_DEPS = ["ma_common.hpp", "ma_index.hpp", "ma_sort.hpp", "ma_onesweep.hpp", "map_body.inc"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    f"-I{INCLUDE}",
]

LIB = HERE / "libmialign.so"
SYNTH_LIB = HERE / "libmasynth.so"


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        msg = "\n".join(
            line
            for line in (proc.stdout + proc.stderr).splitlines()
            if "packed-fp32-ops" not in line
        )
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{msg}")


def build_hip(jobs: int = 8) -> Path:
    BUILD_DIR.mkdir(parents=True, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = BUILD_DIR / (Path(src).stem + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(cmds)))) as ex:
        for fut in [ex.submit(_run, c) for c in cmds]:
            fut.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lz",
          "-pthread"])
    os.replace(tmp, LIB)
    return LIB


def build_synth() -> Path:
    tmp = SYNTH_LIB.with_suffix(".so.tmp")
    _run(["gcc", "-O3", "-fopenmp", "-fPIC", "-shared", "-Wall", "-o", str(tmp), str(CSRC / "synth.c")])
    os.replace(tmp, SYNTH_LIB)
    return SYNTH_LIB


def build_all() -> None:
    build_synth()
    build_hip()


if __name__ == "__main__":
    build_all()
    print("built", LIB, SYNTH_LIB, file=sys.stderr)

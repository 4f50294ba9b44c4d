"""Build recipe for libmrs_amd.so (gfx950) and the CPU oracle.

Compiles every HIP translation unit under csrc/ with hipcc for gfx950, links
them into manifolds-rs_amd/libmrs_amd.so, and builds oracle/liboracle.so with
the oracle's own Makefile.  Used by __graft_entry__.build(); also runnable as
``python manifolds-rs_amd/build.py``.
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
OBJ = PKG / "_build"
LIB = PKG / "libmrs_amd.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: the reference's arithmetic is never fused, and the parity
# tests compare against a CPU restatement that is not fused either.
# -packed-fp32-ops off: no v_pk_{add,mul,fma}_f32 anywhere in this library.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-parameter",
]

# Per-file extras.  knn.hip: MFMA results in ordinary VGPRs (the top-k
# epilogue reads every accumulator element, which from AGPRs costs one move
# each), and no IEEE NaN quieting on min/max (scores of finite data are never
# NaN; IEEE mode makes every fmaxf canonicalise its inputs first).
FILE_FLAGS = {
    "knn.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form", "-mno-amdgpu-ieee", "-fno-honor-nans"],
}

SOURCES = sorted(CSRC.glob("*.hip"))


def _compile(src: Path) -> Path:
    obj = OBJ / (src.stem + ".o")
    cmd = [HIPCC, *HIP_FLAGS, *FILE_FLAGS.get(src.name, []), "-I", str(ROOT / "include"), "-c",
           str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    return obj


def build_library(jobs: int | None = None) -> Path:
    OBJ.mkdir(exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    r = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return ROOT / "oracle" / "liboracle.so"


def main() -> None:
    lib = build_library()
    orc = build_oracle()
    print(f"built {lib} and {orc}")


if __name__ == "__main__":
    sys.exit(main())

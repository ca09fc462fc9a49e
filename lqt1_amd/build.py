"""Build lqt1_amd/liblqt.so (HIP kernels for gfx950 + the C ABI) with hipcc.

    python -m lqt1_amd.build          # or: python lqt1_amd/build.py

Every source is compiled for gfx950 only.  Packed FP32 instructions are
disabled for all device code (``-Xclang -target-feature -Xclang
-packed-fp32-ops``); no fast-math flag is used, so IEEE infinities of the
model's exponentials are kept.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblqt.so")
ARCH = os.environ.get("LQT_OFFLOAD_ARCH", "gfx950")

SOURCES = [
    "kernels_cells.hip",
    "kernels_tissue.hip",
    "kernels_strands.hip",
    "kernels_mlp.hip",
    "api_cells.cpp",
    "api_tissue.cpp",
    "api_strands.cpp",
    "api_mlp.cpp",
    "api_vw.cpp",
    "api_surrogate.cpp",
]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value", "-Wno-unused-result",
          "-I" + os.path.join(HERE, "..", "include")]
# Multiply-adds are fused only within one source expression
# (-ffp-contract=on), never across statements: HIP's default lets the
# backend fuse by context, so the same update inlined into two kernels could
# round differently, and every launch shape of a kernel must give the same
# bits.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=on"]


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    deps = [path] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".hpp")]
    deps.append(os.path.join(HERE, "..", "include", "lqt.h"))
    if os.path.exists(obj) and os.path.getmtime(obj) >= max(os.path.getmtime(d) for d in deps):
        return obj
    lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [HIPCC, *COMMON, *DEVICE, *lang, "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        # hipcc's host pass does not know the packed-fp32 feature; that note is harmless
        lines = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if lines:
            print("\n".join(lines), file=sys.stderr)
    return obj


def build(force: bool = False, jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    if force:
        for f in os.listdir(BUILD):
            os.remove(os.path.join(BUILD, f))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_drivers()
    return LIB


DRIVERS = {"lqt_single_cell": "single_cell.cpp", "lqt_oned": "one_d.cpp", "lqt_oned_vw": "one_d_vw.cpp",
           "lqt_tissue2d": "tissue2d.cpp", "lqt_tissue3d": "tissue3d.cpp"}


def build_drivers() -> None:
    """The C++ programs mirroring SingleCell.cc, OneD.cc, OneD_VW.cc, new_11.cu
    and GPU3D.cu, linked
    against liblqt.so (found next to them through the rpath)."""
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    for exe, src in DRIVERS.items():
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-o", os.path.join(bindir, exe),
               os.path.join(HERE, "drivers", src), "-L" + HERE, "-llqt", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))

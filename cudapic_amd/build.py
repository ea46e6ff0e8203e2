"""Build recipe of the engine: hipcc for gfx950, in-tree outputs.

    python -m cudapic_amd.build          # libcudapic_amd.so + the cudapic CLI

Objects go to cudapic_amd/_build/ (git-ignored); the shared library
cudapic_amd/libcudapic_amd.so and the executable cudapic_amd/cudapic are
written next to this file so that they travel with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcudapic_amd.so")
CLI = os.path.join(HERE, "cudapic")
ARCH = os.environ.get("PIC_OFFLOAD_ARCH", "gfx950")

SOURCES = ["pic_kernels.hip", "pic_fused.hip", "pic_sim.hip", "pic_step.hip", "pic_ensemble.hip", "pic_adjoint.hip",
           "pic_host.cpp"]
HEADERS = ["pic_device.h", "pic_internal.h"]

# -ffp-contract=off: the only fused multiply-adds are the explicit fmaf()
# calls, which makes the push reproducible on the CPU.  Packed-FP32 VALU ops
# are disabled for all device code.
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build cudapic_amd)")


def _compile(src: str) -> str:
    obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *DEVICE_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [hipcc(), "-O3", "-std=c++17", "-fPIC", "-Wall", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile of {src} failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(7, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    cli_src = os.path.join(CSRC, "cudapic_main.cpp")
    if os.path.exists(cli_src):
        cmd = [hipcc(), "-O2", "-std=c++17", "-o", CLI, cli_src, f"-L{HERE}", "-lcudapic_amd",
               f"-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"CLI build failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:  # pragma: no cover
        print(e, file=sys.stderr)
        sys.exit(1)

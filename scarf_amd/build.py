"""Builds libscarfhip.so (the HIP kernels + C ABI) in-tree for gfx950.

Run ``python -m scarf_amd.build`` (or ``__graft_entry__.build()``).  Each ``.hip``
source under ``scarf_amd/csrc`` is compiled to an object with hipcc, in parallel, and
the objects are linked into ``scarf_amd/libscarfhip.so``.  Packed-FP32 instructions
are disabled for all device code (``-packed-fp32-ops``).
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libscarfhip.so")
ARCH = os.environ.get("SCARF_AMD_ARCH", "gfx950")


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: a ROCm toolchain is required to build libscarfhip")


COMMON_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-I", INCLUDE,
    "-I", CSRC,
]


# per-source extra flags: the kNN screen keeps MFMA accumulators in VGPRs (its filter
# reads them every tile) and assumes no NaNs (INF is its only sentinel); the FP64 Gram
# keeps them in VGPRs too (128 accumulator AGPRs on top of its VGPRs would leave one wave
# per SIMD; in VGPRs it runs two)
EXTRA_FLAGS = {
    "sh_knn.hip": ["-mllvm", "--amdgpu-mfma-vgpr-form", "-fno-honor-nans"],
    "sh_gram.hip": ["-mllvm", "--amdgpu-mfma-vgpr-form"],
}


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.basename(src).replace(".hip", ".o"))
    cmd = [_hipcc(), *COMMON_FLAGS, *EXTRA_FLAGS.get(os.path.basename(src), []), "-c", src,
           "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)

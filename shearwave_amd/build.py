"""Build the gfx950 HIP library ``shearwave_amd/libshearwave_hip.so`` in-tree.

Each ``csrc/*.hip`` file is compiled by hipcc for gfx950 only, in parallel, and
linked into one shared library exporting the C ABI of ``include/shearwave.h``.

    python -m shearwave_amd.build
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libshearwave_hip.so")

SOURCES = ("abi.hip", "stencils.hip", "project.hip", "fdtd.hip", "strain.hip", "batch.hip", "adjoint.hip",
           "radforce.hip", "acoustic.hip", "static.hip")

ARCH = "gfx950"
CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # packed FP32 (v_pk_*_f32) instructions stay off: hipcc still vectorises scalar code
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-I",
    INCLUDE,
    "-I",
    CSRC,
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{p.stdout}")
    return p.stdout


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every HIP source for gfx950 and link the shared library; returns its path."""
    cc = hipcc()
    os.makedirs(BUILD, exist_ok=True)
    sources = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(BUILD, s.replace(".hip", ".o")) for s in sources]

    def compile_one(i):
        cmd = [cc, *CFLAGS, "-c", os.path.join(CSRC, sources[i]), "-o", objs[i]]
        out = _run(cmd)
        if verbose:
            print(" ".join(cmd))
            if out.strip():
                print(out)

    n = jobs or min(len(sources), max(1, (os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=n) as ex:
        list(ex.map(compile_one, range(len(sources))))
    tmp = LIB + ".tmp"
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)

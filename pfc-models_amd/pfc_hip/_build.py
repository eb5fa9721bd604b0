"""Build libpfc_hip.so in-tree with hipcc for gfx950.

Run as `python -m pfc_hip._build` (from `pfc-models_amd/`) or through
`__graft_entry__.build()`.  The library lands next to this file, where
`pfc_hip._lib` loads it.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")
OUT = os.path.join(HERE, "libpfc_hip.so")
SOURCES = ["pfc_capi.hip", "pfc_inst_row64.hip", "pfc_inst_row32.hip", "pfc_inst_col64.hip", "pfc_inst_col32.hip",
           "pfc_inst_row64o.hip", "pfc_inst_row32o.hip", "pfc_inst_col64o.hip", "pfc_inst_col32o.hip",
           "pfc_inst_group64.hip", "pfc_inst_group32.hip"]
ARCH = os.environ.get("PFC_HIP_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    # keep hipcc from emitting packed FP32 VALU ops (see the design notes)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode}): {' '.join(cmd)}")


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile each translation unit to an object in parallel, then link."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(os.path.dirname(HERE), "_build")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    compile_flags = [f for f in FLAGS if f != "-shared"]
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([cc, *compile_flags, "-c", os.path.join(CSRC, src), "-o", obj])
    jobs = jobs or min(len(cmds), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = OUT + ".tmp"
    _run([cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs], verbose)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))

"""Builds the HIP library libgitm_vert.so in-tree for gfx950.

    python -m gitm_amd.build          # or __graft_entry__.build()
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libgitm_vert.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GITM_VERT_ARCH", "gfx950")

# FP64 parity with the reference needs IEEE semantics without contraction;
# packed FP32 instructions are disabled throughout (not used by this code).
HIPFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]
# translation units, compiled in parallel: the single-grid step and C ABI,
# the ensemble batches (host side, and the step kernel once per workgroup
# size of 1..4 wavefronts), the adjoint, the sources step, the horizontal
# step; (source, extra flags, object tag)
SOURCES = (("gitm_vert.hip", (), ""), ("gitm_vert_batch.hip", (), ""),
           *(("gitm_vert_ensemble.hip", (f"-DGV_ENS_WAVES={w}",), f"_w{w}")
             for w in (1, 2, 3, 4)),
           ("gitm_vert_adjoint.hip", (), ""), ("gitm_vert_sources.hip", (), ""),
           ("gitm_horiz.hip", (), ""))


def build(verbose: bool = False, extra: list[str] | None = None, out: str | None = None) -> str:
    out = out or LIB
    objs, procs = [], []
    for src, flags, tag in SOURCES:
        obj = out + "." + os.path.splitext(src)[0] + tag + ".o"
        cmd = ([HIPCC] + HIPFLAGS + list(flags) + (extra or []) +
               ["-c", os.path.join(CSRC, src), "-o", obj])
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        objs.append(obj)
    failed = [src for (src, _, _), pr in zip(SOURCES, procs) if pr.wait() != 0]
    if failed:
        raise RuntimeError(f"hipcc failed on {failed}")
    link = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}"] + objs + ["-o", out + ".tmp"]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    for o in objs:
        os.remove(o)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])

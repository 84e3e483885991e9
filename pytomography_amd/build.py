"""Builds libspect_hip.so in-tree with hipcc for gfx950.

    python -m pytomography_amd.build        (or __graft_entry__.build())

Each translation unit under csrc/ is compiled in parallel, then linked into a
plain C-ABI shared object (no torch, no Python headers) that
``pytomography_amd._lib`` loads through ctypes.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["forward.hip", "backward.hip", "misc.hip", "scatter.hip", "shared.hip", "volume.hip", "capi.hip"]
OUT = os.path.join(HERE, "libspect_hip.so")
ARCH = os.environ.get("SPECT_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
    f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build libspect_hip.so")
    return path


def build(verbose: bool = True) -> str:
    cc = hipcc()
    with tempfile.TemporaryDirectory(prefix="spect_build_") as tmp:
        procs, objs = [], []
        for src in SOURCES:
            obj = os.path.join(tmp, src.replace(".hip", ".o"))
            cmd = [cc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((src, subprocess.Popen(cmd)))
            objs.append(obj)
        failed = [src for src, p in procs if p.wait() != 0]
        if failed:
            raise RuntimeError(f"hipcc failed for {failed}")
        tmp_out = OUT + ".tmp"
        link = [cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp_out]
        if verbose:
            print(" ".join(link), flush=True)
        subprocess.run(link, check=True)
        os.replace(tmp_out, OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)

"""Build libfxg.so (the C-ABI HIP library) in-tree for gfx950.

    python -m difx_amd.build          # or: from difx_amd.build import build; build()

Every HIP source is compiled with hipcc for gfx950 only, with packed-FP32 instruction
selection disabled, and linked into difx_amd/libfxg.so next to this file.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
INCLUDE = HERE.parent / "include"
LIB = HERE / "libfxg.so"
SOURCES = ["fxg_fengine.hip", "fxg_fpersist.hip", "fxg_fbig.hip", "fxg_xengine.hip", "fxg_api.cpp",
           "fxg_server.cpp", "fxg_vdifmux.hip", "fxg_control.cpp",
           "fxg_visnorm.hip", "fxg_kurt.hip", "fxg_pcal.hip"]
ARCH = "gfx950"


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> Path:
    cc = hipcc()
    objdir = HERE / "_build"
    objdir.mkdir(exist_ok=True)
    flags = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I", str(INCLUDE), "-I", str(CSRC), "-Wno-unused-result"]
    def compile_one(src):
        obj = objdir / (src + ".o")
        cmd = [cc] + flags + (["-x", "hip"] if src.endswith(".cpp") else []) + ["-c", str(CSRC / src), "-o", str(obj)]
        return src, obj, subprocess.run(cmd, capture_output=True, text=True)

    objs = []
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as pool:
        for src, obj, r in pool.map(compile_one, SOURCES):
            if r.returncode != 0:
                sys.stderr.write(r.stdout + r.stderr)
                raise RuntimeError(f"hipcc failed on {src}")
            if verbose:
                sys.stderr.write(r.stderr)
            objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp)] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libfxg.so failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

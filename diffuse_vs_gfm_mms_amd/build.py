"""Builds in-tree: the gfx950 HIP library diffuse_vs_gfm_mms_amd/_lib/libdimhip.so
and the C++17 case drivers _lib/dim_main_{1,2,3}d that run the reference's DIM
case files on it.

    python -m diffuse_vs_gfm_mms_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
# the C ABI and one translation unit per arithmetic policy, compiled in parallel
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("dimhip.hip", "ops_fp64.hip", "ops_fp64na.hip", "ops_fp32.hip", "ops_fp32i.hip")]
OUT_DIR = os.path.join(HERE, "_lib")
OUT = os.path.join(OUT_DIR, "libdimhip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # the reference's arithmetic is reproduced operation by operation: no FMA contraction
    "-ffp-contract=off",
    "-fno-fast-math",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value",
    "-Wno-unused-result",
]


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    noise = "is not a recognized feature for this target"
    with tempfile.TemporaryDirectory(prefix="dimhip_build_") as tmpdir:
        procs = []
        for src in SOURCES:
            obj = os.path.join(tmpdir, os.path.basename(src) + ".o")
            cmd = [hipcc(), *HIPCC_FLAGS, "-c", "-o", obj, src]
            procs.append((src, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        failed = False
        for src, _, proc in procs:
            out, _ = proc.communicate()
            if proc.returncode != 0:
                sys.stderr.write(out)
                failed = True
            elif verbose:
                sys.stderr.write("".join(l + "\n" for l in out.splitlines() if noise not in l))
        if failed:
            raise RuntimeError("hipcc failed building libdimhip.so")
        tmp = OUT + ".tmp"
        cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *[o for _, o, _ in procs]]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("hipcc failed linking libdimhip.so")
    os.replace(tmp, OUT)
    return OUT


HOST_DIR = os.path.join(HERE, "host")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def driver_path(dim):
    return os.path.join(OUT_DIR, f"dim_main_{dim}d")


def build_drivers():
    """g++ -std=c++17 -O2 (the reference's own host flags), linked to libdimhip.so."""
    if not os.path.exists(OUT):
        raise RuntimeError("build the HIP library first")
    outs = []
    for dim in (1, 2, 3):
        out = driver_path(dim)
        cmd = ["g++", "-std=c++17", "-O2", f"-DDIM_={dim}", "-I", INCLUDE, "-I", HOST_DIR,
               os.path.join(HOST_DIR, "dim_main.cpp"), "-o", out + ".tmp", "-L", OUT_DIR, "-ldimhip",
               "-Wl,-rpath,$ORIGIN", "-pthread"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"g++ failed building {out}")
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(*build_drivers())

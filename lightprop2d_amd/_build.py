"""Builds liblp2d.so (the HIP kernels + C ABI) in-tree for gfx950 with hipcc.

    python -m lightprop2d_amd._build          # or lightprop2d_amd._build.build()

Each translation unit under csrc/ compiles in its own hipcc process (in parallel),
then one hipcc link produces lightprop2d_amd/liblp2d.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
ROOT = os.path.dirname(PKG)
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "liblp2d.so")

SOURCES = ["lp2d_f32_b.hip", "lp2d_f32_c.hip", "lp2d_f64_b.hip", "lp2d_f64_c.hip",
           "lp2d_f32_a.hip", "lp2d_f64_a.hip", "lp2d_api.hip", "lp2d_profiles.hip",
           "lp2d_resample.hip", "lp2d_modes.hip"]

ARCH = os.environ.get("LP2D_ARCH", "gfx950")
# Packed FP32 instructions are disabled for all device code (the compiler would
# otherwise form v_pk_* from scalar float code).
CFLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + CSRC, "-I" + os.path.join(ROOT, "include")]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build liblp2d")


def _compile(src, extra):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), *CFLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{out}")
    return obj, out


def build(verbose: bool = False, extra_flags=()) -> str:
    """Compile every HIP source and link liblp2d.so.  Always rebuilds."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = min(len(SOURCES), max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda s: _compile(s, list(extra_flags)), SOURCES))
    if verbose:
        for obj, out in results:
            if out.strip():
                print(out)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

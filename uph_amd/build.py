"""Build libuph_gpu.so (the HIP kernels and C ABI) in-tree for gfx950."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRCS = [os.path.join(HERE, "csrc", f) for f in
        ("uph_step.hip", "uph_pair.hip", "uph_search.hip", "uph_diag.hip", "uph_graph.hip", "uph_compat.hip",
         "uph_gnn.hip", "uph_surrogate.hip")]
OUT = os.path.join(HERE, "libuph_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-ffp-contract=off",
         # no packed FP32 instructions (v_pk_*_f32) in the generated code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-result"]


DRIVER_SRC = os.path.join(HERE, "host", "uph_run.c")
DRIVER = os.path.join(HERE, "uph_run")


def build(verbose: bool = False) -> str:
    """Compile the translation units side by side (objects under uph_amd/_build/), then link."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", os.path.join(HERE, "..", "include")]
    cflags = [f for f in FLAGS if f != "-shared"]

    def compile_one(src: str) -> str:
        obj = os.path.join(objdir, os.path.splitext(os.path.basename(src))[0] + ".o")
        cmd = [HIPCC, *cflags, *inc, "-c", "-o", obj, src]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        return obj

    jobs = max(1, min(len(SRCS), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_one, SRCS))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    # the C host driver links the C ABI only
    dcmd = ["gcc", "-std=gnu11", "-O2", "-Wall", "-o", DRIVER + ".tmp", DRIVER_SRC,
            "-L", HERE, "-luph_gpu", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib", "-lm"]
    if verbose:
        print(" ".join(dcmd), flush=True)
    subprocess.run(dcmd, check=True)
    os.replace(DRIVER + ".tmp", DRIVER)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)

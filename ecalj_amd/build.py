"""Build libecalj_gpu.so (the HIP library behind include/ecalj_gpu.h) for gfx950.

    python -m ecalj_amd.build          # compile in place
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libecalj_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ECALJ_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed-FP32 instructions (v_pk_{add,mul,fma}_f32) in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    return sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))


def _compile(src: str, obj: str, verbose: bool):
    cmd = [HIPCC, *[f for f in FLAGS if f != "-shared"], "-c", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(HERE, "csrc"), src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln]
    return r.returncode, errs


def build(verbose: bool = False) -> str:
    """Compiles every source (one hipcc per file, in parallel) and links the library."""
    from concurrent.futures import ThreadPoolExecutor

    odir = os.path.join(HERE, "_build")
    os.makedirs(odir, exist_ok=True)
    srcs = sources()
    objs = [os.path.join(odir, os.path.basename(s)[:-4] + ".o") for s in srcs]
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        res = list(ex.map(lambda so: _compile(so[0], so[1], verbose), zip(srcs, objs)))
    failed = False
    for src, (rc, errs) in zip(srcs, res):
        if rc != 0:
            failed = True
            sys.stderr.write(f"--- {os.path.basename(src)}\n" + "\n".join(errs) + "\n")
        elif verbose and errs:
            sys.stderr.write("\n".join(errs) + "\n")
    if failed:
        raise RuntimeError("hipcc failed")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", LIB + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError(f"link failed ({r.returncode})")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))

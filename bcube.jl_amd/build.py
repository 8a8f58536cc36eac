"""Builds libbcubehip.so (the HIP/CDNA4 assembly library) in-tree for gfx950.

    python bcube.jl_amd/build.py          # incremental
    python bcube.jl_amd/build.py --force  # rebuild everything
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "bcubehip", "libbcubehip.so")
BUILD = os.path.join(HERE, "_build")
SOURCES = ["bh_api.hip", "bh_kernels.hip", "bh_mfma.hip", "bh_gp.hip", "bh_dg.hip", "bh_build.hip", "bh_sf.hip",
           "bh_dgb.hip", "bh_euler.hip", "bh_source.hip", "bh_orders.hip", "bh_host.cpp"]
HEADERS = ["bh_internal.h", "bh_kernels.h", "bh_device.h", "bh_dg.h", os.path.join("..", "..", "include", "bcubehip.h")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BH_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         # packed FP32 instructions stay off (the compiler would form them from float code)
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def _newer(src, dst):
    return not os.path.exists(dst) or os.path.getmtime(src) > os.path.getmtime(dst)


def build(force=False, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hdr_time = max(os.path.getmtime(os.path.join(CSRC, h)) for h in HEADERS)
    jobs = []
    objs = []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(BUILD, s + ".o")
        objs.append(obj)
        stale = force or _newer(src, obj) or (os.path.exists(obj) and hdr_time > os.path.getmtime(obj))
        if stale:
            lang = [] if s.endswith(".hip") else ["-x", "hip"]
            jobs.append([HIPCC] + FLAGS + lang + ["-c", src, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        for cmd, r in zip(jobs, ex.map(lambda c: subprocess.run(c, capture_output=True, text=True), jobs)):
            err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
            if r.returncode != 0:
                raise RuntimeError(f"build failed: {' '.join(cmd)}\n{err}")
            if verbose and err.strip():
                print(err)
    if force or jobs or not os.path.exists(OUT):
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", OUT] + objs + ["-lpthread"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stderr}")
    return OUT


if __name__ == "__main__":
    print(build(force="--force" in sys.argv, verbose=True))

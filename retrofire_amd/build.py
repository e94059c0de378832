"""Builds the HIP library (retrofire_amd/librfamd.so) for gfx950 with hipcc.

Float semantics: -ffp-contract=off keeps every a*b+c as two rounded
operations (the reference never fuses), and the default correctly rounded
single-precision division and square root are kept (no fast-math).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librfamd.so")
# (object name, source, extra flags). rf_kernels.hip builds twice: the pipeline
# kernels, and k_raster alone under the max-ilp scheduler (shorter dependent
# chains in the rasterizer; the same scheduler makes k_setup slower).
SOURCES = [
    ("rf_kernels", "rf_kernels.hip", []),
    ("rf_raster", "rf_kernels.hip", ["-DRF_RASTER_TU", "-mllvm", "-amdgpu-sched-strategy=max-ilp"]),
    ("rf_sort", "rf_sort.hip", []),
    ("rf_api", "rf_api.cpp", []),
    ("rf_batch", "rf_batch.cpp", []),
    ("rf_serve", "rf_serve.cpp", []),
    ("rf_dsort", "rf_dsort.cpp", []),
    ("rf_noalloc", "rf_noalloc.cpp", []),
]
HEADERS = ["rf_types.hpp", "rf_math.hpp", "rf_internal.hpp"]
ARCH = os.environ.get("RF_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I", os.path.join(ROOT, "include"),
]


def build(verbose=False, jobs=4, defines=(), out=None, build_dir=None):
    bdir = build_dir or os.path.join(CSRC, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = [os.path.join(bdir, name + ".o") for name, _, _ in SOURCES]
    procs = []
    for (_, src, extra), obj in zip(SOURCES, objs):
        cmd = [HIPCC] + FLAGS + list(extra) + [f"-D{d}" for d in defines] + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = False
    for src, p in procs:
        log, _ = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"--- {src}\n{log.decode(errors='replace')}\n")
        elif verbose and log:
            sys.stderr.write(log.decode(errors="replace"))
    if failed:
        raise RuntimeError("hipcc failed")
    lib = out or LIB
    tmp = lib + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    subprocess.check_call(cmd)
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

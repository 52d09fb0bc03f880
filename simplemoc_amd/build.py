"""Build recipe of the simplemoc_amd native library and host driver (gfx950).

Produces, in-tree:
  simplemoc_amd/_lib/libsimplemoc_hip.so   HIP kernels + C ABI (include/simplemoc.h,
                                           include/simplemoc_dropin.h)
  simplemoc_amd/_lib/simplemoc             C host driver (reference CLI and input file)

Run as `python -m simplemoc_amd.build` or through __graft_entry__.build().
hipcc cross-compiles for gfx950 without a GPU present.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
OUT = os.path.join(PKG, "_lib")
OBJ = os.path.join(OUT, "obj")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Device flags: IEEE division (HIP default), no FMA contraction (the trace and
# table-index arithmetic must round exactly like the reference), hardware FP
# atomics, and no packed-FP32 instructions.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", f"-I{INCLUDE}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", f"-I{INCLUDE}"]

LIB_NAME = "libsimplemoc_hip.so"
DRIVER_NAME = "simplemoc"

HIP_SOURCES = ["moc_device.hip", "moc_batch.hip"]
CXX_SOURCES = ["moc_host.cpp", "moc_dropin.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def build(verbose: bool = True) -> str:
    """Compile every source (always) and return the library path."""
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        jobs.append(([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", o], o))
    for src in CXX_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        o = os.path.join(OBJ, src + ".o")
        jobs.append((["g++", *CXX_FLAGS, "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-c",
                      path, "-o", o], o))
    # the translation units compile independently
    with ThreadPoolExecutor(max_workers=min(4, len(jobs))) as ex:
        list(ex.map(lambda j: _run(j[0], verbose), jobs))
    objs = [o for _, o in jobs]
    lib = os.path.join(OUT, LIB_NAME)
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          "-Wl,--no-undefined", "-lm"], verbose)
    os.replace(tmp, lib)
    # C host driver (links the library by rpath)
    drv_src = os.path.join(HOST, "simplemoc_main.c")
    if os.path.exists(drv_src):
        drv = os.path.join(OUT, DRIVER_NAME)
        _run(["gcc", "-std=gnu99", "-O2", "-Wall", f"-I{INCLUDE}", drv_src, "-o", drv,
              f"-L{OUT}", "-lsimplemoc_hip", "-Wl,-rpath,$ORIGIN", "-lm"], verbose)
    shutil.rmtree(OBJ, ignore_errors=True)
    return lib


if __name__ == "__main__":
    print(build(verbose=True))

"""Builds libyahmc.so (the HIP/CDNA4 hot path + C ABI) in-tree for gfx950.

Every HIP translation unit is compiled with hipcc for gfx950 only, with
packed-FP32 instructions disabled, then linked into one shared library
next to this file (yahmc_amd/lib/libyahmc.so).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "lib", "obj")
LIBNAME = "libyahmc.so"
ARCH = os.environ.get("YAHMC_ARCH", "gfx950")

SOURCES = ["yh_dslash_d8.hip", "yh_dslash_d3.hip", "yh_dslash_d6.hip", "yh_gauge.hip", "yh_dslash_d2.hip",
           "yh_dslash_d4.hip", "yh_dslash.hip", "yh_blas.hip", "yh_fields.hip", "yh_capi.cpp", "yh_solvers.cpp",
           "yh_force.hip", "yh_forces_api.cpp", "yh_mshift.hip", "yh_rational.cpp", "yh_md.hip", "yh_net.hip", "yh_io.cpp", "yh_meson.hip", "yh_bicgstab.hip", "yh_noalloc.cpp"]

CFLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unused-result",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + CFLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    out = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, out


def build(verbose=False, jobs=None):
    os.makedirs(OBJDIR, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, SOURCES))
    errs = []
    objs = []
    for src, obj, out in results:
        objs.append(obj)
        if out.returncode != 0:
            errs.append(f"--- {src}\n{out.stdout}\n{out.stderr}")
        elif verbose and out.stderr.strip():
            print(f"--- {src}\n{out.stderr}", file=sys.stderr)
    if errs:
        raise RuntimeError("HIP compile failed:\n" + "\n".join(errs))
    lib = os.path.join(LIBDIR, LIBNAME)
    tmp = lib + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("link failed:\n" + out.stdout + out.stderr)
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    print(build(verbose=True))

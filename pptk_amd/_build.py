"""Builds the native parts of pptk_amd in-tree.

* ``pptk_amd/lib/libpptk_amd.so``: the HIP kernels and the C ABI declared in
  ``include/pptk_amd.h``, compiled for gfx950 with hipcc.
* ``pptk_amd/kdtree/kdtree*.so`` and
  ``pptk_amd/processing/estimate_normals/estimate_normals*.so``: the CPython
  modules (same module names and signatures as pptk's), compiled with g++ and
  linked against ``libpptk_amd.so`` through an ``$ORIGIN`` rpath.

Usage: ``python -m pptk_amd._build [-j N]``.
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(ROOT, "_build", "obj")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libpptk_amd.so")

HIP_SOURCES = ["runtime.hip", "sort.hip", "build.hip", "query.hip", "normals.hip", "batch.hip", "persist.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # Bit-exact distances: no a*b+c contraction into FMA, no fast-math.
    "-ffp-contract=off",
    "-fno-fast-math",
    # Keep the compiler from emitting packed FP32 VALU ops.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d):\n%s\n%s" % (r.returncode, " ".join(cmd), r.stdout))
    return r.stdout


def _hip_obj(src):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", out])
    return out


def build_lib(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_hip_obj, HIP_SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def _ext_suffix():
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def build_pyext(name, src, outdir):
    import numpy

    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, name + _ext_suffix())
    rel = os.path.relpath(LIBDIR, outdir)
    cmd = [
        os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall",
        "-I" + sysconfig.get_paths()["include"],
        "-I" + numpy.get_include(),
        "-I" + os.path.join(ROOT, "include"),
        os.path.join(CSRC, "pyext", src),
        "-L" + LIBDIR, "-lpptk_amd",
        "-Wl,-rpath,$ORIGIN/" + rel,
        "-o", out + ".tmp",
    ]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build(jobs=8):
    build_lib(jobs)
    outs = [LIB]
    outs.append(build_pyext("kdtree", "kdtree_module.cpp", os.path.join(PKG, "kdtree")))
    outs.append(build_pyext("estimate_normals", "estimate_normals_module.cpp",
                            os.path.join(PKG, "processing", "estimate_normals")))
    return outs


if __name__ == "__main__":
    j = 8
    if "-j" in sys.argv:
        j = int(sys.argv[sys.argv.index("-j") + 1])
    for p in build(j):
        print(p)

"""Build the in-tree HIP library ``shearcuda_amd/libshx.so`` for gfx950.

Every source under ``csrc/`` is compiled with hipcc (device code for gfx950
only, packed-FP32 instructions disabled) into an object under ``_build/`` and
linked into one shared library next to this file.  ``python -m
shearcuda_amd.build`` (or ``__graft_entry__.build()``) runs it.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libshx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["api.hip", "api3d.hip", "pyramid.hip", "spectral.hip", "spectral3d.hip", "rdwt.hip", "select.hip", "dict_dev.hip", "meyer_dev.hip", "maxflat.cpp",
           "dict_host.cpp"]
GEN = os.path.join(BUILD, "gen")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-I" + CSRC, "-I" + GEN, "-I" + os.path.join(ROOT, "include")]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src.replace(".", "_") + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip"] + DEVICE + COMMON + ["-c", path, "-o", obj]
    else:
        cmd = [HIPCC] + COMMON + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def _generate():
    """Compile and run gen_factors (host) to write maxflat_factors.h."""
    os.makedirs(GEN, exist_ok=True)
    exe = os.path.join(BUILD, "gen_factors")
    cmd = [HIPCC, "-O2", "-std=c++17", "-I" + CSRC, os.path.join(CSRC, "gen_factors.cpp"),
           os.path.join(CSRC, "maxflat.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("generator build failed: %s\n%s" % (r.stdout, r.stderr))
    r = subprocess.run([exe, os.path.join(GEN, "maxflat_factors.h")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("maxflat factor generation failed: %s %s" % (r.stdout, r.stderr))


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    _generate()
    jobs = min(len(SOURCES), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    _build_demo()
    if verbose:
        print("built", LIB, "and", DEMO)
    return LIB


DEMO = os.path.join(HERE, "demo", "demo_denoise")


def _build_demo():
    """The C++ denoise demo over include/shearcuda_amd.hpp, linked to libshx.so
    (found next to the package at run time through its rpath)."""
    src = os.path.join(HERE, "demo", "demo_denoise.cpp")
    tmp = DEMO + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), src, "-o", tmp, "-L" + HERE, "-lshx",
           "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("demo build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, DEMO)


if __name__ == "__main__":
    build()
    sys.exit(0)

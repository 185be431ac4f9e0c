"""Build libo3dbiot.so in-tree: every HIP source for gfx950 with hipcc, in parallel.

    python -m omega3d_amd.build            # compile + link
    python -m omega3d_amd.build --clean

Objects go to omega3d_amd/_build/, the library to omega3d_amd/libo3dbiot.so.
"""
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libo3dbiot.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 (v_pk_fma_f32 and friends) is switched off: only scalar FP32 VALU ops.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in (p.stdout + p.stderr).splitlines()
                     if "packed-fp32-ops' is not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{errs}")
    return obj, errs


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    if not srcs:
        raise RuntimeError("no HIP sources found")
    jobs = jobs or min(len(srcs), max(1, min(16, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, srcs))
    for obj, errs in results:
        if errs.strip() and verbose:
            print(errs)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, LIB)
    build_cpp_driver()
    return LIB


DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "drive_influence.cpp")
DRIVER = os.path.join(BUILD, "drive_influence")


def build_cpp_driver():
    """The C++ caller of include/omega3d/Influence.hpp used by the GPU tests: plain
    g++ against the C ABI, linked to libo3dbiot.so by rpath."""
    if not os.path.exists(DRIVER_SRC):
        return None
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"),
           DRIVER_SRC, "-o", DRIVER, "-L" + HERE, "-lo3dbiot", "-Wl,-rpath,$ORIGIN/.."]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("C++ driver build failed:\n" + p.stdout + p.stderr)
    return DRIVER


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)
    if os.path.exists(LIB):
        os.remove(LIB)


if __name__ == "__main__":
    if "--clean" in sys.argv:
        clean()
    else:
        print(build(verbose="-v" in sys.argv))

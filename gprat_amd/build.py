"""Build recipe for the gprat_amd backend (in-tree, no CMake).

1. hipcc compiles the gfx950 kernels and the host engine/C ABI
   (gprat_amd/csrc) into gprat_amd/libgprat_amd.so.
2. g++ compiles the host C++ facade and its pybind11 module `gprat`
   (gprat_amd/host) into gprat_amd/gprat.<EXT_SUFFIX>, linked against
   libgprat_amd.so through an $ORIGIN rpath.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_build")
ARCH = os.environ.get("GPRAT_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC",
    # packed FP32 instructions are kept out of all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I", CSRC, "-I", INCLUDE,
]
KERNEL_SOURCES = ["k_covariance.hip", "k_gemm.hip", "k_gemm32.hip", "k_triangular.hip", "k_batch.hip", "k_calibrate.hip", "k_refine.hip", "k_noalloc.hip"]
HOST_SOURCES = ["engine.cpp", "batch.cpp", "panel_plan.cpp", "server.cpp", "capi.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out, flush=True)


def build_library(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for src in KERNEL_SOURCES + HOST_SOURCES:
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC] + HIP_FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        jobs.append((cmd, obj))
    workers = min(len(jobs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(workers) as ex:
        list(ex.map(lambda j: _run(j[0], verbose), jobs))
    lib = os.path.join(HERE, "libgprat_amd.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", lib + ".tmp"]
         + [o for _, o in jobs] + ["-Wl,-soname,libgprat_amd.so"], verbose)
    os.replace(lib + ".tmp", lib)
    return lib


def build_module(verbose=False):
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    target = os.path.join(HERE, "gprat" + suffix)
    cmd = [
        "g++", "-std=c++20", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra", "-fvisibility=hidden",
        "-I", HOST, "-I", INCLUDE, "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
        os.path.join(HOST, "gprat.cpp"), os.path.join(HOST, "bindings.cpp"),
        "-L", HERE, "-lgprat_amd", "-Wl,-rpath,$ORIGIN",
        "-o", target + ".tmp",
    ]
    _run(cmd, verbose)
    os.replace(target + ".tmp", target)
    return target


def build(verbose=False):
    lib = build_library(verbose)
    mod = build_module(verbose)
    return lib, mod


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print(p)

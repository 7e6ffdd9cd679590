"""Build recipe for the parity oracle (test infrastructure, not product).

Compiles oracle/gprat_oracle.cpp and its pybind face into
oracle/_build/gprat_oracle*.so with the host g++. Default x86-64 code
generation (no -march, no FMA contraction) matches how the reference's own
CPU path is compiled, so the scalar arithmetic of tile generation and the
loss agrees bit for bit; BLAS arithmetic comes from the library bound at run
time (oneMKL or numpy's OpenBLAS).

The reference's own sources are not compiled: they need HPX, CBLAS/LAPACKE
headers, Boost.JSON and Catch2, none of which exist in this image.
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build")


def build(verbose: bool = False) -> str:
    import pybind11

    os.makedirs(OUT, exist_ok=True)
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    target = os.path.join(OUT, "gprat_oracle" + suffix)
    cmd = [
        "g++", "-std=c++20", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra",
        "-fvisibility=hidden",
        "-I", HERE,
        "-I", pybind11.get_include(),
        "-I", sysconfig.get_paths()["include"],
        os.path.join(HERE, "gprat_oracle.cpp"),
        os.path.join(HERE, "oracle_py.cpp"),
        "-o", target + ".tmp",
        "-ldl", "-lpthread",
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(target + ".tmp", target)
    return target


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

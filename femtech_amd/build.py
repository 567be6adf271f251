"""Builds the engine in-tree: the gfx950 library with hipcc, then the
FemTech-compatible C++ layer and its driver with g++.

femtech_amd/_lib/libfemtech_amd.so  HIP kernels, the C ABI declared in
                                    include/femtech_amd.h, host mesh tools
femtech_amd/_lib/libFemTech.so      FemTech's C++ call surface
                                    (include/FemTech.h) over that C ABI
femtech_amd/_lib/explicit_tension   driver with ex9's call sequence
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libfemtech_amd.so")
COMPAT_PATH = os.path.join(LIB_DIR, "libFemTech.so")
DRIVER_PATH = os.path.join(LIB_DIR, "explicit_tension")
INCLUDE = os.path.join(ROOT, "include")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("ft_engine.hip", "ft_surrogate.hip", "ft_point.hip",
                                                    "ft_mesh.cpp")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are switched off for every HIP translation unit.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


HOST_FLAGS = ["-O2", "-std=c++17", "-Wall", "-Wno-format-security", "-I", INCLUDE]


def _run(cmd, target, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("build failed: %s" % target)


def build(verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = LIB_PATH + ".tmp"
    _run([HIPCC] + FLAGS + SOURCES + ["-o", tmp], LIB_PATH, verbose)
    os.replace(tmp, LIB_PATH)
    # FemTech's C++ surface: the driver defines Time, dt and the solver flags,
    # so those stay undefined in the shared library.
    tmp = COMPAT_PATH + ".tmp"
    _run(["g++"] + HOST_FLAGS + ["-fPIC", "-shared", os.path.join(HERE, "csrc", "ft_femtech.cpp"),
                                 "-L", LIB_DIR, "-lfemtech_amd", "-Wl,-rpath,$ORIGIN", "-o", tmp],
         COMPAT_PATH, verbose)
    os.replace(tmp, COMPAT_PATH)
    tmp = DRIVER_PATH + ".tmp"
    _run(["g++"] + HOST_FLAGS + [os.path.join(HERE, "drivers", "explicit_tension.cpp"), "-L", LIB_DIR,
                                 "-lFemTech", "-lfemtech_amd", "-Wl,-rpath,$ORIGIN", "-o", tmp],
         DRIVER_PATH, verbose)
    os.replace(tmp, DRIVER_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))

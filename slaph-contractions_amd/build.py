"""Build the HIP library of the contraction path in-tree for gfx950.

    python slaph-contractions_amd/build.py        # -> slaph-contractions_amd/libslaph_amd.so

Plain hipcc invocations (no cmake): every source compiles to an object in
_build/, then one shared library is linked next to this file.  The package
loader (slaph/__init__.py) loads that library and fails loudly without it.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libslaph_amd.so")
ARCH = os.environ.get("SLAPH_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
HDF5 = os.environ.get("SLAPH_HDF5", "/opt/conda")  # HDF5 1.10 (C API) of this image
H5LIB = os.path.join(HERE, "libslaph_h5.so")
CONTRACT = os.path.join(HERE, "contract")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-I" + SRC]
# Packed FP32 instructions stay off in device code: some of their forms have
# given wrong results on this hardware without any error.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
SOURCES = ["kernels.hip", "gauge.hip", "host.cpp", "lime.cpp", "engine.cpp", "contract.cpp", "capi.cpp"]


def compile_one(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + DEVICE + ["-c", os.path.join(SRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return out


def run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("%s failed: %s\n%s\n%s" % (what, " ".join(cmd), r.stdout, r.stderr))


def build_h5(verbose=True):
    """libslaph_h5.so (HDF5 correlator files, include/slaph_h5.h): a library of
    its own so that libslaph_amd.so does not depend on libhdf5.  Skipped,
    with a message, where no HDF5 installation is found."""
    if not os.path.exists(os.path.join(HDF5, "include", "hdf5.h")):
        if verbose:
            print("HDF5 not found under %s: libslaph_h5.so and the contract program are not built" % HDF5)
        return None
    tmp = H5LIB + ".tmp"
    run([CXX, "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", "-I" + os.path.join(HDF5, "include"),
         os.path.join(SRC, "h5_writer.cpp"), "-o", tmp, "-L" + os.path.join(HDF5, "lib"), "-lhdf5",
         "-Wl,-rpath," + os.path.join(HDF5, "lib")], "libslaph_h5 build")
    os.replace(tmp, H5LIB)
    # the contract program: reference-style driver on the two C ABIs
    tmp = CONTRACT + ".tmp"
    run([CXX, "-O2", "-std=c++17", "-Wall", os.path.join(SRC, "contract_main.cpp"), "-o", tmp, "-L" + HERE,
         "-lslaph_amd", "-lslaph_h5", "-Wl,-rpath,$ORIGIN"], "contract build")
    os.replace(tmp, CONTRACT)
    if verbose:
        print("built", H5LIB, "and", CONTRACT)
    return H5LIB


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"], "link")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_h5(verbose)
    return LIB


if __name__ == "__main__":
    sys.exit(0 if build() else 1)

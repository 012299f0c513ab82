"""In-tree build of libtrixi_hip.so for gfx950 (hipcc, no JIT cache).

Every csrc/*.hip translation unit is compiled to an object in parallel (the
element-kernel instantiations are spread over elem_*.hip), the host-only
csrc/*.cpp units with the host C++ compiler, then all are linked into one
shared library."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(PKG_ROOT, "csrc")
LIB = os.path.join(PKG_ROOT, "libtrixi_hip.so")
OBJ_DIR = os.path.join(PKG_ROOT, "_build")

HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-fPIC", "-std=c++17",
    # packed FP32 (v_pk_*_f32) stays off on this hardware
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-Wall"]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def build_library(verbose=False, extra_flags=(), jobs=None):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sources()
    objs = [os.path.join(OBJ_DIR, os.path.splitext(os.path.basename(s))[0] + ".o") for s in srcs]

    def compile_one(so):
        src, obj = so
        if src.endswith(".cpp"):  # host code only
            cmd = [os.environ.get("CXX", "g++"), *CXX_FLAGS, "-c", "-o", obj, src]
        else:
            cmd = [hipcc, *HIPCC_FLAGS, *extra_flags, "-c", "-o", obj, src]
        if verbose:
            print(" ".join(cmd), flush=True)
        return subprocess.run(cmd, capture_output=True, text=True)

    jobs = jobs or min(len(srcs), max(1, os.cpu_count() or 1), 8)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(compile_one, zip(srcs, objs)))
    for src, r in zip(srcs, results):
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"compiling {os.path.basename(src)} failed")
    tmp = LIB + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("linking libtrixi_hip.so failed")
    os.replace(tmp, LIB)
    return LIB


HDF5_DIR = os.environ.get("TRIXI_HDF5_DIR", "/opt/conda")
IO_LIB = os.path.join(PKG_ROOT, "libtrixi_hip_io.so")


def build_io_library(verbose=False):
    """libtrixi_hip_io.so (solution/restart/mesh files): host C++ against libhdf5.
    Returns None (and builds nothing) when the HDF5 headers are absent."""
    inc = os.path.join(HDF5_DIR, "include")
    libdir = os.path.join(HDF5_DIR, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        print(f"HDF5 headers not found under {inc}: libtrixi_hip_io.so not built", flush=True)
        return None
    repo = os.path.dirname(PKG_ROOT)
    tmp = IO_LIB + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
           "-I" + os.path.join(repo, "include"), "-I" + inc, os.path.join(CSRC, "io", "trixi_hip_io.cpp"),
           "-L" + libdir, "-lhdf5", "-Wl,-rpath," + libdir, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("building libtrixi_hip_io.so failed")
    os.replace(tmp, IO_LIB)
    return IO_LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_io_library(verbose=True))

"""Build recipe of libafar.so (hipcc, gfx950).

Every translation unit is compiled with -ffp-contract=off: the FP64 path must
round exactly like the host oracle.  Packed FP32 instructions are disabled
for the device code.  Objects go to _build/ beside this file; the shared
library lands next to this file so that it travels with the package.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libafar.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["mesh.cpp", "precompute.cpp", "layout.cpp", "foam_io.cpp", "solver.cpp", "capi.cpp", "kernels.hip",
           "precompute_gpu.hip", "noalloc.hip"]
HOST_ONLY = {"mesh.cpp", "precompute.cpp", "layout.cpp", "foam_io.cpp"}  # no HIP runtime use: plain C++

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-result"]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + ["-I", CSRC, "-I", os.path.join(HERE, "..", "include")]
    if src in HOST_ONLY:
        cmd += ["-x", "c++"]
    else:
        cmd += DEVICE + ["-x", "hip"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), msgs))
    return obj, msgs


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    for obj, msgs in results:
        if verbose and msgs.strip():
            print(msgs, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stdout + r.stderr))
    os.replace(tmp, LIB)
    build_driver()
    return LIB


DRIVER = os.path.join(HERE, "cudaEuler2DmixO3RK3solver")
HOST_SOURCES = ["foamDict.cpp", "basicCUDAsolver.cpp", "cudaEuler2DmixO3RK3solver.cpp"]


def build_driver():
    """The C++ application over the C-ABI (g++, linked against libafar.so)."""
    srcs = [os.path.join(HERE, "host", s) for s in HOST_SOURCES]
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(HERE, "host"), "-I",
           os.path.join(HERE, "..", "include")] + srcs + ["-L", HERE, "-lafar", "-Wl,-rpath,$ORIGIN", "-o",
                                                         DRIVER + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed: %s\n%s" % (" ".join(cmd), r.stdout + r.stderr))
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))

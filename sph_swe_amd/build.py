"""Build recipe for libsphswe.so (gfx950) and the host program.

    python -m sph_swe_amd.build          # compile every HIP source, link the library,
                                         # build the host program

Every HIP translation unit is compiled with hipcc for gfx950 only; objects go to
sph_swe_amd/lib/obj, the shared library to sph_swe_amd/lib/libsphswe.so.  The
host program (host/main.cpp with the C++ classes of host/include) is plain C++
built with g++ against the C ABI into sph_swe_amd/lib/sphswe_app.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(LIBDIR, "obj")
LIB = os.path.join(LIBDIR, "libsphswe.so")
HOST = os.path.join(HERE, "host")
APP = os.path.join(LIBDIR, "sphswe_app")
CXX = os.environ.get("CXX", "g++")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC] + HIPFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    build_app()
    if verbose:
        print("built", LIB)
        print("built", APP)
    return LIB


def build_app():
    """The host program: C++17, the C ABI header, linked to libsphswe.so with an
    rpath to its directory."""
    tmp = APP + ".tmp"
    cmd = [CXX, "-O2", "-std=c++17", "-Wall", "-Wno-unused-function",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HOST, "include"),
           os.path.join(HOST, "main.cpp"), "-L" + LIBDIR, "-lsphswe", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host program build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, APP)
    return APP


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)

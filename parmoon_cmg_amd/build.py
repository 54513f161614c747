"""Build libcd3d_hip.so for gfx950 in-tree with hipcc (no GPU needed)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcd3d_hip.so")
SOURCES = ["setup.cpp", "kernels.hip", "lattice.hip", "diag.hip", "capi.cpp", "server.cpp", "noalloc.hip", "geo.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
         # packed FP32 instructions stay off (v_pk_* FP32 forms are avoided on this platform)
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose=False):
    objs = []
    os.makedirs(os.path.join(HERE, "_obj"), exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(HERE, "_obj", src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [HIPCC] + FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"], check=True)
    os.replace(tmp, LIB)
    build_examples(verbose)
    return LIB


def build_examples(verbose=False):
    """The reference-style main over the drop-in C++ classes
    (include/parmoon_cd3d.hpp), linked against the library in-tree."""
    os.makedirs(os.path.join(HERE, "_bin"), exist_ok=True)
    src = os.path.join(HERE, "examples", "cd3d_shim_main.cpp")
    exe = os.path.join(HERE, "_bin", "cd3d_shim_main")
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", "-Wextra",
           "-I" + os.path.join(os.path.dirname(HERE), "include"), src, "-o", exe + ".tmp",
           "-L" + HERE, "-lcd3d_hip", "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(exe + ".tmp", exe)
    return exe


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

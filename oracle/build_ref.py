#!/usr/bin/env python3
"""Build recipe for ``oracle/_ref``: the ParMooN reference compiled from its own
sources, used only as a checker (fixture generation and the CPU baseline).

The recipe compiles the 3D sequential (``_SEQ``) translation units of the
reference in place (sources are read from ``/root/reference``; nothing is copied
into this repository) with ``g++``/``gcc`` directly -- the reference's CMake
build is not used. Objects go to ``oracle/_ref/obj`` and the linked driver to
``oracle/_ref/cd3d_ref``.

Source selection mirrors the reference's CMake lists for ``AParMooN_GEO=3D``,
``AParMooN_PARALLEL_TYPE=SEQUENTIAL``:
  src/General, src/QuadFormulas, src/Refinement  -- every ``*.C`` (GLOB in CMakeLists.txt)
  src/FE, src/Geometry, src/System               -- ``list(APPEND ...)`` entries outside
                                                    the 2D branch (src/FE/CMakeLists.txt:84-170)
The four translation units that need MKL/cblas/pardiso headers the image lacks
(DirectSolver.C, FE3D_ALE.C, SystemTNSE3D_ALE.C, SystemTNSE3D.C) are skipped;
none of them is on the CD3D/TCD3D path.  The link ignores unresolved symbols of
those absent third-party libraries (UMFPACK, LAPACK, pardiso, tecplot, tetgen):
no function on the GMG/FGMRES path with coarse smoother 1/2/3 calls them.  No
stand-in is written for any of them.

Usage: python oracle/build_ref.py [-j N]
"""
import argparse
import concurrent.futures as cf
import os
import re
import subprocess
import sys

REF = os.environ.get("PARMOON_REF", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
OBJ = os.path.join(OUT, "obj")

SKIP = {"DirectSolver.C", "FE3D_ALE.C", "SystemTNSE3D_ALE.C", "SystemTNSE3D.C"}
INC_DIRS = ["include/General", "include/FE", "include/Geometry", "include/QuadFormulas",
            "include/FE1D", "include/FE2D", "include/FE3D", "include/AMG", "include/PBE",
            "include/Refinement", "include/System", "include/Parallel", "Examples",
            "include/Solvers", "EXT_LIB/UMFPACK/Include", "EXT_LIB/tetgen",
            "EXT_LIB/GridGen", "EXT_LIB/tecplot/include"]
DEFS = ["-D__3D__", "-D__LINUX64__", "-DTRILIBRARY", "-DTETLIBRARY", "-D_SEQ", "-D__PRIVATE__"]


def cmake_list(subdir):
    """Sources appended in src/<subdir>/CMakeLists.txt for the 3D build."""
    path = os.path.join(REF, "src", subdir, "CMakeLists.txt")
    out, branch = [], None
    for line in open(path, encoding="latin-1"):
        s = line.strip()
        if s.startswith("#"):
            continue
        m = re.match(r'(if|elseif)\("\$\{AParMooN_GEO\}" STREQUAL "(\w+)"\)', s)
        if m:
            branch = m.group(2)
            continue
        if s.startswith('endif("${AParMooN_GEO}"'):
            branch = None
            continue
        if 'AParMooN_PARALLEL_TYPE' in s and s.startswith("if("):
            branch = "PAR"
            continue
        if s.startswith('endif("${AParMooN_PARALLEL_TYPE}"'):
            branch = None
            continue
        m = re.match(r'list\(APPEND \w+ "\$\{PROJECT_SOURCE_DIR\}/(src/[^"]+)"\)', s)
        if m and branch in (None, "3D"):
            out.append(m.group(1))
    return out


def sources():
    srcs = []
    for d in ("General", "QuadFormulas", "Refinement"):
        for f in sorted(os.listdir(os.path.join(REF, "src", d))):
            if f.endswith(".C"):
                srcs.append(f"src/{d}/{f}")
    for d in ("FE", "Geometry", "System"):
        srcs += cmake_list(d)
    seen, res = set(), []
    for s in srcs:
        if os.path.basename(s) in SKIP or s in seen:
            continue
        if not os.path.exists(os.path.join(REF, s)):
            continue
        seen.add(s)
        res.append(s)
    return res


def cxx_flags():
    return ["-O2", "-w", "-std=c++11", "-fopenmp"] + DEFS + [f"-I{os.path.join(REF, i)}" for i in INC_DIRS]


def compile_one(src):
    obj = os.path.join(OBJ, src.replace("/", "__") + ".o")
    full = os.path.join(REF, src)
    if os.path.exists(obj) and os.path.getmtime(obj) >= os.path.getmtime(full):
        return src, 0, ""
    cmd = ["g++"] + cxx_flags() + ["-c", full, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    return src, p.returncode, p.stderr[-2000:]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    args = ap.parse_args()
    if not os.path.isdir(os.path.join(REF, "src")):
        print(f"reference sources not found at {REF}; skipping oracle/_ref build")
        return 0
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    failed = []
    with cf.ThreadPoolExecutor(args.j) as ex:
        for src, rc, err in ex.map(compile_one, srcs):
            if rc:
                failed.append(src)
                print(f"FAILED {src}\n{err}", file=sys.stderr)
    if failed:
        print(f"{len(failed)} reference translation units failed", file=sys.stderr)
        return 1
    drv = os.path.join(HERE, "ref_driver.C")
    drv_obj = os.path.join(OBJ, "ref_driver.o")
    p = subprocess.run(["g++"] + cxx_flags() + ["-c", drv, "-o", drv_obj], capture_output=True, text=True)
    if p.returncode:
        print(p.stderr[-4000:], file=sys.stderr)
        return 1
    objs = [os.path.join(OBJ, s.replace("/", "__") + ".o") for s in srcs]
    exe = os.path.join(OUT, "cd3d_ref")
    p = subprocess.run(["g++", "-fopenmp", "-o", exe, drv_obj] + objs +
                       ["-Wl,--unresolved-symbols=ignore-all"], capture_output=True, text=True)
    if p.returncode:
        print(p.stderr[-4000:], file=sys.stderr)
        return 1
    print(f"built {exe} from {len(srcs)} reference translation units")
    return 0


if __name__ == "__main__":
    sys.exit(main())

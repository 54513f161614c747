"""The C-ABI library loads on a CPU-only host and exports every entry point
that include/cd3d.h and include/cd3d_server.h declare (no compute call: there is no GPU here)."""
import ctypes
import os
import re
import subprocess

import parmoon_cmg_amd as pm
from tests.conftest import REPO


def declared_functions():
    names = set()
    for h in ("cd3d.h", "cd3d_server.h"):
        src = open(os.path.join(REPO, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(cd3d_\w+)\s*\(", src))
    return sorted(names)


def test_header_matches_python_mirror():
    assert declared_functions() == sorted(pm.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = pm.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", pm.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(l.split()[-1] for l in out.stdout.splitlines() if l.strip())
    for name in declared_functions():
        assert name in exported, name


def test_default_solver_needs_no_gpu():
    sp = pm.solver_params()
    assert (sp.pre, sp.post, sp.restart, sp.coarse_smoother) == (2, 2, 20, 17)
    assert abs(sp.omega - 0.8) < 1e-15 and abs(sp.rel_tol - 1e-10) < 1e-25


def test_library_carries_gfx950_code_object():
    # the device code sits in the .hip_fatbin section, tagged with its target
    out = subprocess.run(["readelf", "-S", pm.LIB_PATH], capture_output=True, text=True, check=True)
    assert ".hip_fatbin" in out.stdout
    assert b"amdgcn-amd-amdhsa--gfx950" in open(pm.LIB_PATH, "rb").read()

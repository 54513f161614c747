"""The drop-in C++ classes (include/parmoon_cd3d.hpp) driven by a
reference-style main program (parmoon_cmg_amd/examples/cd3d_shim_main.cpp):
TSystemCD3D, the hand-wired TMultiGrid3D / TMGLevel3D / TMultiGridScaIte /
TFgmresIte stack and TSystemTCD3D reproduce the reference runs of the golden
fixtures (iteration counts equal, solutions within 1e-10)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from tests.cases import CD_CASES, REF_BOX, TCD_CASES, tcd_solver

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "parmoon_cmg_amd", "_bin", "cd3d_shim_main")
HEADER = os.path.join(REPO, "include", "parmoon_cd3d.hpp")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def case_args(c):
    co = c["co"]
    box = c.get("box", REF_BOX)
    a = [f"order={c['order']}", f"m={c.get('m', 1)}", f"k0={c['k0']}", f"levels={c['levels']}",
         "box=" + ",".join(map(repr, box))]
    a += [f"eps={co.get('eps', 1.0)!r}", "b=" + ",".join(map(repr, co.get("b", (0.0, 0.0, 0.0)))),
          f"c={co.get('c', 0.0)!r}", f"f_kind={co.get('f_kind', 0)}", f"f_val={co.get('f_val', 1.0)!r}",
          f"supg={co.get('supg', 0)}", f"conv_kind={co.get('conv_kind', 0)}"]
    if "rot_c" in co:
        a.append("rot=" + ",".join(map(repr, co["rot_c"])))
    if "pulse_c" in co:
        a.append("pulse=" + ",".join(map(repr, co["pulse_c"])))
    return a


def run_shim(args, tmp_path, tag):
    out = str(tmp_path / tag)
    p = subprocess.run([EXE] + args + [f"out={out}"], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    x = np.fromfile(out + ".sol", dtype=np.float64)
    meta = dict(line.split() for line in open(out + ".txt"))
    return x, int(meta["its"]), float(meta["r0"])


def test_shim_header_compiles_standalone(tmp_path):
    src = tmp_path / "inc.cpp"
    src.write_text('#include "parmoon_cd3d.hpp"\nint main() { return 0; }\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                    "-I" + os.path.dirname(HEADER), str(src)], check=True)


def test_shim_example_is_built_against_the_library():
    assert os.path.exists(EXE), "run __graft_entry__.build() first"
    ldd = subprocess.run(["ldd", EXE], capture_output=True, text=True).stdout
    assert "libcd3d_hip.so" in ldd and "not found" not in ldd.split("libcd3d_hip.so")[1].splitlines()[0]


def test_shim_reports_errors_like_the_reference(tmp_path):
    # an invalid element order is refused before any device work: the shim
    # prints the message and exits with 4711 (seen as 4711 & 255 = 103)
    p = subprocess.run([EXE, "order=3", f"out={tmp_path / 'x'}"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 4711 & 255
    assert "cd3d_create failed" in p.stderr and "order must be 1 or 2" in p.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["q1_c1", "q2_l2_supg", "q1_l3_supg"])
@pytest.mark.parametrize("mode", ["cd", "mg"])
def test_shim_cd3d_matches_reference(name, mode, golden, tmp_path):
    d = golden(name)
    say(f"[gpu] shim {mode} {name}")
    x, its, _ = run_shim([f"mode={mode}"] + case_args(CD_CASES[name]), tmp_path, f"{mode}_{name}")
    err = np.linalg.norm(x - d["sol"]) / np.linalg.norm(d["sol"])
    say(f"[gpu] shim {mode} {name}: its {its} (ref {int(d['its'])}) err {err:.2e}")
    assert its == int(d["its"])
    assert err < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["unit_q1_var_gal", "unit_q1_var_supg"])
def test_shim_coefficient_function_matches_reference(name, golden, tmp_path):
    # TSystemCD3D::Init with a CoeffFct3D (the reference's signature): the
    # callback runs on the host at every quadrature point, the device
    # assembles with its values; the reference run used the same function
    c = CD_CASES[name]
    d = golden(name)
    q = c["qp"]
    say(f"[gpu] shim cdfct {name}")
    args = [a for a in case_args(c) if not a.startswith(("eps=", "b=", "c=", "f_val="))]
    args += [f"eps={q['eps']!r}", "b=" + ",".join(map(repr, q["b"])), f"c={q['c']!r}", f"f_val={q['f']!r}"]
    sv = c.get("solver", {})
    args += [f"{k}={v!r}" for k, v in sv.items()]
    x, its, _ = run_shim(["mode=cdfct"] + args, tmp_path, "cdfct")
    err = np.linalg.norm(x - d["sol"]) / np.linalg.norm(d["sol"])
    say(f"[gpu] shim cdfct {name}: its {its} (reference {int(d['its'])}), err {err:.2e}")
    assert its == int(d["its"]) and err < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["q1_tcd", "unit_q1_tcd_supg_const"])
def test_shim_tcd3d_matches_reference(name, golden, tmp_path):
    c = TCD_CASES[name]
    d = golden(name)
    say(f"[gpu] shim tcd {name}")
    sv = tcd_solver(c)
    extra = [f"omega={sv['omega']!r}", f"pre={sv['pre']}", f"post={sv['post']}"]
    if c["co"].get("supg"):
        extra.append("tcd_stop=step")  # 1e-10 of each step's initial residual
    x, _, _ = run_shim(["mode=tcd", f"tau={c['tau']!r}", f"nsteps={c['nsteps']}"] + extra + case_args(c), tmp_path,
                       "tcd")
    states = x.reshape(c["nsteps"], -1)
    for k in range(c["nsteps"]):
        ref = d[f"u{k + 1}"]
        assert np.abs(states[k] - ref).max() <= 1e-10 * np.abs(ref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["m3w_q1_poisson", "m3w_q2_supg"])
def test_shim_geo_domain_matches_reference(name, golden, tmp_path):
    # the reference main's own mesh set-up: TDomain::Init with the GEO file
    # the reference run read, RegRefineAll, GetCollection(It_Finest, 0) and
    # TFESpace3D(coll, name, description, BoundCondition, order) per level
    from tests.cases import case_solver, geo_text
    c = CD_CASES[name]
    d = golden(name)
    geo = tmp_path / "mesh.GEO"
    geo.write_text(geo_text(c))
    sv = case_solver(name)
    args = [a for a in case_args(c) if not a.startswith(("box=", "m="))] + [f"geo={geo}"]
    args += [f"omega={sv['omega']!r}", f"pre={sv['pre']}", f"post={sv['post']}"]
    say(f"[gpu] shim geo {name}")
    x, its, _ = run_shim(["mode=cd"] + args, tmp_path, "geo")
    err = np.linalg.norm(x - d["sol"]) / np.linalg.norm(d["sol"])
    say(f"[gpu] shim geo {name}: its {its} (reference {int(d['its'])}), err {err:.2e}")
    assert its == int(d["its"]) and err < 1e-10


def test_shim_geo_domain_refuses_parametrised_vertices(tmp_path):
    # a GEO file whose vertices are boundary parameters (KNPR != 0, as in
    # ParMooN's Wuerfel.GEO) needs the PRM file: refused with the reader's reason
    from tests.cases import _box_geo_module
    lines = _box_geo_module().box_geo(1).split("\n")
    k = lines.index("KNPR")
    lines[k + 1] = " " + " ".join(["  1"] * 8)
    geo = tmp_path / "param.GEO"
    geo.write_text("\n".join(lines))
    p = subprocess.run([EXE, f"geo={geo}", f"out={tmp_path / 'x'}"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 4711 & 255
    assert "TDomain::Init" in p.stderr and "KNPR" in p.stderr

"""Parity of the HIP path (through the C ABI) with the reference and the oracle.

Bars: DOF maps, CSR pattern and row pointers bit-exact; assembled entries
within 1e-12 of the row's largest entry; V-cycle probe 1e-12 relative;
FGMRES the same iteration count as the reference and the solution within
1e-10 relative L2; mixed precision: relative residual 1e-10 in FP64 and
solution within 1e-9 (relative L2) of the FP64 oracle's solution.

Cases (tests/cases.py): ParMooN's Wuerfel cube, the unit cube split into
m^3 macro bricks (m = 2, 3, 5: the batched and serving workloads' meshes)
and deformed m^3 macro meshes of general hexahedra read from GEO files,
Q1 and Q2, Galerkin and SUPG (constant and rotating convection), V-, W- and
F-cycles, and the benchmark's smoother; each runs through the class-stencil
assembly the library picks and, for the SUPG cases, through the cell-by-cell
element kernel as well."""
import sys

import numpy as np
import pytest

import parmoon_cmg_amd as pm
from oracle import pyoracle as po
from tests.cases import (CD_CASES, REF_BOX, REF_SOLVER, TCD_CASE, TCD_CASES, case_box, case_solver, geo_text,
                         oracle_hier, qp_coeffs, tcd_solver)

pytestmark = pytest.mark.gpu

ENTRY_TOL = 1e-12
SOL_TOL = 1e-10
MIXED_TOL = 1e-9


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def row_scaled_err(e, er, rp):
    rowmax = np.maximum.reduceat(np.abs(er), rp[:-1])
    return (np.abs(e - er) / np.repeat(rowmax, np.diff(rp))).max()


def make_problem(c):
    if "geo" in c:  # general hexahedra: the GEO file the reference run read, through cd3d_read_geo
        import os
        import tempfile
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "mesh.GEO")
            with open(path, "w") as f:
                f.write(geo_text(c))
            m, xyz = pm.read_geo(path)
        assert m == c["m"]
        P = pm.Problem(c["order"], m, c["k0"], c["levels"], vertices=xyz)
    else:
        P = pm.Problem(c["order"], c.get("m", 1), c["k0"], c["levels"], box=case_box(c))
    if "qp" in c:  # the reference run's CoeffFct3D at every quadrature point
        for l in range(c["levels"]):
            P.set_coeffs_qp(l, qp_coeffs(c, P.quad_points(l)))
    return P


_problems = {}


def problem(name, golden):
    if name not in _problems:
        c = CD_CASES[name]
        say(f"[gpu] create {name} order={c['order']} m={c.get('m', 1)} k0={c['k0']} levels={c['levels']}")
        P = make_problem(c)
        P.assemble(pm.coeffs(**c["co"]))
        _problems[name] = P
    return _problems[name]


_oracles = {}


def oracle_solution(name):
    """FP64 oracle solution with the exact coarse solve, relative residual 1e-10."""
    if name not in _oracles:
        c = CD_CASES[name]
        H = oracle_hier(c)
        if "qp" in c:
            for l in range(c["levels"]):
                H.set_coeffs_qp(l, qp_coeffs(c, H.quad_points(l)))
        H.assemble(po.coeffs(**c["co"]))
        b = H.rhs(c["levels"] - 1)
        _, r0 = H.defect(c["levels"] - 1, np.zeros_like(b), b)
        xo, its_o, _ = H.fgmres(po.solver(**dict(case_solver(name), coarse_smoother=17, res_norm_min=1e-10 * r0,
                                                 red_factor=0.0)), b)
        _oracles[name] = (xo, its_o)
    return _oracles[name]


@pytest.mark.parametrize("name", list(CD_CASES))
def test_structure_bit_exact(name, golden):
    d = golden(name)
    P = problem(name, golden)
    for l in range(CD_CASES[name]["levels"]):
        nc, nbf, ndof, act, hb, ndir, nnz = d[f"L{l}_meta"].tolist()
        i = P.info(l)
        assert (i["ncell"], i["ndof"], i["nact"], i["nnz"]) == (nc, ndof, act, nnz)
        assert np.array_equal(P.gn(l), d[f"L{l}_gn"])
        # vertex coordinates: the GEO files of m = 3, 5 hold i/m rounded once
        assert np.abs(P.cells(l) - d[f"L{l}_cells"]).max() <= 2.3e-16
        rp, kc, _ = P.csr(l)
        assert np.array_equal(rp, d[f"L{l}_rowptr"]) and np.array_equal(kc, d[f"L{l}_kcol"])


def check_entries(P, name, d, tag):
    for l in range(CD_CASES[name]["levels"]):
        rp, _, e = P.csr(l)
        err = row_scaled_err(e, d[f"L{l}_entries"], rp)
        rr = d[f"L{l}_rhs"]
        rerr = np.abs(P.rhs(l) - rr).max() / np.abs(rr).max()
        say(f"[gpu] {name} {tag} level {l} entry err {err:.2e} rhs err {rerr:.2e}")
        assert err < ENTRY_TOL
        assert rerr <= ENTRY_TOL


@pytest.mark.parametrize("name", list(CD_CASES))
def test_assembly_matches_reference(name, golden):
    check_entries(problem(name, golden), name, golden(name), "assembly")


@pytest.mark.parametrize("name", list(CD_CASES))
def test_vcycle_matches_reference(name, golden):
    d = golden(name)
    P = problem(name, golden)
    sp = pm.solver_params(**case_solver(name))
    z = P.vcycle(sp, d["vcycle_in"])
    err = np.abs(z - d["vcycle_out"]).max() / np.abs(d["vcycle_out"]).max()
    say(f"[gpu] {name} vcycle err {err:.2e}")
    assert err < 1e-12


def check_fgmres(P, name, d, tag):
    sp = pm.solver_params(**case_solver(name), rel_tol=1e-10)
    x, st = P.solve(sp)
    err = np.linalg.norm(x - d["sol"]) / np.linalg.norm(d["sol"])
    say(f"[gpu] {name} {tag} its {st.iterations} (ref {int(d['its'])}) sol err {err:.2e} "
        f"rel res {st.rfinal / st.r0:.2e}")
    assert st.status == 0
    assert st.iterations == int(d["its"])
    assert err < SOL_TOL
    assert st.rfinal <= 1.0001e-10 * st.r0


@pytest.mark.parametrize("name", list(CD_CASES))
def test_fgmres_matches_reference(name, golden):
    check_fgmres(problem(name, golden), name, golden(name), "fgmres")


SUPG_CASES = [n for n, c in CD_CASES.items() if c["co"].get("supg")]


@pytest.mark.parametrize("name", SUPG_CASES)
def test_cellwise_assembly_matches_reference(name, golden, monkeypatch):
    # the per-element quadrature kernel (colour-scattered into CSR) instead of
    # the class stencils: the same entries, rhs and FGMRES run as the reference
    d = golden(name)
    monkeypatch.setenv("CD3D_CELLWISE_ASSEMBLY", "1")
    P = make_problem(CD_CASES[name])
    P.assemble(pm.coeffs(**CD_CASES[name]["co"]))
    check_entries(P, name, d, "cellwise")
    check_fgmres(P, name, d, "cellwise")
    P.close()


@pytest.mark.parametrize("name", list(CD_CASES))
def test_fgmres_exact_coarse_matches_oracle(name, golden):
    P = problem(name, golden)
    xo, its_o = oracle_solution(name)
    x, st = P.solve(pm.solver_params(**dict(case_solver(name), coarse_smoother=17), rel_tol=1e-10))
    err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    say(f"[gpu] {name} exact coarse: its {st.iterations} (oracle {its_o}) err {err:.2e}")
    assert st.iterations == its_o
    assert err < SOL_TOL


MIXED_CASES = ["q2_l2_supg", "q1_l3_supg", "q1_c1", "unit_c2_n8_bench", "m2_q1_supg", "m3_q1_supg",
               "m5_q1_supg", "m2_q2_supg", "m3_q2_supg", "unit_q1_supg_rot", "m2_q2_supg_rot",
               "m2_q1_wcycle", "m2_q1_fcycle", "unit_q1_var_supg", "m2_q2_var_supg",
               "m3w_q1_poisson", "m3w_q2_supg", "m2w_q1_var_supg", "m2w_q2_supg_rot"]


@pytest.mark.parametrize("name", MIXED_CASES)
def test_mixed_precision_against_oracle(name, golden):
    # FP32 V-cycle inside FP64 FGMRES: relative residual 1e-10 (recomputed in
    # FP64 with the CSR matrix) and the solution within 1e-9 of the FP64
    # oracle's solution (exact coarse solve, relative residual 1e-10)
    P = problem(name, golden)
    xo, its_o = oracle_solution(name)
    xm, stm = P.solve(pm.solver_params(**dict(case_solver(name), coarse_smoother=17), rel_tol=1e-10,
                                       precision=pm.MIXED))
    err = np.linalg.norm(xm - xo) / np.linalg.norm(xo)
    say(f"[gpu] {name} mixed its {stm.iterations} vs oracle fp64 {its_o}; rel res {stm.rfinal / stm.r0:.2e};"
        f" err vs oracle {err:.2e}")
    assert stm.status == 0
    assert stm.rfinal <= 1.0001e-10 * stm.r0
    assert err < MIXED_TOL


def test_maxit_inside_a_restart_cycle_returns_the_arnoldi_iterate(golden):
    # maxit 25 with restart 20 ends the iteration 5 columns into the second
    # cycle: the returned iterate is x + Z_5 y_5 of that cycle, so its true
    # residual is the Arnoldi estimate |s_25| (the oracle's residual history,
    # which the reference's back-substitution through restart - 1 columns
    # would not reproduce); status NOT_CONVERGED
    name = "m2_q1_supg"
    c = CD_CASES[name]
    P = problem(name, golden)
    H = po.Hierarchy(c["order"], c["m"], c["k0"], c["levels"], box=case_box(c))
    H.assemble(po.coeffs(**c["co"]))
    b = H.rhs(c["levels"] - 1)
    sv = dict(case_solver(name), maxit=25)
    _, its_o, hist = H.fgmres(po.solver(**sv, rel_tol=1e-10), b)
    x, st = P.solve(pm.solver_params(**sv, rel_tol=1e-10), check=False)
    say(f"[gpu] maxit 25: status {st.status}, its {st.iterations}, true residual {st.rfinal:.6e}, "
        f"Arnoldi estimate {hist[24]:.6e}")
    assert its_o == 25 and st.iterations == 25 and st.status == pm.CD3D_NOT_CONVERGED
    assert abs(st.rfinal / hist[24] - 1) < 1e-6


@pytest.mark.parametrize("order", [1, 2])
def test_transfers_and_matvec_match_oracle(order):
    say(f"[gpu] transfers order {order}")
    P = pm.Problem(order, 1, 1, 3, box=REF_BOX)
    H = po.Hierarchy(order, 1, 1, 3, box=REF_BOX)
    co = dict(eps=0.01, b=(1.0, -0.5, 0.3), c=0.2, f_val=1.0, supg=1)
    P.assemble(pm.coeffs(**co))
    H.assemble(po.coeffs(**co))
    rng = np.random.default_rng(7)
    for l in (1, 2):
        nc, nf = H.info(l - 1)["ndof"], H.info(l)["ndof"]
        u, v = rng.standard_normal(nc), rng.standard_normal(nf)
        assert np.abs(P.prolongate(l, u) - H.prolongate(l, u)).max() < 1e-14 * np.abs(u).max() * 8
        rg, ro = P.restrict(l, v), H.restrict(l, v)
        assert np.abs(rg - ro).max() < 1e-13 * np.abs(ro).max()
        x = rng.standard_normal(nf)
        yg, yo = P.matvec(l, x), H.matvec(l, x)
        assert np.abs(yg - yo).max() < 1e-12 * np.abs(yo).max()


@pytest.mark.parametrize("name", list(TCD_CASES))
def test_tcd_crank_nicolson_matches_reference(name, golden):
    c = TCD_CASES[name]
    d = golden(name)
    say(f"[gpu] tcd {name}")
    P = make_problem(c)
    co = pm.coeffs(**c["co"])
    u0 = P.interp_pulse(c["levels"] - 1, co)
    sp = pm.solver_params(**tcd_solver(c))
    _, states, st = P.tcd_run(co, sp, 0.0, c["tau"], c["nsteps"], u0)
    for k in range(c["nsteps"]):
        ref = d[f"u{k + 1}"]
        err = np.abs(states[k] - ref).max() / np.abs(ref).max()
        say(f"[gpu] tcd {name} step {k + 1} err {err:.2e}")
        assert err < SOL_TOL
    if "its_steps" in d:
        say(f"[gpu] tcd {name} its {st.iterations} (reference {int(d['its_steps'].sum())})")
        assert st.iterations == int(d["its_steps"].sum())
    P.close()


@pytest.mark.parametrize("name", [n for n, c in TCD_CASES.items() if c["co"].get("supg")])
def test_tcd_supg_mixed_matches_reference(name, golden):
    # the SUPG Crank-Nicolson steps with the FP32 V-cycle: every state within
    # the mixed form's bar of the reference's FP64 states
    c = TCD_CASES[name]
    d = golden(name)
    P = make_problem(c)
    co = pm.coeffs(**c["co"])
    u0 = P.interp_pulse(c["levels"] - 1, co)
    sp = pm.solver_params(**tcd_solver(c), precision=pm.MIXED)
    _, states, st = P.tcd_run(co, sp, 0.0, c["tau"], c["nsteps"], u0)
    for k in range(c["nsteps"]):
        ref = d[f"u{k + 1}"]
        err = np.abs(states[k] - ref).max() / np.abs(ref).max()
        say(f"[gpu] tcd supg mixed {name} step {k + 1} err {err:.2e}")
        assert err < MIXED_TOL
    P.close()


def test_tcd_stepwise_matches_run_and_backward_euler_residual(golden):
    import scipy.sparse as sps
    c = TCD_CASE
    P = pm.Problem(c["order"], 1, c["k0"], c["levels"], box=REF_BOX)
    co = pm.coeffs(**c["co"])
    u0 = P.interp_pulse(c["levels"] - 1, co)
    sp = pm.solver_params(**REF_SOLVER, rel_tol=0.0, res_norm_min=1e-30, red_factor=1e-10)
    _, states, _ = P.tcd_run(co, sp, 0.0, c["tau"], c["nsteps"], u0)
    P.tcd_init(co, 0.0, u0)
    for k in range(c["nsteps"]):
        u, _ = P.tcd_step(co, sp, c["tau"])
        assert np.array_equal(u, states[k])
    assert np.array_equal(P.tcd_state(), states[-1])
    # backward Euler (theta 1): (M + tau A(tau)) u1 = M u0 on the active rows (f = 0)
    P.tcd_init(co, 0.0, u0)
    u1, _ = P.tcd_step(co, sp, c["tau"], theta=1.0)
    lf = c["levels"] - 1
    rp, kc, M = P.csr(lf)  # the mass matrix, restored after the step
    Q = pm.Problem(c["order"], 1, c["k0"], c["levels"], box=REF_BOX)
    Q.assemble(pm.coeffs(**dict(c["co"], t=c["tau"])))
    _, _, A = Q.csr(lf)
    n, na = len(rp) - 1, P.info(lf)["nact"]
    Mm = sps.csr_matrix((M, kc, rp), shape=(n, n))
    Am = sps.csr_matrix((A, kc, rp), shape=(n, n))
    u0d = u0.copy()
    u0d[na:] = 0.0  # Dirichlet value g = 0 from the rhs
    r = (Mm @ u1 + c["tau"] * (Am @ u1) - Mm @ u0d)[:na]
    rel = np.linalg.norm(r) / np.linalg.norm((Mm @ u0d)[:na])
    say(f"[gpu] tcd backward Euler residual {rel:.2e}")
    assert rel < 1e-9


@pytest.mark.parametrize("name", ["q2_l2_supg", "q1_l3_supg", "q1_c1"])
def test_mixed_precision_against_fp64_form(name, golden):
    P = problem(name, golden)
    x64, st64 = P.solve(pm.solver_params(**dict(REF_SOLVER, coarse_smoother=17), rel_tol=1e-10))
    xm, stm = P.solve(pm.solver_params(**dict(REF_SOLVER, coarse_smoother=17), rel_tol=1e-10,
                                       precision=pm.MIXED))
    err = np.linalg.norm(xm - x64) / np.linalg.norm(x64)
    say(f"[gpu] {name} mixed its {stm.iterations} vs fp64 {st64.iterations}; rel res {stm.rfinal / stm.r0:.2e};"
        f" err vs fp64 {err:.2e}")
    assert stm.rfinal <= 1.0001e-10 * stm.r0
    assert err < 1e-9


@pytest.mark.parametrize("order,k0,levels", [(1, 2, 4), (2, 2, 3)])
def test_structure_closed_forms_larger(order, k0, levels):
    # size-independent check at sizes the reference fixtures do not cover
    P = pm.Problem(order, 1, k0, levels, box=(1.0, 1.0, 1.0))
    for l in range(levels):
        n = 2 ** (k0 + l)
        i = P.info(l)
        if order == 1:
            assert i["nnz"] == (3 * (n - 1)) ** 3 + (n + 1) ** 3 - (n - 1) ** 3
        else:
            assert i["nnz"] == (8 * n - 5) ** 3 + (2 * n + 1) ** 3 - (2 * n - 1) ** 3
    rp, kc, _ = P.csr(levels - 1)
    assert np.all(np.diff(rp) > 0)
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    same_row = rows[1:] == rows[:-1]
    assert np.all(kc[1:][same_row] > kc[:-1][same_row])  # sorted, unique columns


@pytest.mark.parametrize("order,m,k0,levels,conv_kind", [(2, 1, 1, 3, 0), (1, 2, 1, 3, 0), (2, 3, 0, 3, 0),
                                                          (1, 1, 2, 4, 0), (2, 1, 1, 3, 1), (1, 2, 1, 3, 1)])
def test_mixed_vcycle_against_fp64(order, m, k0, levels, conv_kind):
    # the FP32 lattice-stencil V-cycle is the FP64 CSR V-cycle up to rounding;
    # constant coefficients give the same stencil at every node of a class
    # (compact form), the rotating field a different one at every node
    P = pm.Problem(order, m, k0, levels, box=(1.0, 0.7, 0.5))
    P.assemble(pm.coeffs(eps=1e-2, b=(1.0, 0.5, 0.25), supg=1, f_val=1.0, conv_kind=conv_kind, t=0.1))
    i = P.info(levels - 1)
    r = np.random.default_rng(3).standard_normal(i["ndof"])
    r[i["nact"]:] = 0.0
    sp = dict(REF_SOLVER, coarse_smoother=17)
    z64 = P.vcycle(pm.solver_params(**sp), r)
    z32 = P.vcycle(pm.solver_params(**sp, precision=pm.MIXED), r)
    err = np.linalg.norm(z32 - z64) / np.linalg.norm(z64)
    say(f"[gpu] mixed vcycle order={order} m={m} k0={k0} levels={levels} conv_kind={conv_kind} err vs fp64 {err:.2e}")
    assert err < 1e-5
    assert np.all(z32[i["nact"]:] == 0.0)
    x, st = P.solve(pm.solver_params(**sp, rel_tol=1e-10, precision=pm.MIXED))
    x64, st64 = P.solve(pm.solver_params(**sp, rel_tol=1e-10))
    assert st.rfinal <= 1.0001e-10 * st.r0
    assert np.linalg.norm(x - x64) / np.linalg.norm(x64) < 1e-9


@pytest.mark.parametrize("order,m,k0,levels,conv_kind", [(2, 1, 1, 3, 0), (2, 1, 1, 3, 1), (1, 1, 2, 4, 0),
                                                          (1, 2, 1, 3, 1), (2, 1, 2, 4, 0)])
def test_small_level_program_matches_launch_per_step(order, m, k0, levels, conv_kind, monkeypatch):
    # the bottom of the mixed V-cycle as one workgroup program (levels of at
    # most CD3D_SMALL_NODES lattice nodes and the exact coarse solve) against
    # the launch-per-step V-cycle: the same arithmetic, so bitwise equal.
    # Cases: the whole V-cycle in one launch (Q2 compact / per-node stencils),
    # three Q1 levels, the Q1 rotating field, and a fused bottom below
    # launch-per-step levels
    P = pm.Problem(order, m, k0, levels, box=(1.0, 0.7, 0.5))
    P.assemble(pm.coeffs(eps=1e-2, b=(1.0, 0.5, 0.25), supg=1, f_val=1.0, conv_kind=conv_kind, t=0.1))
    i = P.info(levels - 1)
    r = np.random.default_rng(9).standard_normal(i["ndof"])
    r[i["nact"]:] = 0.0
    sp = pm.solver_params(**dict(REF_SOLVER, omega=0.5, pre=3, post=3, coarse_smoother=17, rel_tol=1e-10),
                          precision=pm.MIXED)
    out = {}
    for bound in ("0", "100000"):
        monkeypatch.setenv("CD3D_SMALL_NODES", bound)
        z = P.vcycle(sp, r)
        x, st = P.solve(sp)
        out[bound] = (z, x, st.iterations)
    say(f"[gpu] small-level program order={order} m={m} k0={k0} levels={levels} conv_kind={conv_kind}: "
        f"its {out['0'][2]} / {out['100000'][2]}")
    assert np.array_equal(out["0"][0], out["100000"][0])
    assert np.array_equal(out["0"][1], out["100000"][1])
    assert out["0"][2] == out["100000"][2]


@pytest.mark.parametrize("m,k0,levels", [(1, 1, 5), (3, 0, 4), (11, 0, 2)])
def test_pair_groups_match_single_pairs(m, k0, levels, monkeypatch):
    # the Q2 compact-stencil sweep with 2 or 4 vertex/midpoint pairs per
    # thread against one pair per thread: the same per-node sums, so the
    # V-cycle and the mixed solve are bitwise equal.  32, 24 and 22 cells per
    # side: 22 pairs per line do not split into groups of 4 (that size falls
    # back to single pairs)
    P = pm.Problem(2, m, k0, levels, box=(1.0, 0.7, 0.5))
    P.assemble(pm.coeffs(eps=1e-3, b=(1.0, 0.5, 0.25), supg=1, f_val=1.0))
    i = P.info(levels - 1)
    r = np.random.default_rng(5).standard_normal(i["ndof"])
    r[i["nact"]:] = 0.0
    sp = pm.solver_params(**dict(REF_SOLVER, omega=0.5, pre=3, post=3, coarse_smoother=17, rel_tol=1e-10),
                          precision=pm.MIXED)
    # groups run as quads of lines (lat_quad_gk, 1, 2 or 4 pairs on each of
    # the 4 lines: CD3D_QUAD_G) or, with CD3D_LAT_QUADS=0, by pair class
    # (lat_pair_gk)
    out = {}
    # ("2r": the restriction's pair classes interleaved, CD3D_RESTRICT_INTER)
    for g in ("1", "2", "4", "2c", "4c", "2q1", "4q4", "2r"):
        monkeypatch.setenv("CD3D_PAIR_GROUP", g[0])
        monkeypatch.setenv("CD3D_LAT_QUADS", "0" if g.endswith("c") else "1")
        monkeypatch.setenv("CD3D_QUAD_G", g[2] if "q" in g else "2")
        monkeypatch.setenv("CD3D_RESTRICT_INTER", "1" if g.endswith("r") else "0")
        z = P.vcycle(sp, r)
        x, st = P.solve(sp)
        out[g] = (z, x, st.iterations)
    say(f"[gpu] pair groups m={m} k0={k0} levels={levels}: its {[out[g][2] for g in out]}")
    for g in ("2", "4", "2c", "4c", "2q1", "4q4", "2r"):
        assert np.array_equal(out[g][0], out["1"][0])
        assert np.array_equal(out[g][1], out["1"][1])
        assert out[g][2] == out["1"][2]


@pytest.mark.parametrize("order,m,k0,levels,pre,cycle", [
    (2, 1, 2, 5, 3, 1), (1, 1, 2, 5, 3, 1), (2, 3, 0, 4, 1, 2), (2, 1, 1, 4, 2, 0), (1, 3, 0, 3, 0, 1)])
def test_restriction_zero_sweep_matches_the_sweep(order, m, k0, levels, pre, cycle, monkeypatch):
    # the mixed V-cycle's restriction also writes the coarse level's first
    # pre-smoothing sweep from zero (omega D^-1 f, pointwise) when that
    # level's stencils are uniform, and in the lattice FGMRES the kernel that
    # writes each Krylov vector's FP32 copy writes the finest level's; the
    # sweeps' launches are skipped.  The V-cycle (coarse levels) and the
    # solve (all levels) are bitwise those with the sweeps launched
    # (CD3D_FUSE_ZERO_SWEEP=0).  Q2 64^3 cells puts a pair-form level
    # (65^3 nodes) under the fine one; V (1), W (2) and F (0) cycles; pre = 0
    # runs no such sweep
    P = pm.Problem(order, m, k0, levels, box=(1.0, 1.0, 1.0))
    P.assemble(pm.coeffs(eps=1e-3, b=(1.0, 0.5, 0.25), supg=1, f_val=1.0))
    i = P.info(levels - 1)
    r = np.random.default_rng(9).standard_normal(i["ndof"])
    r[i["nact"]:] = 0.0
    sp = pm.solver_params(**dict(REF_SOLVER, omega=0.5, pre=pre, post=3, coarse_smoother=17, rel_tol=1e-10,
                                 cycle=cycle), precision=pm.MIXED)
    out = {}
    for fz in ("1", "0"):
        monkeypatch.setenv("CD3D_FUSE_ZERO_SWEEP", fz)
        z = P.vcycle(sp, r)
        x, st = P.solve(sp, check=False)
        out[fz] = (z, x, st.iterations, st.status)
    say(f"[gpu] zero-start sweep in the restriction order={order} cells={m * 2 ** (k0 + levels - 1)} pre={pre} "
        f"cycle={cycle}: its {out['1'][2]} / {out['0'][2]}, status {out['1'][3]}")
    assert out["1"][3] == out["0"][3]
    assert np.array_equal(out["1"][0], out["0"][0])
    assert np.array_equal(out["1"][1], out["0"][1])
    assert out["1"][2] == out["0"][2]
    P.close()


@pytest.mark.parametrize("order,m,k0,levels", [(2, 1, 2, 5), (2, 3, 0, 4), (2, 11, 0, 2), (1, 1, 2, 4)])
def test_fp32_krylov_columns_match_fp64_columns(order, m, k0, levels, monkeypatch):
    # the mixed lattice FGMRES keeps its columns z_i = M^-1 v_i in FP32 (the
    # V-cycle's output, exact): the FP64 lattice operator reads them as FP32
    # and writes the column, and x += Z y converts them back.  Every value
    # is converted exactly, so the solve is bitwise the one with FP64
    # columns (CD3D_FGMRES_Z32=0).  Q1 takes no quad form and keeps FP64
    # columns in both settings
    P = pm.Problem(order, m, k0, levels, box=(1.0, 1.0, 1.0))
    P.assemble(pm.coeffs(eps=1e-3, b=(1.0, 0.5, 0.25), supg=1, f_val=1.0))
    sp = pm.solver_params(**dict(REF_SOLVER, omega=0.5, pre=3, post=3, coarse_smoother=17, rel_tol=1e-10),
                          precision=pm.MIXED)
    out = {}
    for z32 in ("1", "0"):
        monkeypatch.setenv("CD3D_FGMRES_Z32", z32)
        x, st = P.solve(sp, check=False)
        out[z32] = (x, st.iterations, st.status, st.rfinal)
    say(f"[gpu] FP32 Krylov columns order={order} cells={m * 2 ** (k0 + levels - 1)}: its {out['1'][1]} / "
        f"{out['0'][1]}, status {out['1'][2]}")
    assert out["1"][2] == out["0"][2] and out["1"][1] == out["0"][1]
    assert np.array_equal(out["1"][0], out["0"][0]) and out["1"][3] == out["0"][3]
    P.close()


@pytest.mark.parametrize("order,m,k0,levels", [(1, 1, 2, 4), (2, 1, 1, 3)])
def test_rotating_field_lattice_fgmres(order, m, k0, levels):
    # Galerkin rotating field: the CSR values are the affine class stencils,
    # so the mixed form runs FGMRES in lattice order with the FP64 lattice
    # operator; the residual reported is recomputed with the CSR matrix
    P = pm.Problem(order, m, k0, levels, box=(1.0, 1.0, 1.0))
    P.assemble(pm.coeffs(eps=1e-2, conv_kind=1, rot_c=(0.45, 0.55), c=0.5, f_val=1.0, t=0.2))
    sp = dict(REF_SOLVER, omega=0.5, pre=3, post=3, coarse_smoother=17, rel_tol=1e-10)
    x, st = P.solve(pm.solver_params(**sp, precision=pm.MIXED))
    x64, st64 = P.solve(pm.solver_params(**sp))
    err = np.linalg.norm(x - x64) / np.linalg.norm(x64)
    say(f"[gpu] rotating-field lattice FGMRES order={order}: its {st.iterations} / {st64.iterations}, "
        f"rfinal/r0 {st.rfinal / st.r0:.2e}, err vs fp64 {err:.2e}")
    assert st.status == 0 and st.rfinal <= 1.0001e-10 * st.r0
    assert err < 1e-9
    # TCD3D: mixed steps (lattice FGMRES on M + tau th1 A) against FP64 steps
    co = pm.coeffs(eps=1e-3, conv_kind=1, rot_c=(0.45, 0.55), pulse_c=(0.3, 0.4, 0.5), f_val=0.0)
    states = []
    for prec in (pm.MIXED, pm.FP64):
        Q = pm.Problem(order, m, k0, levels, box=(1.0, 1.0, 1.0))
        Q.tcd_init(co, 0.0, Q.interp_pulse(levels - 1, co))
        spt = pm.solver_params(**dict(sp, rel_tol=0.0, res_norm_min=1e-30, red_factor=1e-10), precision=prec)
        for _ in range(3):
            Q.tcd_step(co, spt, 1e-3)
        states.append(Q.tcd_state())
        Q.close()
    terr = np.linalg.norm(states[0] - states[1]) / np.linalg.norm(states[1])
    say(f"[gpu] rotating-field TCD3D mixed vs fp64 states: {terr:.2e}")
    assert terr < 1e-8


@pytest.mark.parametrize("order,m,k0,levels", [(1, 1, 2, 4), (2, 1, 1, 3)])
def test_affine_sweeps_match_per_node_values(order, m, k0, levels, monkeypatch):
    # rotating field, mixed form: on levels of more than CD3D_SMALL_NODES
    # lattice nodes the FP32 sweeps evaluate the affine class stencils node by
    # node, on the others they read per-node values filled from the same
    # formula (and the bottom levels run as one workgroup): the same bits
    co = pm.coeffs(eps=1e-3, conv_kind=1, rot_c=(0.45, 0.55), pulse_c=(0.3, 0.4, 0.5), c=0.0, f_val=0.0)
    sp = pm.solver_params(**dict(REF_SOLVER, omega=0.5, pre=3, post=3, coarse_smoother=17, rel_tol=0.0,
                                 res_norm_min=1e-30, red_factor=1e-10), precision=pm.MIXED)
    out = {}
    for bound in ("0", "1000000000"):
        monkeypatch.setenv("CD3D_SMALL_NODES", bound)
        P = pm.Problem(order, m, k0, levels, box=(1.0, 1.0, 1.0))
        u0 = P.interp_pulse(levels - 1, co)
        P.tcd_init(co, 0.0, u0)
        its = [P.tcd_step(co, sp, 1e-3)[1].iterations for _ in range(3)]
        out[bound] = (P.tcd_state(), its)
        P.close()
    say(f"[gpu] affine sweeps order={order} levels={levels}: its {out['0'][1]} / {out['1000000000'][1]}")
    assert out["0"][1] == out["1000000000"][1]
    assert np.array_equal(out["0"][0], out["1000000000"][0])


@pytest.mark.parametrize("order,m,k0", [(1, 1, 2), (2, 3, 0), (1, 1, 3), (2, 1, 1), (1, 12, 0), (2, 7, 0),
                                         (1, 16, 0)])
def test_coarse_inverse_solves_the_coarse_system(order, m, k0):
    # one level: the preconditioner is the exact coarse solve (SolveExact);
    # 27 and 125 unknowns take the single-workgroup LDS elimination, 343 the
    # blocked elimination (panels of 16 pivots, one panel row per thread),
    # 1331 / 2197 / 3375 the blocked elimination with panels of 8 pivots and
    # 2 / 3 / 4 rows per thread
    P = pm.Problem(order, m, k0, 1, box=(1.0, 0.8, 0.6))
    P.assemble(pm.coeffs(eps=1e-3, b=(1.0, -0.4, 0.3), c=0.5, supg=1, f_val=1.0))
    i = P.info(0)
    r = np.random.default_rng(5).standard_normal(i["ndof"])
    r[i["nact"]:] = 0.0
    z = P.vcycle(pm.solver_params(coarse_smoother=17), r)
    res = np.linalg.norm(P.matvec(0, z) - r) / np.linalg.norm(r)
    say(f"[gpu] coarse inverse order={order} na={i['nact']}: |A z - r| / |r| = {res:.2e}")
    assert res < 1e-12


@pytest.mark.parametrize("order,m,k0", [(1, 1, 2), (2, 1, 1), (1, 2, 1), (2, 3, 0)])
def test_rotating_field_assembly_matches_oracle(order, m, k0):
    # rotating convection, Galerkin: assembled through affine class stencils
    # on the device, cell by cell in the oracle (TCD3D.C:242-295)
    co = dict(eps=1e-2, conv_kind=1, rot_c=(0.45, 0.55), c=0.3, f_val=2.0, t=0.3)
    P = pm.Problem(order, m, k0, 2, box=(1.0, 0.8, 0.6))
    P.assemble(pm.coeffs(**co))
    H = po.Hierarchy(order, m, k0, 2, box=(1.0, 0.8, 0.6))
    H.assemble(po.coeffs(**co))
    for l in range(2):
        rp, kc, e = P.csr(l)
        rpo, kco, eo = H.csr(l)
        assert np.array_equal(rp, rpo) and np.array_equal(kc, kco)
        err = row_scaled_err(e, eo, rp)
        rerr = np.abs(P.rhs(l) - H.rhs(l)).max() / np.abs(H.rhs(l)).max()
        say(f"[gpu] rotating field order={order} m={m} level {l}: entry err {err:.2e}, rhs err {rerr:.2e}")
        assert err < ENTRY_TOL and rerr < ENTRY_TOL


def test_kernel_bench_covers_every_kernel():
    # every case of cd3d_kernel_bench runs on an assembled Q2 problem (the
    # timings bench.py reports), with the stated algorithmic bytes (flops for
    # the element kernel) per launch, and an out-of-range case is refused
    P = pm.Problem(2, 1, 1, 3)
    P.assemble(pm.coeffs(eps=1e-3, b=(1.0, 0.5, 0.25), supg=1, f_val=1.0))
    l = 2
    i = P.info(l)
    N, nnz = i["ndof"], i["nnz"]
    rp, kc, e0 = P.csr(l)
    nbytes = {}
    for which in range(10):
        ms, b = P.kernel_bench(which, l, 3)
        assert ms > 0 and b > 0
        nbytes[which] = b
    say(f"[gpu] kernel bench bytes/flops {nbytes}")
    assert nbytes[0] == 12 * nnz + 20 * N
    assert nbytes[2] == 16 * N and nbytes[6] == 16 * N and nbytes[7] == 32 * N
    assert nbytes[8] == i["ncell"] * 64 * 8 * 27 * 27
    assert nbytes[9] == (2 * 10 + 5) * 8 * N
    # the element kernel's timing leaves the level's matrix as it was
    assert np.array_equal(P.csr(l)[2], e0)
    with pytest.raises(pm.CD3DError):  # the Gram-Schmidt step runs on the finest level only
        P.kernel_bench(9, 1, 1)
    with pytest.raises(pm.CD3DError):
        P.kernel_bench(10, l, 1)


# multicolour SSOR smoother (13): the HIP V-cycle and FGMRES against the
# oracle's statement of it (itself pinned against its numpy definition), on a
# moderate case and on a convection-dominated Q2 problem (eps 2.7e-6 on the
# 5^3-brick mesh, 20^3 cells) where the damped Jacobi smoother stalls near 4e-3
MC_CASES = {
    "m2_q1": (1, 2, 1, 3, dict(eps=1e-2, b=(0.3, 0.4, -0.5), f_val=1.0, supg=1)),
    "m2_q2": (2, 2, 0, 3, dict(eps=1e-3, b=(0.6, 0.5, -0.4), f_val=1.0, supg=1)),
    "m5_q2_pe600": (2, 5, 0, 3, dict(eps=2.6659e-06, b=(0.039539, -0.036416, -0.032403), f_val=1.0, supg=1)),
}
MC_SOLVER = dict(smoother=13, omega=1.0, pre=2, post=2, coarse_smoother=17, restart=20, maxit=400)


@pytest.mark.parametrize("name", list(MC_CASES))
def test_multicolour_ssor_vcycle_and_fgmres_match_oracle(name):
    order, m, k0, levels, co = MC_CASES[name]
    say(f"[gpu] multicolour SSOR {name}")
    P = pm.Problem(order, m, k0, levels, box=(1.0, 1.0, 1.0))
    P.assemble(pm.coeffs(**co))
    H = po.Hierarchy(order, m, k0, levels, box=(1.0, 1.0, 1.0))
    H.assemble(po.coeffs(**co))
    n = H.info(levels - 1)["ndof"]
    na = H.info(levels - 1)["nact"]
    # the oracle's solve first: its V-cycle works in the finest level's rhs array
    b = H.rhs(levels - 1)
    _, r0 = H.defect(levels - 1, np.zeros_like(b), b)
    xo, its_o, _ = H.fgmres(po.solver(**dict(MC_SOLVER, res_norm_min=1e-10 * r0, red_factor=0.0)), b)
    r = np.random.default_rng(11).standard_normal(n)
    r[na:] = 0.0
    z = P.vcycle(pm.solver_params(**MC_SOLVER), r)
    zo = H.vcycle(po.solver(**MC_SOLVER), r)
    verr = np.abs(z - zo).max() / np.abs(zo).max()
    for prec, tol in ((pm.FP64, SOL_TOL), (pm.MIXED, MIXED_TOL)):
        x, st = P.solve(pm.solver_params(**MC_SOLVER, rel_tol=1e-10, precision=prec))
        err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
        say(f"[gpu] {name} precision {prec}: vcycle err {verr:.2e}, its {st.iterations} (oracle {its_o}), "
            f"sol err {err:.2e}, rel res {st.rfinal / st.r0:.2e}")
        assert st.status == 0 and st.rfinal <= 1.0001e-10 * st.r0
        assert err < tol
        if prec == pm.FP64:
            assert st.iterations == its_o
    assert verr < 1e-12
    P.close()


@pytest.mark.parametrize("name", list(CD_CASES))
def test_diagnostics_and_output_match_reference(name, golden, tmp_path):
    # on the device: the nodal interpolant of the sine of Laplace.h and the
    # L2/H1 errors (TFEFunction3D::Interpolate / GetErrors) of the
    # reference's solution and interpolant; the errors of the library's own
    # solution; the VTK file (TOutput3D::WriteVtk: same points, hexahedra and
    # point data, vertices matched by coordinates) and the binary matrix
    # (TSquareMatrix::Write: same header, size, pattern; entries 1e-12)
    from tests.golden.make_ref_fixtures import parse_vtk
    c = CD_CASES[name]
    d = golden(name)
    P = problem(name, golden)
    lf = c["levels"] - 1
    ui = P.interpolate(lf, P.FN_SINE)
    ierr = np.abs(ui - d["interp_sine"]).max()
    errs = []
    for u, ref in ((d["sol"], d["errors_sol"]), (d["interp_sine"], d["errors_interp"])):
        e = P.errors(lf, P.FN_SINE, u)
        errs.append(np.abs(e[:3] - ref).max() / np.abs(ref).max())
    x, st = P.solve(pm.solver_params(**case_solver(name), rel_tol=1e-10))
    own = P.errors(lf, P.FN_SINE)  # the solution just computed, on the device
    own_err = np.abs(own[:2] - d["errors_sol"][:2]).max() / np.abs(d["errors_sol"][:2]).max()
    say(f"[gpu] {name} diagnostics: interp {ierr:.1e}, errors {errs[0]:.1e} / {errs[1]:.1e}, own solve {own_err:.1e}")
    # (deformed meshes: the refined vertices agree with the reference's to an ulp or two)
    assert ierr <= (4e-15 if "geo" in c else 1e-15) and max(errs) <= 1e-12 and own_err <= 1e-6
    P.write_vtk(lf, d["sol"], tmp_path / "u.vtk")
    v = parse_vtk(str(tmp_path / "u.vtk"))
    assert v["vtk_text_head"] == d["vtk_text_head"]
    key = lambda pts: [tuple(np.round(q, 5)) for q in pts]
    ref_idx = {k: i for i, k in enumerate(key(d["vtk_points"]))}
    perm = np.array([ref_idx[k] for k in key(v["vtk_points"])])  # our point -> reference point
    assert len(set(perm.tolist())) == len(perm) == len(d["vtk_points"])
    assert np.array_equal(perm[v["vtk_cells"]], d["vtk_cells"])
    assert np.array_equal(v["vtk_cell_types"], d["vtk_cell_types"])
    assert np.array_equal(d["vtk_values"][perm], v["vtk_values"])
    assert np.array_equal(v["vtk_region"], d["vtk_region"])
    P.write_matrix(lf, tmp_path / "A.bin")
    raw = np.fromfile(tmp_path / "A.bin", dtype=np.uint8)
    hdr = raw[:12].view(np.int32)
    assert np.array_equal(hdr, d["matrix_header"]) and len(raw) == int(d["matrix_bytes"])
    n, nnz = int(hdr[0]), int(hdr[2])
    rp = raw[12:12 + 4 * (n + 1)].view(np.int32)
    kc = raw[12 + 4 * (n + 1):12 + 4 * (n + 1) + 4 * nnz].view(np.int32)
    ent = raw[12 + 4 * (n + 1) + 4 * nnz:].view(np.float64)
    assert np.array_equal(rp, d[f"L{lf}_rowptr"]) and np.array_equal(kc, d[f"L{lf}_kcol"])
    assert row_scaled_err(ent, d[f"L{lf}_entries"], rp) < ENTRY_TOL


@pytest.mark.parametrize("name", [n for n in CD_CASES if "qp" in CD_CASES[n] or "geo" in CD_CASES[n]])
def test_quad_points_match_oracle(name):
    # the points handed to the caller's coefficient function are the
    # oracle's (which restates GetOrig's affine map) to rounding
    c = CD_CASES[name]
    P = make_problem(c)
    H = oracle_hier(c)
    for l in range(c["levels"]):
        assert np.abs(P.quad_points(l) - H.quad_points(l)).max() <= 1e-15
    P.close()


@pytest.mark.parametrize("name", ["unit_q1_var_supg", "m2_q2_supg_rot", "m3w_q2_supg", "m2w_q1_var_supg"])
def test_two_pass_assembly_equals_the_scatter(name, golden, monkeypatch):
    # the cell-by-cell assembly in two passes (element rows written out, CSR
    # entries gathered in colour order) against the colour-by-colour scatter:
    # the same additions in the same order, so the same bits
    c = CD_CASES[name]
    monkeypatch.setenv("CD3D_CELLWISE_ASSEMBLY", "1")
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("CD3D_ASM_SCATTER", mode)
        P = make_problem(c)
        P.assemble(pm.coeffs(**c["co"]))
        out[mode] = [(P.csr(l)[2], P.rhs(l)) for l in range(c["levels"])]
        P.close()
    say(f"[gpu] two-pass assembly {name}: {len(out['0'])} levels compared")
    for (e0, r0), (e1, r1) in zip(out["0"], out["1"]):
        assert np.array_equal(e0, e1) and np.array_equal(r0, r1)

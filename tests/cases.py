"""The golden cases written by tests/golden/make_ref_fixtures.py, described the
way both the oracle and the HIP library take them.

Each case: order, macro bricks per side m (default 1: ParMooN's Wuerfel
cube; m >= 2: the unit cube split into m^3 bricks by
oracle/data/make_box_geo.py), coarsest refinement k0 (n0 = m 2^k0 cells per
side), number of levels, the problem coefficients (the .dat P-parameters of
the reference run, decoded) and, where the run changed them, the solver
parameters that differ from REF_SOLVER."""

REF_BOX = (1.0, 0.346, 0.346)
UNIT_BOX = (1.0, 1.0, 1.0)

CD_CASES = {
    "q1_l2_poisson": dict(order=1, k0=2, levels=2,
                          co=dict(eps=1.0, f_kind=1)),
    "q1_c1": dict(order=1, k0=2, levels=3,
                  co=dict(eps=1.0, f_kind=1)),
    "q2_l2_supg": dict(order=2, k0=1, levels=2,
                       co=dict(eps=1e-3, b=(1.0, 0.5, 0.25), f_val=1.0, supg=1)),
    "q1_l3_supg": dict(order=1, k0=2, levels=3,
                       co=dict(eps=1e-6, b=(-0.3, 0.7, 0.4), c=0.5, f_val=1.0, supg=1)),
    "q2_l2_galerkin_conv": dict(order=2, k0=1, levels=2,
                                co=dict(eps=0.1, b=(1.0, 0.5, 0.25), c=1.0, f_kind=1)),
    "unit_c1": dict(order=1, k0=2, levels=3, box=UNIT_BOX,
                    co=dict(eps=1.0, f_kind=1)),
    "unit_c2_n8": dict(order=2, k0=1, levels=3, box=UNIT_BOX,
                       co=dict(eps=1e-3, b=(1.0, 0.5, 0.25), f_val=1.0, supg=1)),
    # macro-brick meshes (the batched and serving workloads' meshes)
    "m2_q1_supg": dict(order=1, m=2, k0=1, levels=3, box=UNIT_BOX,
                       co=dict(eps=1e-6, b=(0.48, -0.36, 0.64), f_val=1.0, supg=1)),
    "m3_q1_supg": dict(order=1, m=3, k0=0, levels=3, box=UNIT_BOX,
                       co=dict(eps=1e-2, b=(-0.5, 0.2, 0.6), f_val=1.0, supg=1)),
    "m5_q1_supg": dict(order=1, m=5, k0=0, levels=2, box=UNIT_BOX,
                       co=dict(eps=3e-4, b=(0.1, 0.7, -0.2), f_val=1.0, supg=1)),
    "m2_q2_supg": dict(order=2, m=2, k0=0, levels=3, box=UNIT_BOX,
                       co=dict(eps=1e-3, b=(0.6, 0.5, -0.4), f_val=1.0, supg=1)),
    "m3_q2_supg": dict(order=2, m=3, k0=0, levels=2, box=UNIT_BOX,
                       co=dict(eps=1e-5, b=(-0.2, -0.3, 0.5), f_val=1.0, supg=1)),
    # SUPG with the rotating field (per-cell convection and SUPG delta)
    "unit_q1_supg_rot": dict(order=1, k0=2, levels=3, box=UNIT_BOX, solver=dict(omega=0.5, pre=3, post=3),
                             co=dict(eps=1e-3, conv_kind=1, rot_c=(0.45, 0.55), f_val=1.0, supg=1)),
    "m2_q2_supg_rot": dict(order=2, m=2, k0=0, levels=3, box=UNIT_BOX, solver=dict(omega=0.5, pre=3, post=3),
                           co=dict(eps=1e-4, conv_kind=1, rot_c=(0.45, 0.55), c=0.2, f_val=1.0, supg=1)),
    # SOR / SSOR smoothers in the reference's row order (SC_SMOOTHER_SCALAR 2, 3)
    "m2_q1_ssor": dict(order=1, m=2, k0=0, levels=3, box=UNIT_BOX, solver=dict(smoother=3, pre=3, post=3, omega=1.0),
                       co=dict(eps=1e-2, b=(0.3, 0.4, -0.5), f_val=1.0, supg=1)),
    "unit_q2_sor": dict(order=2, k0=1, levels=3, box=UNIT_BOX, solver=dict(smoother=2, coarse_smoother=2, omega=1.0),
                        co=dict(eps=1e-3, b=(1.0, 0.5, 0.25), f_val=1.0, supg=1)),
    "m2_q2_ssor_w12": dict(order=2, m=2, k0=0, levels=2, box=UNIT_BOX, solver=dict(smoother=3, omega=1.2),
                           co=dict(eps=1e-2, b=(-0.4, 0.6, 0.3), c=0.5, f_val=1.0, supg=1)),
    # W- and F-cycle preconditioners
    "m2_q1_wcycle": dict(order=1, m=2, k0=0, levels=4, box=UNIT_BOX, solver=dict(cycle=2),
                         co=dict(eps=1e-2, b=(0.3, 0.4, -0.5), f_val=1.0, supg=1)),
    "m2_q1_fcycle": dict(order=1, m=2, k0=0, levels=4, box=UNIT_BOX, solver=dict(cycle=0),
                         co=dict(eps=1e-2, b=(0.3, 0.4, -0.5), f_val=1.0, supg=1)),
    # coefficients varying in space, handed in per quadrature point
    # (qp_coeffs below evaluates the reference run's CoeffFct3D)
    "unit_q1_var_gal": dict(order=1, k0=2, levels=3, box=UNIT_BOX, qp=dict(eps=0.05, b=(0.8, -0.6, 0.5), c=0.3, f=1.0),
                            co=dict(eps=0.05, f_val=1.0)),
    "unit_q1_var_supg": dict(order=1, k0=2, levels=3, box=UNIT_BOX, solver=dict(omega=0.5, pre=3, post=3),
                             qp=dict(eps=1e-3, b=(0.8, -0.6, 0.5), c=0.3, f=1.0), co=dict(eps=1e-3, f_val=1.0, supg=1)),
    "m2_q2_var_supg": dict(order=2, m=2, k0=0, levels=3, box=UNIT_BOX, solver=dict(omega=0.5, pre=3, post=3),
                           qp=dict(eps=1e-3, b=(0.8, -0.6, 0.5), c=0.3, f=1.0), co=dict(eps=1e-3, f_val=1.0, supg=1)),
    # general hexahedra: the interior macro vertices moved (geo kind "m<m>w",
    # oracle/data/make_box_geo.py), every cell mapped trilinearly
    "m3w_q1_poisson": dict(order=1, m=3, k0=0, levels=3, box=UNIT_BOX, geo="m3w",
                           co=dict(eps=1.0, f_kind=1)),
    "m3w_q2_supg": dict(order=2, m=3, k0=0, levels=2, box=UNIT_BOX, geo="m3w", solver=dict(omega=0.5, pre=3, post=3),
                        co=dict(eps=1e-3, b=(1.0, 0.5, 0.25), f_val=1.0, supg=1)),
    "m2w_q1_var_supg": dict(order=1, m=2, k0=1, levels=3, box=UNIT_BOX, geo="m2w", solver=dict(omega=0.5, pre=3, post=3),
                            qp=dict(eps=1e-3, b=(0.8, -0.6, 0.5), c=0.3, f=1.0), co=dict(eps=1e-3, f_val=1.0, supg=1)),
    "m2w_q2_supg_rot": dict(order=2, m=2, k0=0, levels=3, box=UNIT_BOX, geo="m2w", solver=dict(omega=0.5, pre=3, post=3),
                            co=dict(eps=1e-4, conv_kind=1, rot_c=(0.45, 0.55), c=0.2, f_val=1.0, supg=1)),
    # the benchmark's smoother: damped Jacobi omega 0.5, 3 + 3 sweeps
    "unit_c2_n8_bench": dict(order=2, k0=1, levels=3, box=UNIT_BOX, solver=dict(omega=0.5, pre=3, post=3),
                             co=dict(eps=1e-3, b=(1.0, 0.5, 0.25), f_val=1.0, supg=1)),
}

TCD_CASES = {
    "q1_tcd": dict(order=1, k0=2, levels=2, tau=1e-3, nsteps=3, box=REF_BOX,
                   co=dict(eps=1e-4, conv_kind=1, rot_c=(0.5, 0.173), pulse_c=(0.3, 0.173, 0.173), f_val=0.0)),
    "m2_q1_tcd": dict(order=1, m=2, k0=1, levels=2, tau=1e-3, nsteps=3, box=UNIT_BOX,
                      co=dict(eps=1e-4, conv_kind=1, rot_c=(0.5, 0.5), pulse_c=(0.3, 0.5, 0.5), f_val=0.0)),
    # SOR 2 + 2 (omega 1), as the reference's ReadIn/tcd3d.dat
    "m2_q1_tcd_sor": dict(solver=dict(smoother=2, omega=1.0), order=1, m=2, k0=1, levels=2, tau=1e-3, nsteps=3,
                          box=UNIT_BOX,
                          co=dict(eps=1e-4, conv_kind=1, rot_c=(0.5, 0.5), pulse_c=(0.3, 0.5, 0.5), f_val=0.0)),
    "m2_q2_tcd": dict(order=2, m=2, k0=0, levels=2, tau=1e-3, nsteps=3, box=UNIT_BOX,
                      co=dict(eps=1e-4, conv_kind=1, rot_c=(0.5, 0.5), pulse_c=(0.3, 0.5, 0.5), f_val=0.0)),
    # SUPG-stabilised Crank-Nicolson (TCD3D.C:48,102): rotating field with
    # f = 1, Q1 and Q2, and a constant field with reaction
    "m2_q1_tcd_supg": dict(solver=dict(omega=0.5, pre=3, post=3), order=1, m=2, k0=1, levels=2, tau=1e-3, nsteps=3, box=UNIT_BOX,
                           co=dict(eps=1e-4, conv_kind=1, rot_c=(0.5, 0.5), pulse_c=(0.3, 0.5, 0.5), f_val=1.0,
                                   supg=1)),
    "m2_q2_tcd_supg": dict(solver=dict(omega=0.5, pre=3, post=3), order=2, m=2, k0=0, levels=2, tau=1e-3, nsteps=3, box=UNIT_BOX,
                           co=dict(eps=1e-4, conv_kind=1, rot_c=(0.5, 0.5), pulse_c=(0.3, 0.5, 0.5), f_val=1.0,
                                   supg=1)),
    "m2w_q2_tcd_supg": dict(solver=dict(omega=0.5, pre=3, post=3), order=2, m=2, k0=0, levels=2, tau=1e-3, nsteps=3,
                            box=UNIT_BOX, geo="m2w",
                            co=dict(eps=1e-4, conv_kind=1, rot_c=(0.5, 0.5), pulse_c=(0.3, 0.5, 0.5), f_val=1.0,
                                    supg=1)),
    "unit_q1_tcd_supg_const": dict(solver=dict(omega=0.5, pre=3, post=3), order=1, k0=1, levels=3, tau=1e-2, nsteps=3, box=UNIT_BOX,
                                   co=dict(eps=1e-3, b=(0.6, -0.3, 0.4), c=0.1, pulse_c=(0.3, 0.5, 0.5),
                                           f_val=1.0, supg=1)),
}
TCD_CASE = dict(TCD_CASES["q1_tcd"], name="q1_tcd")

# solver of every reference run: Jacobi(0.8) 2+2, V-cycle, SSOR coarse smoother
# (10 its / 0.1 reduction), FGMRES(20)
REF_SOLVER = dict(smoother=1, pre=2, post=2, omega=0.8, coarse_smoother=3, coarse_maxit=10,
                  coarse_red=0.1, gmg_damp=1.0, restart=20, maxit=200)


def case_solver(name):
    """REF_SOLVER with the case's own solver changes."""
    return dict(REF_SOLVER, **CD_CASES[name].get("solver", {}))


def case_box(c):
    return c.get("box", REF_BOX)


def _box_geo_module():
    import os
    import sys
    d = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "data")
    if d not in sys.path:
        sys.path.insert(0, d)
    import make_box_geo
    return make_box_geo


def geo_text(c):
    """GEO file text of a deformed-mesh case (what the reference run read)."""
    return _box_geo_module().geo_for_kind(c["geo"])[1]


def case_vertices(c):
    """(m+1)^3 x 3 macro vertex coordinates of a deformed-mesh case, None for a box."""
    if "geo" not in c:
        return None
    return _box_geo_module().macro_vertices(geo_text(c))


def oracle_hier(c, levels=None):
    """The oracle hierarchy of a case (box or mapped)."""
    from oracle import pyoracle as po
    lv = c["levels"] if levels is None else levels
    return po.Hierarchy(c["order"], c.get("m", 1), c["k0"], lv, box=case_box(c), vertices=case_vertices(c))


def tcd_solver(c):
    """Solver keywords of a TCD case's reference run: REF_SOLVER with the
    case's changes; Galerkin runs stop on the reference's restart-relative
    1e-10 (TSystemTCD3D::Solve), SUPG runs on 1e-10 of each step's initial
    residual."""
    kw = dict(REF_SOLVER, **c.get("solver", {}))
    if c["co"].get("supg"):
        return dict(kw, rel_tol=1e-10, res_norm_min=1e-30, red_factor=0.0)
    return dict(kw, rel_tol=0.0, res_norm_min=1e-30, red_factor=1e-10)


def qp_coeffs(c, xyz):
    """The coefficients of a per-point case at the quadrature points xyz
    (n_cells x n_q x 3): eps (1 + xyz), b = (b1 sin(pi y), b2 cos(pi x), b3 z),
    c x, f (1 + x) -- the CoeffFct3D of the reference runs (ref_driver.C, P7 2)."""
    import numpy as np
    q = c["qp"]
    x, y, z = xyz[..., 0], xyz[..., 1], xyz[..., 2]
    out = np.empty(xyz.shape[:-1] + (6,))
    out[..., 0] = q["eps"] * (1 + x * y * z)
    out[..., 1] = q["b"][0] * np.sin(np.pi * y)
    out[..., 2] = q["b"][1] * np.cos(np.pi * x)
    out[..., 3] = q["b"][2] * z
    out[..., 4] = q["c"] * x
    out[..., 5] = q["f"] * (1 + x)
    return out

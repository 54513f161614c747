"""The golden cases written by tests/golden/make_ref_fixtures.py, described the
way both the oracle and the HIP library take them.

Each case: order, coarsest refinement k0 (n0 = 2^k0 cells per side), number of
levels, and the problem coefficients (the .dat P-parameters of the reference
run, decoded)."""

REF_BOX = (1.0, 0.346, 0.346)

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
    "unit_c1": dict(order=1, k0=2, levels=3, box=(1.0, 1.0, 1.0),
                    co=dict(eps=1.0, f_kind=1)),
    "unit_c2_n8": dict(order=2, k0=1, levels=3, box=(1.0, 1.0, 1.0),
                       co=dict(eps=1e-3, b=(1.0, 0.5, 0.25), f_val=1.0, supg=1)),
}

TCD_CASE = dict(name="q1_tcd", order=1, k0=2, levels=2, tau=1e-3, nsteps=3,
                co=dict(eps=1e-4, conv_kind=1, rot_c=(0.5, 0.173), pulse_c=(0.3, 0.173, 0.173), f_val=0.0))

# solver of every reference run: Jacobi(0.8) 2+2, V-cycle, SSOR coarse smoother
# (10 its / 0.1 reduction), FGMRES(20)
REF_SOLVER = dict(smoother=1, pre=2, post=2, omega=0.8, coarse_smoother=3, coarse_maxit=10,
                  coarse_red=0.1, gmg_damp=1.0, restart=20, maxit=200)

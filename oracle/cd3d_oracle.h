/* CPU restatement of ParMooN's CD3D/TCD3D hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This oracle is the parity checker for the HIP implementation.  Only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it; product code never links, loads or calls it.
 *
 * Every function restates one reference routine (file:line in its comment).
 * Pinned by: tests/test_oracle_golden.py (reference fixtures written by the
 * reference build, tests/golden/ (npz files)) and tests/test_oracle_math.py
 * (closed forms, invariants).  See the per-function notes for what pins what.
 */
#ifndef CD3D_ORACLE_H
#define CD3D_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* problem coefficients: eps, convection, reaction, source, stabilisation */
typedef struct {
  double eps;      /* diffusion */
  double b[3];     /* constant convection (conv_kind 0) */
  double c;        /* reaction */
  int conv_kind;   /* 0: constant b; 1: rotating b = (-(y-rc1), x-rc0, 0.1 sin(2 pi t)) */
  double rot_c[2]; /* rotation centre (rc0, rc1) of the rotating field */
  double pulse_c[3]; /* centre of the Gaussian pulse initial condition (sigma 0.05) */
  int f_kind;      /* 0: f = f_val; 1: manufactured u = sin(pi x) sin(pi y) sin(pi z) */
  double f_val;
  int supg;        /* 0: Galerkin; 1: SDFEM_TYPE 2 (CELL_MEASURE 4, DELTA0 1, degree = order) */
  double t;        /* time at which the coefficients are evaluated */
} orc_coeffs;

typedef struct {
  int smoother;          /* 1 Jacobi, 2 SOR, 3 SSOR (SC_SMOOTHER_SCALAR); 13 multicolour SSOR (this build's own) */
  int pre, post;         /* SC_PRE/POST_SMOOTH_SCALAR */
  double omega;          /* SC_SMOOTH_DAMP_FACTOR_SCALAR */
  int coarse_smoother;   /* 1 Jacobi, 2 SOR, 3 SSOR, 17 exact (SC_COARSE_SMOOTHER_SCALAR) */
  int coarse_maxit;      /* SC_COARSE_MAXIT_SCALAR */
  double coarse_red;     /* SC_COARSE_RED_FACTOR_SCALAR */
  double gmg_damp;       /* SC_GMG_DAMP_FACTOR_SCALAR */
  int restart;           /* SC_GMRES_RESTART */
  int maxit;             /* SC_LIN_MAXIT_SCALAR */
  double res_norm_min;   /* SC_LIN_RES_NORM_MIN_SCALAR (absolute) */
  double red_factor;     /* SC_LIN_RED_FACTOR_SCALAR (relative to the restart cycle start) */
  double div_factor;     /* SC_DIV_FACTOR */
  int cycle;             /* SC_MG_CYCLE_SCALAR: 1 V, 2 W, 0 F */
  double rel_tol;        /* > 0: stop at |b - A x| <= rel_tol |b - A x0| of this solve (the reference's
                            SC_LIN_RES_NORM_MIN_SCALAR set to rel_tol |r0| before the solve, red_factor 0) */
} orc_solver;

typedef struct orc_hier orc_hier;

/* Hierarchy of `levels` uniformly refined meshes of the box
 * [0,box0] x [0,box1] x [0,box2], order 1 or 2.  Coarsest level: m^3 macro
 * cells refined k0 times (n0 = m * 2^k0 cells per side); m = 1 is ParMooN's
 * Wuerfel cube (which the reference reads as the box (1, 0.346, 0.346)). */
orc_hier *orc_create(int order, int m, int k0, int levels, const double *box);
/* The same hierarchy over general hexahedra: the m^3 macro cells (x fastest,
 * identity local frames) have the given vertices ((m+1)^3 x 3, x fastest),
 * refinement places new vertices at edge, face and cell means, and every cell
 * is mapped trilinearly (THexaTrilinear). */
orc_hier *orc_create_mapped(int order, int m, int k0, int levels, const double *macro_xyz);
void orc_destroy(orc_hier *h);

/* sizes of level l: returns n_dof; fills n_active, nnz, n_cells */
int orc_level_info(const orc_hier *h, int l, int *n_active, long *nnz, int *n_cells);
/* copies of the level data (caller-allocated) */
void orc_get_cells(const orc_hier *h, int l, double *xyz /* n_cells*8*3 */);
void orc_get_gn(const orc_hier *h, int l, int *gn /* n_cells*nb */);
void orc_get_csr(const orc_hier *h, int l, int *rowptr, int *kcol, double *entries);
void orc_get_rhs(const orc_hier *h, int l, double *rhs);
/* lattice node of every DOF (node = X + nn (Y + nn Z), nn = order n + 1) */
void orc_get_dof_nodes(const orc_hier *h, int l, int *node);
void orc_set_entries(orc_hier *h, int l, const double *entries);

/* quadrature points of level l (n_cells * n_q * 3, cell order, Gauss rule
 * order) and per-point coefficients for it (n_cells * n_q * 6: eps, b1, b2,
 * b3, c, f -- what a CoeffFct3D returns there; NULL: the descriptor again) */
void orc_quad_points(const orc_hier *h, int l, double *xyz);
void orc_set_coeffs_qp(orc_hier *h, int l, const double *qp);

/* assemble A and rhs on every level (SystemCD3D::Assemble) */
void orc_assemble(orc_hier *h, const orc_coeffs *co);
/* assemble only level l into caller arrays: what = 0 conv-diff (A), 1 mass (M) */
void orc_assemble_level(const orc_hier *h, int l, const orc_coeffs *co, int what,
                        double *entries, double *rhs);

/* one preconditioner application z = V(r) (TMultiGridScaIte::Iterate) */
void orc_vcycle(orc_hier *h, const orc_solver *s, const double *r, double *z);
/* FGMRES + GMG on the finest level; returns the iteration count and writes
 * the residual history (|s_{i+1}| per iteration, up to maxit entries) */
int orc_fgmres(orc_hier *h, const orc_solver *s, const double *rhs, double *sol, double *hist);
/* y = A x on level l (MatVect) and d = f - A x (ScalarDefect), returns |d| */
void orc_matvec(const orc_hier *h, int l, const double *x, double *y);
double orc_defect(const orc_hier *h, int l, const double *x, const double *f, double *d);
/* prolongation / restriction between level l-1 and l (MGComponents3D.C) */
void orc_prolongate(const orc_hier *h, int l, const double *coarse, double *fine);
void orc_restrict(const orc_hier *h, int l, const double *fine, double *coarse);

/* Crank-Nicolson TCD3D run (SystemTCD3D + TCD3D_ParMooN time loop):
 * u (finest level) in: initial state, out: state after nsteps; writes the
 * states after each step into states (nsteps * N) when non-NULL */
int orc_tcd_run(orc_hier *h, const orc_coeffs *co, const orc_solver *s, double t0, double tau,
                int nsteps, double *u, double *states);
/* nodal interpolation of the Gaussian pulse initial condition on level l */
void orc_interp_pulse(const orc_hier *h, int l, const orc_coeffs *co, double *u);

/* known functions: 0 zero, 1 sin(pi x) sin(pi y) sin(pi z), 2 the Gaussian pulse at co->pulse_c,
 * 3 x + 2y + 3z */
/* nodal interpolation on level l (TFEFunction3D::Interpolate) */
void orc_interpolate(const orc_hier *h, int l, int fn, const orc_coeffs *co, double *u);
/* TFEFunction3D::GetErrors with L2H1Errors: err[0] L2, err[1] H1 semi-norm of
 * u - fn, err[2] max |u - fn| at the last cell's quadrature points, err[3]
 * the same over every cell */
void orc_errors(const orc_hier *h, int l, const double *u, int fn, const orc_coeffs *co, double *err);

/* one SOR (kind 2) or SSOR (kind 3) sweep in row order on level l: x in/out, f the rhs */
void orc_gs_sweep(orc_hier *h, int l, int kind, double omega, const double *f, double *x);
/* one multicolour SSOR sweep (smoother 13) on level l: x in/out, f the rhs */
void orc_mcssor_sweep(orc_hier *h, int l, double omega, const double *f, double *x);

/* SUPG parameter (Compute_SDFEM_delta, SDFEM_TYPE 2) for tests */
double orc_sdfem_delta(double hconv, double eps, const double *b, int order);
/* mesh size in convection direction for an axis-aligned brick with sides h */
double orc_hconv_brick(const double *h, const double *b);

#ifdef __cplusplus
}
#endif
#endif

/* cd3d.h -- C ABI of libcd3d_hip.so, the MI355X implementation of ParMooN's
 * 3D scalar convection-diffusion hot path (CD3D / TCD3D): Q1/Q2 (SUPG)
 * assembly into CSR, CSR SpMV, the geometric multigrid V-cycle and the
 * FGMRES iteration that drives it.
 *
 * Plain C types only.  Every call returns a cd3d_status; no call exits the
 * process.  A handle owns one multilevel problem resident in device memory
 * (mesh, DOF maps, CSR matrices of every level, transfer operators, Krylov
 * workspace); calls on one handle are serialised by the handle's mutex, calls
 * on different handles may run concurrently from different threads.
 *
 * Pointer arguments named *_host are host memory; *_dev are device memory
 * (hipMalloc'd, or obtained from cd3d_device_ptr) and are used on the
 * handle's stream.
 */
#ifndef CD3D_H
#define CD3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  CD3D_OK = 0,
  CD3D_ERR_ARG = 1,          /* invalid argument (ParMooN: OutPut + exit(4711)) */
  CD3D_ERR_HIP = 2,          /* HIP runtime error */
  CD3D_ERR_NOMEM = 3,        /* device or host allocation failed / budget exceeded */
  CD3D_NOT_CONVERGED = 4,    /* FGMRES reached maxit (ParMooN returns maxit) */
  CD3D_DIVERGED = 5,         /* res > div_factor * res0 (ParMooN: exit(4711), FgmresIte.C:509) */
  CD3D_NONFINITE = 6,        /* NaN/Inf in data or residual */
  CD3D_ERR_STATE = 7         /* call out of order (e.g. solve before assemble) */
} cd3d_status;

typedef struct cd3d_problem cd3d_problem; /* opaque handle */

/* Mesh hierarchy: box [0,box[0]] x [0,box[1]] x [0,box[2]] split into m^3
 * macro bricks, refined k0 times for the coarsest level, `levels` levels in
 * total (level l has n_l = m 2^(k0+l) cells per side).  m = 1 is ParMooN's
 * data/Wuerfel.GEO refined by RegRefineAll (Domain.C:1037); order 1 (Q1) or
 * 2 (Q2) conforming elements, Dirichlet boundary on every face. */
typedef struct {
  int order;
  int m;
  int k0;
  int levels;
  double box[3];
} cd3d_mesh_desc;

/* Problem coefficients (replaces the CoeffFct3D callback, Constants.h:123,
 * which cannot run on the device): coeff[0] = eps, coeff[1..3] = b,
 * coeff[4] = c, coeff[5] = f, coeff[6] = |b| at every quadrature point. */
typedef struct {
  double eps;
  double b[3];          /* conv_kind 0: constant convection */
  double c;             /* reaction */
  int conv_kind;        /* 0 constant; 1 rotating b = (-(y-rot_c[1]), x-rot_c[0], 0.1 sin(2 pi t)) */
  double rot_c[2];
  double pulse_c[3];    /* centre of the Gaussian pulse initial condition (sigma 0.05) */
  int f_kind;           /* 0 f = f_val; 1 manufactured u = sin(pi x) sin(pi y) sin(pi z) */
  double f_val;
  int supg;             /* 0 Galerkin (DISCTYPE 1); 1 SDFEM_TYPE 2 (DISCTYPE 2, CELL_MEASURE 4, DELTA0 1) */
  double t;             /* time at which the coefficients are evaluated */
} cd3d_coeffs;

typedef enum { CD3D_FP64 = 0, CD3D_MIXED = 1 } cd3d_precision;

/* Solver parameters: the ParamDB SC_* knobs of the GMG/FGMRES path
 * (Database.C:648-757, FgmresIte.C:71-85, MultiGrid3D.C). */
typedef struct {
  int smoother;         /* SC_SMOOTHER_SCALAR: 1 damped Jacobi (MGLevel3D.C:443); 2 SOR and 3 SSOR in the
                           reference's row order (MGLevel3D.C:254-441: sequential by nature, run as
                           wavefronts of independent rows by one workgroup; FP64 in both precision
                           forms); 13 multicolour SSOR (no reference counterpart: relaxed Gauss-Seidel
                           over node colours, forward then backward, for convection-dominated problems;
                           FP64 in both precision forms) */
  int pre, post;        /* SC_PRE_SMOOTH_SCALAR, SC_POST_SMOOTH_SCALAR */
  double omega;         /* SC_SMOOTH_DAMP_FACTOR_SCALAR */
  int coarse_smoother;  /* SC_COARSE_SMOOTHER_SCALAR: 17 exact (default), 3 SSOR, 2 SOR, 1 Jacobi */
  int coarse_maxit;     /* SC_COARSE_MAXIT_SCALAR */
  double coarse_red;    /* SC_COARSE_RED_FACTOR_SCALAR */
  double gmg_damp;      /* SC_GMG_DAMP_FACTOR_SCALAR */
  int restart;          /* SC_GMRES_RESTART */
  int maxit;            /* SC_LIN_MAXIT_SCALAR */
  double res_norm_min;  /* SC_LIN_RES_NORM_MIN_SCALAR (absolute) */
  double red_factor;    /* SC_LIN_RED_FACTOR_SCALAR (relative to the restart-cycle start) */
  double div_factor;    /* SC_DIV_FACTOR */
  double rel_tol;       /* > 0: stop at |b - A x| <= rel_tol |b - A x0| (overrides the two above) */
  int precision;        /* cd3d_precision */
  int cycle;            /* SC_MG_CYCLE_SCALAR: 1 V (default), 2 W, 0 F */
} cd3d_solver_params;

typedef struct {
  int iterations;
  int status;           /* cd3d_status of the solve */
  double r0;            /* |b - A x0| */
  double rfinal;        /* |b - A x| recomputed in FP64 after the solve */
  double ms_assemble;
  double ms_solve;
} cd3d_stats;

/* defaults = ParMooN defaults of the configuration-1 solver (Jacobi 0.8,
 * 2+2, V-cycle, exact coarse solve, FGMRES(20), maxit 200, rel_tol 1e-10) */
void cd3d_default_solver(cd3d_solver_params *p);

/* Build the multilevel problem on the current HIP device:
 * TDomain::Init + RegRefineAll (Domain.C:818,1037), TFESpace3D (FESpace3D.C:55),
 * TSquareStructure3D + Sort (SquareStructure3D.C:32, SquareStructure.C:65),
 * TSquareMatrix3D (SquareMatrix3D.C), TMGLevel3D/TMultiGrid3D::AddLevel
 * (MultiGrid3D.C:61-68), TFgmresIte ctor (FgmresIte.C:44). */
cd3d_status cd3d_create(const cd3d_mesh_desc *mesh, cd3d_problem **out);
/* The same hierarchy over general hexahedra (TDomain::Init with a GEO file of
 * explicit vertex coordinates, Domain.C:818 / ReadGeo.C:1790-1812; every Q1/Q2
 * cell mapped by THexaTrilinear, HexaTrilinear.C): the m^3 macro cells have
 * the vertices macro_xyz ((m+1)^3 x 3 doubles, x fastest) and the identity
 * local frame; refinement puts new vertices at edge, face and cell means
 * (Refinement.C).  mesh->box is not used.  Every macro cell must be oriented
 * like the reference cube (positive Jacobian at its 8 corners), else
 * CD3D_ERR_ARG.  Assembly runs cell by cell on every level (the class-stencil
 * paths need bricks); both precision forms and TCD3D work as on the box. */
cd3d_status cd3d_create_mapped(const cd3d_mesh_desc *mesh, const double *macro_xyz, cd3d_problem **out);
/* TDomain::ReadGeo (ReadGeo.C:46-200) for structured macro meshes: reads a
 * GEO file whose NEL = m^3 hexahedra and NVT = (m+1)^3 vertices are numbered
 * along the lattice (x fastest, cell vertices in the reference's hexahedron
 * order, every KNPR 0); writes m and, when xyz is non-NULL, the vertex
 * coordinates (cap: doubles available, at least 3 (m+1)^3).  Other meshes:
 * CD3D_ERR_ARG with the reason in msg (msg_len bytes, may be NULL). */
cd3d_status cd3d_read_geo(const char *path, int *m, double *xyz, size_t cap, char *msg, size_t msg_len);
cd3d_status cd3d_destroy(cd3d_problem *p);
const char *cd3d_last_error(const cd3d_problem *p);
/* bytes of device memory held by the handle */
size_t cd3d_device_bytes(const cd3d_problem *p);

/* sizes of level l (TFESpace3D::GetN_DegreesOfFreedom / GetActiveBound,
 * TSquareMatrix::GetN_Entries, TCollection::GetN_Cells) */
cd3d_status cd3d_level_info(const cd3d_problem *p, int level, int *n_dof, int *n_active, int64_t *nnz,
                            int *n_cells);
/* host copies: TFESpace3D::GetGlobalNumbers (n_cells x n_basis),
 * TSquareMatrix::GetRowPtr/GetKCol/GetEntries, rhs and vertex coordinates
 * (n_cells x 8 x 3) */
cd3d_status cd3d_get_dofmap(const cd3d_problem *p, int level, int *gn_host);
cd3d_status cd3d_get_cells(const cd3d_problem *p, int level, double *xyz_host);
cd3d_status cd3d_get_csr(cd3d_problem *p, int level, int *rowptr_host, int *kcol_host, double *entries_host);
cd3d_status cd3d_get_rhs(cd3d_problem *p, int level, double *rhs_host);
cd3d_status cd3d_set_entries(cd3d_problem *p, int level, const double *entries_host);

/* Coefficients of any CoeffFct3D (Constants.h:123).  The callback cannot run
 * on the device, so the caller evaluates it where the reference would: at
 * the quadrature points of every cell (TFEDatabase3D::GetOrig,
 * FEDatabase3D.C:1116; DiscreteForm3D.C:187-270).
 * cd3d_quad_points: their physical coordinates, n_cells x n_q x 3 doubles in
 * the reference's cell order and Gauss-point order (n_q = 27 for Q1, 64 for Q2).
 * cd3d_set_coeffs_qp: the callback's values there, n_cells x n_q x 6 doubles
 * (coeff[0..5] = eps, b1, b2, b3, c, f; |b| is recomputed), used by every
 * later assembly of the level instead of the descriptor's functions (the
 * descriptor still selects Galerkin or SUPG); NULL returns to the descriptor.
 * The values are used as they are at every assembly: a time-dependent
 * callback needs new values before each TCD3D step. */
cd3d_status cd3d_quad_points(const cd3d_problem *p, int level, double *xyz_host);
cd3d_status cd3d_set_coeffs_qp(cd3d_problem *p, int level, const double *coeffs_host);

/* TSystemCD3D::Assemble (SystemCD3D.C:314): TAssembleMat3D::Assemble3D
 * (AssembleMat3D.C:230) on every level with BilinearAssemble /
 * BilinearAssemble_SD (ConvDiff3D.C:61,124); rhs on Dirichlet DOFs = 0. */
cd3d_status cd3d_assemble(cd3d_problem *p, const cd3d_coeffs *co);

/* MatVect / ScalarDefect on level l (Blas2.C:60,11), host vectors */
cd3d_status cd3d_matvec(cd3d_problem *p, int level, const double *x_host, double *y_host);
/* same on device buffers, on the handle's stream (asynchronous) */
cd3d_status cd3d_matvec_dev(cd3d_problem *p, int level, const double *x_dev, double *y_dev);
/* cd3d_matvec without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc).  x and y are staged through a 64 MiB array
 * that the library holds on each device for the life of the process, so the
 * level may have at most 4 194 304 unknowns (CD3D_ERR_ARG beyond that).
 * Calls are serialised process-wide (they share the array). */
cd3d_status cd3d_matvec_noalloc(cd3d_problem *p, int level, const double *x_host, double *y_host);
/* Prolongate / DefectRestriction between level l-1 and l (MGComponents3D.C:182,706) */
cd3d_status cd3d_prolongate(cd3d_problem *p, int level, const double *coarse_host, double *fine_host);
cd3d_status cd3d_restrict(cd3d_problem *p, int level, const double *fine_host, double *coarse_host);

/* one preconditioner application z = V(r), TMultiGridScaIte::Iterate
 * (MultiGridScaIte.C:63) with TMultiGrid3D::Cycle (MultiGrid3D.C:423) */
cd3d_status cd3d_vcycle(cd3d_problem *p, const cd3d_solver_params *sp, const double *r_host, double *z_host);

/* TSystemCD3D::Solve -> TFgmresIte::Iterate (SystemCD3D.C:342, FgmresIte.C:219).
 * rhs_host NULL: the assembled rhs; sol_host NULL: keep the solution on the
 * device (read it with cd3d_get_solution).  Initial guess: 0 on active DOFs,
 * the rhs (Dirichlet values) on Dirichlet DOFs, or sol_host's values when
 * use_guess != 0. */
cd3d_status cd3d_solve(cd3d_problem *p, const cd3d_solver_params *sp, const double *rhs_host, double *sol_host,
                       int use_guess, cd3d_stats *st);
/* Same-level batching (no reference counterpart: the reference solves one
 * problem per process): `count` assembled problems with the same number of
 * levels on one device -- of any sizes, element orders and coefficients --
 * solved together, each as cd3d_solve(p, sp, NULL, NULL, 0, &stats[b])
 * would: the V-cycle (damped Jacobi, exact coarse solve) inside FP64
 * FGMRES, in lockstep, one launch per step for all problems still iterating
 * (kernels index each problem's arrays from device tables) and one host
 * transfer of all H columns per iteration.  Each problem keeps its own
 * stopping test.  FP64: iteration counts equal a lone solve's, solutions
 * agree to rounding (the reduction trees differ).  Mixed precision: the
 * V-cycle runs on FP32 level vectors over the CSR matrices rounded to FP32
 * as they are read (the lone mixed solve uses the lattice form instead), to
 * the same FP64 residual test; iteration counts may differ from a lone
 * mixed solve's by the rounding of the preconditioner.  Other parameters
 * (smoothers other than 1, W/F cycles, coarse smoothers other than 17, no
 * pre-smoothing, restart > 200) or count 1: the problems are solved one
 * after another.
 * Runs on the first handle's stream; every handle is locked meanwhile.
 * stats (count entries, may be NULL) receives each problem's result; the
 * return value is CD3D_OK or the first non-OK status. */
cd3d_status cd3d_solve_batch(cd3d_problem **problems, int count, const cd3d_solver_params *sp, cd3d_stats *stats);
cd3d_status cd3d_get_solution(cd3d_problem *p, double *sol_host);

/* TCD3D Crank-Nicolson (3DPrograms/TCD3D_ParMooN.C:539-637 over
 * TSystemTCD3D, SystemTCD3D.C:176-376): the mass matrix and rhs at t0
 * (AssembleMRhs), then per step AssembleARhs at t+tau, AssembleSystMat,
 * Solve, RestoreMassMat.  u_host: in = initial state (finest level), out =
 * final state; states_host (nsteps x N) optional. */
cd3d_status cd3d_tcd_run(cd3d_problem *p, const cd3d_coeffs *co, const cd3d_solver_params *sp, double t0,
                         double tau, int nsteps, double *u_host, double *states_host, cd3d_stats *st);
/* Step-wise TCD3D with the theta-scheme of SetTimeDiscParameters
 * (TimeDiscRout.C; theta 0.5 Crank-Nicolson, 1 backward Euler).
 * cd3d_tcd_init: TSystemTCD3D::AssembleMRhs at t0 (SystemTCD3D.C:176-202)
 * and the initial state u_host (finest level; Dirichlet values from the rhs).
 * co->supg: the SUPG forms of TCD3D.C:48,102 -- the rhs with SUPG test
 * functions and, per step, the time-derivative matrix K beside M (system
 * M + K + tau theta A, rhs matrix M + K - tau (1 - theta) A).
 * cd3d_tcd_step: one step t -> t + tau: AssembleARhs at t + tau,
 * AssembleSystMat, Solve, RestoreMassMat (SystemTCD3D.C:204-376, driven as
 * in TCD3D_ParMooN.C:569-622); u_host (optional) receives the new state.
 * The coefficients' t is ignored (the handle keeps the time).  After
 * cd3d_tcd_init the level matrices hold the mass matrices: cd3d_solve needs a
 * new cd3d_assemble. */
cd3d_status cd3d_tcd_init(cd3d_problem *p, const cd3d_coeffs *co, double t0, const double *u_host);
cd3d_status cd3d_tcd_step(cd3d_problem *p, const cd3d_coeffs *co, const cd3d_solver_params *sp, double tau,
                          double theta, double *u_host, cd3d_stats *st);
cd3d_status cd3d_tcd_get_state(cd3d_problem *p, double *u_host);
/* nodal interpolation of the Gaussian pulse (TFEFunction3D::Interpolate, FEFunction3D.C:1446) */
cd3d_status cd3d_interp_pulse(cd3d_problem *p, int level, const cd3d_coeffs *co, double *u_host);

/* Diagnostics and output.  Known functions fn: 0 zero, 1 the sine
 * sin(pi x) sin(pi y) sin(pi z) of Examples/CD_3D/Laplace.h:12-19 (value and
 * gradient), 2 the Gaussian pulse centred at co->pulse_c (sigma 0.05; co may
 * be NULL for fn 0, 1 and 3), 3 the linear function x + 2y + 3z (it lies in
 * every Q1/Q2 space, also on trilinear cells, so its errors vanish). */
/* TFEFunction3D::Interpolate (FEFunction3D.C:1446): nodal interpolation of fn
 * on level l, evaluated on the device */
cd3d_status cd3d_interpolate(cd3d_problem *p, int level, int fn, const cd3d_coeffs *co, double *u_host);
/* TFEFunction3D::GetErrors with L2H1Errors (FEFunction3D.C:57,
 * MainUtilities.C:1623) on the device, at the quadrature points of the
 * assembly: errors[0] = |u - fn|_L2, errors[1] = |u - fn|_H1 (semi-norm),
 * errors[2] = max |u - fn| at the last cell's points (what the reference
 * leaves in its third entry), errors[3] = max |u - fn| at all points.
 * u_host NULL: the level's current solution on the device (finest level: the
 * last cd3d_solve, or the TCD3D state after cd3d_tcd_init). */
cd3d_status cd3d_errors(cd3d_problem *p, int level, const double *u_host, int fn, const cd3d_coeffs *co,
                        double *errors);
/* TOutput3D::WriteVtk (Output3D.C:1747): legacy ASCII VTK of level l with the
 * scalar u_host as point data `name` (the FE function at the cell vertices,
 * averaged over the cells, as the reference writes it); vertices are numbered
 * along the lattice (x fastest), not in the reference's pointer order */
cd3d_status cd3d_write_vtk(cd3d_problem *p, int level, const double *u_host, const char *name, const char *path);
/* TSquareMatrix::Write (SquareMatrix.C:156): binary file of int32 n_rows,
 * n_cols, nnz, the row pointers, column indices and FP64 entries of level l */
cd3d_status cd3d_write_matrix(cd3d_problem *p, int level, const char *path);

/* Measurement: launch one kernel of the path `reps` times on the handle's
 * stream between two HIP events (after one untimed launch) and return the
 * mean duration per launch and the kernel's algorithmic bytes per launch.
 * which: 0 FP64 CSR SpMV (12 nnz + 20 N bytes); 1 FP32 lattice Jacobi sweep
 * (compact class stencils: 12 bytes per node; per-node stencils: + 4 per
 * nonzero + 4 per node); 2 FP64 lattice operator (16 bytes per node);
 * 3 FP32 lattice restriction to level-1 (4 N_fine + 4 N_coarse);
 * 4 FP32 lattice prolongation + update from level-1 (8 N_fine + 4 N_coarse);
 * 5 CSR values from the class stencils (10 bytes per active-row nonzero); 6 copy of an
 * FP64 level vector (16 N: the traffic-counter calibration); 7 one fused
 * Gram-Schmidt step of FGMRES, w -= h v_k and the partial sums of <w, v_k+1>
 * (32 N: v_k, w, v_k+1 read, w written); 8 the cell-by-cell element kernel
 * (BilinearAssemble / BilinearAssemble_SD per quadrature point) over all cells
 * of the level with the coefficients of the last assembly, element rows
 * written out and gathered into a scratch CSR value array (CD3D_ASM_SCATTER=1:
 * the 8 colour-by-colour scatters): *bytes_per_launch then holds the
 * algorithmic FP64 flops of one pass, n_cells n_q 8 n_b^2; 9 the mixed form's
 * low-synchronisation Gram-Schmidt step at i = 10 on the finest level (dot
 * pass, coefficients, update pass: (2 i + 5) 8 N bytes).
 * Needs an assembled problem; the level vectors are overwritten. */
cd3d_status cd3d_kernel_bench(cd3d_problem *p, int which, int level, int reps, double *ms_per_launch,
                              double *bytes_per_launch);

/* device pointers of level vectors for timing and zero-copy callers:
 * which 0 = solution, 1 = rhs, 2 = scratch vector, 3 = scratch vector */
cd3d_status cd3d_device_ptr(cd3d_problem *p, int level, int which, double **dev);
/* the handle's HIP stream (hipStream_t) */
void *cd3d_stream(cd3d_problem *p);
cd3d_status cd3d_synchronize(cd3d_problem *p);

#ifdef __cplusplus
}
#endif
#endif

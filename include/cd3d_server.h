/* cd3d_server.h -- serving loop over libcd3d_hip.so: CD3D / TCD3D solve
 * requests of uneven sizes and coefficients, admitted by an HBM budget and run
 * concurrently on one GPU.
 *
 * The reference has no counterpart (its mains run one problem per process,
 * 3DPrograms/CD3D_ParMooN.C, TCD3D_ParMooN.C); a request is exactly one such
 * run: TSystemCD3D::Assemble + Solve, or TSystemTCD3D's AssembleMRhs and
 * nsteps time steps.
 *
 * Requests are queued on submit and admitted in arrival order while the
 * in-flight count is below max_inflight and their device memory fits the
 * budget (an idle resident hierarchy of the same mesh is reused; idle ones
 * are evicted, least recently used first, to make room).  A request that
 * does not fit may be overtaken by later ones that do, at most
 * max_bypass times.  Admitted requests run on worker threads, each request on
 * its own hierarchy and HIP stream, so their kernels overlap on the device.
 * Every result is the same as the request run alone: no state is shared
 * between requests, and every reduction is deterministic.
 *
 * All calls are thread-safe.
 */
#ifndef CD3D_SERVER_H
#define CD3D_SERVER_H

#include "cd3d.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cd3d_server cd3d_server;

typedef struct {
  size_t hbm_budget;   /* bytes of device memory for hierarchies and work space; 0: 85% of the free memory */
  int max_inflight;    /* requests on the device at once (slots) */
  int workers;         /* host threads driving the device; 0: max_inflight */
  int keep_resident;   /* keep idle hierarchies for later requests with the same mesh */
  int max_bypass;      /* times a waiting request may be overtaken by later ones that fit */
  int simulate;        /* 1: do not touch the device; a request takes sim_us_per_dof * DOFs microseconds
                          (admission and statistics only, for capacity planning) */
  double sim_us_per_dof;
  int robust_after;    /* a CD solve with the damped Jacobi smoother that has not converged after this many
                          FGMRES iterations continues from its iterate with the multicolour SSOR smoother
                          (13, omega 1, 4 + 4) on the rest of its iteration budget; 0: never */
  int max_batch;       /* a worker takes up to this many admitted CD requests with the same number of levels
                          and identical FP64 solver parameters (damped Jacobi, V-cycle, exact coarse solve)
                          and solves them together (cd3d_solve_batch); 1: every request alone */
} cd3d_server_params;

typedef enum { CD3D_REQ_CD = 0, CD3D_REQ_TCD = 1 } cd3d_request_kind;

typedef struct {
  int kind;                    /* cd3d_request_kind */
  cd3d_mesh_desc mesh;
  cd3d_coeffs coeffs;
  cd3d_solver_params solver;
  double t0, tau, theta;       /* TCD3D: start time, step, theta (0.5 Crank-Nicolson) */
  int nsteps;                  /* TCD3D: time steps */
  int want_solution;           /* keep the finest-level solution (CD) / final state (TCD) for cd3d_server_wait,
                                  also when the solve ends CD3D_NOT_CONVERGED */
} cd3d_request;

typedef struct {
  uint64_t id;
  int status;                  /* cd3d_status of the request */
  int iterations;              /* FGMRES iterations (TCD: summed over the steps) */
  double rel_residual;         /* |b - A x| / |b - A x0| (TCD: the largest over the steps run) */
  double queue_ms;             /* submit -> admitted */
  double run_ms;               /* admitted -> done */
  double latency_ms;           /* submit -> done */
  size_t device_bytes;         /* bytes of the hierarchy it ran on */
  int reused_hierarchy;        /* 1: an idle resident hierarchy was reused */
  int n_dof;                   /* finest-level DOFs (length of the solution) */
  int steps_done;              /* TCD: time steps completed; a step that does not converge ends the run
                                  (status CD3D_NOT_CONVERGED) and the state after it is returned */
  int escalated;               /* 1: the solve was continued with the multicolour SSOR smoother (robust_after) */
  int batch_size;              /* requests solved together with this one (max_batch), itself included */
} cd3d_result;

typedef struct {
  uint64_t submitted, completed, failed, rejected;
  double p50_ms, p99_ms, mean_ms;   /* latency over completed requests */
  double solves_per_s;              /* completed / (last completion - first submit) */
  size_t bytes_in_use, peak_bytes, budget;
  int inflight, queued, resident;
} cd3d_server_stats;

void cd3d_server_default_params(cd3d_server_params *p);
/* device bytes one problem of this mesh needs, as the admission control
 * counts them: hierarchy, lattice form, Krylov space of the given restart;
 * cellwise_assembly: variable coefficients (slot maps, per-node FP32
 * stencils); tcd: the TCD3D matrices and vectors (1 Galerkin, 2 SUPG: also
 * the time-derivative matrix K) */
size_t cd3d_estimate_bytes(const cd3d_mesh_desc *mesh, int restart, int cellwise_assembly, int tcd);

cd3d_status cd3d_server_create(const cd3d_server_params *p, cd3d_server **out);
/* non-blocking; CD3D_ERR_NOMEM when the request can never fit the budget,
 * CD3D_ERR_ARG for an invalid request */
cd3d_status cd3d_server_submit(cd3d_server *s, const cd3d_request *req, uint64_t *id);
/* blocks until request id is done (timeout_ms < 0: no limit; CD3D_ERR_STATE
 * on timeout); sol_host (optional, sol_len doubles) receives the solution of
 * a want_solution request.  The result is released by this call. */
cd3d_status cd3d_server_wait(cd3d_server *s, uint64_t id, double timeout_ms, cd3d_result *res, double *sol_host,
                             int64_t sol_len);
/* blocks until every submitted request is done */
cd3d_status cd3d_server_drain(cd3d_server *s);
cd3d_status cd3d_server_get_stats(cd3d_server *s, cd3d_server_stats *st);
/* finishes the requests on the device, drops the queued ones, frees everything */
cd3d_status cd3d_server_destroy(cd3d_server *s);

#ifdef __cplusplus
}
#endif
#endif

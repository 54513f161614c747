// Internal data structures of libcd3d_hip.so.
//
// Layout in device memory (HBM), per multigrid level l:
//   CSR  rowptr int32[N+1], col int32[nnz], val f64[nnz] (+ val32 f32[nnz] for
//        the mixed-precision V-cycle), diag int32[N] (position of a_ii)
//   vectors f64[N]: sol, rhs, aux (defect / correction), tmp (Jacobi buffer)
//            f32[N]: the same four for the mixed-precision V-cycle
//   transfer from level l-1: P (fine rows, CSR) and R = P^T (coarse rows, CSR)
//   assembly: cell geometry int4[ncell] (origin in node units + frame code),
//            slot map int32[ncell][nb][nb] (CSR position of each local entry,
//            -1 for Dirichlet rows), DOF map int32[ncell][nb], cells grouped by
//            the 8 parity colours (no two cells of one colour share a DOF)
// DOF numbering, CSR pattern and row pointers are those of the reference
// (FESpace3D.C / SquareStructure3D.C); vectors are in that DOF order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cd3d.h"

namespace cd3d {

// Lattice-stencil form of a level (mixed-precision V-cycle): interior nodes
// of the node lattice split into parity classes; class c holds the box of
// nodes with X % 2 == px (Q2; one class for Q1), each with the same stencil
// (per axis offsets -p..p on cell-face planes, the cell's p+1 nodes inside a
// cell).  Values are stored [class][stencil entry][node of the class].
struct LatClass {
  int px[3] = {0, 0, 0};   // parity per axis (Q2), 0 for Q1
  int lo[3] = {0, 0, 0};   // first lattice coordinate per axis
  int cnt[3] = {0, 0, 0};  // nodes per axis (step 2 for Q2, 1 for Q1)
  int step = 1;
  int64_t nodes = 0;       // cnt0 cnt1 cnt2
  int nsten = 0;           // stencil entries
  int64_t off = 0;         // offset of the class in the stencil value array
  int64_t first = 0;       // index of the class's first node in the interior-node numbering
};

struct HostLevel {
  int n = 0;        // cells per side
  int ncell = 0;
  int nn = 0;       // nodes per side (p n + 1)
  int ndof = 0, nact = 0;
  int64_t nnz = 0;
  std::vector<int> org;         // 3 per cell, node-lattice units
  std::vector<double> vx;       // mapped hierarchy: (n+1)^3 x 3 vertex coordinates, x fastest
  std::vector<int8_t> ax;       // 9 per cell: ax[9c + 3a + d]
  std::vector<int> gn;          // nb per cell
  std::vector<int> dof_node;    // DOF -> lattice node
  std::vector<int> rowptr, kcol, diag;
  std::vector<int> blk;  // row blocks for the stream kernels
  // transfer from the next coarser level
  std::vector<int> p_rowptr, p_col;
  std::vector<double> p_val;
  std::vector<int> r_rowptr, r_col;
  std::vector<double> r_val;
  std::vector<int> colour_start;  // 9 offsets into colour_cells
  std::vector<int> colour_cells;
  // multicolour SSOR: active rows grouped by node colour X%k + k(Y%k + k Z%k),
  // k = 2 (Q1) or 3 (Q2), ascending within a colour
  std::vector<int> mc_start, mc_rows;
  // lattice-stencil form
  std::vector<LatClass> lat;
  int64_t lat_nval = 0;             // sum over classes of nsten * nodes
  std::vector<int> lat_perm;        // CSR position of every stencil value
  std::vector<uint16_t> lat_sid;    // per active-row CSR entry: class * 128 + stencil entry
  std::vector<int> lat_first;       // CSR position of entry s of class c's first node (c * 128 + s)
  std::vector<int> lat_node2dof;    // interior node (class numbering) -> DOF
};

struct DevLevel {
  int ndof = 0, nact = 0;
  int64_t nnz = 0;
  int *rowptr = nullptr, *col = nullptr, *diag = nullptr;
  double *val = nullptr;
  // row blocks of the stream kernels (at most 2048 nonzeros and 256 rows)
  int *blk = nullptr;
  int nblk = 0;
  // per row block: first row, end row, first nonzero, nonzero count
  int4 *bt = nullptr;
  double *invdiag = nullptr;  // 1 / a_ii (valid when invdiag_ok)
  bool invdiag_ok = false;    // computed on first use after each change of val
  double *sol = nullptr, *rhs = nullptr, *aux = nullptr, *tmp = nullptr;
  double *rhs_asm = nullptr;  // assembled rhs (the multigrid works in `rhs`)
  // TCD3D: conv-diff matrix next to the system (mass) matrix in `val`; with
  // SUPG also the time-derivative matrix K (M + K once a step has formed it)
  double *aconv = nullptr, *ksupg = nullptr;
  // transfer (level >= 1)
  int *p_rowptr = nullptr, *p_col = nullptr, *r_rowptr = nullptr, *r_col = nullptr;
  double *p_val = nullptr, *r_val = nullptr;
  int64_t p_nnz = 0, r_nnz = 0;
  // coefficients per quadrature point handed in by the caller (cd3d_set_coeffs_qp), or null
  double *qp = nullptr;
  // multicolour SSOR row lists (uploaded on first use)
  int *mc_rows = nullptr;
  std::vector<int> mc_start;
  // batched mixed V-cycle: sol, rhs, aux, tmp and 1/a_ii in FP32 (5 ndof
  // floats, allocated on first use)
  float *b32 = nullptr;
  // SOR / SSOR in row order: forward [0] and backward [1] wavefront schedules
  // (build_gs_schedule; uploaded on first use)
  int *gs_rows[2] = {nullptr, nullptr}, *gs_start[2] = {nullptr, nullptr};
  int gs_nlev[2] = {0, 0};
  // assembly
  int4 *cells = nullptr;
  int *slots = nullptr;
  // two-pass cell-by-cell assembly: element rows E (ncell nb x nb), their rhs
  // entries Er (ncell nb) and the gather lists (allocated on first use)
  double *E = nullptr, *Er = nullptr;
  int *gptr = nullptr, *gidx = nullptr, *rptr = nullptr, *ridx = nullptr;
  int *gn = nullptr;
  int *colour_cells = nullptr;
  std::vector<int> colour_start;
  int ncell = 0;
  // mapped hierarchy: the vertex lattice ((nv)^3 x 3, nv = n + 1) and the
  // per-cell SUPG mesh size in convection direction (allocated on first use)
  double *vx = nullptr;
  int nv = 0;
  double *hconv = nullptr;
  double h[3] = {0, 0, 0};   // cell sides
  double dn[3] = {0, 0, 0};  // node spacing
  // lattice-stencil form (mixed-precision V-cycle)
  int nn = 0;                  // nodes per side
  int nclass = 0;
  LatClass cls[8];
  int64_t lat_nval = 0, lat_nint = 0;
  float *lval = nullptr;       // stencil values f32
  int *lperm = nullptr;        // CSR position of each stencil value (per-node FP32 path only)
  uint16_t *lsid = nullptr;    // stencil id of each active-row CSR entry
  int *lfirst = nullptr;       // CSR positions of each class's first node (8 x 128)
  float *linv = nullptr;       // 1/a_ii per interior node (class numbering)
  int *lnode2dof = nullptr;    // interior node -> DOF
  int *ldof2node = nullptr;    // DOF -> lattice node
  float *lx = nullptr, *lf = nullptr, *la = nullptr, *lt = nullptr;  // lattice vectors, nn^3 each
  // classes whose stencil is the same at every node (constant coefficients on
  // the uniform mesh) use one compact stencil: 128 floats per class, the last 1/a_ii
  int uni[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  float *lcst = nullptr;
  double *lscale = nullptr;
  int *lflag = nullptr;
  // assembly through class stencils: element matrices (3 NB x NB + NB) and
  // FP64 class stencils: [S0 8x128][class rhs 8][SX 8x128][SY 8x128]
  // [mass stencils 8x128][centre entry per class, 8 ints] (kLst64 doubles)
  double *lelem = nullptr, *lst64 = nullptr;
  // what the last assembly left for the FP32 lattice form of `val`:
  // 0 nothing (detect on the device), 1 compact stencils in lcst,
  // 2 affine stencils: lval = lat_alpha_m * mass + lat_alpha * (S0 + X SX + Y SY)
  int lat_fresh = 0;
  double lat_alpha_m = 0.0, lat_alpha = 1.0;
  bool lat_exact64 = false;  // lst64 holds exactly the CSR values of the interior rows
  // the FP32 sweeps evaluate the affine stencils node by node (lat_fresh 2
  // on a level too large for the one-workgroup program): lval / linv unused
  bool lat_affine = false;
  int lat_centre[8] = {0};  // stencil index of the offset (0, 0, 0) per class
  // the CSR values of the interior rows are exactly the affine stencils
  // (alpha (S0 + X SX + Y SY), or fma(alpha, that, M) with lat_alpha_m 1):
  // the FP64 lattice operator can stand for the matrix
  bool lat_aff64 = false;
  bool tcd_mass_uniform = false;  // TCD3D: the mass class stencils are in lst64 (mass slot)
};

struct BatchLevel;
struct BatchProblem;

struct Problem {
  cd3d_mesh_desc mesh{};
  // general hexahedra: the macro vertices were given (cd3d_create_mapped);
  // every cell is mapped trilinearly and assembled cell by cell
  bool mapped = false;
  int nb = 0;
  std::vector<HostLevel> hl;
  std::vector<DevLevel> dl;
  hipStream_t stream = nullptr;
  // the coarse inverse runs on a stream of its own, beside the assembly of
  // the finer levels; the main stream waits for it before the first coarse
  // solve and at the end of every call (coarse_join)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool coarse_pending = false, fork_recorded = false;
  // the coarse inverse may run on `side` (off for handles of the serving
  // loop: with many handles in flight, side streams share the hardware
  // queues with other handles' main streams and hold them up)
  bool side_ok = true;
  int device = 0;
  size_t dev_bytes = 0;
  std::vector<void *> allocs;
  std::vector<size_t> alloc_bytes;
  std::string err;
  std::mutex mu;
  bool assembled = false;
  // exact coarse solve: explicit inverse of the coarsest matrix (f64)
  double *cinv = nullptr;
  double *gj_scratch = nullptr;
  bool cinv_valid = false;
  // FGMRES workspace
  int krylov_restart = 0;
  int64_t krylov_ld = 0;  // column stride of V and Z
  double *V = nullptr, *Z = nullptr, *w = nullptr, *xs = nullptr, *bs = nullptr;
  double *red = nullptr;      // reduction partials
  double *dscal = nullptr;    // device scalars: H column at [0, restart], y at [yoff, yoff + restart)
  // low-synchronisation Gram-Schmidt of the mixed form: dot partials and the
  // strictly lower part of V^T V (allocated on first use)
  double *ls_part = nullptr, *ls_L = nullptr;
  int ls_blocks = 0;          // chunks ls_part holds partials for
  double *hscal = nullptr;    // pinned host mirror of dscal
  int yoff = 64, nscal = 256;
  int *hflag = nullptr;       // pinned, lattice uniformity flags
  double *xl = nullptr, *bl = nullptr;  // lattice-ordered iterate and rhs of the lattice FGMRES
  int red_blocks = 0;
  // TCD3D state: f(t_old), u_old, a defect buffer; the current state is in xs
  double *tcd_oldrhs = nullptr, *tcd_oldsol = nullptr, *tcd_def = nullptr;
  double tcd_t = 0;
  bool tcd_ready = false;
  // device tables for assembly (reference basis at quadrature points)
  double *tab = nullptr;
  // coefficients of the last assembly (the kernel bench re-runs its element kernel)
  cd3d_coeffs last_co{};
  // workspace of the same-level batches this handle leads (grown on demand)
  struct BatchWs {
    int cap_b = 0, cap_l = 0;
    BatchLevel *dlev = nullptr;
    BatchProblem *dprob = nullptr;
    int *dactive = nullptr;
    double *dred = nullptr, *dscal = nullptr, *dcoef = nullptr;
    double *hscal = nullptr, *hcoef = nullptr;  // pinned mirrors
    int *hactive = nullptr;                     // pinned
  } bws;
};

// ---------------------------------------------------------------- setup (setup.cpp)
void build_host_level(HostLevel &L, int order, int m, int depth);
// one regular refinement of an n^3 vertex lattice (edge, face and cell means)
void refine_vertices(const std::vector<double> &src, int n, std::vector<double> &dst);
void build_transfer(HostLevel &fine, const HostLevel &coarse, int order);
void build_slots(const HostLevel &L, int order, std::vector<int> &slots);
void build_gather(const HostLevel &L, int order, const std::vector<int> &slots, std::vector<int> &gptr,
                  std::vector<int> &gidx, std::vector<int> &rptr, std::vector<int> &ridx);
void build_colours(HostLevel &L, int order);
void build_row_blocks(const std::vector<int> &rowptr, int max_nnz, int max_rows, std::vector<int> &blk);
void build_mc_colours(HostLevel &L, int order);
void build_gs_schedule(const HostLevel &L, bool backward, std::vector<int> &rows, std::vector<int> &start);
void build_lattice(HostLevel &L, int order);
void basis_tables(int order, std::vector<double> &tab);
int quad_points_1d(int order);

// ---------------------------------------------------------------- kernels (kernels.hip)
// Workgroup b of a launch runs on XCD b mod 8 (round-robin dispatch over the
// 8 XCDs of an MI355X, each with its own L2).  xcd_block maps b to a logical
// block so that XCD x takes the contiguous range of logical blocks
// [x q + min(x, r), ...) of its q or q + 1 blocks (n = 8 q + r): a
// bijection of 0..n-1 onto itself.
__host__ __device__ inline int xcd_block(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

#ifndef CD3D_RED_BLOCKS
#define CD3D_RED_BLOCKS 1024
#endif
constexpr int kRedBlocks = CD3D_RED_BLOCKS;
// CSR-stream row blocks: at most kStreamNnz nonzeros and kStreamRows rows,
// processed by kStreamBlock threads (with XCD-contiguous block ranges, on the
// 130 M-nonzero Q2 matrix: 2048 x 512 342 us, 2048 x 256 352 us, 4096 x 1024
// 350 us, 4096 x 256 397 us, 1024 x 512 405 us)
#ifndef CD3D_STREAM_NNZ
#define CD3D_STREAM_NNZ 2048
#endif
#ifndef CD3D_STREAM_BLOCK
#define CD3D_STREAM_BLOCK 512
#endif
constexpr int kStreamNnz = CD3D_STREAM_NNZ;
constexpr int kStreamBlock = CD3D_STREAM_BLOCK;
constexpr int kStreamRows = 256;

void k_spmv(hipStream_t s, const DevLevel &L, const double *x, double *y);
void k_defect(hipStream_t s, const DevLevel &L, const double *x, const double *f, double *d);
void k_jacobi(hipStream_t s, const DevLevel &L, const double *x, const double *f, double *out, double omega);
void k_jacobi_zero(hipStream_t s, const DevLevel &L, const double *f, double *out, double omega);
// one multicolour SSOR sweep in place (smoother 13): x_D = f_D, relaxed
// Gauss-Seidel colour by colour forward, then backward (needs mc_rows)
void k_mcssor(hipStream_t s, const DevLevel &L, double *x, const double *f, double omega);
// low-synchronisation Gram-Schmidt step i (restart length <= k_ls_max()):
// h = the MGS column of w against V_0 .. V_i into h[0 .. i], w -= V h, and
// the partials of |w|^2 (kRedBlocks); part holds (2 k_ls_max() + 1)
// k_ls_blocks(n) doubles, Lm k_ls_max()^2
int k_ls_blocks(int n);
int k_ls_max();
void k_ls_step(hipStream_t s, int n, int i, const double *V, int64_t ld, double *w, double *part, double *Lm,
               double *h, double *partials);
// one SOR (kind 2) or SSOR (kind 3) sweep in the reference's row order, in
// place: x_D = f_D, then the wavefront schedule(s) of the level
void k_gs_sweep(hipStream_t s, const DevLevel &L, double *x, const double *f, double omega, int kind);
void k_invdiag(hipStream_t s, const DevLevel &L);
void k_restrict(hipStream_t s, const DevLevel &F, const DevLevel &C, const double *fine, double *coarse);
void k_prolong_update(hipStream_t s, const DevLevel &F, const DevLevel &C, const double *coarse, double *x,
                      double damp);
void k_prolong(hipStream_t s, const DevLevel &F, const DevLevel &C, const double *coarse, double *fine);
void k_gemv_dense(hipStream_t s, int n, const double *A, const double *x, double *y);
void k_coarse_smooth(hipStream_t s, const DevLevel &L, double *x, const double *f, double *d, int smoother,
                     double omega, int maxit, double red);
// BLAS-1 with deterministic two-stage reductions
void k_dot(hipStream_t s, int n, const double *a, const double *b, double *partials, double *out);
void k_axpy_dot(hipStream_t s, int n, const double *alpha_partials, double sign, const double *x, double *y,
                const double *z, double *partials, double *alpha_out);
// y32 (optional): also (float) y
// The finest lattice level's first pre-smoothing sweep from zero, written
// beside the FP32 copy of a Krylov vector (which is that V-cycle's rhs f):
// x[e] = omega f[e] d[c dstride] at the interior lattice nodes e (1 <= X,
// Y, Z <= nn - 2), c the node's parity class (X & 1) + 2 (Y & 1) + 4 (Z & 1)
// for Q2, 0 for Q1; other entries untouched.  x == nullptr: none.
struct TopZeroSweep {
  float *x = nullptr;
  const float *d = nullptr;
  int dstride = 0;
  float omega = 0.f;
  int nn = 0, order = 0;
};
// zt: the finest level's zero-start sweep beside y32 (y32 in lattice order)
void k_scale_copy(hipStream_t s, int n, const double *partials, double *norm_out, const double *x, double *y,
                  float *y32 = nullptr, const TopZeroSweep *zt = nullptr);
// x *= alpha, and y32 = (float) x when y32 is given
void k_scale_cast(hipStream_t s, int n, double alpha, double *x, float *y32, const TopZeroSweep *zt = nullptr);
void k_axpy(hipStream_t s, int n, double alpha, const double *x, double *y);
void k_scale(hipStream_t s, int n, double alpha, double *x);
void k_axpy_dev(hipStream_t s, int n, const double *alpha_dev, int idx, const double *x, double *y);
void k_copy(hipStream_t s, int n, const double *x, double *y);
void k_fill(hipStream_t s, int64_t n, double v, double *y);
void k_matadd(hipStream_t s, int64_t n, double *C, const double *D, double alpha);
// C = D + alpha C
void k_matadd_into(hipStream_t s, int64_t n, double *C, const double *D, double alpha);
// y = alpha (A x) + beta (B x) on the active rows (A, B: value arrays of L's pattern)
void k_matvec_pair_active(hipStream_t s, const DevLevel &L, const double *a, const double *b, double alpha,
                          double beta, const double *x, double *y);
void k_matvec_active(hipStream_t s, const DevLevel &L, const double *x, double *y);
// assembly
struct AsmParams {
  int what;  // 0 conv-diff, 1 mass (SUPG rhs when supg), 2 SUPG time-derivative matrix K
  double eps, b[3], c, f_val;
  int conv_kind, f_kind, supg, order;
  double rot_c[2], sin2pit;
  double h[3], dn[3];
  double delta_const;  // SUPG delta for constant b (conv_kind 0)
  // per quadrature point coefficients (ncell x nq x 6: eps, b1, b2, b3, c, f)
  // replacing the descriptor's functions, or null
  const double *qp;
  // mapped hierarchy: vertex lattice (nv^3 x 3) and per-cell h_conv, or null
  const double *vx;
  int nv;
  const double *hconv;
  // two-pass assembly: the element rows and rhs entries go here (cells in
  // natural order, no colours) instead of being scattered into CSR
  double *E, *Er;
};
void k_assemble(hipStream_t s, const DevLevel &L, const double *tab, const AsmParams &ap, double *val,
                double *rhs, int colour_begin, int colour_count);
void k_dirichlet_rows(hipStream_t s, const DevLevel &L, double *val, double *rhs);
// the element kernel over every cell at once into ap.E / ap.Er, then the
// CSR entries and rhs entries of the active rows gathered from them
void k_assemble_rows(hipStream_t s, const DevLevel &L, const double *tab, const AsmParams &ap);
void k_gather_assembly(hipStream_t s, const DevLevel &L, int64_t nnz_act, double *val, double *rhs);
// mapped levels: per-cell SUPG mesh size in convection direction into hconv (ncell)
void k_hconv_geo(hipStream_t s, const DevLevel &L, const double *tab, const AsmParams &ap, double *hconv);
// ---------------------------------------------------------------- same-level batches (kernels.hip)
// B problems with the same number of levels run one FP64 V-cycle/FGMRES in
// lockstep: every launch serves all of them (blockIdx.y = problem), each
// problem reading its own arrays through these tables in device memory.
// Problems that have finished are masked (active[b] = 0).
constexpr int kBatchRed = 32;         // reduction workgroups per problem
constexpr int kBatchScal = 256;       // device scalars per problem (H column)
constexpr int kBatchCoef = 256;       // host-written coefficients per problem: [count, c_0, c_1, ...]
constexpr int kBatchMaxRestart = 200;
// row operations of the batched stream kernel (the values of kernels.hip's RowMode)
constexpr int kSpmvMode = 0, kDefectMode = 1, kJacobiMode = 2;
// vector ids: the level's own vectors, then the finest level's Krylov space
enum BatchVec { kBSol = 0, kBRhs = 1, kBAux = 2, kBTmp = 3, kBXs = 4, kBBs = 5, kBW = 6, kBV0 = 8,
                kBZ0 = 8 + kBatchMaxRestart + 2 };
struct BatchLevel {
  int ndof = 0, nact = 0, nblk = 0;
  const int4 *bt = nullptr;
  const int *rowptr = nullptr, *col = nullptr, *diag = nullptr;
  const double *val = nullptr, *invdiag = nullptr;
  double *vec[4] = {nullptr, nullptr, nullptr, nullptr};  // sol, rhs, aux, tmp
  // the mixed form's FP32 V-cycle: the same four vectors and 1/a_ii in FP32
  float *vec32[4] = {nullptr, nullptr, nullptr, nullptr};
  float *invdiag32 = nullptr;
  // prolongation (rows of this level) and restriction (rows of the coarser level)
  const int *p_rowptr = nullptr, *p_col = nullptr, *r_rowptr = nullptr, *r_col = nullptr;
  const double *p_val = nullptr, *r_val = nullptr;
  int ndof_c = 0, nact_c = 0;
};
struct BatchProblem {
  int n = 0, na0 = 0;             // finest unknowns, coarsest active unknowns
  const double *cinv = nullptr;   // explicit inverse of the coarsest active block
  double *V = nullptr, *Z = nullptr, *xs = nullptr, *bs = nullptr, *w = nullptr;
  int64_t ld = 0;                 // column stride of V and Z
  double *red = nullptr;          // 2 x kBatchRed reduction partials
  double *scal = nullptr;         // kBatchScal device scalars
  const double *coef = nullptr;   // kBatchCoef host-written coefficients
  // the mixed form's low-synchronisation Gram-Schmidt: dot partials and L
  double *ls_part = nullptr, *ls_L = nullptr;
};
// level kernels: lev = this level's table (B entries); grid sizes from the
// largest problem.  f32: the level vectors are the FP32 ones (vec32,
// invdiag32) of the mixed form's V-cycle; the matrices are read in FP64
void kb_stream(hipStream_t s, int mode, int B, int maxblk, const BatchLevel *lev, const BatchProblem *prob,
               const int *active, int x, int f, int out, double omega, bool f32 = false);
void kb_jacobi_zero(hipStream_t s, int B, int maxn, const BatchLevel *lev, const BatchProblem *prob,
                    const int *active, int f, int out, double omega, bool f32 = false);
void kb_restrict(hipStream_t s, int B, int maxn_c, const BatchLevel *fine, const BatchProblem *prob,
                 const int *active, int x, int out_coarse, const BatchLevel *coarse, bool f32 = false);
void kb_prolong_update(hipStream_t s, int B, int maxn, const BatchLevel *fine, const BatchProblem *prob,
                       const int *active, const BatchLevel *coarse, int x_coarse, int out, double damp,
                       bool f32 = false);
// out = x on [from, ndof) (from 0: the whole vector) or, with x < 0, out = 0
void kb_copy(hipStream_t s, int B, int maxn, const BatchLevel *lev, const BatchProblem *prob, const int *active,
             int x, int out, bool dirichlet_only, bool f32 = false);
void kb_fill_zero(hipStream_t s, int B, int maxn, const BatchLevel *lev, const BatchProblem *prob, const int *active,
                  int out, bool f32 = false);
// coarsest level: out_A += Ainv d_A (accumulated in FP64)
void kb_gemv_coarse(hipStream_t s, int B, int maxna, const BatchLevel *lev, const BatchProblem *prob,
                    const int *active, int d, int out, bool f32 = false);
// FP64 level vector x -> FP32 level vector out32, and back; invdiag32 = (float) invdiag
void kb_to32(hipStream_t s, int B, int maxn, const BatchLevel *lev, const BatchProblem *prob, const int *active,
             int x, int out32);
void kb_from32(hipStream_t s, int B, int maxn, const BatchLevel *lev, const BatchProblem *prob, const int *active,
               int x32, int out);
void kb_invdiag32(hipStream_t s, int B, int maxn, const BatchLevel *lev);
// finest level reductions: partials slot 0/1; scalars at scal[idx]
void kb_dot(hipStream_t s, int B, const BatchLevel *fine, const BatchProblem *prob, const int *active, int x, int y,
            int slot);
void kb_finish(hipStream_t s, int B, const BatchProblem *prob, const int *active, int slot, int idx, bool root);
void kb_axpy_dot(hipStream_t s, int B, const BatchLevel *fine, const BatchProblem *prob, const int *active,
                 int slot_in, int alpha_idx, double sign, int x, int y, int z, int slot_out);
void kb_scale_copy(hipStream_t s, int B, const BatchLevel *fine, const BatchProblem *prob, const int *active,
                   int slot_in, int norm_idx, int x, int y);
// the low-synchronisation Gram-Schmidt step i (i < k_ls_max()) of every
// active problem (w = kBW against kBV0 .. kBV0 + i; maxn its largest n): the
// column h to scal[0 .. i], w -= V h, the partials of |w|^2 in slot `slot`
// (kb_scale_copy then normalises)
void kb_ls_step(hipStream_t s, int B, int maxn, const BatchProblem *prob, const int *active, int i, int slot);
// x *= coef[1] (problems whose coef[0] != 0)
void kb_scale(hipStream_t s, int B, const BatchLevel *fine, const BatchProblem *prob, const int *active, int x);
// xs += sum_k coef[1 + k] Z_k over k < coef[0] (problems whose coef[0] != 0; ignores the mask)
void kb_update(hipStream_t s, int B, const BatchLevel *fine, const BatchProblem *prob);
// ---------------------------------------------------------------- diagnostics (diag.hip)
// known function of the interpolation and error norms: fn 0 zero, 1 the sine
// of Examples/CD_3D/Laplace.h, 2 the Gaussian pulse centred at pc
struct FnParams {
  int fn = 0;
  double pc[3] = {0, 0, 0};
};
struct ErrGeom {
  double h[3], dn[3];
  const double *vx = nullptr;  // mapped level: vertex lattice (trilinear cells)
  int nv = 0;
};
void k_interpolate(hipStream_t s, const DevLevel &L, const double *box, int order, int n, const FnParams &f,
                   double *u);
// workgroups of the error kernel (partials: 3 per workgroup + 1)
int k_errors_blocks(int ncell);
// out (device, 4 doubles): L2, H1 semi-norm, last cell's max |e|, max |e|
void k_errors(hipStream_t s, const DevLevel &L, int order, const double *tab, const double *u, const FnParams &f,
              double *partials, double *out);
// ---------------------------------------------------------------- lattice form (lattice.hip)
void kl_jacobi(hipStream_t s, const DevLevel &L, int order, const float *x, const float *f, float *out,
               float omega);
void kl_jacobi_zero(hipStream_t s, const DevLevel &L, int order, const float *f, float *out, float omega);
void kl_defect(hipStream_t s, const DevLevel &L, int order, const float *x, const float *f, float *d);
// zero_sweep: also write the coarse level's first pre-smoothing sweep from a
// zero start, omega D^-1 coarse (needs kl_zero_sweep_fusable(C))
void kl_restrict(hipStream_t s, const DevLevel &F, const DevLevel &C, int order, const float *fine, float *coarse,
                 float *zero_sweep = nullptr, float omega = 0.f);
// every class of C has a uniform stencil (and no affine field)
bool kl_zero_sweep_fusable(const DevLevel &C);
// zt for the finest level F (x: its lt) when its classes allow it
bool kl_top_zero_sweep(const DevLevel &F, int order, float omega, float *x, TopZeroSweep &zt);
// y = A x on the lattice with FP32 x (converted exactly: the products of
// kl_spmv64 on the FP64 copy of x); xcopy (optional) receives x at every
// lattice node.  Only where kl_spmv64 takes its quad form
// (kl_spmv64_f32_ok)
bool kl_spmv64_f32_ok(const DevLevel &L, int order);
void kl_spmv64_f32(hipStream_t s, const DevLevel &L, int order, const float *x, double *y, float *xcopy);
void kl_prolong_update(hipStream_t s, const DevLevel &F, const DevLevel &C, int order, const float *coarse, float *x,
                       float damp);
// classes whose stencil is the same at every node (within 1e-12 of the
// largest entry) get the compact form; synchronises the stream (flag_host:
// pinned, 8 ints).  all_uniform false: kl_convert must fill lval / linv.
hipError_t kl_detect_uniform(hipStream_t s, DevLevel &L, int *flag_host, bool *all_uniform);
// per-node FP32 stencil values and 1/a_ii from the FP64 CSR matrix
hipError_t kl_convert(hipStream_t s, const DevLevel &L);
// constant-coefficient assembly of the active rows into val / rhs through
// the class stencils (also writes the FP32 compact stencils)
void kl_assemble_uniform(hipStream_t s, DevLevel &L, int order, const double *tab, const AsmParams &ap,
                         double *val, double *rhs);
// the interior-row CSR values from the FP64 class stencils lst64
void kl_fill_csr(hipStream_t s, const DevLevel &L, double *val);
// rotating field (affine in space), Galerkin, constant f: affine class
// stencils S0 + X SX + Y SY, CSR values of the active rows and the rhs
// csr: also write the CSR values (else the class stencils only)
void kl_assemble_affine(hipStream_t s, DevLevel &L, int order, const double *tab, const AsmParams &ap,
                        double *val, double *rhs, bool csr = true);
// FP32 per-node lattice values and 1/a_ii from the affine stencils
// (alpha_m times the mass class stencils plus alpha times the affine ones)
hipError_t kl_fill_lattice_affine(hipStream_t s, DevLevel &L, double alpha_m, double alpha);
// the same operator without per-node values: the sweeps evaluate the affine
// stencils at each node (bitwise the values the fill would store)
hipError_t kl_set_lattice_affine(hipStream_t s, DevLevel &L, double alpha_m, double alpha);
constexpr int kLst64 = 4112;
// FP64 operator in lattice order (class stencils lst64, identity rows on the
// boundary) and the order / precision changes of the lattice FGMRES
void kl_spmv64(hipStream_t s, const DevLevel &L, int order, const double *x, double *y);
// y = B x in lattice order, B entries alpha (S0 + X SX + Y SY) or, with_mass,
// fma(alpha, S0 + X SX + Y SY, M) from the level's affine and mass stencils
// (identity rows on the boundary)
void kl_affine64_apply(hipStream_t s, const DevLevel &L, int order, double alpha, bool with_mass, const double *x,
                       double *y);
void kl_defect64(hipStream_t s, const DevLevel &L, int order, const double *x, const double *f, double *d);
void kl_dof_to_lat64(hipStream_t s, const DevLevel &L, const double *v, double *l);
void kl_lat64_to_dof(hipStream_t s, const DevLevel &L, const double *l, double *v);
void kl_cast_d2f(hipStream_t s, int64_t n, const double *a, float *b);
void kl_cast_f2d(hipStream_t s, int64_t n, const float *a, double *b);
// DOF-ordered f64 <-> lattice f32 (Dirichlet DOFs map to zero)
void kl_from_dof(hipStream_t s, const DevLevel &L, const double *v, float *l);
// zero-start mixed V-cycle of levels top..1 and the exact coarse solve in
// one launch (levels[0..top]; 1 <= top <= 3)
void kl_small_vcycle(hipStream_t s, const DevLevel *levels, int top, int order, int pre, int post, float omega,
                     float damp, const double *Ainv);
void kl_to_dof(hipStream_t s, const DevLevel &L, const float *l, double *v);
// scratch: gj_scratch_doubles(na) doubles; inv: na^2 doubles
void k_invert_active(hipStream_t s, const DevLevel &L, double *scratch, double *inv);
int64_t gj_scratch_doubles(int na);
// x_A = A_AA^-1 f_A on the FP32 lattice vectors of the coarsest level (zero
// iterate, zero Dirichlet data)
void k_coarse_from_lattice(hipStream_t s, const DevLevel &C, const double *Ainv, const float *lf, float *lx);
// the inversion takes more than one launch (larger than the one-workgroup LDS form)
bool k_invert_multi_launch(int na);
void k_multi_axpy(hipStream_t s, int n, int m, const double *coef, const double *Z, int64_t ldz, double *x);
// the same with FP32 columns (converted exactly)
void k_multi_axpy32(hipStream_t s, int n, int m, const double *coef, const float *Z, int64_t ldz, double *x);
void k_copy_range(hipStream_t s, int b, int e, const double *src, double *dst);

}  // namespace cd3d

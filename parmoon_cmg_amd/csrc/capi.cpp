// C ABI of libcd3d_hip.so and the host drivers behind it: problem setup and
// upload, device assembly, the multigrid V-cycle (FP64 and mixed FP32) and the
// flexible GMRES iteration, plus the TCD3D Crank-Nicolson step.
//
// The host only sequences kernels on the handle's stream; it reads back one
// column of the Hessenberg matrix per FGMRES iteration (the Givens rotations
// and the stopping test run on the host, as in TFgmresIte::Iterate).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <new>
#include <stdexcept>

#include "../../include/cd3d_server.h"
#include "cd3d_internal.hpp"

using namespace cd3d;

struct cd3d_problem : public Problem {};

// for the serving loop's handles (not part of the C ABI): no side stream
void cd3d_internal_serving_handle(cd3d_problem *p) {
  if (p) p->side_ok = false;
}

namespace {

struct HipError : std::runtime_error {
  hipError_t code;
  HipError(hipError_t c, const char *what) : std::runtime_error(what), code(c) {}
};

#define HIPCHK(x)                                                                                   \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess) {                                                                         \
      char b_[256];                                                                                 \
      snprintf(b_, sizeof b_, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
      throw HipError(e_, b_);                                                                       \
    }                                                                                               \
  } while (0)

struct ArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

template <class T>
T *dalloc(Problem &P, size_t count) {
  if (count == 0) count = 1;
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, count * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw std::bad_alloc();
  }
  P.allocs.push_back(p);
  P.alloc_bytes.push_back(count * sizeof(T));
  P.dev_bytes += count * sizeof(T);
  return (T *)p;
}

void dfree(Problem &P, void *ptr) {
  for (size_t i = 0; i < P.allocs.size(); i++)
    if (P.allocs[i] == ptr) {
      (void)hipFree(ptr);
      P.dev_bytes -= P.alloc_bytes[i];
      P.allocs.erase(P.allocs.begin() + i);
      P.alloc_bytes.erase(P.alloc_bytes.begin() + i);
      return;
    }
}

template <class T>
void upload(Problem &P, T *dst, const std::vector<T> &src) {
  if (!src.empty()) HIPCHK(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, P.stream));
}

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// frame code of a cell: local axis a -> (global axis, sign) in 3 bits
int frame_code(const int8_t *E) {
  int code = 0;
  for (int a = 0; a < 3; a++)
    for (int d = 0; d < 3; d++)
      if (E[3 * a + d] != 0) code |= (d | (E[3 * a + d] < 0 ? 4 : 0)) << (3 * a);
  return code;
}

cd3d_status run(cd3d_problem *p, const std::function<void()> &f, bool join_before = true, bool join_after = true);

}  // namespace

// ============================================================== setup

// the macro vertices of a mapped hierarchy: finite, and every macro cell's
// trilinear map with a positive Jacobian determinant at its 8 corners (the
// hexahedra are oriented like the reference cube)
static void check_macro_vertices(int m, const double *xyz) {
  const int64_t nv = m + 1;
  for (int64_t k = 0; k < 3 * nv * nv * nv; k++)
    if (!std::isfinite(xyz[k])) throw ArgError("non-finite macro vertex coordinate");
  auto X = [&](int64_t i, int64_t j, int64_t k, int d) { return xyz[3 * (i + nv * (j + nv * k)) + d]; };
  for (int k = 0; k < m; k++)
    for (int j = 0; j < m; j++)
      for (int i = 0; i < m; i++)
        for (int t = 0; t < 8; t++) {
          const int a = t & 1, b = (t >> 1) & 1, c = (t >> 2) & 1;
          double J[3][3];
          for (int d = 0; d < 3; d++) {
            // one-sided differences along the three edges leaving corner t
            J[d][0] = (X(i + 1, j + b, k + c, d) - X(i, j + b, k + c, d));
            J[d][1] = (X(i + a, j + 1, k + c, d) - X(i + a, j, k + c, d));
            J[d][2] = (X(i + a, j + b, k + 1, d) - X(i + a, j + b, k, d));
          }
          const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                             J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                             J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
          if (!(det > 0))
            throw ArgError("macro cell (" + std::to_string(i) + "," + std::to_string(j) + "," + std::to_string(k) +
                           ") is degenerate or inverted");
        }
}

static void setup_problem(cd3d_problem &P, const cd3d_mesh_desc &m, const double *macro_xyz = nullptr) {
  if (m.order != 1 && m.order != 2) throw ArgError("order must be 1 or 2");
  if (m.m < 1 || m.k0 < 0 || m.levels < 1 || m.levels > 12) throw ArgError("invalid mesh hierarchy");
  for (int d = 0; d < 3; d++)
    if (!(m.box[d] > 0) && !macro_xyz) throw ArgError("box sides must be positive");
  if (macro_xyz) check_macro_vertices(m.m, macro_xyz);
  int64_t nfine = (int64_t)m.m << (m.k0 + m.levels - 1);
  if (nfine * m.order + 1 > 1290) throw ArgError("mesh too large (nodes per side > 1290)");
  {
    // CSR row pointers and column indices are int32: bound the finest nnz
    // (interior rows hold 27 (Q1) or on average 64 (Q2) entries)
    const double nn = (double)(nfine * m.order + 1);
    if (nn * nn * nn * (m.order == 1 ? 27 : 64) >= 2147483647.0)
      throw ArgError("mesh too large (finest level exceeds 2^31 matrix entries)");
  }
  P.mesh = m;
  P.mapped = macro_xyz != nullptr;
  if (P.mapped)
    for (int d = 0; d < 3; d++) P.mesh.box[d] = 1.0;  // unused: the geometry is the vertex lattice
  P.nb = (m.order + 1) * (m.order + 1) * (m.order + 1);
  HIPCHK(hipGetDevice(&P.device));
  HIPCHK(hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking));
  const int L = m.levels;
  P.hl.resize(L);
  P.dl.resize(L);
  for (int l = 0; l < L; l++) build_host_level(P.hl[l], m.order, m.m, m.k0 + l);
  if (P.mapped) {
    std::vector<double> vx(macro_xyz, macro_xyz + 3 * (size_t)(m.m + 1) * (m.m + 1) * (m.m + 1)), t;
    int nv = m.m;
    for (int r = 0; r < m.k0; r++, nv *= 2) {
      refine_vertices(vx, nv, t);
      vx.swap(t);
    }
    for (int l = 0; l < L; l++) {
      if (l > 0) {
        refine_vertices(P.hl[l - 1].vx, P.hl[l - 1].n, t);
        vx.swap(t);
      }
      P.hl[l].vx = vx;
    }
  }
  for (int l = 1; l < L; l++) build_transfer(P.hl[l], P.hl[l - 1], m.order);
  std::vector<double> tab;
  basis_tables(m.order, tab);
  P.tab = dalloc<double>(P, tab.size());
  upload(P, P.tab, tab);
  for (int l = 0; l < L; l++) {
    HostLevel &H = P.hl[l];
    DevLevel &D = P.dl[l];
    D.ndof = H.ndof;
    D.nact = H.nact;
    D.nnz = H.nnz;
    D.ncell = H.ncell;
    for (int d = 0; d < 3; d++) {
      D.h[d] = P.mesh.box[d] / H.n;
      D.dn[d] = P.mesh.box[d] / ((double)m.order * H.n);
    }
    if (P.mapped) {
      D.nv = H.n + 1;
      D.vx = dalloc<double>(P, H.vx.size());
      upload(P, D.vx, H.vx);
    }
    D.rowptr = dalloc<int>(P, H.ndof + 1);
    D.col = dalloc<int>(P, H.nnz);
    D.diag = dalloc<int>(P, H.ndof);
    D.val = dalloc<double>(P, H.nnz);
    upload(P, D.rowptr, H.rowptr);
    upload(P, D.col, H.kcol);
    upload(P, D.diag, H.diag);
    D.nblk = (int)H.blk.size() - 1;
    D.blk = dalloc<int>(P, H.blk.size());
    upload(P, D.blk, H.blk);
    {
      std::vector<int> bt(4 * (size_t)D.nblk);
      for (int b = 0; b < D.nblk; b++) {
        const int r0 = H.blk[b], r1 = H.blk[b + 1];
        bt[4 * b] = r0;
        bt[4 * b + 1] = r1;
        bt[4 * b + 2] = H.rowptr[r0];
        bt[4 * b + 3] = H.rowptr[r1] - H.rowptr[r0];
      }
      D.bt = (int4 *)dalloc<int>(P, bt.size());
      upload(P, (int *)D.bt, bt);
    }
    D.invdiag = dalloc<double>(P, H.ndof);
    for (double **v : {&D.sol, &D.rhs, &D.aux, &D.tmp, &D.rhs_asm}) {
      *v = dalloc<double>(P, H.ndof);
      HIPCHK(hipMemsetAsync(*v, 0, sizeof(double) * H.ndof, P.stream));
    }
    if (l > 0) {
      D.p_nnz = (int64_t)H.p_col.size();
      D.r_nnz = (int64_t)H.r_col.size();
      D.p_rowptr = dalloc<int>(P, H.p_rowptr.size());
      D.p_col = dalloc<int>(P, H.p_col.size());
      D.p_val = dalloc<double>(P, H.p_val.size());
      D.r_rowptr = dalloc<int>(P, H.r_rowptr.size());
      D.r_col = dalloc<int>(P, H.r_col.size());
      D.r_val = dalloc<double>(P, H.r_val.size());
      upload(P, D.p_rowptr, H.p_rowptr);
      upload(P, D.p_col, H.p_col);
      upload(P, D.p_val, H.p_val);
      upload(P, D.r_rowptr, H.r_rowptr);
      upload(P, D.r_col, H.r_col);
      upload(P, D.r_val, H.r_val);
    }
    // assembly data
    std::vector<int4> cells(H.ncell);
    for (int c = 0; c < H.ncell; c++)
      cells[c] = make_int4(H.org[3 * c], H.org[3 * c + 1], H.org[3 * c + 2], frame_code(&H.ax[9 * c]));
    D.cells = dalloc<int4>(P, H.ncell);
    upload(P, D.cells, cells);
    D.gn = dalloc<int>(P, H.gn.size());
    upload(P, D.gn, H.gn);
    D.colour_cells = dalloc<int>(P, H.colour_cells.size());
    upload(P, D.colour_cells, H.colour_cells);
    D.colour_start = H.colour_start;
    // lattice-stencil form of the mixed-precision V-cycle
    D.nn = H.nn;
    D.nclass = (int)H.lat.size();
    for (int c = 0; c < D.nclass; c++) D.cls[c] = H.lat[c];
    D.lat_nval = H.lat_nval;
    D.lat_nint = (int64_t)H.lat_node2dof.size();
    D.ldof2node = dalloc<int>(P, H.dof_node.size());
    upload(P, D.ldof2node, H.dof_node);
    D.lsid = dalloc<uint16_t>(P, H.lat_sid.size());
    upload(P, D.lsid, H.lat_sid);
    D.lfirst = dalloc<int>(P, H.lat_first.size());
    upload(P, D.lfirst, H.lat_first);
    D.lnode2dof = dalloc<int>(P, H.lat_node2dof.size());
    upload(P, D.lnode2dof, H.lat_node2dof);
    D.lcst = dalloc<float>(P, 8 * 128);
    D.lscale = dalloc<double>(P, 8);
    D.lflag = dalloc<int>(P, 8);
    D.lelem = dalloc<double>(P, (size_t)3 * P.nb * P.nb + P.nb);
    D.lst64 = dalloc<double>(P, kLst64);
    const size_t nn3 = (size_t)H.nn * H.nn * H.nn;
    for (float **v : {&D.lx, &D.lf, &D.la, &D.lt}) {
      *v = dalloc<float>(P, nn3);  // boundary entries stay zero for good
      HIPCHK(hipMemsetAsync(*v, 0, sizeof(float) * nn3, P.stream));
    }
    HIPCHK(hipStreamSynchronize(P.stream));  // host vectors above go out of scope
  }
  // exact coarse solve storage
  const int na = P.dl[0].nact;
  P.cinv = dalloc<double>(P, (size_t)na * na);
  // Krylov workspace (allocated for restart 20; grown on demand)
  P.red = dalloc<double>(P, 4 * kRedBlocks);
  P.dscal = dalloc<double>(P, P.nscal);
  HIPCHK(hipHostMalloc((void **)&P.hscal, P.nscal * sizeof(double), hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void **)&P.hflag, 64 * sizeof(int), hipHostMallocDefault));
}

// basis column stride: a multiple of 32 doubles, so that every column is
// 256-byte aligned (the vector kernels stream 16-byte pairs)
static int64_t krylov_ld(int64_t N) { return (N + 31) / 32 * 32; }

static void ensure_krylov(cd3d_problem &P, int restart) {
  if (restart <= P.krylov_restart) return;
  const int N = P.dl.back().ndof;
  // the basis grows; w, xs (which carries the TCD3D state) and bs stay
  if (P.V) {
    dfree(P, P.V);
    dfree(P, P.Z);
  }
  P.krylov_ld = krylov_ld(N);
  P.V = dalloc<double>(P, (size_t)(restart + 1) * P.krylov_ld);
  P.Z = dalloc<double>(P, (size_t)(restart + 1) * P.krylov_ld);
  // the H column (restart + 1 entries) and y (restart entries) share dscal
  if (restart + 2 > P.yoff) {
    dfree(P, P.dscal);
    (void)hipHostFree(P.hscal);
    P.hscal = nullptr;
    P.yoff = restart + 2;
    P.nscal = 2 * P.yoff;
    P.dscal = dalloc<double>(P, P.nscal);
    HIPCHK(hipHostMalloc((void **)&P.hscal, P.nscal * sizeof(double), hipHostMallocDefault));
  }
  if (!P.w) {
    P.w = dalloc<double>(P, N);
    P.xs = dalloc<double>(P, N);
    P.bs = dalloc<double>(P, N);
  }
  P.krylov_restart = restart;
}

// ============================================================== assembly

// CD3D_CELLWISE_ASSEMBLY=1 assembles constant coefficients cell by cell too
// (the general kernel), for comparisons; read at every assembly
// the class-stencil assembly paths need the uniform lattice of bricks: off on
// mapped hierarchies and under CD3D_CELLWISE_ASSEMBLY=1
static bool uniform_assembly_off(const cd3d_problem &P) {
  const char *e = getenv("CD3D_CELLWISE_ASSEMBLY");
  return P.mapped || (e && atoi(e) != 0);
}

static double sdfem_delta_host(double hconv, double eps, const double *b, int order) {
  if (fabs(eps) < 1e-20) eps = 1e-20;
  const double nb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  if (!(nb > 0)) return 0.0;
  const double nu = 1.0 / order, alpha = nu * nb * hconv / (2 * eps);
  return nu * hconv * (1 / tanh(alpha) - 1 / alpha) / (2 * nb);
}

static double hconv_host(const double *h, const double *b) {
  const double nb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  double den = 0;
  for (int s = 0; s < 8; s++) {
    const double v = ((s & 1) ? -b[0] : b[0]) / h[0] + ((s & 2) ? -b[1] : b[1]) / h[1] + ((s & 4) ? -b[2] : b[2]) / h[2];
    den += fabs(v) / 4;
  }
  if (den < 1e-10) return sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
  return 2 * nb / den;
}

static void check_coeffs(const cd3d_coeffs &co) {
  const double v[] = {co.eps, co.b[0], co.b[1], co.b[2], co.c, co.f_val, co.t, co.rot_c[0], co.rot_c[1]};
  for (double x : v)
    if (!std::isfinite(x)) throw ArgError("non-finite coefficient");
  if (co.conv_kind != 0 && co.conv_kind != 1) throw ArgError("conv_kind must be 0 or 1");
  if (co.f_kind != 0 && co.f_kind != 1) throw ArgError("f_kind must be 0 or 1");
}

// kernel parameters of the element integration on level D (coefficients,
// cell sizes, the SUPG delta of a constant convection field)
static AsmParams asm_params(const cd3d_problem &P, const DevLevel &D, const cd3d_coeffs &co, int what) {
  AsmParams ap{};
  ap.what = what;
  ap.eps = co.eps;
  for (int d = 0; d < 3; d++) ap.b[d] = co.b[d];
  ap.c = co.c;
  ap.f_val = co.f_val;
  ap.conv_kind = co.conv_kind;
  ap.f_kind = co.f_kind;
  ap.supg = co.supg;
  ap.order = P.mesh.order;
  ap.rot_c[0] = co.rot_c[0];
  ap.rot_c[1] = co.rot_c[1];
  ap.sin2pit = sin(2 * M_PI * co.t);
  for (int d = 0; d < 3; d++) {
    ap.h[d] = D.h[d];
    ap.dn[d] = D.dn[d];
  }
  ap.delta_const = 0;
  if (co.supg && co.conv_kind == 0)
    ap.delta_const = sdfem_delta_host(hconv_host(D.h, co.b), co.eps, co.b, P.mesh.order);
  ap.qp = D.qp;
  ap.vx = P.mapped ? D.vx : nullptr;
  ap.nv = D.nv;
  ap.hconv = D.hconv;
  return ap;
}

// CD3D_ASM_SCATTER=1: the cell-by-cell assembly scatters element rows into
// CSR colour by colour (the one-pass form) instead of writing them out and
// gathering (read at every assembly; the two give the same bits)
static bool scatter_assembly(const cd3d_problem &P, int l) {
  const char *e = getenv("CD3D_ASM_SCATTER");
  if (e && atoi(e) != 0) return true;
  // the gather lists index the element rows with int32
  return (int64_t)P.dl[l].ncell * P.nb * P.nb >= ((int64_t)1 << 31);
}

// element-row buffers and gather lists of level l (first use)
static void ensure_gather(cd3d_problem &P, int l) {
  DevLevel &D = P.dl[l];
  if (D.E) return;
  std::vector<int> slots, gptr, gidx, rptr, ridx;
  build_slots(P.hl[l], P.mesh.order, slots);
  build_gather(P.hl[l], P.mesh.order, slots, gptr, gidx, rptr, ridx);
  slots.clear();
  slots.shrink_to_fit();
  D.gptr = dalloc<int>(P, gptr.size());
  D.gidx = dalloc<int>(P, gidx.size());
  D.rptr = dalloc<int>(P, rptr.size());
  D.ridx = dalloc<int>(P, ridx.size());
  upload(P, D.gptr, gptr);
  upload(P, D.gidx, gidx);
  upload(P, D.rptr, rptr);
  upload(P, D.ridx, ridx);
  D.E = dalloc<double>(P, (size_t)D.ncell * P.nb * P.nb);
  D.Er = dalloc<double>(P, (size_t)D.ncell * P.nb);
  HIPCHK(hipStreamSynchronize(P.stream));
}

static void ensure_slots(cd3d_problem &P, int l) {
  DevLevel &D = P.dl[l];
  if (D.slots) return;  // slot maps only for the cell-by-cell path (nb^2 ints per cell)
  std::vector<int> slots;
  build_slots(P.hl[l], P.mesh.order, slots);
  D.slots = dalloc<int>(P, slots.size());
  upload(P, D.slots, slots);
  HIPCHK(hipStreamSynchronize(P.stream));
}

// assemble level l: what 0 -> conv-diff into `val`, 1 -> mass into `val`;
// rhs into D.rhs_asm (TAssembleMat3D::Reset + Assemble3D, AssembleMat3D.C:204,230).
// Returns which path ran: 1 constant coefficients (it leaves the FP32 compact
// stencils of `val` behind), 2 rotating field (affine class stencils in
// lst64), 0 cell by cell.
// csr false: a level assembled through affine class stencils gets the
// stencils (and rhs) only, not its CSR values
// what 2 (the SUPG time-derivative matrix K) writes no rhs: rhs names a
// scratch vector then (default: the level's assembled rhs)
static int assemble_level(cd3d_problem &P, int l, const cd3d_coeffs &co, int what, double *val, bool csr = true,
                          double *rhs = nullptr) {
  if (val == P.dl[l].val) P.dl[l].invdiag_ok = false;
  DevLevel &D = P.dl[l];
  if (!rhs) rhs = D.rhs_asm;
  const AsmParams ap = asm_params(P, D, co, what);
  if (what == 0) P.last_co = co;
  // constant coefficients (and the plain mass matrix under any convection):
  // one element matrix per level through the class stencils
  const bool uniform = (co.conv_kind == 0 || (what == 1 && !co.supg)) && !D.qp;
  if (uniform && co.f_kind == 0 && D.nclass > 0 && !uniform_assembly_off(P)) {
    // every active row is written: no clearing pass
    kl_assemble_uniform(P.stream, D, P.mesh.order, P.tab, ap, val, rhs);
    k_dirichlet_rows(P.stream, D, val, rhs);
    HIPCHK(hipGetLastError());
    return 1;
  }
  if (what == 0 && co.conv_kind == 1 && co.f_kind == 0 && !co.supg && !D.qp && D.nclass > 0 &&
      !uniform_assembly_off(P)) {
    kl_assemble_affine(P.stream, D, P.mesh.order, P.tab, ap, val, rhs, csr);
    k_dirichlet_rows(P.stream, D, val, rhs);
    HIPCHK(hipGetLastError());
    return 2;
  }
  AsmParams apc = ap;
  if (P.mapped && co.supg) {  // per-cell h_conv of the general hexahedra
    if (!D.hconv) D.hconv = dalloc<double>(P, (size_t)D.ncell);
    apc.hconv = D.hconv;
    k_hconv_geo(P.stream, D, P.tab, apc, D.hconv);
  }
  if (!scatter_assembly(P, l)) {
    // two passes: every element row written out in one launch, then each CSR
    // entry and rhs entry of the active rows summed from its contributions
    // (colour order); Dirichlet rows: zero, then the identity row
    ensure_gather(P, l);
    apc.E = D.E;
    apc.Er = D.Er;
    k_assemble_rows(P.stream, D, P.tab, apc);
    const int64_t nnz_act = P.hl[l].rowptr[D.nact];
    k_gather_assembly(P.stream, D, nnz_act, val, rhs);
    if (D.nnz > nnz_act) HIPCHK(hipMemsetAsync(val + nnz_act, 0, sizeof(double) * (D.nnz - nnz_act), P.stream));
    k_dirichlet_rows(P.stream, D, val, rhs);
    HIPCHK(hipGetLastError());
    return 0;
  }
  ensure_slots(P, l);
  HIPCHK(hipMemsetAsync(val, 0, sizeof(double) * D.nnz, P.stream));
  HIPCHK(hipMemsetAsync(rhs, 0, sizeof(double) * D.ndof, P.stream));
  for (int k = 0; k < 8; k++)
    k_assemble(P.stream, D, P.tab, apc, val, rhs, D.colour_start[k], D.colour_start[k + 1] - D.colour_start[k]);
  k_dirichlet_rows(P.stream, D, val, rhs);
  HIPCHK(hipGetLastError());
  return 0;
}

// explicit inverse of the coarsest matrix on the side stream.  fork_coarse
// marks the point of the main stream where the level-0 matrix is final;
// launch_coarse_inverse enqueues the inversion after that point (called once
// the main stream has other work queued, so that it runs beside that work).
// The coarse inverse goes to the side stream unless the handle belongs to
// the serving loop or CD3D_SIDE_STREAM=0.  Measured (one default bench run
// each): one handle, configuration 2 13.29 vs 13.47 ms and configuration 3
// 3.32 vs 3.50 ms per step with the side stream; the serving stream at 400
// requests/s 54 vs 79 mixed solves/s.
static bool side_stream_on(const cd3d_problem &P) {
  const char *e = getenv("CD3D_SIDE_STREAM");
  return P.side_ok && !(e && e[0] == '0');
}

static void fork_coarse(cd3d_problem &P) {
  if (!k_invert_multi_launch(P.dl[0].nact) || !side_stream_on(P)) return;  // inverted on the main stream
  if (!P.side) {
    // created on first need only: HIP hands streams to its hardware queues
    // round-robin, and a side stream per handle would leave every other
    // queue to the side streams of concurrently served handles
    HIPCHK(hipStreamCreateWithFlags(&P.side, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&P.ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&P.ev_join, hipEventDisableTiming));
  }
  HIPCHK(hipEventRecord(P.ev_fork, P.stream));
  P.fork_recorded = true;
}

static void launch_coarse_inverse(cd3d_problem &P) {
  DevLevel &C = P.dl[0];
  const int na = C.nact;
  if (!P.fork_recorded) fork_coarse(P);
  P.fork_recorded = false;  // (fork_coarse records nothing when the inverse stays on the main stream)
  if (na > 0 && (!k_invert_multi_launch(na) || !side_stream_on(P))) {
    // one short launch: the fork and join would cost more than they hide
    if (!P.gj_scratch) P.gj_scratch = dalloc<double>(P, (size_t)gj_scratch_doubles(na));
    k_invert_active(P.stream, C, P.gj_scratch, P.cinv);
    HIPCHK(hipGetLastError());
  } else if (na > 0) {
    if (!P.gj_scratch) P.gj_scratch = dalloc<double>(P, (size_t)gj_scratch_doubles(na));
    HIPCHK(hipStreamWaitEvent(P.side, P.ev_fork, 0));
    k_invert_active(P.side, C, P.gj_scratch, P.cinv);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(P.ev_join, P.side));
    P.coarse_pending = true;
  }
  P.cinv_valid = true;
}

// the main stream waits for the coarse inverse (before it is used, before
// the level-0 matrix changes again, and at the end of every call)
static void coarse_join(cd3d_problem &P) {
  if (!P.coarse_pending) return;
  HIPCHK(hipStreamWaitEvent(P.stream, P.ev_join, 0));
  P.coarse_pending = false;
}

// CD3D_SMALL_NODES overrides the node bound of the one-workgroup V-cycle
// program (0: off), read once per solve and at each refresh
static long long small_nodes_bound() {
  const char *e = getenv("CD3D_SMALL_NODES");
  return e ? atoll(e) : 3000LL;
}

// after a change of the level matrices: the FP32 lattice values of every
// level (1/a_ii of the FP64 form is recomputed on its first use)
static void refresh_f32(cd3d_problem &P) {
  for (size_t l = 0; l < P.dl.size(); l++) {
    DevLevel &D = P.dl[l];
    D.invdiag_ok = false;  // only the FP64 smoother reads it: computed there on first use
    if (l == 0) D.lat_fresh = 0;  // the coarsest level has no FP32 stencils
    if (l > 0) {
      auto per_node = [&](bool with_map) {  // per-node FP32 stencils (4 bytes per nonzero), on first need
        if (!D.lval) {
          D.lval = dalloc<float>(P, (size_t)D.lat_nval);
          D.linv = dalloc<float>(P, (size_t)D.lat_nint);
        }
        if (with_map && !D.lperm) {
          D.lperm = dalloc<int>(P, P.hl[l].lat_perm.size());
          upload(P, D.lperm, P.hl[l].lat_perm);
        }
      };
      const long long nn = D.nn;
      if (D.lat_fresh == 2 && nn * nn * nn > small_nodes_bound()) {
        // affine class stencils evaluated by the sweeps node by node (a level
        // this large never runs in the one-workgroup program, which reads
        // per-node values)
        HIPCHK(kl_set_lattice_affine(P.stream, D, D.lat_alpha_m, D.lat_alpha));
        D.lat_fresh = 0;
        D.lat_exact64 = false;
        D.lat_aff64 = D.lat_alpha_m == 0.0 || D.lat_alpha_m == 1.0;
      } else if (D.lat_fresh == 2) {  // straight from the affine class stencils
        per_node(false);
        HIPCHK(kl_fill_lattice_affine(P.stream, D, D.lat_alpha_m, D.lat_alpha));
        D.lat_fresh = 0;
        D.lat_exact64 = false;
        D.lat_aff64 = D.lat_alpha_m == 0.0 || D.lat_alpha_m == 1.0;
      } else {
        D.lat_affine = false;
        D.lat_aff64 = false;
        bool all_uniform = false;
        HIPCHK(kl_detect_uniform(P.stream, D, P.hflag, &all_uniform));
        if (!all_uniform) {
          per_node(true);
          HIPCHK(kl_convert(P.stream, D));
        }
      }
    }
  }
}

// ============================================================== V-cycle

namespace {

// Mixed V-cycles: the bottom levels, up to the highest level with at most
// this many lattice nodes (and at most kSmallMaxLevels of them above an exact
// coarse solve of at most kSmallCoarseMax unknowns), run as one launch
// (kl_small_vcycle).  One workgroup pays off only while a level's nodes take
// about one round of the CU (measured: 4913-node Q2 levels and a 343-unknown
// coarse solve are several times slower than the launch-per-step form).
// The node bound is small_nodes_bound().
constexpr int kSmallMaxLevels = 3, kSmallCoarseMax = 128;

static int small_top(const cd3d_problem &P, const cd3d_solver_params &sp) {
  if (sp.cycle != 1 || sp.coarse_smoother != 17 || P.dl.empty() || P.dl[0].nact > kSmallCoarseMax) return 0;
  int top = 0;
  for (int i = 1; i < (int)P.dl.size() && i <= kSmallMaxLevels; i++) {
    const long long nn = P.dl[i].nn;
    // the program reads compact or per-node FP32 values (not the node-by-node affine form)
    if (nn * nn * nn > small_nodes_bound() || P.dl[i].nclass <= 0 || P.dl[i].lat_affine) break;
    top = i;
  }
  return top;
}

struct VC {
  cd3d_problem &P;
  const cd3d_solver_params &sp;
  int rec[32];     // mg_recursions (TMultiGrid3D::SetRecursion, MultiGrid3D.C:690-703)
  int small = -1;  // the level whose zero-start visit runs as one launch (mixed V-cycle; 0: none)
  // mixed V-cycle: bit i set when the restriction into level i also wrote
  // its first pre-smoothing sweep from zero (into lt); CD3D_FUSE_ZERO_SWEEP=0
  // launches that sweep instead
  unsigned zero_ready = 0;
  bool fuse_zero = true;
  VC(cd3d_problem &p, const cd3d_solver_params &s) : P(p), sp(s) {
    const char *e = getenv("CD3D_FUSE_ZERO_SWEEP");
    fuse_zero = !(e && e[0] == '0');
    reset();
  }
  void reset() {
    zero_ready = 0;
    rec[1] = 1;
    for (int k = 2; k < 32; k++) rec[k] = sp.cycle > 0 ? sp.cycle : 2;
  }
};

static void refresh_coarse_and_f32(cd3d_problem &P) {
  launch_coarse_inverse(P);
  refresh_f32(P);
}

// coarse level: x starts at zero (zero_start) or holds an iterate
void coarse64(VC &v, bool zero_start) {
  DevLevel &C = v.P.dl[0];
  coarse_join(v.P);
  if (zero_start) HIPCHK(hipMemsetAsync(C.sol, 0, sizeof(double) * C.ndof, v.P.stream));
  if (v.sp.coarse_smoother == 17) {
    // SolveExact once: x_A += A_AA^-1 (f - A x)_A, x_D = f_D; the loop of
    // TMultiGrid3D::Smooth ends after one exact solve (MultiGrid3D.C:205-258)
    k_copy_range(v.P.stream, C.nact, C.ndof, C.rhs, C.sol);
    k_defect(v.P.stream, C, C.sol, C.rhs, C.aux);
    k_gemv_dense(v.P.stream, C.nact, v.P.cinv, C.aux, C.sol);
  } else {
    k_coarse_smooth(v.P.stream, C, C.sol, C.rhs, C.aux, v.sp.coarse_smoother, v.sp.omega, v.sp.coarse_maxit,
                    v.sp.coarse_red);
  }
}

// the multicolour SSOR row lists of level l on the device (first use)
static void ensure_mc(cd3d_problem &P, int l) {
  DevLevel &D = P.dl[l];
  if (D.mc_rows) return;
  HostLevel &H = P.hl[l];
  if (H.mc_start.empty()) build_mc_colours(H, P.mesh.order);
  D.mc_rows = dalloc<int>(P, H.mc_rows.size());
  upload(P, D.mc_rows, H.mc_rows);
  D.mc_start = H.mc_start;
  HIPCHK(hipStreamSynchronize(P.stream));
}

// the wavefront schedules of the row-order SOR (and, for SSOR, the backward
// sweep) of level l on the device (first use)
static void ensure_gs(cd3d_problem &P, int l, bool backward) {
  DevLevel &D = P.dl[l];
  for (int b = 0; b <= (backward ? 1 : 0); b++) {
    if (D.gs_rows[b]) continue;
    std::vector<int> rows, start;
    build_gs_schedule(P.hl[l], b == 1, rows, start);
    D.gs_rows[b] = dalloc<int>(P, rows.size());
    if (!rows.empty()) upload(P, D.gs_rows[b], rows);
    D.gs_start[b] = dalloc<int>(P, start.size());
    upload(P, D.gs_start[b], start);
    D.gs_nlev[b] = (int)start.size() - 1;
    HIPCHK(hipStreamSynchronize(P.stream));  // the host vectors end here
  }
}

// TMultiGrid3D::Cycle (MultiGrid3D.C:423-669), V-cycle, FP64.  On entry the
// level's rhs is set; zero_start: the iterate is zero and the Dirichlet part
// of the rhs is zero, so the first Jacobi sweep needs only the diagonal.
// Smoothers 2 (SOR), 3 (SSOR) and 13 (multicolour SSOR) relax L.sol in place.
void cycle64(VC &v, int i, bool zero_start) {
  cd3d_problem &P = v.P;
  if (i == 0) {
    coarse64(v, zero_start);
    return;
  }
  DevLevel &L = P.dl[i], &C = P.dl[i - 1];
  const int sm = v.sp.smoother;
  if (sm == 13 || sm == 2 || sm == 3) {
    if (sm == 13) ensure_mc(P, i);
    else ensure_gs(P, i, sm == 3);
    auto relax = [&]() {
      if (sm == 13) k_mcssor(P.stream, L, L.sol, L.rhs, v.sp.omega);
      else k_gs_sweep(P.stream, L, L.sol, L.rhs, v.sp.omega, sm);
    };
    if (zero_start) HIPCHK(hipMemsetAsync(L.sol, 0, sizeof(double) * L.ndof, P.stream));
    for (int k = 0; k < v.sp.pre; k++) relax();
    if (v.sp.pre == 0 && zero_start) k_copy_range(P.stream, L.nact, L.ndof, L.rhs, L.sol);
    k_defect(P.stream, L, L.sol, L.rhs, L.aux);
    k_restrict(P.stream, L, C, L.aux, C.rhs);
    for (int r = 0; r < v.rec[i]; r++) cycle64(v, i - 1, r == 0);
    if (v.sp.cycle < 1) v.rec[i] = 1;  // F-cycle
    k_prolong_update(P.stream, L, C, C.sol, L.sol, v.sp.gmg_damp);
    for (int k = 0; k < v.sp.post; k++) relax();
    return;
  }
  double *x = L.sol, *t = L.tmp;
  if (!L.invdiag_ok) {
    k_invdiag(P.stream, L);
    L.invdiag_ok = true;
  }
  for (int k = 0; k < v.sp.pre; k++) {
    if (k == 0 && zero_start) k_jacobi_zero(P.stream, L, L.rhs, t, v.sp.omega);
    else k_jacobi(P.stream, L, x, L.rhs, t, v.sp.omega);
    std::swap(x, t);
  }
  if (v.sp.pre == 0 && zero_start) {
    HIPCHK(hipMemsetAsync(x, 0, sizeof(double) * L.ndof, P.stream));
    k_copy_range(P.stream, L.nact, L.ndof, L.rhs, x);
  }
  k_defect(P.stream, L, x, L.rhs, L.aux);
  k_restrict(P.stream, L, C, L.aux, C.rhs);
  for (int r = 0; r < v.rec[i]; r++) cycle64(v, i - 1, r == 0);
  if (v.sp.cycle < 1) v.rec[i] = 1;  // F-cycle
  k_prolong_update(P.stream, L, C, C.sol, x, v.sp.gmg_damp);
  for (int k = 0; k < v.sp.post; k++) {
    k_jacobi(P.stream, L, x, L.rhs, t, v.sp.omega);
    std::swap(x, t);
  }
  if (x != L.sol) k_copy(P.stream, L.ndof, x, L.sol);
}

// mixed precision on the lattice-stencil form: the same cycle, with vectors
// in node-lattice order whose boundary entries are zero (the Dirichlet part
// of every preconditioner rhs and iterate vanishes)
void coarseL(VC &v, bool zero_start) {
  DevLevel &C = v.P.dl[0];
  if (zero_start && v.sp.coarse_smoother == 17) {
    // inside the preconditioner the Dirichlet data are zero, so the defect of
    // the zero iterate is the rhs: one launch, same arithmetic as below
    coarse_join(v.P);
    k_coarse_from_lattice(v.P.stream, C, v.P.cinv, C.lf, C.lx);
    return;
  }
  kl_to_dof(v.P.stream, C, C.lf, C.rhs);
  if (!zero_start) kl_to_dof(v.P.stream, C, C.lx, C.sol);
  coarse64(v, zero_start);
  kl_from_dof(v.P.stream, C, C.sol, C.lx);
}

void cycleL(VC &v, int i, bool zero_start) {
  cd3d_problem &P = v.P;
  if (i == 0) {
    coarseL(v, zero_start);
    return;
  }
  if (v.small < 0) v.small = small_top(P, v.sp);
  if (zero_start && i == v.small) {
    coarse_join(P);
    kl_small_vcycle(P.stream, P.dl.data(), i, P.mesh.order, v.sp.pre, v.sp.post, (float)v.sp.omega,
                    (float)v.sp.gmg_damp, P.cinv);
    return;
  }
  DevLevel &L = P.dl[i], &C = P.dl[i - 1];
  const int p = P.mesh.order;
  float *x = L.lx, *t = L.lt;
  const float om = (float)v.sp.omega;
  const bool ready = zero_start && (v.zero_ready >> i & 1u);
  v.zero_ready &= ~(1u << i);
  for (int k = 0; k < v.sp.pre; k++) {
    if (k == 0 && zero_start) {
      if (!ready) kl_jacobi_zero(P.stream, L, p, L.lf, t, om);
    } else {
      kl_jacobi(P.stream, L, p, x, L.lf, t, om);
    }
    std::swap(x, t);
  }
  const size_t nn3 = (size_t)L.nn * L.nn * L.nn;
  if (v.sp.pre == 0 && zero_start) HIPCHK(hipMemsetAsync(x, 0, sizeof(float) * nn3, P.stream));
  kl_defect(P.stream, L, p, x, L.lf, L.la);
  // the coarse level's zero-start sweep rides on the restriction when that
  // level runs the launch-per-sweep cycle (not the direct solve of level 0 or
  // the one-launch bottom) with uniform stencils
  const bool fuse = v.fuse_zero && v.sp.pre > 0 && i - 1 > 0 && i - 1 != v.small && i - 1 < 32 &&
                    kl_zero_sweep_fusable(C);
  kl_restrict(P.stream, L, C, p, L.la, C.lf, fuse ? C.lt : nullptr, om);
  if (fuse) v.zero_ready |= 1u << (i - 1);
  for (int r = 0; r < v.rec[i]; r++) cycleL(v, i - 1, r == 0);
  if (v.sp.cycle < 1) v.rec[i] = 1;  // F-cycle
  kl_prolong_update(P.stream, L, C, p, C.lx, x, (float)v.sp.gmg_damp);
  for (int k = 0; k < v.sp.post; k++) {
    kl_jacobi(P.stream, L, p, x, L.lf, t, om);
    std::swap(x, t);
  }
  if (x != L.lx) HIPCHK(hipMemcpyAsync(L.lx, x, sizeof(float) * nn3, hipMemcpyDeviceToDevice, P.stream));
}

// TMultiGridScaIte::Iterate (MultiGridScaIte.C:63-104): z = V(r), zero start.
// rhs_dirichlet_zero: r vanishes on Dirichlet DOFs (always true for the
// Krylov vectors of FGMRES), which lets the first sweep skip the matrix.
void precondition(VC &v, const double *r, double *z, bool rhs_dirichlet_zero) {
  cd3d_problem &P = v.P;
  v.reset();
  const int top = (int)P.dl.size() - 1;
  DevLevel &F = P.dl[top];
  if (v.sp.precision == CD3D_MIXED && top > 0 && rhs_dirichlet_zero && v.sp.smoother == 1) {
    kl_from_dof(P.stream, F, r, F.lf);
    cycleL(v, top, true);
    kl_to_dof(P.stream, F, F.lx, z);
    return;
  }
  k_copy(P.stream, F.ndof, r, F.rhs);
  if (rhs_dirichlet_zero) {
    cycle64(v, top, true);
  } else {
    // SetDirichletNodes (MultiGrid3D.C:671-688): x = 0, x_D = r_D
    HIPCHK(hipMemsetAsync(F.sol, 0, sizeof(double) * F.ndof, P.stream));
    k_copy_range(P.stream, F.nact, F.ndof, F.rhs, F.sol);
    cycle64(v, top, false);
  }
  k_copy(P.stream, F.ndof, F.sol, z);
}

// GeneratePlaneRotation / ApplyPlaneRotation (FgmresIte.C:168-195)
void gen_rot(double dx, double dy, double &cs, double &sn) {
  if (dy == 0.0) { cs = 1.0; sn = 0.0; }
  else if (fabs(dy) > fabs(dx)) { double t = dx / dy; sn = 1.0 / sqrt(1.0 + t * t); cs = t * sn; }
  else { double t = dy / dx; cs = 1.0 / sqrt(1.0 + t * t); sn = t * cs; }
}
void app_rot(double &dx, double &dy, double cs, double sn) {
  const double t = cs * dx + sn * dy;
  dy = -sn * dx + cs * dy;
  dx = t;
}

double dev_norm(cd3d_problem &P, int n, const double *a) {
  k_dot(P.stream, n, a, a, P.red, P.dscal);
  HIPCHK(hipMemcpyAsync(P.hscal, P.dscal, sizeof(double), hipMemcpyDeviceToHost, P.stream));
  HIPCHK(hipStreamSynchronize(P.stream));
  return sqrt(P.hscal[0]);
}

// TFgmresIte::Iterate, flexible branch (FgmresIte.C:219-595) on the finest
// level: x (P.xs) and b (P.bs) on the device.  Returns the iteration count.
// The vector space FGMRES runs in: DOF order with the CSR matrix (the plain
// form), or node-lattice order with the class stencils (mixed form on a
// constant-coefficient level); x and b are device vectors of that space.
struct Space {
  int n;
  double *x, *b;
  std::function<void(const double *, const double *, double *)> defect;
  std::function<void(const double *, double *)> matvec;
  // precond(v, r, z, r32_state): z = M^-1 r; r32_state 1: r32 already holds
  // (float) r, written by the kernel that produced r; 2: that kernel also
  // wrote the finest level's first pre-smoothing sweep from zero (into lt);
  // 0: neither
  std::function<void(VC &, const double *, double *, int)> precond;
  float *r32 = nullptr;  // FP32 copy of the preconditioner input (lattice space)
  DevLevel *lat_top = nullptr;  // the finest level of the lattice V-cycle
  int lat_order = 0;
  // the Krylov columns z_i kept in FP32 (the V-cycle's FP32 output, exact):
  // precond leaves z in lat_top->lx (z == nullptr) and the product
  // kl_spmv64_f32 reads it there and writes the column
  bool z32 = false;
  // orthogonalise by the low-synchronisation form of MGS (k_ls_step)
  // instead of MGS step by step (the FP64 form keeps the reference's MGS)
  bool lowsync = false;
};

Space dof_space(cd3d_problem &P) {
  DevLevel &F = P.dl.back();
  Space S;
  S.n = F.ndof;
  S.x = P.xs;
  S.b = P.bs;
  S.defect = [&P, &F](const double *x, const double *f, double *d) { k_defect(P.stream, F, x, f, d); };
  S.matvec = [&P, &F](const double *x, double *y) { k_spmv(P.stream, F, x, y); };
  S.precond = [](VC &v, const double *r, double *z, int) { precondition(v, r, z, true); };
  return S;
}

Space lattice_space(cd3d_problem &P) {
  DevLevel &F = P.dl.back();
  const int p = P.mesh.order, top = (int)P.dl.size() - 1;
  Space S;
  S.n = F.ndof;  // every lattice node is a DOF of the finest level
  S.x = P.xl;
  S.b = P.bl;
  S.defect = [&P, &F, p](const double *x, const double *f, double *d) { kl_defect64(P.stream, F, p, x, f, d); };
  S.matvec = [&P, &F, p](const double *x, double *y) { kl_spmv64(P.stream, F, p, x, y); };
  // the Krylov vectors vanish on the boundary, as the V-cycle's rhs must
  S.r32 = F.lf;
  // CD3D_FGMRES_MGS=1: MGS step by step as in the FP64 form
  const char *em = getenv("CD3D_FGMRES_MGS");
  S.lowsync = !(em && em[0] == '1');
  S.lat_top = &F;
  S.lat_order = p;
  // CD3D_FGMRES_Z32=0: FP64 columns, written by a cast after each V-cycle
  const char *ez = getenv("CD3D_FGMRES_Z32");
  S.z32 = !(ez && ez[0] == '0') && kl_spmv64_f32_ok(F, p);
  S.precond = [&P, &F, top](VC &v, const double *r, double *z, int r32_state) {
    v.reset();
    if (r32_state == 0) kl_cast_d2f(P.stream, F.ndof, r, F.lf);
    if (r32_state == 2 && top < 32) v.zero_ready |= 1u << top;
    cycleL(v, top, true);
    if (z) kl_cast_f2d(P.stream, F.ndof, F.lx, z);
  };
  return S;
}

int fgmres_iterate(cd3d_problem &P, const cd3d_solver_params &sp, const Space &S, double &r0, cd3d_status &status);

// the buffers of the low-synchronisation Gram-Schmidt step (first use)
void ensure_ls(cd3d_problem &P, int n) {
  if (!P.ls_L) P.ls_L = dalloc<double>(P, (size_t)k_ls_max() * k_ls_max());
  if (P.ls_blocks >= k_ls_blocks(n)) return;
  if (P.ls_part) dfree(P, P.ls_part);
  P.ls_blocks = k_ls_blocks(n);
  P.ls_part = dalloc<double>(P, (size_t)(2 * k_ls_max() + 1) * P.ls_blocks);
}

// rfinal (optional): |b - A x| recomputed in the space after the iteration
int fgmres(cd3d_problem &P, const cd3d_solver_params &sp, const Space &S, double &r0, cd3d_status &status,
           double *rfinal = nullptr) {
  const int its = fgmres_iterate(P, sp, S, r0, status);
  if (rfinal) {
    S.defect(S.x, S.b, P.w);
    *rfinal = dev_norm(P, S.n, P.w);
  }
  return its;
}

int fgmres_iterate(cd3d_problem &P, const cd3d_solver_params &sp, const Space &S, double &r0, cd3d_status &status) {
  const int N = S.n;
  int R = std::min(sp.restart, sp.maxit);
  if (R <= 0) R = 10;
  ensure_krylov(P, R);
  VC v{P, sp};
  double *V = P.V, *Z = P.Z, *w = P.w;
  const int64_t ld = P.krylov_ld;
  if (ld < N) throw std::logic_error("Krylov basis smaller than the space");
  auto Vk = [&](int k) { return V + (int64_t)k * ld; };
  auto Zk = [&](int k) { return Z + (int64_t)k * ld; };
  // FP32 columns at the same stride, in the same buffer
  float *Z32 = S.z32 ? (float *)Z : nullptr;
  std::vector<double> H((size_t)(R + 1) * R, 0.0), s(R + 1), cs(R + 1), sn(R + 1);
  double *pa = P.red, *pb = P.red + kRedBlocks;
  const bool lowsync = S.lowsync && R <= k_ls_max() && ld % 2 == 0 && ((uintptr_t)V & 15) == 0 &&
                       ((uintptr_t)w & 15) == 0;
  if (lowsync) ensure_ls(P, N);
  // the finest lattice level's zero-start sweep written beside each Krylov
  // vector's FP32 copy (TopZeroSweep); CD3D_FUSE_ZERO_SWEEP=0: launched
  TopZeroSweep zt{};
  bool topz = false;
  if (S.lat_top && S.r32 && sp.pre > 0) {
    const char *ez = getenv("CD3D_FUSE_ZERO_SWEEP");
    if (!(ez && ez[0] == '0'))
      topz = kl_top_zero_sweep(*S.lat_top, S.lat_order, (float)sp.omega, S.lat_top->lt, zt);
  }
  const TopZeroSweep *ztp = topz ? &zt : nullptr;
  const int r32_state = S.r32 ? (topz ? 2 : 1) : 0;
  S.defect(S.x, S.b, Vk(0));
  double beta = dev_norm(P, N, Vk(0));
  r0 = beta;
  if (!std::isfinite(beta)) { status = CD3D_NONFINITE; return 0; }
  double res_norm_min = sp.res_norm_min, red_factor = sp.red_factor;
  if (sp.rel_tol > 0) { res_norm_min = sp.rel_tol * beta; red_factor = 0.0; }
  status = CD3D_OK;
  if (beta <= res_norm_min) return 0;
  double res0 = beta;
  int j = 1;
  auto update_x = [&](int k) {
    // back substitution (UpdateGmresIterate, FgmresIte.C:198-216), then x += Z y
    std::vector<double> y(s.begin(), s.begin() + k + 1);
    for (int i = k; i >= 0; i--) {
      y[i] /= H[(size_t)i * R + i];
      for (int jj = i - 1; jj >= 0; jj--) y[jj] -= H[(size_t)jj * R + i] * y[i];
    }
    memcpy(P.hscal + P.yoff, y.data(), sizeof(double) * (k + 1));
    HIPCHK(hipMemcpyAsync(P.dscal + P.yoff, P.hscal + P.yoff, sizeof(double) * (k + 1), hipMemcpyHostToDevice,
                          P.stream));
    if (Z32) k_multi_axpy32(P.stream, N, k + 1, P.dscal + P.yoff, Z32, ld, S.x);
    else k_multi_axpy(P.stream, N, k + 1, P.dscal + P.yoff, Z, ld, S.x);
    HIPCHK(hipStreamSynchronize(P.stream));  // pinned y is reused next time
  };
  while (j <= sp.maxit) {
    k_scale_cast(P.stream, N, 1.0 / beta, Vk(0), S.r32, ztp);  // Dscal(N, 1/beta, v[0]), and its FP32 copy
    std::fill(s.begin(), s.end(), 0.0);
    s[0] = beta;
    int i;
    for (i = 0; i < R && j <= sp.maxit; i++, j++) {
      if (Z32) {
        S.precond(v, Vk(i), nullptr, r32_state);
        kl_spmv64_f32(P.stream, *S.lat_top, S.lat_order, S.lat_top->lx, w, Z32 + (int64_t)i * ld);
      } else {
        S.precond(v, Vk(i), Zk(i), r32_state);
        S.matvec(Zk(i), w);
      }
      double *last;
      if (lowsync) {
        // the MGS column in two passes over the basis (k_ls_step)
        k_ls_step(P.stream, N, i, V, ld, w, P.ls_part, P.ls_L, P.dscal, pa);
        last = pa;
      } else {
        k_dot(P.stream, N, w, Vk(0), pa, nullptr);
        // modified Gram-Schmidt: H[k][i] = <w, v_k>; w -= H[k][i] v_k, fused with the next dot
        for (int k = 0; k <= i; k++) {
          double *in = (k % 2 == 0) ? pa : pb, *out = (k % 2 == 0) ? pb : pa;
          k_axpy_dot(P.stream, N, in, -1.0, Vk(k), w, k < i ? Vk(k + 1) : nullptr, out, P.dscal + k);
        }
        last = (i % 2 == 0) ? pb : pa;
      }
      k_scale_copy(P.stream, N, last, P.dscal + i + 1, w, Vk(i + 1), S.r32, ztp);
      HIPCHK(hipMemcpyAsync(P.hscal, P.dscal, sizeof(double) * (i + 2), hipMemcpyDeviceToHost, P.stream));
      HIPCHK(hipStreamSynchronize(P.stream));
      for (int k = 0; k <= i + 1; k++) H[(size_t)k * R + i] = P.hscal[k];
      if (!std::isfinite(H[(size_t)(i + 1) * R + i])) { status = CD3D_NONFINITE; return j; }
      for (int k = 0; k < i; k++) app_rot(H[(size_t)k * R + i], H[(size_t)(k + 1) * R + i], cs[k], sn[k]);
      gen_rot(H[(size_t)i * R + i], H[(size_t)(i + 1) * R + i], cs[i], sn[i]);
      app_rot(H[(size_t)i * R + i], H[(size_t)(i + 1) * R + i], cs[i], sn[i]);
      app_rot(s[i], s[i + 1], cs[i], sn[i]);
      const double res = fabs(s[i + 1]);
      if (res < res_norm_min || res < red_factor * res0) {
        update_x(i);
        return j;
      }
      if (res > sp.div_factor * res0) { status = CD3D_DIVERGED; return j; }
    }
    // i columns were built in this cycle: i == R after a full cycle; i < R
    // when maxit ends the loop inside a cycle.  The reference back-substitutes
    // through restart - 1 columns in both cases (FgmresIte.C:520), which adds
    // the previous cycle's leftover columns to the returned iterate; here
    // only the columns of this cycle enter.
    if (i > 0) update_x(i - 1);
    S.defect(S.x, S.b, Vk(0));
    beta = dev_norm(P, N, Vk(0));
    res0 = beta;
    if (!std::isfinite(beta)) { status = CD3D_NONFINITE; return j; }
  }
  status = CD3D_NOT_CONVERGED;
  return sp.maxit;
}

// ============================================================== same-level batches

// A batch: B assembled problems with the same number of levels, solved by
// the FP64 V-cycle (damped Jacobi, exact coarse solve) inside FGMRES in
// lockstep.  Every launch serves every problem still iterating
// (blockIdx.y = problem, arrays through the tables); the host runs the
// Givens rotations and stopping tests of all problems after one transfer of
// their H columns per iteration.  Per problem the iteration follows
// fgmres_iterate step by step (same stopping rules, same iteration count
// semantics); only the reduction trees differ, so results agree with a
// problem solved alone to rounding.
struct BatchRun {
  std::vector<cd3d_problem *> ps;
  const cd3d_solver_params &sp;
  cd3d_problem &lead;
  hipStream_t s;
  int B, NL, R;
  std::vector<int> maxblk, maxn, maxn_c;
  int maxna = 0;
  BatchRun(std::vector<cd3d_problem *> p, const cd3d_solver_params &sp_)
      : ps(std::move(p)), sp(sp_), lead(*ps[0]), s(ps[0]->stream), B((int)ps.size()),
        NL((int)ps[0]->dl.size()) {
    R = std::min(sp.restart, sp.maxit);
    if (R <= 0) R = 10;
  }
  Problem::BatchWs &ws() { return lead.bws; }
  const BatchLevel *lev(int l) { return ws().dlev + (size_t)l * ws().cap_b; }
  const BatchProblem *prob() { return ws().dprob; }
  const int *act() { return ws().dactive; }
};

static void batch_workspace(cd3d_problem &P, int B, int NL) {
  Problem::BatchWs &w = P.bws;
  if (w.cap_b >= B && w.cap_l >= NL) return;
  const int cb = std::max(B, w.cap_b), cl = std::max(NL, w.cap_l);
  for (void *d : {(void *)w.dlev, (void *)w.dprob, (void *)w.dactive, (void *)w.dred, (void *)w.dscal,
                  (void *)w.dcoef})
    if (d) dfree(P, d);
  for (void *h : {(void *)w.hscal, (void *)w.hcoef, (void *)w.hactive})
    if (h) (void)hipHostFree(h);
  w = Problem::BatchWs{};
  w.dlev = dalloc<BatchLevel>(P, (size_t)cb * cl);
  w.dprob = dalloc<BatchProblem>(P, cb);
  w.dactive = dalloc<int>(P, cb);
  w.dred = dalloc<double>(P, (size_t)cb * 2 * kBatchRed);
  w.dscal = dalloc<double>(P, (size_t)cb * kBatchScal);
  w.dcoef = dalloc<double>(P, (size_t)cb * kBatchCoef);
  HIPCHK(hipHostMalloc((void **)&w.hscal, sizeof(double) * cb * kBatchScal, hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void **)&w.hcoef, sizeof(double) * cb * kBatchCoef, hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void **)&w.hactive, sizeof(int) * cb, hipHostMallocDefault));
  w.cap_b = cb;
  w.cap_l = cl;
}

// one V-cycle of all active problems, z = V(r) with the finest level's rhs r
// (zero on Dirichlet rows), as precondition / cycle64 / coarse64; in the
// mixed form on the FP32 level vectors (the matrices rounded to FP32 as they
// are read, the coarse solve accumulated in FP64)
static void batch_cycle(BatchRun &br, int i) {
  const cd3d_solver_params &sp = br.sp;
  const int B = br.B;
  const bool f32 = sp.precision == CD3D_MIXED;
  if (i == 0) {
    kb_fill_zero(br.s, B, br.maxn[0], br.lev(0), br.prob(), br.act(), kBSol, f32);
    kb_copy(br.s, B, br.maxn[0], br.lev(0), br.prob(), br.act(), kBRhs, kBSol, true, f32);
    kb_stream(br.s, kDefectMode, B, br.maxblk[0], br.lev(0), br.prob(), br.act(), kBSol, kBRhs, kBAux, 0.0, f32);
    kb_gemv_coarse(br.s, B, br.maxna, br.lev(0), br.prob(), br.act(), kBAux, kBSol, f32);
    return;
  }
  int x = kBSol, t = kBTmp;
  for (int k = 0; k < sp.pre; k++) {
    if (k == 0) kb_jacobi_zero(br.s, B, br.maxn[i], br.lev(i), br.prob(), br.act(), kBRhs, t, sp.omega, f32);
    else kb_stream(br.s, kJacobiMode, B, br.maxblk[i], br.lev(i), br.prob(), br.act(), x, kBRhs, t, sp.omega, f32);
    std::swap(x, t);
  }
  kb_stream(br.s, kDefectMode, B, br.maxblk[i], br.lev(i), br.prob(), br.act(), x, kBRhs, kBAux, 0.0, f32);
  kb_restrict(br.s, B, br.maxn_c[i], br.lev(i), br.prob(), br.act(), kBAux, kBRhs, br.lev(i - 1), f32);
  batch_cycle(br, i - 1);
  kb_prolong_update(br.s, B, br.maxn[i], br.lev(i), br.prob(), br.act(), br.lev(i - 1), kBSol, x, sp.gmg_damp, f32);
  for (int k = 0; k < sp.post; k++) {
    kb_stream(br.s, kJacobiMode, B, br.maxblk[i], br.lev(i), br.prob(), br.act(), x, kBRhs, t, sp.omega, f32);
    std::swap(x, t);
  }
  if (x != kBSol) kb_copy(br.s, B, br.maxn[i], br.lev(i), br.prob(), br.act(), x, kBSol, false, f32);
}

// the stopping state of one problem of a batch
struct BatchState {
  std::vector<double> H, s, cs, sn;
  double beta = 0, res0 = 0, r0 = 0, res_norm_min = 0, red_factor = 0;
  int its = 0;
  cd3d_status status = CD3D_OK;
  bool active = false;
};

static void batch_solve_impl(BatchRun &br, std::vector<BatchState> &st) {
  const cd3d_solver_params &sp = br.sp;
  const int B = br.B, NL = br.NL, R = br.R, top = NL - 1;
  if (R > kBatchMaxRestart) throw ArgError("batched solves take restart <= 200");
  batch_workspace(br.lead, B, NL);
  Problem::BatchWs &w = br.ws();
  // per problem: Krylov space, 1/a_ii, initial guess (0 with the Dirichlet
  // values of the rhs) and rhs on its own stream, then the tables
  std::vector<BatchLevel> hl((size_t)w.cap_b * NL);
  std::vector<BatchProblem> hp(B);
  br.maxblk.assign(NL, 0);
  br.maxn.assign(NL, 0);
  br.maxn_c.assign(NL, 0);
  // the mixed form orthogonalises in two passes over the basis per Arnoldi
  // step (kb_ls_step) unless CD3D_FGMRES_MGS=1; the FP64 form keeps MGS
  // vector by vector
  const char *em = getenv("CD3D_FGMRES_MGS");
  bool lowsync = sp.precision == CD3D_MIXED && R <= k_ls_max() && !(em && em[0] == '1');
  for (int b = 0; b < B; b++) {
    cd3d_problem &P = *br.ps[b];
    coarse_join(P);
    if (!P.cinv_valid) {
      launch_coarse_inverse(P);
      coarse_join(P);
    }
    ensure_krylov(P, R);
    DevLevel &F = P.dl.back();
    for (int l = 0; l < NL; l++) {
      DevLevel &D = P.dl[l];
      if (!D.invdiag_ok) {
        k_invdiag(P.stream, D);
        D.invdiag_ok = true;
      }
      BatchLevel &L = hl[(size_t)l * w.cap_b + b];
      L.ndof = D.ndof;
      L.nact = D.nact;
      L.nblk = D.nblk;
      L.bt = D.bt;
      L.rowptr = D.rowptr;
      L.col = D.col;
      L.diag = D.diag;
      L.val = D.val;
      L.invdiag = D.invdiag;
      L.vec[0] = D.sol;
      L.vec[1] = D.rhs;
      L.vec[2] = D.aux;
      L.vec[3] = D.tmp;
      if (sp.precision == CD3D_MIXED) {
        if (!D.b32) D.b32 = dalloc<float>(P, 5 * (size_t)D.ndof);
        for (int k = 0; k < 4; k++) L.vec32[k] = D.b32 + (size_t)k * D.ndof;
        L.invdiag32 = D.b32 + 4 * (size_t)D.ndof;
      }
      if (l > 0) {
        L.p_rowptr = D.p_rowptr;
        L.p_col = D.p_col;
        L.p_val = D.p_val;
        L.r_rowptr = D.r_rowptr;
        L.r_col = D.r_col;
        L.r_val = D.r_val;
        L.ndof_c = P.dl[l - 1].ndof;
        L.nact_c = P.dl[l - 1].nact;
        br.maxn_c[l] = std::max(br.maxn_c[l], L.ndof_c);
      }
      br.maxblk[l] = std::max(br.maxblk[l], D.nblk);
      br.maxn[l] = std::max(br.maxn[l], D.ndof);
    }
    BatchProblem &Q = hp[b];
    Q.n = F.ndof;
    Q.na0 = P.dl[0].nact;
    Q.cinv = P.cinv;
    Q.V = P.V;
    Q.Z = P.Z;
    Q.ld = P.krylov_ld;
    Q.xs = P.xs;
    Q.bs = P.bs;
    Q.w = P.w;
    Q.red = w.dred + (size_t)b * 2 * kBatchRed;
    Q.scal = w.dscal + (size_t)b * kBatchScal;
    Q.coef = w.dcoef + (size_t)b * kBatchCoef;
    br.maxna = std::max(br.maxna, Q.na0);
    // 16-byte pairs in the update pass
    if (Q.ld % 2 || ((uintptr_t)Q.V & 15) || ((uintptr_t)Q.w & 15)) lowsync = false;
    k_copy(P.stream, F.ndof, F.rhs_asm, P.bs);
    HIPCHK(hipMemsetAsync(P.xs, 0, sizeof(double) * F.ndof, P.stream));
    k_copy_range(P.stream, F.nact, F.ndof, P.bs, P.xs);
    HIPCHK(hipStreamSynchronize(P.stream));
  }
  if (lowsync)
    for (int b = 0; b < B; b++) {
      cd3d_problem &P = *br.ps[b];
      ensure_ls(P, hp[b].n);
      hp[b].ls_part = P.ls_part;
      hp[b].ls_L = P.ls_L;
    }
  HIPCHK(hipMemcpyAsync(w.dlev, hl.data(), sizeof(BatchLevel) * hl.size(), hipMemcpyHostToDevice, br.s));
  HIPCHK(hipMemcpyAsync(w.dprob, hp.data(), sizeof(BatchProblem) * B, hipMemcpyHostToDevice, br.s));
  if (sp.precision == CD3D_MIXED)
    for (int l = 0; l < NL; l++) kb_invdiag32(br.s, B, br.maxn[l], br.lev(l));
  const BatchLevel *FL = br.lev(top);
  auto upload_active = [&] {
    for (int b = 0; b < B; b++) w.hactive[b] = st[b].active ? 1 : 0;
    HIPCHK(hipMemcpyAsync(w.dactive, w.hactive, sizeof(int) * B, hipMemcpyHostToDevice, br.s));
  };
  auto upload_coef = [&] {
    HIPCHK(hipMemcpyAsync(w.dcoef, w.hcoef, sizeof(double) * kBatchCoef * B, hipMemcpyHostToDevice, br.s));
  };
  // betas of the active problems: |b - A x| into scal[0]
  auto residual_norms = [&](int out_vec) {
    kb_stream(br.s, kDefectMode, B, br.maxblk[top], FL, br.prob(), br.act(), kBXs, kBBs, out_vec, 0.0);
    kb_dot(br.s, B, FL, br.prob(), br.act(), out_vec, out_vec, 0);
    kb_finish(br.s, B, br.prob(), br.act(), 0, 0, true);
    HIPCHK(hipMemcpyAsync(w.hscal, w.dscal, sizeof(double) * kBatchScal * B, hipMemcpyDeviceToHost, br.s));
    HIPCHK(hipStreamSynchronize(br.s));
  };
  auto Hm = [&](BatchState &q, int r, int c) -> double & { return q.H[(size_t)r * R + c]; };
  // x += Z y for the problems listed (back substitution through k + 1 columns)
  auto update = [&](const std::vector<std::pair<int, int>> &which) {
    if (which.empty()) return;
    for (int b = 0; b < B; b++) w.hcoef[(size_t)b * kBatchCoef] = 0.0;
    for (auto [b, k] : which) {
      BatchState &q = st[b];
      std::vector<double> y(q.s.begin(), q.s.begin() + k + 1);
      for (int i = k; i >= 0; i--) {
        y[i] /= Hm(q, i, i);
        for (int jj = i - 1; jj >= 0; jj--) y[jj] -= Hm(q, jj, i) * y[i];
      }
      double *c = w.hcoef + (size_t)b * kBatchCoef;
      c[0] = k + 1;
      for (int i = 0; i <= k; i++) c[1 + i] = y[i];
    }
    upload_coef();
    kb_update(br.s, B, FL, br.prob());
    HIPCHK(hipStreamSynchronize(br.s));  // the pinned coefficients are rewritten next
  };
  for (int b = 0; b < B; b++) {
    BatchState &q = st[b];
    q.H.assign((size_t)(R + 1) * R, 0.0);
    q.s.assign(R + 1, 0.0);
    q.cs.assign(R + 1, 0.0);
    q.sn.assign(R + 1, 0.0);
    q.active = true;
  }
  upload_active();
  residual_norms(kBV0);
  for (int b = 0; b < B; b++) {
    BatchState &q = st[b];
    q.beta = q.r0 = q.res0 = w.hscal[(size_t)b * kBatchScal];
    q.res_norm_min = sp.res_norm_min;
    q.red_factor = sp.red_factor;
    if (sp.rel_tol > 0) {
      q.res_norm_min = sp.rel_tol * q.beta;
      q.red_factor = 0.0;
    }
    if (!std::isfinite(q.beta)) {
      q.status = CD3D_NONFINITE;
      q.active = false;
    } else if (q.beta <= q.res_norm_min) {
      q.active = false;
    }
  }
  upload_active();
  int j = 1;
  auto any_active = [&] {
    for (auto &q : st)
      if (q.active) return true;
    return false;
  };
  while (j <= sp.maxit && any_active()) {
    // v_0 /= beta (problems that finished at the restart boundary keep their coef 0)
    for (int b = 0; b < B; b++) {
      double *c = w.hcoef + (size_t)b * kBatchCoef;
      c[0] = st[b].active ? 1.0 : 0.0;
      c[1] = st[b].active ? 1.0 / st[b].beta : 0.0;
      if (st[b].active) {
        std::fill(st[b].s.begin(), st[b].s.end(), 0.0);
        st[b].s[0] = st[b].beta;
      }
    }
    upload_coef();
    kb_scale(br.s, B, FL, br.prob(), br.act(), kBV0);
    int i;
    for (i = 0; i < R && j <= sp.maxit && any_active(); i++, j++) {
      // z_i = V(v_i), w = A z_i, modified Gram-Schmidt, v_{i+1} = w / |w|
      if (sp.precision == CD3D_MIXED) {
        kb_to32(br.s, B, br.maxn[top], FL, br.prob(), br.act(), kBV0 + i, kBRhs);
        batch_cycle(br, top);
        kb_from32(br.s, B, br.maxn[top], FL, br.prob(), br.act(), kBSol, kBZ0 + i);
      } else {
        kb_copy(br.s, B, br.maxn[top], FL, br.prob(), br.act(), kBV0 + i, kBRhs, false);
        batch_cycle(br, top);
        kb_copy(br.s, B, br.maxn[top], FL, br.prob(), br.act(), kBSol, kBZ0 + i, false);
      }
      kb_stream(br.s, kSpmvMode, B, br.maxblk[top], FL, br.prob(), br.act(), kBZ0 + i, -1, kBW, 0.0);
      int last = 0;  // reduction slot holding the partials of |w|^2
      if (lowsync) {
        kb_ls_step(br.s, B, br.maxn[top], br.prob(), br.act(), i, 0);
      } else {
        kb_dot(br.s, B, FL, br.prob(), br.act(), kBW, kBV0, 0);
        for (int k = 0; k <= i; k++)
          kb_axpy_dot(br.s, B, FL, br.prob(), br.act(), k % 2, k, -1.0, kBV0 + k, kBW, k < i ? kBV0 + k + 1 : -1,
                      1 - k % 2);
        last = (i + 1) % 2;
      }
      kb_scale_copy(br.s, B, FL, br.prob(), br.act(), last, i + 1, kBW, kBV0 + i + 1);
      HIPCHK(hipMemcpyAsync(w.hscal, w.dscal, sizeof(double) * kBatchScal * B, hipMemcpyDeviceToHost, br.s));
      HIPCHK(hipStreamSynchronize(br.s));
      std::vector<std::pair<int, int>> done;
      bool changed = false;
      for (int b = 0; b < B; b++) {
        BatchState &q = st[b];
        if (!q.active) continue;
        const double *hs = w.hscal + (size_t)b * kBatchScal;
        for (int k = 0; k <= i + 1; k++) Hm(q, k, i) = hs[k];
        if (!std::isfinite(Hm(q, i + 1, i))) {
          q.status = CD3D_NONFINITE;
          q.its = j;
          q.active = false;
          changed = true;
          continue;
        }
        for (int k = 0; k < i; k++) app_rot(Hm(q, k, i), Hm(q, k + 1, i), q.cs[k], q.sn[k]);
        gen_rot(Hm(q, i, i), Hm(q, i + 1, i), q.cs[i], q.sn[i]);
        app_rot(Hm(q, i, i), Hm(q, i + 1, i), q.cs[i], q.sn[i]);
        app_rot(q.s[i], q.s[i + 1], q.cs[i], q.sn[i]);
        const double res = fabs(q.s[i + 1]);
        if (res < q.res_norm_min || res < q.red_factor * q.res0) {
          done.push_back({b, i});
          q.its = j;
          q.active = false;
          changed = true;
        } else if (res > sp.div_factor * q.res0) {
          q.status = CD3D_DIVERGED;
          q.its = j;
          q.active = false;
          changed = true;
        }
      }
      update(done);
      if (changed) upload_active();
    }
    // end of a cycle (or maxit inside one): the active problems take this
    // cycle's columns and restart from their new residual
    std::vector<std::pair<int, int>> cyc;
    for (int b = 0; b < B; b++)
      if (st[b].active && i > 0) cyc.push_back({b, i - 1});
    update(cyc);
    if (!any_active()) break;
    residual_norms(kBV0);
    for (int b = 0; b < B; b++) {
      BatchState &q = st[b];
      if (!q.active) continue;
      q.beta = q.res0 = w.hscal[(size_t)b * kBatchScal];
      if (!std::isfinite(q.beta)) {
        q.status = CD3D_NONFINITE;
        q.its = j;
        q.active = false;
      }
    }
    upload_active();
  }
  for (auto &q : st)
    if (q.active) {
      q.status = CD3D_NOT_CONVERGED;
      q.its = sp.maxit;
      q.active = false;
    }
}

void check_params(const cd3d_solver_params &sp) {
  if (sp.smoother != 1 && sp.smoother != 2 && sp.smoother != 3 && sp.smoother != 13)
    throw ArgError("smoother: 1 (damped Jacobi), 2 (SOR), 3 (SSOR) or 13 (multicolour SSOR); the reference's ILU "
                   "and reordered SOR variants (4-7) are not provided");
  if (sp.coarse_smoother != 17 && sp.coarse_smoother != 3 && sp.coarse_smoother != 2 && sp.coarse_smoother != 1)
    throw ArgError("coarse_smoother must be 17, 3, 2 or 1");
  if (sp.pre < 0 || sp.post < 0 || sp.restart < 1 || sp.maxit < 1) throw ArgError("invalid solver counts");
  if (sp.precision != CD3D_FP64 && sp.precision != CD3D_MIXED) throw ArgError("precision must be 0 or 1");
  if (sp.cycle < 0 || sp.cycle > 2) throw ArgError("cycle must be 0 (F), 1 (V) or 2 (W)");
  if (!(sp.omega > 0) || !(sp.gmg_damp > 0)) throw ArgError("damping factors must be positive");
}

// join_before / join_after: the main stream waits for a pending coarse
// inverse at the start / end of the call.  cd3d_assemble leaves its inverse
// running on the side stream and cd3d_solve, which joins it only before its
// first coarse solve, does not wait at its start, so the inversion runs
// beside the solve's first defect, Krylov vector and pre-smoothing; every
// other call waits at both ends.
cd3d_status run(cd3d_problem *p, const std::function<void()> &f, bool join_before, bool join_after) {
  if (!p) return CD3D_ERR_ARG;
  try {
    p->fork_recorded = false;  // a fork point never outlives its call
    if (join_before) coarse_join(*p);  // (also: left pending by a call that failed)
    f();
    if (join_after) coarse_join(*p);
    return CD3D_OK;
  } catch (const ArgError &e) {
    p->err = e.what();
    return CD3D_ERR_ARG;
  } catch (const HipError &e) {
    p->err = e.what();
    return CD3D_ERR_HIP;
  } catch (const std::bad_alloc &) {
    p->err = "allocation failed";
    return CD3D_ERR_NOMEM;
  } catch (const std::exception &e) {
    p->err = e.what();
    return CD3D_ERR_STATE;
  }
}

}  // namespace

// ============================================================== C ABI

extern "C" {

void cd3d_default_solver(cd3d_solver_params *p) {
  if (!p) return;
  p->smoother = 1;
  p->pre = 2;
  p->post = 2;
  p->omega = 0.8;
  p->coarse_smoother = 17;
  p->coarse_maxit = 10;
  p->coarse_red = 0.1;
  p->gmg_damp = 1.0;
  p->restart = 20;
  p->maxit = 200;
  p->res_norm_min = 1e-30;
  p->red_factor = 0.0;
  p->div_factor = 1e10;
  p->rel_tol = 1e-10;
  p->precision = CD3D_FP64;
  p->cycle = 1;
}

size_t cd3d_estimate_bytes(const cd3d_mesh_desc *mesh, int restart, int cellwise_assembly, int tcd) {
  if (!mesh || (mesh->order != 1 && mesh->order != 2) || mesh->m < 1 || mesh->k0 < 0 || mesh->levels < 1 ||
      mesh->levels > 12 || restart < 1)
    return 0;
  const double p = mesh->order, nb = (p + 1) * (p + 1) * (p + 1);
  double total = 0, N_fine = 0, prevN = 0;
  for (int l = 0; l < mesh->levels; l++) {
    const double n = (double)mesh->m * std::ldexp(1.0, mesh->k0 + l), nn = p * n + 1, N = nn * nn * nn;
    const double na = (nn - 2) * (nn - 2) * (nn - 2), ncell = n * n * n;
    const double row = p == 1 ? 3 * (n - 1) : 8 * n - 5;  // active nonzeros per row and axis, summed
    const double nact_nnz = row * row * row, nnz = nact_nnz + (N - na);
    double b = 0;
    b += 4 * (N + 1) + 4 * nnz + 4 * N + 8 * nnz;  // rowptr, col, diag, val
    b += (4 + 16) * (nnz / 2048 + N / 256 + 2);    // row blocks and their extents
    b += 8 * N + 5 * 8 * N;                        // 1/a_ii, sol rhs aux tmp rhs_asm
    if (l > 0) {
      const double pnnz = (p == 1 ? 3.375 : 8.0) * N;
      b += 4 * (N + 1) + 4 * (prevN + 1) + 2 * 12 * pnnz;  // P and R = P^T
    }
    b += 16 * ncell + 4 * nb * ncell + 4 * ncell;   // cell geometry, DOF map, colours
    b += 4 * N + 2 * nact_nnz + 4 * na + 4 * 4 * N;  // lattice: maps, stencil ids, vectors
    b += 8 * (nb * nb + nb) + 8 * 8 * 129 + 4 * 8 * 128 * 2;
    if (cellwise_assembly) {
      // two-pass assembly: element rows, their gather list and pointers, rhs lists
      b += 12.0 * nb * nb * ncell + 4 * nact_nnz + 12 * nb * ncell + 4 * N;
      if (l > 0) b += 8 * nact_nnz + 4 * na;          // per-node FP32 stencils and their CSR positions
    }
    if (tcd) b += 8 * nnz * (tcd == 2 ? 2 : 1);      // conv-diff matrix (and K with SUPG) beside the mass matrix
    total += b + (2 << 20);                         // allocation granularity of ~20 buffers
    if (l == 0) total += 8.0 * na * na + 8.0 * (double)gj_scratch_doubles(na);  // coarse inverse + Gauss-Jordan scratch
    prevN = N;
    N_fine = N;
  }
  total += 8.0 * (2.0 * (restart + 1) * krylov_ld(N_fine) + 3.0 * N_fine);  // Krylov basis V, Z and w, xs, bs
  if (tcd) total += 3 * 8.0 * N_fine;
  total += 8.0 * (4 * 1024 + 256) + 8.0 * 4096;     // reductions, scalars, basis tables
  return (size_t)total;
}

cd3d_status cd3d_create(const cd3d_mesh_desc *mesh, cd3d_problem **out) {
  if (!mesh || !out) return CD3D_ERR_ARG;
  *out = nullptr;
  cd3d_problem *p = new (std::nothrow) cd3d_problem();
  if (!p) return CD3D_ERR_NOMEM;
  cd3d_status st = run(p, [&]() { setup_problem(*p, *mesh); });
  if (st != CD3D_OK) {
    fprintf(stderr, "cd3d_create: %s\n", p->err.c_str());
    cd3d_destroy(p);
    return st;
  }
  *out = p;
  return CD3D_OK;
}

cd3d_status cd3d_create_mapped(const cd3d_mesh_desc *mesh, const double *macro_xyz, cd3d_problem **out) {
  if (!mesh || !macro_xyz || !out) return CD3D_ERR_ARG;
  *out = nullptr;
  cd3d_problem *p = new (std::nothrow) cd3d_problem();
  if (!p) return CD3D_ERR_NOMEM;
  cd3d_status st = run(p, [&]() { setup_problem(*p, *mesh, macro_xyz); });
  if (st != CD3D_OK) {
    fprintf(stderr, "cd3d_create_mapped: %s\n", p->err.c_str());
    cd3d_destroy(p);
    return st;
  }
  *out = p;
  return CD3D_OK;
}

cd3d_status cd3d_destroy(cd3d_problem *p) {
  if (!p) return CD3D_ERR_ARG;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  if (p->side) (void)hipStreamSynchronize(p->side);
  for (void *a : p->allocs) (void)hipFree(a);
  if (p->hscal) (void)hipHostFree(p->hscal);
  for (void *h : {(void *)p->bws.hscal, (void *)p->bws.hcoef, (void *)p->bws.hactive})
    if (h) (void)hipHostFree(h);
  if (p->hflag) (void)hipHostFree(p->hflag);
  if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
  if (p->ev_join) (void)hipEventDestroy(p->ev_join);
  if (p->side) (void)hipStreamDestroy(p->side);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return CD3D_OK;
}

const char *cd3d_last_error(const cd3d_problem *p) { return p ? p->err.c_str() : "null handle"; }
size_t cd3d_device_bytes(const cd3d_problem *p) { return p ? p->dev_bytes : 0; }

cd3d_status cd3d_level_info(const cd3d_problem *p, int level, int *n_dof, int *n_active, int64_t *nnz, int *n_cells) {
  if (!p || level < 0 || level >= (int)p->hl.size()) return CD3D_ERR_ARG;
  const HostLevel &H = p->hl[level];
  if (n_dof) *n_dof = H.ndof;
  if (n_active) *n_active = H.nact;
  if (nnz) *nnz = H.nnz;
  if (n_cells) *n_cells = H.ncell;
  return CD3D_OK;
}

cd3d_status cd3d_get_dofmap(const cd3d_problem *p, int level, int *gn) {
  if (!p || !gn || level < 0 || level >= (int)p->hl.size()) return CD3D_ERR_ARG;
  memcpy(gn, p->hl[level].gn.data(), sizeof(int) * p->hl[level].gn.size());
  return CD3D_OK;
}

// ---------------------------------------------------------------- host geometry of mapped levels

static const int kHexV[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

// vertex-lattice index of local vertex v of cell c
static int64_t cell_vertex_index(const cd3d_problem &P, const HostLevel &H, int c, int v) {
  const int8_t *E = &H.ax[9 * c];
  int64_t I[3];
  for (int d = 0; d < 3; d++)
    I[d] = H.org[3 * c + d] / P.mesh.order + kHexV[v][0] * E[d] + kHexV[v][1] * E[3 + d] + kHexV[v][2] * E[6 + d];
  const int64_t nv = H.n + 1;
  return I[0] + nv * (I[1] + nv * I[2]);
}

// the trilinear map of cell c of a mapped level at the reference point r
// (THexaTrilinear::GetOrigFromRef, HexaTrilinear.C)
static void cell_map_point(const cd3d_problem &P, const HostLevel &H, int c, const double *r, double *X) {
  for (int d = 0; d < 3; d++) X[d] = 0.0;
  for (int v = 0; v < 8; v++) {
    double w = 1.0;
    for (int a = 0; a < 3; a++) w *= kHexV[v][a] ? 0.5 * (1 + r[a]) : 0.5 * (1 - r[a]);
    const double *x = &H.vx[3 * cell_vertex_index(P, H, c, v)];
    for (int d = 0; d < 3; d++) X[d] += w * x[d];
  }
}

cd3d_status cd3d_get_cells(const cd3d_problem *p, int level, double *xyz) {
  if (!p || !xyz || level < 0 || level >= (int)p->hl.size()) return CD3D_ERR_ARG;
  static const int hexv[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  const HostLevel &H = p->hl[level];
  const int S = p->mesh.order;
  if (p->mapped) {
    for (int c = 0; c < H.ncell; c++)
      for (int v = 0; v < 8; v++)
        for (int d = 0; d < 3; d++) xyz[24 * (size_t)c + 3 * v + d] = H.vx[3 * cell_vertex_index(*p, H, c, v) + d];
    return CD3D_OK;
  }
  for (int c = 0; c < H.ncell; c++)
    for (int v = 0; v < 8; v++)
      for (int d = 0; d < 3; d++) {
        const int8_t *E = &H.ax[9 * c];
        const int X = H.org[3 * c + d] + S * (hexv[v][0] * E[d] + hexv[v][1] * E[3 + d] + hexv[v][2] * E[6 + d]);
        xyz[24 * c + 3 * v + d] = X * (p->mesh.box[d] / ((double)S * H.n));
      }
  return CD3D_OK;
}

cd3d_status cd3d_get_csr(cd3d_problem *p, int level, int *rowptr, int *kcol, double *entries) {
  if (!p || level < 0 || level >= (int)p->hl.size()) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  return run(p, [&]() {
    const HostLevel &H = p->hl[level];
    if (rowptr) memcpy(rowptr, H.rowptr.data(), sizeof(int) * (H.ndof + 1));
    if (kcol) memcpy(kcol, H.kcol.data(), sizeof(int) * H.nnz);
    if (entries) {
      HIPCHK(hipMemcpyAsync(entries, p->dl[level].val, sizeof(double) * H.nnz, hipMemcpyDeviceToHost, p->stream));
      HIPCHK(hipStreamSynchronize(p->stream));
    }
  });
}

cd3d_status cd3d_get_rhs(cd3d_problem *p, int level, double *rhs) {
  if (!p || !rhs || level < 0 || level >= (int)p->hl.size()) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  return run(p, [&]() {
    HIPCHK(hipMemcpyAsync(rhs, p->dl[level].rhs_asm, sizeof(double) * p->dl[level].ndof, hipMemcpyDeviceToHost,
                          p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
  });
}

cd3d_status cd3d_set_entries(cd3d_problem *p, int level, const double *entries) {
  if (!p || !entries || level < 0 || level >= (int)p->hl.size()) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  return run(p, [&]() {
    HIPCHK(hipMemcpyAsync(p->dl[level].val, entries, sizeof(double) * p->dl[level].nnz, hipMemcpyHostToDevice,
                          p->stream));
    refresh_coarse_and_f32(*p);
    HIPCHK(hipStreamSynchronize(p->stream));
  });
}

cd3d_status cd3d_assemble(cd3d_problem *p, const cd3d_coeffs *co) {
  if (!p || !co) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  return run(p, [&]() {
    check_coeffs(*co);
    // the coarse inverse is enqueued right after level 0's assembly, so it
    // starts while the host still enqueues the finer levels (measured:
    // mixed solve 10.21-10.23 vs 10.27-10.28 ms when enqueued after the
    // finest level; CD3D_COARSE_EARLY=0 gives that order)
    const char *ee = getenv("CD3D_COARSE_EARLY");
    const bool early = !(ee && ee[0] == '0');
    for (int l = 0; l < (int)p->dl.size(); l++) {
      DevLevel &D = p->dl[l];
      D.lat_fresh = assemble_level(*p, l, *co, 0, D.val);
      D.lat_alpha_m = 0.0;
      D.lat_alpha = 1.0;
      // SystemCD3D.C:330: Dirichlet part of the rhs into the level solution
      k_copy_range(p->stream, D.nact, D.ndof, D.rhs_asm, D.sol);
      if (l == 0) {
        fork_coarse(*p);
        if (early) launch_coarse_inverse(*p);
      }
    }
    if (!early) launch_coarse_inverse(*p);
    refresh_f32(*p);
    HIPCHK(hipStreamSynchronize(p->stream));
    p->assembled = true;
    p->tcd_ready = false;  // the level matrices no longer hold the mass matrices
  }, true, false);  // the coarse inverse may still run: the next call joins it
}

static void host_vec_op(cd3d_problem *p, int level, const double *in, double *out, int nin, int nout,
                        const std::function<void(double *, double *)> &op) {
  DevLevel &D = p->dl[level];
  (void)D;
  double *din = dalloc<double>(*p, nin), *dout = dalloc<double>(*p, nout);
  HIPCHK(hipMemcpyAsync(din, in, sizeof(double) * nin, hipMemcpyHostToDevice, p->stream));
  HIPCHK(hipMemsetAsync(dout, 0, sizeof(double) * nout, p->stream));
  op(din, dout);
  HIPCHK(hipMemcpyAsync(out, dout, sizeof(double) * nout, hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  dfree(*p, din);
  dfree(*p, dout);
}

cd3d_status cd3d_matvec(cd3d_problem *p, int level, const double *x, double *y) {
  if (!p || !x || !y || level < 0 || level >= (int)p->dl.size()) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  return run(p, [&]() {
    const int n = p->dl[level].ndof;
    host_vec_op(p, level, x, y, n, n, [&](double *a, double *b) { k_spmv(p->stream, p->dl[level], a, b); });
  });
}

}  // extern "C"

int cd3d_internal_device(const cd3d_problem *p) { return p ? p->device : -1; }

// cd3d_matvec staged through a device buffer the caller holds (noalloc.hip):
// no allocation; nothing in the buffer needs clearing (x is copied in whole,
// every row of y is written)
cd3d_status cd3d_internal_matvec_staged(cd3d_problem *p, int level, const double *x, double *y, double *base,
                                        size_t capacity) {
  if (!p || !x || !y || !base || level < 0 || level >= (int)p->dl.size()) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  return run(p, [&]() {
    const DevLevel &L = p->dl[level];
    const size_t n = (size_t)L.ndof, ld = (n + 31) / 32 * 32;
    if (2 * ld > capacity) {
      char b[200];
      snprintf(b, sizeof b,
               "cd3d_matvec_noalloc: level %d has %zu unknowns; the staging array holds x and y of at most %zu",
               level, n, capacity / 2);
      throw ArgError(b);
    }
    double *xd = base, *yd = base + ld;
    HIPCHK(hipMemcpyAsync(xd, x, sizeof(double) * n, hipMemcpyHostToDevice, p->stream));
    k_spmv(p->stream, L, xd, yd);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(y, yd, sizeof(double) * n, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
  });
}

extern "C" {

cd3d_status cd3d_matvec_dev(cd3d_problem *p, int level, const double *x, double *y) {
  if (!p || !x || !y || level < 0 || level >= (int)p->dl.size()) return CD3D_ERR_ARG;
  return run(p, [&]() {
    k_spmv(p->stream, p->dl[level], x, y);
    HIPCHK(hipGetLastError());
  });
}

cd3d_status cd3d_prolongate(cd3d_problem *p, int level, const double *coarse, double *fine) {
  if (!p || !coarse || !fine || level < 1 || level >= (int)p->dl.size()) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  return run(p, [&]() {
    host_vec_op(p, level, coarse, fine, p->dl[level - 1].ndof, p->dl[level].ndof, [&](double *a, double *b) {
      k_prolong(p->stream, p->dl[level], p->dl[level - 1], a, b);
    });
  });
}

cd3d_status cd3d_restrict(cd3d_problem *p, int level, const double *fine, double *coarse) {
  if (!p || !coarse || !fine || level < 1 || level >= (int)p->dl.size()) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  return run(p, [&]() {
    // plain R = P^T (without zeroing the Dirichlet rows of the result)
    DevLevel C = p->dl[level - 1];
    C.nact = C.ndof;
    host_vec_op(p, level, fine, coarse, p->dl[level].ndof, C.ndof, [&](double *a, double *b) {
      k_restrict(p->stream, p->dl[level], C, a, b);
    });
  });
}

cd3d_status cd3d_vcycle(cd3d_problem *p, const cd3d_solver_params *sp, const double *r, double *z) {
  if (!p || !sp || !r || !z) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  return run(p, [&]() {
    check_params(*sp);
    if (!p->assembled) throw std::logic_error("assemble before vcycle");
    const int n = p->dl.back().ndof;
    VC v{*p, *sp};
    host_vec_op(p, (int)p->dl.size() - 1, r, z, n, n, [&](double *a, double *b) {
      // a general rhs may carry Dirichlet data: SetDirichletNodes path
      std::vector<double> rd(r + p->dl.back().nact, r + n);
      bool dz = std::all_of(rd.begin(), rd.end(), [](double x) { return x == 0.0; });
      precondition(v, a, b, dz);
    });
  });
}

cd3d_status cd3d_solve(cd3d_problem *p, const cd3d_solver_params *sp, const double *rhs, double *sol, int use_guess,
                       cd3d_stats *st) {
  if (!p || !sp) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  cd3d_status solve_status = CD3D_OK;
  cd3d_status rc = run(p, [&]() {
    check_params(*sp);
    if (!p->assembled) throw std::logic_error("assemble before solve");
    DevLevel &F = p->dl.back();
    const int N = F.ndof;
    ensure_krylov(*p, std::min(sp->restart, sp->maxit));
    const double t0 = now_ms();
    if (rhs) HIPCHK(hipMemcpyAsync(p->bs, rhs, sizeof(double) * N, hipMemcpyHostToDevice, p->stream));
    else k_copy(p->stream, N, F.rhs_asm, p->bs);
    // initial guess: 0 on active DOFs, Dirichlet values from the rhs
    // (SystemCD3D.C:330), or the caller's guess with its Dirichlet part reset
    if (use_guess && sol) HIPCHK(hipMemcpyAsync(p->xs, sol, sizeof(double) * N, hipMemcpyHostToDevice, p->stream));
    else HIPCHK(hipMemsetAsync(p->xs, 0, sizeof(double) * N, p->stream));
    k_copy_range(p->stream, F.nact, N, p->bs, p->xs);
    double r0 = 0;
    int its;
    const int top = (int)p->dl.size() - 1;
    if (sp->precision == CD3D_MIXED && top > 0 && sp->smoother == 1 && (F.lat_exact64 || F.lat_aff64)) {
      if (!p->xl) {
        p->xl = dalloc<double>(*p, N);
        p->bl = dalloc<double>(*p, N);
      }
      kl_dof_to_lat64(p->stream, F, p->bs, p->bl);
      kl_dof_to_lat64(p->stream, F, p->xs, p->xl);
      its = fgmres(*p, *sp, lattice_space(*p), r0, solve_status);
      kl_lat64_to_dof(p->stream, F, p->xl, p->xs);
    } else {
      its = fgmres(*p, *sp, dof_space(*p), r0, solve_status);
    }
    double rf = 0;  // recomputed with the CSR matrix in DOF order
    if (st) {
      k_defect(p->stream, F, p->xs, p->bs, p->w);
      rf = dev_norm(*p, N, p->w);
    }
    k_copy(p->stream, N, p->xs, F.sol);
    if (sol) HIPCHK(hipMemcpyAsync(sol, p->xs, sizeof(double) * N, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    if (st) {
      st->iterations = its;
      st->status = solve_status;
      st->r0 = r0;
      st->rfinal = rf;
      st->ms_solve = now_ms() - t0;
    }
  }, false, true);  // a coarse inverse left by cd3d_assemble is joined before its first use
  return rc != CD3D_OK ? rc : solve_status;
}

cd3d_status cd3d_solve_batch(cd3d_problem **ps, int count, const cd3d_solver_params *sp, cd3d_stats *stats) {
  if (!ps || count < 1 || !sp) return CD3D_ERR_ARG;
  std::vector<cd3d_problem *> v(ps, ps + count);
  for (auto *p : v)
    if (!p || p->device != v[0]->device || p->dl.size() != v[0]->dl.size()) return CD3D_ERR_ARG;
  std::vector<cd3d_problem *> order(v);
  std::sort(order.begin(), order.end());
  if (std::adjacent_find(order.begin(), order.end()) != order.end()) return CD3D_ERR_ARG;  // a handle twice
  const int R = std::min(sp->restart, sp->maxit);
  const bool batchable = count > 1 && (sp->precision == CD3D_FP64 || sp->precision == CD3D_MIXED) &&
                         sp->smoother == 1 && sp->coarse_smoother == 17 &&
                         sp->cycle == 1 && sp->pre >= 1 && R >= 1 && R <= kBatchMaxRestart;
  if (!batchable) {  // one after another, each with its own stream
    cd3d_status worst = CD3D_OK;
    for (int b = 0; b < count; b++) {
      const cd3d_status st = cd3d_solve(v[b], sp, nullptr, nullptr, 0, stats ? &stats[b] : nullptr);
      if (worst == CD3D_OK) worst = st;
    }
    return worst;
  }
  std::vector<std::unique_lock<std::mutex>> locks;
  for (auto *p : order) locks.emplace_back(p->mu);
  std::vector<BatchState> st(count);
  double ms = 0;
  const cd3d_status rc = run(v[0], [&]() {
    check_params(*sp);
    for (auto *p : v)
      if (!p->assembled) throw std::logic_error("assemble every problem before cd3d_solve_batch");
    const double t0 = now_ms();
    BatchRun br(v, *sp);
    batch_solve_impl(br, st);
    // |b - A x| of every problem, and x into the finest level's solution
    Problem::BatchWs &w = br.ws();
    for (int b = 0; b < count; b++) w.hactive[b] = 1;
    HIPCHK(hipMemcpyAsync(w.dactive, w.hactive, sizeof(int) * count, hipMemcpyHostToDevice, br.s));
    const BatchLevel *FL = br.lev(br.NL - 1);
    kb_stream(br.s, kDefectMode, count, br.maxblk[br.NL - 1], FL, br.prob(), br.act(), kBXs, kBBs, kBW, 0.0);
    kb_dot(br.s, count, FL, br.prob(), br.act(), kBW, kBW, 0);
    kb_finish(br.s, count, br.prob(), br.act(), 0, 0, true);
    kb_copy(br.s, count, br.maxn[br.NL - 1], FL, br.prob(), br.act(), kBXs, kBSol, false);
    HIPCHK(hipMemcpyAsync(w.hscal, w.dscal, sizeof(double) * kBatchScal * count, hipMemcpyDeviceToHost, br.s));
    HIPCHK(hipStreamSynchronize(br.s));
    ms = now_ms() - t0;
    if (stats)
      for (int b = 0; b < count; b++) {
        stats[b].iterations = st[b].its;
        stats[b].status = st[b].status;
        stats[b].r0 = st[b].r0;
        stats[b].rfinal = w.hscal[(size_t)b * kBatchScal];
        stats[b].ms_solve = ms;
      }
  });
  if (rc != CD3D_OK) {
    for (int b = 1; b < count; b++) v[b]->err = v[0]->err;
    return rc;
  }
  for (auto &q : st)
    if (q.status != CD3D_OK) return q.status;
  return CD3D_OK;
}

cd3d_status cd3d_get_solution(cd3d_problem *p, double *sol) {
  if (!p || !sol) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  return run(p, [&]() {
    DevLevel &F = p->dl.back();
    HIPCHK(hipMemcpyAsync(sol, F.sol, sizeof(double) * F.ndof, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
  });
}

cd3d_status cd3d_interp_pulse(cd3d_problem *p, int level, const cd3d_coeffs *co, double *u) {
  if (!p || !co || !u || level < 0 || level >= (int)p->hl.size()) return CD3D_ERR_ARG;
  const HostLevel &H = p->hl[level];
  const double *pc = co->pulse_c;
  if (p->mapped) {  // cell by cell, the nodes' reference points mapped (the last cell wins)
    const int q = p->mesh.order + 1, nb = p->nb;
    for (int c = 0; c < H.ncell; c++)
      for (int b = 0; b < nb; b++) {
        const double r[3] = {-1.0 + 2.0 * (b % q) / p->mesh.order, -1.0 + 2.0 * ((b / q) % q) / p->mesh.order,
                             -1.0 + 2.0 * (b / (q * q)) / p->mesh.order};
        double X[3];
        cell_map_point(*p, H, c, r, X);
        const double r2 = (X[0] - pc[0]) * (X[0] - pc[0]) + (X[1] - pc[1]) * (X[1] - pc[1]) + (X[2] - pc[2]) * (X[2] - pc[2]);
        u[H.gn[(size_t)c * nb + b]] = exp(-r2 / (2 * 0.05 * 0.05));
      }
    return CD3D_OK;
  }
  for (int d = 0; d < H.ndof; d++) {
    const int node = H.dof_node[d];
    const double den = (double)p->mesh.order * H.n;
    const double x = (node % H.nn) * (p->mesh.box[0] / den), y = ((node / H.nn) % H.nn) * (p->mesh.box[1] / den),
                 z = (node / (H.nn * H.nn)) * (p->mesh.box[2] / den);
    const double r2 = (x - pc[0]) * (x - pc[0]) + (y - pc[1]) * (y - pc[1]) + (z - pc[2]) * (z - pc[2]);
    u[d] = exp(-r2 / (2 * 0.05 * 0.05));
  }
  return CD3D_OK;
}

// ---------------------------------------------------------------- TCD3D
// TSystemTCD3D (SystemTCD3D.C:176-376) over the theta-scheme of
// SetTimeDiscParameters (TimeDiscRout.C): theta1 = theta, theta2 = 1 - theta,
// theta3 = 1 - theta, theta4 = theta (Crank-Nicolson 0.5, backward Euler 1).

static void tcd_init_impl(cd3d_problem &P, const cd3d_coeffs &co0, double t0, const double *u) {
  check_coeffs(co0);
  const int Lf = (int)P.dl.size() - 1;
  DevLevel &F = P.dl[Lf];
  const int N = F.ndof, NA = F.nact;
  for (auto &D : P.dl)
    if (!D.aconv) D.aconv = dalloc<double>(P, D.nnz);
  if (!P.tcd_oldrhs) {
    P.tcd_oldrhs = dalloc<double>(P, N);
    P.tcd_oldsol = dalloc<double>(P, N);
    P.tcd_def = dalloc<double>(P, N);
  }
  ensure_krylov(P, 1);
  // AssembleMRhs at t0 on every level (SystemTCD3D.C:176-202): the level
  // matrices hold the mass matrices from here on; with SUPG the rhs carries
  // the SUPG test functions (MatrixMRhsAssemble_SUPG, TCD3D.C:48)
  cd3d_coeffs co = co0;
  co.t = t0;
  P.assembled = false;
  for (int l = 0; l <= Lf; l++) {
    DevLevel &D = P.dl[l];
    D.tcd_mass_uniform = assemble_level(P, l, co, 1, D.val) == 1;
    if (D.tcd_mass_uniform)  // keep the mass class stencils beside the conv-diff ones
      HIPCHK(hipMemcpyAsync(D.lst64 + 3 * 8 * 128 + 8, D.lst64, sizeof(double) * 8 * 128, hipMemcpyDeviceToDevice,
                            P.stream));
  }
  HIPCHK(hipMemcpyAsync(P.xs, u, sizeof(double) * N, hipMemcpyHostToDevice, P.stream));
  k_copy_range(P.stream, NA, N, F.rhs_asm, P.xs);
  k_copy(P.stream, N, F.rhs_asm, P.tcd_oldrhs);
  HIPCHK(hipStreamSynchronize(P.stream));
  P.tcd_t = t0;
  P.tcd_ready = true;
}

// one step t -> t + tau; returns the FGMRES iterations
// r0_out / rf_out: |b - A x| of the step's system before and after the solve
static int tcd_step_impl(cd3d_problem &P, const cd3d_coeffs &co0, const cd3d_solver_params &sp, double tau,
                         double theta, cd3d_status &solve_status, double *r0_out = nullptr,
                         double *rf_out = nullptr) {
  check_coeffs(co0);
  if (!P.tcd_ready) throw std::logic_error("cd3d_tcd_init before cd3d_tcd_step");
  const double th1 = theta, th2 = 1 - theta, th3 = 1 - theta, th4 = theta;
  const int Lf = (int)P.dl.size() - 1;
  DevLevel &F = P.dl[Lf];
  const int N = F.ndof, NA = F.nact;
  ensure_krylov(P, std::min(sp.restart, sp.maxit));
  k_copy(P.stream, N, P.xs, P.tcd_oldsol);
  cd3d_coeffs co = co0;
  co.t = P.tcd_t + tau;
  // The mixed form works on the lattice wherever the assembled matrices are
  // affine class stencils with a uniform mass matrix (rotating field): there
  // the system matrix fma(tau th1, A, M) and the rhs matrix
  // fma(-tau th2, A, M) are evaluated entry by entry from the stencils by the
  // FP32 sweeps and the FP64 lattice operator, with the arithmetic of the CSR
  // path, so levels above the coarsest get no CSR values.  The coarsest level
  // keeps its CSR system matrix for the exact coarse solve.
  const bool affine_asm = co.conv_kind == 1 && co.f_kind == 0 && !co.supg && !uniform_assembly_off(P);
  bool lattice = sp.precision == CD3D_MIXED && Lf > 0 && affine_asm && sp.smoother == 1;
  for (int l = 1; l <= Lf; l++) lattice = lattice && P.dl[l].tcd_mass_uniform && P.dl[l].nclass > 0;
  // AssembleARhs at t_new (SystemTCD3D.C:204-226); with SUPG the A and rhs of
  // MatricesAKRhsAssemble_SUPG (TCD3D.C:102) and its time-derivative matrix
  // K, kept as M + K: the reference adds K to M for the rhs product and the
  // system and removes it again after the solve (the SUPG branch of
  // Main_Users/Sashi/3DPrograms/TCD_3D.C:1184-1188,1314-1317); here M itself
  // stays untouched
  std::vector<int> conv_path(Lf + 1);
  for (int l = 0; l <= Lf; l++) conv_path[l] = assemble_level(P, l, co, 0, P.dl[l].aconv, !(lattice && l > 0));
  if (co.supg)
    for (int l = 0; l <= Lf; l++) {
      DevLevel &D = P.dl[l];
      if (!D.ksupg) D.ksupg = dalloc<double>(P, D.nnz);
      assemble_level(P, l, co, 2, D.ksupg, true, D.tmp);
      k_matadd(P.stream, P.hl[l].rowptr[D.nact], D.ksupg, D.val, 1.0);  // M + K on the active rows
    }
  if (lattice)
    for (int l = 1; l <= Lf; l++)
      if (conv_path[l] != 2) throw std::logic_error("TCD3D lattice step: a level was not assembled affinely");
  k_copy_range(P.stream, NA, N, F.rhs_asm, P.xs);
  // AssembleSystMat (SystemTCD3D.C:228-294): B = tau th3 f_old + tau th4 f_new
  // + (M - tau th2 A) u_old on the active rows, system matrix M + tau th1 A.
  // The reference forms M - tau th2 A in place, adds tau (th1 + th2) A and
  // after the solve restores M (RestoreMassMat).  Here (M - tau th2 A) u_old
  // is one pass over both matrices (or over the stencils), and the system
  // matrix is written over A (assembled again next step), so M is never
  // changed and needs no restore: the same operators up to rounding.
  HIPCHK(hipMemsetAsync(P.bs, 0, sizeof(double) * N, P.stream));
  k_axpy(P.stream, NA, tau * th3, P.tcd_oldrhs, P.bs);
  k_axpy(P.stream, NA, tau * th4, F.rhs_asm, P.bs);
  HIPCHK(hipMemsetAsync(P.tcd_def, 0, sizeof(double) * N, P.stream));
  if (lattice) {
    if (!P.xl) {
      P.xl = dalloc<double>(P, N);
      P.bl = dalloc<double>(P, N);
    }
    kl_dof_to_lat64(P.stream, F, P.tcd_oldsol, P.xl);
    kl_affine64_apply(P.stream, F, P.mesh.order, -tau * th2, true, P.xl, P.bl);
    kl_lat64_to_dof(P.stream, F, P.bl, P.tcd_def);
  } else {
    // (M - tau th2 A) u_old, or (M + K - tau th2 A) u_old with SUPG
    k_matvec_pair_active(P.stream, F, co.supg ? F.ksupg : F.val, F.aconv, 1.0, -tau * th2, P.tcd_oldsol,
                         P.tcd_def);
  }
  k_axpy(P.stream, NA, 1.0, P.tcd_def, P.bs);
  k_copy_range(P.stream, NA, N, F.rhs_asm, P.bs);
  k_copy_range(P.stream, NA, N, F.rhs_asm, P.xs);
  // the level matrices are the system matrices until the guard swaps back
  struct SystemSwap {
    cd3d_problem &P;
    std::vector<int> levels;
    explicit SystemSwap(cd3d_problem &p) : P(p) {}
    void swap_level(int l) {
      std::swap(P.dl[l].val, P.dl[l].aconv);
      P.dl[l].invdiag_ok = false;
    }
    ~SystemSwap() {
      for (int l : levels) swap_level(l);
    }
  } sys(P);
  for (int l = 0; l <= Lf; l++) {
    DevLevel &D = P.dl[l];
    if (!(lattice && l > 0)) {
      // A <- M + tau th1 A, or (M + K) + tau th1 A with SUPG
      k_matadd_into(P.stream, P.hl[l].rowptr[D.nact], D.aconv, co.supg ? D.ksupg : D.val, tau * th1);
      sys.swap_level(l);
      sys.levels.push_back(l);
    }
    if (l == 0) fork_coarse(P);
  }
  launch_coarse_inverse(P);  // beside the finer levels' updates
  k_copy(P.stream, N, F.rhs_asm, P.tcd_oldrhs);
  for (int l = 0; l <= Lf; l++) {
    // system matrix M + tau th1 A: the FP32 lattice values come straight from
    // the class stencils when M is uniform and A affine
    DevLevel &D = P.dl[l];
    const bool affine = D.tcd_mass_uniform && conv_path[l] == 2;
    D.lat_fresh = affine ? 2 : 0;
    D.lat_alpha_m = 1.0;
    D.lat_alpha = tau * th1;
  }
  refresh_f32(P);
  // Solve (SystemTCD3D.C:317-376); the mixed form runs FGMRES in lattice
  // order when the FP64 lattice operator is exactly the system matrix
  double r0;
  int its;
  if (sp.precision == CD3D_MIXED && Lf > 0 && sp.smoother == 1 && (F.lat_exact64 || F.lat_aff64)) {
    if (!P.xl) {
      P.xl = dalloc<double>(P, N);
      P.bl = dalloc<double>(P, N);
    }
    kl_dof_to_lat64(P.stream, F, P.bs, P.bl);
    kl_dof_to_lat64(P.stream, F, P.xs, P.xl);
    its = fgmres(P, sp, lattice_space(P), r0, solve_status, rf_out);
    kl_lat64_to_dof(P.stream, F, P.xl, P.xs);
  } else {
    its = fgmres(P, sp, dof_space(P), r0, solve_status, rf_out);
  }
  if (r0_out) *r0_out = r0;
  // the level matrices are the mass matrices again (RestoreMassMat,
  // SystemTCD3D.C:296-315, without arithmetic: M was not changed)
  coarse_join(P);
  P.tcd_t += tau;
  return its;
}

static void check_theta(double theta) {
  if (!(theta >= 0.5 && theta <= 1.0)) throw ArgError("theta must lie in [0.5, 1] (Crank-Nicolson .. backward Euler)");
}

cd3d_status cd3d_tcd_init(cd3d_problem *p, const cd3d_coeffs *co, double t0, const double *u_host) {
  if (!p || !co || !u_host) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  return run(p, [&]() { tcd_init_impl(*p, *co, t0, u_host); });
}

cd3d_status cd3d_tcd_step(cd3d_problem *p, const cd3d_coeffs *co, const cd3d_solver_params *sp, double tau,
                          double theta, double *u_host, cd3d_stats *st) {
  if (!p || !co || !sp || !(tau > 0)) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  cd3d_status solve_status = CD3D_OK;
  const double t0 = now_ms();
  int its = 0;
  double r0 = 0, rf = 0;
  cd3d_status rc = run(p, [&]() {
    check_params(*sp);
    check_theta(theta);
    its = tcd_step_impl(*p, *co, *sp, tau, theta, solve_status, &r0, st ? &rf : nullptr);
    const int N = p->dl.back().ndof;
    if (u_host) HIPCHK(hipMemcpyAsync(u_host, p->xs, sizeof(double) * N, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
  });
  if (st) {
    st->iterations = its;
    st->status = rc != CD3D_OK ? rc : solve_status;
    st->r0 = r0;
    st->rfinal = rf;
    st->ms_assemble = 0;
    st->ms_solve = now_ms() - t0;
  }
  return rc != CD3D_OK ? rc : solve_status;
}

cd3d_status cd3d_tcd_get_state(cd3d_problem *p, double *u_host) {
  if (!p || !u_host) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  return run(p, [&]() {
    if (!p->tcd_ready) throw std::logic_error("cd3d_tcd_init before cd3d_tcd_get_state");
    HIPCHK(hipMemcpyAsync(u_host, p->xs, sizeof(double) * p->dl.back().ndof, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
  });
}

cd3d_status cd3d_tcd_run(cd3d_problem *p, const cd3d_coeffs *co, const cd3d_solver_params *sp, double t0, double tau,
                         int nsteps, double *u, double *states, cd3d_stats *st) {
  if (!p || !co || !sp || !u || nsteps < 0 || !(tau > 0)) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  cd3d_status solve_status = CD3D_OK;
  int total_its = 0;
  const double tstart = now_ms();
  cd3d_status rc = run(p, [&]() {
    check_params(*sp);
    tcd_init_impl(*p, *co, t0, u);
    const int N = p->dl.back().ndof;
    for (int m = 0; m < nsteps && solve_status == CD3D_OK; m++) {
      total_its += tcd_step_impl(*p, *co, *sp, tau, 0.5, solve_status);
      if (states)
        HIPCHK(hipMemcpyAsync(states + (size_t)m * N, p->xs, sizeof(double) * N, hipMemcpyDeviceToHost, p->stream));
    }
    HIPCHK(hipMemcpyAsync(u, p->xs, sizeof(double) * N, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
  });
  if (st) {
    st->iterations = total_its;
    st->status = rc != CD3D_OK ? rc : solve_status;
    st->r0 = st->rfinal = 0;
    st->ms_assemble = 0;
    st->ms_solve = now_ms() - tstart;
  }
  return rc != CD3D_OK ? rc : solve_status;
}

// ---------------------------------------------------------------- per-point coefficients

cd3d_status cd3d_quad_points(const cd3d_problem *p, int level, double *xyz) {
  if (!p || !xyz || level < 0 || level >= (int)p->hl.size()) return CD3D_ERR_ARG;
  const HostLevel &H = p->hl[level];
  const int P = p->mesh.order, n1 = quad_points_1d(P), nq = n1 * n1 * n1;
  std::vector<double> tab;
  basis_tables(P, tab);
  const int stride = 4 + 4 * p->nb;
  double dn[3];
  for (int d = 0; d < 3; d++) dn[d] = p->mesh.box[d] / ((double)P * H.n);
  if (p->mapped) {
    for (int c = 0; c < H.ncell; c++)
      for (int q = 0; q < nq; q++)
        cell_map_point(*p, H, c, &tab[(size_t)q * stride + 1], xyz + ((size_t)c * nq + q) * 3);
    return CD3D_OK;
  }
  // the point of the element kernels: X[dax[a]] = dn (O + P sg (xi_a + 1) / 2)
  for (int c = 0; c < H.ncell; c++)
    for (int q = 0; q < nq; q++) {
      double *X = xyz + ((size_t)c * nq + q) * 3;
      for (int a = 0; a < 3; a++)
        for (int d = 0; d < 3; d++) {
          const int e = H.ax[9 * c + 3 * a + d];
          if (e != 0) X[d] = dn[d] * (H.org[3 * c + d] + P * e * (tab[(size_t)q * stride + 1 + a] + 1) * 0.5);
        }
    }
  return CD3D_OK;
}

cd3d_status cd3d_set_coeffs_qp(cd3d_problem *p, int level, const double *coeffs) {
  if (!p || level < 0 || level >= (int)p->hl.size()) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  return run(p, [&]() {
    DevLevel &D = p->dl[level];
    if (!coeffs) {
      if (D.qp) dfree(*p, D.qp);
      D.qp = nullptr;
      return;
    }
    const int n1 = quad_points_1d(p->mesh.order);
    const size_t n = (size_t)D.ncell * n1 * n1 * n1 * 6;
    for (size_t k = 0; k < n; k++)
      if (!std::isfinite(coeffs[k])) throw ArgError("non-finite coefficient");
    if (!D.qp) D.qp = dalloc<double>(*p, n);
    HIPCHK(hipMemcpyAsync(D.qp, coeffs, sizeof(double) * n, hipMemcpyHostToDevice, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
  });
}

// ---------------------------------------------------------------- diagnostics and output

static FnParams fn_params(int fn, const cd3d_coeffs *co) {
  if (fn < 0 || fn > 3) throw ArgError("fn must be 0 (zero), 1 (sine), 2 (pulse) or 3 (linear)");
  FnParams f;
  f.fn = fn;
  if (co)
    for (int d = 0; d < 3; d++) f.pc[d] = co->pulse_c[d];
  return f;
}

cd3d_status cd3d_interpolate(cd3d_problem *p, int level, int fn, const cd3d_coeffs *co, double *u_host) {
  if (!p || !u_host || level < 0 || level >= (int)p->dl.size()) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  return run(p, [&]() {
    DevLevel &D = p->dl[level];
    k_interpolate(p->stream, D, p->mesh.box, p->mesh.order, p->hl[level].n, fn_params(fn, co), D.tmp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(u_host, D.tmp, sizeof(double) * D.ndof, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
  });
}

cd3d_status cd3d_errors(cd3d_problem *p, int level, const double *u_host, int fn, const cd3d_coeffs *co,
                        double *errors) {
  if (!p || !errors || level < 0 || level >= (int)p->dl.size()) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  return run(p, [&]() {
    DevLevel &D = p->dl[level];
    const FnParams f = fn_params(fn, co);
    const double *u = D.sol;
    if (u_host) {
      HIPCHK(hipMemcpyAsync(D.tmp, u_host, sizeof(double) * D.ndof, hipMemcpyHostToDevice, p->stream));
      u = D.tmp;
    } else if (level == (int)p->dl.size() - 1 && p->tcd_ready) {
      u = p->xs;  // the current TCD3D state
    }
    k_errors(p->stream, D, p->mesh.order, p->tab, u, f, p->red, p->dscal);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(p->hscal, p->dscal, sizeof(double) * 4, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    for (int k = 0; k < 4; k++) errors[k] = p->hscal[k];
  });
}

// Local basis values at the 8 cell vertices in the order of the reference's
// hexahedron (Output3D.C:1791-1800; the cell's vertex j sits at these
// reference coordinates in its own frame): bv[j * nb + l]
static void vertex_basis(int order, std::vector<double> &bv) {
  static const int hv[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                               {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  const int p1 = order + 1, nb = p1 * p1 * p1;
  auto lag = [order](double x, int i) {
    if (order == 1) return i == 0 ? 0.5 - 0.5 * x : 0.5 + 0.5 * x;
    return i == 0 ? 0.5 * x * (x - 1) : i == 1 ? 1 - x * x : 0.5 * x * (x + 1);
  };
  bv.assign(8 * nb, 0.0);
  for (int j = 0; j < 8; j++)
    for (int k = 0; k < p1; k++)
      for (int jj = 0; jj < p1; jj++)
        for (int i = 0; i < p1; i++)
          bv[j * nb + i + p1 * (jj + p1 * k)] = lag(hv[j][0], i) * lag(hv[j][1], jj) * lag(hv[j][2], k);
}

cd3d_status cd3d_write_vtk(cd3d_problem *p, int level, const double *u_host, const char *name, const char *path) {
  if (!p || !u_host || !path || level < 0 || level >= (int)p->hl.size()) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  return run(p, [&]() {
    const HostLevel &H = p->hl[level];
    const int S = p->mesh.order, n = H.n, nv = n + 1, nb = p->nb;
    static const int hexv[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                   {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    // vertices: the cell-corner lattice, x fastest; cells: their 8 vertices
    // in the cell's own order; point data: the FE function at every cell's
    // vertices, averaged over the cells (Output3D.C:1747-2100)
    std::vector<int> cv((size_t)H.ncell * 8);
    for (int c = 0; c < H.ncell; c++)
      for (int v = 0; v < 8; v++) {
        const int8_t *E = &H.ax[9 * c];
        int X[3];
        for (int d = 0; d < 3; d++)
          X[d] = (H.org[3 * c + d] + S * (hexv[v][0] * E[d] + hexv[v][1] * E[3 + d] + hexv[v][2] * E[6 + d])) / S;
        cv[(size_t)c * 8 + v] = X[0] + nv * (X[1] + nv * X[2]);
      }
    const size_t npt = (size_t)nv * nv * nv;
    std::vector<double> val(npt, 0.0), bv;
    std::vector<int> cnt(npt, 0);
    vertex_basis(S, bv);
    for (int c = 0; c < H.ncell; c++)
      for (int v = 0; v < 8; v++) {
        double s = 0.0;
        for (int l = 0; l < nb; l++) s += bv[v * nb + l] * u_host[H.gn[(size_t)c * nb + l]];
        val[cv[(size_t)c * 8 + v]] += s;
        cnt[cv[(size_t)c * 8 + v]]++;
      }
    FILE *f = fopen(path, "w");
    if (!f) throw ArgError(std::string("cannot open ") + path);
    const double hx = p->mesh.box[0] / n, hy = p->mesh.box[1] / n, hz = p->mesh.box[2] / n;
    fprintf(f, "# vtk DataFile Version 4.2\nfile created by ParMooN\nASCII\nDATASET UNSTRUCTURED_GRID\n\n");
    fprintf(f, "POINTS %zu double\n", npt);
    for (size_t k = 0; k < npt; k++) {
      if (p->mapped)
        fprintf(f, "%.6e %.6e %.6e\n", H.vx[3 * k], H.vx[3 * k + 1], H.vx[3 * k + 2]);
      else
        fprintf(f, "%.6e %.6e %.6e\n", (k % nv) * hx, ((k / nv) % nv) * hy, (k / ((size_t)nv * nv)) * hz);
    }
    fprintf(f, "\nCELLS %d %d\n", H.ncell, 9 * H.ncell);
    for (int c = 0; c < H.ncell; c++) {
      fprintf(f, "8 ");
      for (int v = 0; v < 8; v++) fprintf(f, "%d ", cv[(size_t)c * 8 + v]);
      fprintf(f, "\n");
    }
    fprintf(f, "\nCELL_TYPES %d\n", H.ncell);
    for (int c = 0; c < H.ncell; c++) fprintf(f, "12 ");
    fprintf(f, "\n\nPOINT_DATA %zu\nSCALARS %s double\nLOOKUP_TABLE default\n", npt, name ? name : "u");
    for (size_t k = 0; k < npt; k++) fprintf(f, "%.6e\n", cnt[k] ? val[k] / cnt[k] : 0.0);
    fprintf(f, "\n\n\nCELL_DATA %d\nSCALARS RegionID int  1\nLOOKUP_TABLE default  \n", H.ncell);
    for (int c = 0; c < H.ncell; c++) fprintf(f, "0\n");
    const bool ok = !ferror(f);
    if (fclose(f) != 0 || !ok) throw ArgError(std::string("write failed: ") + path);
  });
}

cd3d_status cd3d_write_matrix(cd3d_problem *p, int level, const char *path) {
  if (!p || !path || level < 0 || level >= (int)p->hl.size()) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  return run(p, [&]() {
    const HostLevel &H = p->hl[level];
    std::vector<double> e(H.nnz);
    HIPCHK(hipMemcpyAsync(e.data(), p->dl[level].val, sizeof(double) * H.nnz, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    FILE *f = fopen(path, "wb");
    if (!f) throw ArgError(std::string("cannot open ") + path);
    const int hdr[3] = {H.ndof, H.ndof, (int)H.nnz};
    bool ok = fwrite(hdr, sizeof(int), 3, f) == 3;
    ok = ok && fwrite(H.rowptr.data(), sizeof(int), H.rowptr.size(), f) == H.rowptr.size();
    ok = ok && fwrite(H.kcol.data(), sizeof(int), H.kcol.size(), f) == H.kcol.size();
    ok = ok && fwrite(e.data(), sizeof(double), e.size(), f) == e.size();
    if (fclose(f) != 0 || !ok) throw ArgError(std::string("write failed: ") + path);
  });
}

cd3d_status cd3d_device_ptr(cd3d_problem *p, int level, int which, double **dev) {
  if (!p || !dev || level < 0 || level >= (int)p->dl.size() || which < 0 || which > 3) return CD3D_ERR_ARG;
  DevLevel &D = p->dl[level];
  double *v[4] = {D.sol, D.rhs, D.aux, D.tmp};
  *dev = v[which];
  return CD3D_OK;
}

cd3d_status cd3d_kernel_bench(cd3d_problem *p, int which, int level, int reps, double *ms, double *bytes) {
  if (!p || !ms || !bytes || reps < 1 || level < 0 || level >= (int)p->dl.size() || which < 0 || which > 9)
    return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(p->mu);
  return run(p, [&]() {
    if (!p->assembled) throw std::logic_error("assemble before cd3d_kernel_bench");
    DevLevel &L = p->dl[level];
    const int o = p->mesh.order;
    const double N = L.ndof, nnz = (double)L.nnz, nint = (double)L.lat_nint;
    std::function<void()> launch;
    double b = 0;
    switch (which) {
      case 0:
        launch = [&] { k_spmv(p->stream, L, L.sol, L.aux); };
        b = 12 * nnz + 20 * N;
        break;
      case 1: {
        if (level == 0) throw ArgError("the coarsest level has no FP32 lattice stencils");
        bool uni = true;
        for (int c = 0; c < L.nclass; c++) uni = uni && L.uni[c];
        launch = [&] { kl_jacobi(p->stream, L, o, L.lx, L.lf, L.lt, 0.5f); };
        b = uni || L.lat_affine ? 12 * N : 12 * N + 4 * (double)L.lat_nval + 4 * nint;
        break;
      }
      case 2:
        if (!L.lat_exact64 && !L.lat_aff64)
          throw ArgError("the FP64 lattice operator needs assembly through class stencils");
        launch = [&] { kl_spmv64(p->stream, L, o, L.sol, L.aux); };
        b = 16 * N;
        break;
      case 3:
      case 4: {
        if (level == 0) throw ArgError("transfers need level >= 1");
        DevLevel &C = p->dl[level - 1];
        const double nc = (double)C.nn * C.nn * C.nn;
        if (which == 3) {
          launch = [&] { kl_restrict(p->stream, L, C, o, L.la, C.lf); };
          b = 4 * N + 4 * nc;
        } else {
          launch = [&] { kl_prolong_update(p->stream, L, C, o, C.lx, L.lt, 1.0f); };
          b = 8 * N + 4 * nc;
        }
        break;
      }
      case 5:
        if (!L.lat_exact64) throw ArgError("the class stencils need constant-coefficient assembly");
        launch = [&] { kl_fill_csr(p->stream, L, L.val); };
        b = 10 * (double)L.lat_nval;  // 2-byte stencil id in, 8-byte value out, active rows
        break;
      case 6:
        launch = [&] { k_copy(p->stream, L.ndof, L.sol, L.aux); };
        b = 16 * N;
        break;
      case 7:
        // zero alpha partials: w stays finite however many launches run
        HIPCHK(hipMemsetAsync(p->red, 0, sizeof(double) * 2 * kRedBlocks, p->stream));
        launch = [&] {
          k_axpy_dot(p->stream, L.ndof, p->red, -1.0, L.sol, L.aux, L.rhs, p->red + kRedBlocks, p->dscal);
        };
        b = 32 * N;
        break;
      case 9: {
        // the mixed form's low-synchronisation Gram-Schmidt step at i = 10
        // (dot pass, coefficients, update pass) on the finest level, over a
        // zero basis: (i + 2) + (i + 3) vectors of 8 N bytes
        if (level != (int)p->dl.size() - 1) throw ArgError("the Gram-Schmidt step runs on the finest level");
        constexpr int ii = 10;
        ensure_krylov(*p, 2 * ii);
        ensure_ls(*p, L.ndof);
        if (p->krylov_ld % 2) throw std::logic_error("Krylov basis stride not a multiple of 2");
        HIPCHK(hipMemsetAsync(p->V, 0, sizeof(double) * (ii + 1) * p->krylov_ld, p->stream));
        HIPCHK(hipMemsetAsync(p->ls_L, 0, sizeof(double) * k_ls_max() * k_ls_max(), p->stream));
        launch = [&] {
          k_ls_step(p->stream, L.ndof, ii, p->V, p->krylov_ld, p->w, p->ls_part, p->ls_L, p->dscal, p->red);
        };
        b = (2 * ii + 5) * 8 * N;
        break;
      }
    }
    double *scratch_val = nullptr;
    if (which == 8) {
      // the cell-by-cell assembly of the level as assemble_level runs it
      // (element rows of every cell, then the gather of the active rows'
      // entries; CD3D_ASM_SCATTER=1: the 8 colour-by-colour scatters) with
      // the coefficients of the last assembly, into a scratch value array
      // (the level's matrix stays as it is); FP64 flops per (cell, q, i, j): 8
      scratch_val = dalloc<double>(*p, L.nnz);
      double *rhs_scratch = L.tmp;
      AsmParams ap = asm_params(*p, L, p->last_co, 0);
      if (p->mapped && p->last_co.supg) {
        if (!L.hconv) L.hconv = dalloc<double>(*p, (size_t)L.ncell);
        ap.hconv = L.hconv;
      }
      HIPCHK(hipMemsetAsync(scratch_val, 0, sizeof(double) * L.nnz, p->stream));
      if (scatter_assembly(*p, level)) {
        ensure_slots(*p, level);
        launch = [&, ap, rhs_scratch] {
          if (ap.hconv && ap.supg) k_hconv_geo(p->stream, L, p->tab, ap, L.hconv);
          for (int k = 0; k < 8; k++)
            k_assemble(p->stream, L, p->tab, ap, scratch_val, rhs_scratch, L.colour_start[k],
                       L.colour_start[k + 1] - L.colour_start[k]);
        };
      } else {
        ensure_gather(*p, level);
        ap.E = L.E;
        ap.Er = L.Er;
        const int64_t nnz_act = p->hl[level].rowptr[L.nact];
        launch = [&, ap, rhs_scratch, nnz_act] {
          if (ap.hconv && ap.supg) k_hconv_geo(p->stream, L, p->tab, ap, L.hconv);
          k_assemble_rows(p->stream, L, p->tab, ap);
          k_gather_assembly(p->stream, L, nnz_act, scratch_val, rhs_scratch);
        };
      }
      const double nq = std::pow((double)quad_points_1d(o), 3), nbd = p->nb;
      b = (double)L.ncell * nq * 8.0 * nbd * nbd;
    }
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    launch();
    HIPCHK(hipEventRecord(e0, p->stream));
    for (int r = 0; r < reps; r++) launch();
    HIPCHK(hipEventRecord(e1, p->stream));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(hipGetLastError());
    if (scratch_val) dfree(*p, scratch_val);
    *ms = t / reps;
    *bytes = b;
  });
}

void *cd3d_stream(cd3d_problem *p) { return p ? (void *)p->stream : nullptr; }

cd3d_status cd3d_synchronize(cd3d_problem *p) {
  if (!p) return CD3D_ERR_ARG;
  return hipStreamSynchronize(p->stream) == hipSuccess ? CD3D_OK : CD3D_ERR_HIP;
}

}  // extern "C"

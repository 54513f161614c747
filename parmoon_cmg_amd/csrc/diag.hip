// Diagnostics on the device: nodal interpolation of known functions and the
// L2 / H1 error norms of a level vector against them
// (TFEFunction3D::Interpolate, FEFunction3D.C:1446; TFEFunction3D::GetErrors
// with L2H1Errors, FEFunction3D.C:57, MainUtilities.C:1623).
#include <hip/hip_runtime.h>

#include <cmath>

#include "cd3d_internal.hpp"
#include "trilinear.hpp"

namespace cd3d {

namespace {

constexpr int kDiagBlock = 256;

// known functions: 0 zero; 1 sin(pi x) sin(pi y) sin(pi z) (Examples/CD_3D/Laplace.h:12-19);
// 2 the Gaussian pulse exp(-|x - c|^2 / (2 0.05^2)); 3 x + 2y + 3z (in every
// Q1/Q2 space on trilinear cells: its errors vanish); v = value, gradient
__device__ inline void known_fn(const FnParams &f, double x, double y, double z, double *v) {
  v[0] = v[1] = v[2] = v[3] = 0.0;
  if (f.fn == 1) {
    const double sx = sin(M_PI * x), sy = sin(M_PI * y), sz = sin(M_PI * z);
    v[0] = sx * sy * sz;
    v[1] = M_PI * cos(M_PI * x) * sy * sz;
    v[2] = M_PI * sx * cos(M_PI * y) * sz;
    v[3] = M_PI * sx * sy * cos(M_PI * z);
  } else if (f.fn == 2) {
    const double s2 = 0.05 * 0.05, dx = x - f.pc[0], dy = y - f.pc[1], dz = z - f.pc[2];
    v[0] = exp(-(dx * dx + dy * dy + dz * dz) / (2 * s2));
    v[1] = -dx / s2 * v[0];
    v[2] = -dy / s2 * v[0];
    v[3] = -dz / s2 * v[0];
  } else if (f.fn == 3) {
    v[0] = x + 2 * y + 3 * z;
    v[1] = 1.0;
    v[2] = 2.0;
    v[3] = 3.0;
  }
}

// one thread per DOF: the Lagrange nodal functional is the value at the
// node.  On a mapped level (vx non-null) the node is the trilinear map of a
// lattice cell containing it at the node's reference point (cells sharing
// the node agree on it: their maps coincide on shared faces).
__global__ __launch_bounds__(kDiagBlock) void interp_k(int ndof, const int *__restrict__ dof_node, int nn, double dx,
                                                      double dy, double dz, const double *__restrict__ vx, int P,
                                                      FnParams f, double *__restrict__ u) {
  const int d = blockIdx.x * kDiagBlock + threadIdx.x;
  if (d >= ndof) return;
  const int node = dof_node[d];
  const int I[3] = {node % nn, (node / nn) % nn, node / (nn * nn)};
  double X[3], v[4];
  if (vx) {
    const int n = (nn - 1) / P;
    int cc[3];
    double r[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      cc[a] = min(I[a] / P, n - 1);
      r[a] = 2.0 * (I[a] - P * cc[a]) / P - 1.0;
    }
    TriMap M;
    // the identity frame: local axis a along lattice axis a
    tri_load(vx, n + 1, make_int4(P * cc[0], P * cc[1], P * cc[2], 8 | (2 << 6)), P, M);
    tri_point(M, r[0], r[1], r[2], X);
  } else {
    X[0] = I[0] * dx;
    X[1] = I[1] * dy;
    X[2] = I[2] * dz;
  }
  known_fn(f, X[0], X[1], X[2], v);
  u[d] = v[0];
}

// One thread per cell (grid-stride): at the quadrature points of the
// assembly, e = u_h - u and grad e in physical coordinates (the brick's
// J^-T: a signed axis permutation scaled by 2/h).  Per workgroup the sums of
// w|J| e^2, w|J| |grad e|^2 and the maximum |e| go to partials[3 b ..];
// last[0] receives the maximum of the last cell.
template <int P>
__global__ __launch_bounds__(kDiagBlock) void errors_k(int ncell, const int4 *__restrict__ cells,
                                                      const int *__restrict__ gn, const double *__restrict__ tab,
                                                      ErrGeom g, const double *__restrict__ u, FnParams f,
                                                      double *__restrict__ partials, double *__restrict__ last) {
  constexpr int NB = (P + 1) * (P + 1) * (P + 1), N1 = P + 2, NQ = N1 * N1 * N1, STRIDE = 4 + 4 * NB;
  __shared__ double sh[3][kDiagBlock];
  double e0 = 0.0, e1 = 0.0, emax = 0.0;
  const double detj = (g.h[0] / 2) * (g.h[1] / 2) * (g.h[2] / 2);
  for (int c = blockIdx.x * kDiagBlock + threadIdx.x; c < ncell; c += gridDim.x * kDiagBlock) {
    const int4 cg = cells[c];
    TriMap M;
    if (g.vx) tri_load(g.vx, g.nv, cg, P, M);
    const int O[3] = {cg.x, cg.y, cg.z};
    int dax[3], sg[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      dax[a] = (cg.w >> (3 * a)) & 3;
      sg[a] = ((cg.w >> (3 * a + 2)) & 1) ? -1 : 1;
    }
    double uc[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) uc[b] = u[gn[(int64_t)c * NB + b]];
    double cmax = 0.0;
    for (int q = 0; q < NQ; q++) {
      const double *T = tab + q * STRIDE;
      double w = T[0] * detj;
      double uh = 0.0, gr[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int b = 0; b < NB; b++) {
        uh = fma(uc[b], T[4 + 4 * b], uh);
        gr[0] = fma(uc[b], T[5 + 4 * b], gr[0]);
        gr[1] = fma(uc[b], T[6 + 4 * b], gr[1]);
        gr[2] = fma(uc[b], T[7 + 4 * b], gr[2]);
      }
      double X[3], G[3];
      if (g.vx) {  // trilinear cell: grad = J^-T grad_ref = C grad_ref / det
        double C[3][3];
        const double det = tri_jac(M, T[1], T[2], T[3], C);
        w = T[0] * fabs(det);
        tri_point(M, T[1], T[2], T[3], X);
#pragma unroll
        for (int d = 0; d < 3; d++) G[d] = (C[d][0] * gr[0] + C[d][1] * gr[1] + C[d][2] * gr[2]) / det;
      } else {
#pragma unroll
        for (int a = 0; a < 3; a++) {
          const int d = dax[a];
          const double xa = g.dn[d] * (O[d] + P * sg[a] * (T[1 + a] + 1) * 0.5);
          const double ga = sg[a] * (2.0 / g.h[d]) * gr[a];
          X[0] = d == 0 ? xa : X[0];
          X[1] = d == 1 ? xa : X[1];
          X[2] = d == 2 ? xa : X[2];
          G[0] = d == 0 ? ga : G[0];
          G[1] = d == 1 ? ga : G[1];
          G[2] = d == 2 ? ga : G[2];
        }
      }
      double v[4];
      known_fn(f, X[0], X[1], X[2], v);
      const double t = uh - v[0];
      e0 += w * t * t;
      e1 += w * ((G[0] - v[1]) * (G[0] - v[1]) + (G[1] - v[2]) * (G[1] - v[2]) + (G[2] - v[3]) * (G[2] - v[3]));
      cmax = fmax(cmax, fabs(t));
    }
    emax = fmax(emax, cmax);
    if (c == ncell - 1) last[0] = cmax;
  }
  sh[0][threadIdx.x] = e0;
  sh[1][threadIdx.x] = e1;
  sh[2][threadIdx.x] = emax;
  __syncthreads();
  for (int s = kDiagBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
      sh[2][threadIdx.x] = fmax(sh[2][threadIdx.x], sh[2][threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 3) partials[3 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// the workgroup partials in a fixed order: out = [sqrt(L2^2), sqrt(H1^2), last-cell max, max]
__global__ void errors_finish_k(int nblk, const double *__restrict__ partials, const double *__restrict__ last,
                                double *__restrict__ out) {
  if (threadIdx.x != 0) return;
  double a = 0.0, b = 0.0, m = 0.0;
  for (int k = 0; k < nblk; k++) {
    a += partials[3 * k];
    b += partials[3 * k + 1];
    m = fmax(m, partials[3 * k + 2]);
  }
  out[0] = sqrt(a);
  out[1] = sqrt(b);
  out[2] = last[0];
  out[3] = m;
}

}  // namespace

void k_interpolate(hipStream_t s, const DevLevel &L, const double *box, int order, int n, const FnParams &f,
                   double *u) {
  if (L.ndof <= 0) return;
  const double den = (double)order * n;
  interp_k<<<(L.ndof + kDiagBlock - 1) / kDiagBlock, kDiagBlock, 0, s>>>(
      L.ndof, L.ldof2node, L.nn, box[0] / den, box[1] / den, box[2] / den, L.vx, order, f, u);
}

int k_errors_blocks(int ncell) {
  const int b = (ncell + kDiagBlock - 1) / kDiagBlock;
  return b < 1 ? 1 : (b > 1024 ? 1024 : b);
}

void k_errors(hipStream_t s, const DevLevel &L, int order, const double *tab, const double *u, const FnParams &f,
              double *partials, double *out) {
  ErrGeom g;
  for (int d = 0; d < 3; d++) {
    g.h[d] = L.h[d];
    g.dn[d] = L.dn[d];
  }
  g.vx = L.vx;
  g.nv = L.nv;
  const int nblk = k_errors_blocks(L.ncell);
  double *last = partials + 3 * nblk;
  if (order == 1) errors_k<1><<<nblk, kDiagBlock, 0, s>>>(L.ncell, L.cells, L.gn, tab, g, u, f, partials, last);
  else errors_k<2><<<nblk, kDiagBlock, 0, s>>>(L.ncell, L.cells, L.gn, tab, g, u, f, partials, last);
  errors_finish_k<<<1, 64, 0, s>>>(nblk, partials, last, out);
}

}  // namespace cd3d

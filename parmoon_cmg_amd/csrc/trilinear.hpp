// Trilinear cell maps on the device (THexaTrilinear, HexaTrilinear.C:884-948
// and GetOrigValues / GetOrigFromRef): general hexahedra of a mapped
// hierarchy, whose vertices live in a per-level vertex lattice
// ((n+1)^3 x 3 doubles, x fastest).
//
// A cell is a lattice cell; its local frame (origin in node-lattice units
// and the signed axis permutation of its local axes, the same int4 the brick
// kernels use) only decides which lattice corner is local vertex v.  The map
// is x(r) = sum_k xc_k m_k(r) over the monomials m_k of (xi, eta, zeta)
// indexed by bit masks (bit 0 xi, bit 1 eta, bit 2 zeta: k = 3 is xi eta,
// k = 7 xi eta zeta), xc_k = 1/8 sum_v m_k(r_v) x_v.
#pragma once

#include <hip/hip_runtime.h>

namespace cd3d {

struct TriMap {
  double xc[3][8];
};

// local axis a of a cell runs along lattice axis dax[a] with sign sg[a]
__device__ inline void tri_frame(int code, int *dax, int *sg) {
#pragma unroll
  for (int a = 0; a < 3; a++) {
    dax[a] = (code >> (3 * a)) & 3;
    sg[a] = ((code >> (3 * a + 2)) & 1) ? -1 : 1;
  }
}

// the 8 vertices of a cell in tensor order (t = i + 2 j + 4 k for the local
// corner (i, j, k) in {0,1}^3), gathered from the vertex lattice, and the
// coefficients of its map by a Hadamard butterfly over the three local axes
__device__ inline void tri_load(const double *__restrict__ vx, int nv, int4 cg, int P, TriMap &M) {
  int dax[3], sg[3];
  tri_frame(cg.w, dax, sg);
  const int O[3] = {cg.x / P, cg.y / P, cg.z / P};
  double c[3][8];
#pragma unroll
  for (int t = 0; t < 8; t++) {
    int I[3] = {O[0], O[1], O[2]};
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const int step = ((t >> a) & 1) * sg[a];
      I[0] += dax[a] == 0 ? step : 0;
      I[1] += dax[a] == 1 ? step : 0;
      I[2] += dax[a] == 2 ? step : 0;
    }
    const double *x = vx + 3 * ((int64_t)I[0] + (int64_t)nv * (I[1] + (int64_t)nv * I[2]));
    c[0][t] = x[0];
    c[1][t] = x[1];
    c[2][t] = x[2];
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {
#pragma unroll
    for (int bit = 1; bit < 8; bit <<= 1)
#pragma unroll
      for (int t = 0; t < 8; t++)
        if (!(t & bit)) {
          const double lo = c[d][t], hi = c[d][t | bit];
          c[d][t] = lo + hi;
          c[d][t | bit] = hi - lo;
        }
#pragma unroll
    for (int k = 0; k < 8; k++) M.xc[d][k] = 0.125 * c[d][k];
  }
}

// the physical point of the reference point (xi, eta, zeta)
__device__ inline void tri_point(const TriMap &M, double xi, double eta, double zeta, double *X) {
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double *c = M.xc[d];
    X[d] = fma(zeta, fma(eta, fma(c[7], xi, c[6]), fma(c[5], xi, c[4])), fma(eta, fma(c[3], xi, c[2]), fma(c[1], xi, c[0])));
  }
}

// the cofactors of the Jacobian J[d][a] = dx_d / dr_a at the point, so that
// grad phi = C grad_ref phi / det (J^-T = C / det); returns det J
__device__ inline double tri_jac(const TriMap &M, double xi, double eta, double zeta, double (*C)[3]) {
  double J[3][3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double *c = M.xc[d];
    J[d][0] = fma(zeta, fma(c[7], eta, c[5]), fma(c[3], eta, c[1]));
    J[d][1] = fma(zeta, fma(c[7], xi, c[6]), fma(c[3], xi, c[2]));
    J[d][2] = fma(eta, fma(c[7], xi, c[6]), fma(c[5], xi, c[4]));
  }
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const int d1 = (d + 1) % 3, d2 = (d + 2) % 3, a1 = (a + 1) % 3, a2 = (a + 2) % 3;
      C[d][a] = J[d1][a1] * J[d2][a2] - J[d1][a2] * J[d2][a1];
    }
  return J[0][0] * C[0][0] + J[0][1] * C[0][1] + J[0][2] * C[0][2];
}

// Mesh size in convection direction of a general hexahedron
// (Mesh_size_in_convection_direction_without_storing, hexahedron branch,
// ConvDiff.C:411-469): 2|b| / sum_i |b . grad p_i(barycentre)|, p_i the
// function in span{1, x, y, z, xy, xz, yz, xyz} that is 1 at vertex i and 0
// at the others; the longest space diagonal where the sum is below 1e-10.
// The fit is set up in coordinates centred at the barycentre and scaled by
// the cell's extent, which leaves p_i unchanged and keeps the 8 x 8 vertex
// matrix well conditioned (in absolute coordinates its condition grows like
// h^-3).  At the barycentre only the linear coefficients of p_i survive:
// b . grad p_i = (A^-T g)_i / L with g = (0, b1, b2, b3, 0, 0, 0, 0), one
// solve with partial pivoting.
__device__ inline double tri_hconv(const TriMap &M, double b0, double b1, double b2) {
  double V[8][3];
#pragma unroll
  for (int t = 0; t < 8; t++) {
    const double xi = (t & 1) ? 1.0 : -1.0, eta = (t & 2) ? 1.0 : -1.0, zeta = (t & 4) ? 1.0 : -1.0;
    // vertex minus barycentre (xc_0 is the barycentre)
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const double *c = M.xc[d];
      V[t][d] = fma(zeta, fma(eta, fma(c[7], xi, c[6]), fma(c[5], xi, c[4])), fma(eta, fma(c[3], xi, c[2]), c[1] * xi));
    }
  }
  double L = 0.0;
#pragma unroll
  for (int t = 0; t < 8; t++) L = fmax(L, fmax(fabs(V[t][0]), fmax(fabs(V[t][1]), fabs(V[t][2]))));
  const double iL = 1.0 / L;
  // AT[k][t]: monomial k at vertex t (the transposed vertex matrix)
  double AT[8][8], v[8] = {0.0, b0, b1, b2, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < 8; t++) {
    const double x = V[t][0] * iL, y = V[t][1] * iL, z = V[t][2] * iL;
    AT[0][t] = 1.0;
    AT[1][t] = x;
    AT[2][t] = y;
    AT[3][t] = z;
    AT[4][t] = x * y;
    AT[5][t] = x * z;
    AT[6][t] = y * z;
    AT[7][t] = x * y * z;
  }
#pragma unroll
  for (int k = 0; k < 8; k++) {
    // pivot row: largest |AT[r][k]| for r >= k (the lowest index on ties)
    int p = k;
    double best = fabs(AT[k][k]);
#pragma unroll
    for (int r = k + 1; r < 8; r++) {
      const double a = fabs(AT[r][k]);
      if (a > best) {
        best = a;
        p = r;
      }
    }
#pragma unroll
    for (int r = k + 1; r < 8; r++)
      if (r == p) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const double t0 = AT[k][j];
          AT[k][j] = AT[r][j];
          AT[r][j] = t0;
        }
        const double t1 = v[k];
        v[k] = v[r];
        v[r] = t1;
      }
    const double ip = 1.0 / AT[k][k];
#pragma unroll
    for (int r = k + 1; r < 8; r++) {
      const double f = AT[r][k] * ip;
#pragma unroll
      for (int j = k + 1; j < 8; j++) AT[r][j] = fma(-f, AT[k][j], AT[r][j]);
      v[r] = fma(-f, v[k], v[r]);
    }
  }
  double den = 0.0;
#pragma unroll
  for (int k = 7; k >= 0; k--) {
    double s = v[k];
#pragma unroll
    for (int j = k + 1; j < 8; j++) s = fma(-AT[k][j], v[j], s);
    v[k] = s / AT[k][k];
    den += fabs(v[k]);
  }
  den *= iL;
  if (den < 1e-10) {
    // THexahedron::GetDiameter (Hexahedron.C:78-104): local vertex pairs
    // 0-6, 1-7, 2-4, 3-5 are the tensor corners t and 7 - t
    double dm = 0.0;
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const double dx = V[t][0] - V[7 - t][0], dy = V[t][1] - V[7 - t][1], dz = V[t][2] - V[7 - t][2];
      dm = fmax(dm, dx * dx + dy * dy + dz * dz);
    }
    return sqrt(dm);
  }
  return 2.0 * sqrt(b0 * b0 + b1 * b1 + b2 * b2) / den;
}

}  // namespace cd3d

// Lattice-stencil kernels of the mixed-precision V-cycle (FP32 values, FP32
// accumulation).
//
// A level's unknowns are stored in node-lattice order (X fastest), boundary
// nodes included and kept at zero: inside the preconditioner the Dirichlet
// part of every right-hand side and iterate is zero.  Interior nodes are split
// into parity classes that share one stencil shape (Q2: 8 classes with
// 125/75/45/27 points, Q1: one class with 27), and the matrix is stored as
// [class][stencil entry][node of the class]: a thread per node reads its
// stencil values with unit stride across the wavefront, and the column of
// every entry is implicit in the lattice offset -- 4 bytes per nonzero
// instead of 12 for FP64 CSR.  Restriction and prolongation are the
// tensor-product nodal interpolation weights (Q1: 1/2 1 1/2; Q2: -1/8 3/8 1
// 3/8 -1/8 and 3/4 1 3/4), so they need no matrix at all.
#include <hip/hip_runtime.h>

#include "cd3d_internal.hpp"

namespace cd3d {

namespace {

constexpr int kB = 256;

struct LatDev {
  int px[3], lo[3], cnt[3], step, nsten;
  long long nodes, off, first;
};

struct LatArgs {
  int nn, nclass;
  LatDev c[8];
  int bstart[9];  // first block of each class
};

enum LatMode { kLJacobi = 0, kLJacobiZero = 1, kLDefect = 2 };

template <int P, int PXS>
__host__ __device__ inline constexpr int lo_off() { return (P == 1 || PXS == 0) ? -P : -1; }
template <int P, int PXS>
__host__ __device__ inline constexpr int hi_off() { return (P == 1 || PXS == 0) ? P : 1; }

template <int P, int PX, int PY, int PZ, int MODE>
__host__ __device__ inline void lat_node(const LatArgs &a, const LatDev &C, long long k, const float *__restrict__ val,
                                const float *__restrict__ inv, const float *__restrict__ x,
                                const float *__restrict__ f, float *__restrict__ out, float omega) {
  const int kx = (int)(k % C.cnt[0]), ky = (int)((k / C.cnt[0]) % C.cnt[1]), kz = (int)(k / ((long long)C.cnt[0] * C.cnt[1]));
  const int X = C.lo[0] + C.step * kx, Y = C.lo[1] + C.step * ky, Z = C.lo[2] + C.step * kz;
  const long long nn = a.nn, node = X + nn * (Y + nn * Z);
  if (MODE == kLJacobiZero) {
    out[node] = omega * f[node] * inv[C.first + k];
    return;
  }
  const float *__restrict__ v = val + C.off + k;
  float s = 0.f;
  int si = 0;
#pragma unroll
  for (int dz = lo_off<P, PZ>(); dz <= hi_off<P, PZ>(); dz++)
#pragma unroll
    for (int dy = lo_off<P, PY>(); dy <= hi_off<P, PY>(); dy++)
#pragma unroll
      for (int dx = lo_off<P, PX>(); dx <= hi_off<P, PX>(); dx++, si++)
        s += v[(long long)si * C.nodes] * x[node + dx + nn * (dy + nn * dz)];
  if (MODE == kLDefect) out[node] = f[node] - s;
  else out[node] = x[node] + omega * (f[node] - s) * inv[C.first + k];
}

// node k of class c (the class fixes the parities, hence the stencil shape)
template <int P, int MODE>
__host__ __device__ inline void lat_class_node(const LatArgs &a, int c, long long k, const float *__restrict__ val,
                                               const float *__restrict__ inv, const float *__restrict__ x,
                                               const float *__restrict__ f, float *__restrict__ out, float omega) {
  const LatDev &C = a.c[c];
  if (P == 1) {
    lat_node<1, 0, 0, 0, MODE>(a, C, k, val, inv, x, f, out, omega);
  } else {
    switch (c) {
      case 0: lat_node<2, 0, 0, 0, MODE>(a, C, k, val, inv, x, f, out, omega); break;
      case 1: lat_node<2, 1, 0, 0, MODE>(a, C, k, val, inv, x, f, out, omega); break;
      case 2: lat_node<2, 0, 1, 0, MODE>(a, C, k, val, inv, x, f, out, omega); break;
      case 3: lat_node<2, 1, 1, 0, MODE>(a, C, k, val, inv, x, f, out, omega); break;
      case 4: lat_node<2, 0, 0, 1, MODE>(a, C, k, val, inv, x, f, out, omega); break;
      case 5: lat_node<2, 1, 0, 1, MODE>(a, C, k, val, inv, x, f, out, omega); break;
      case 6: lat_node<2, 0, 1, 1, MODE>(a, C, k, val, inv, x, f, out, omega); break;
      default: lat_node<2, 1, 1, 1, MODE>(a, C, k, val, inv, x, f, out, omega); break;
    }
  }
}

template <int P, int MODE>
__global__ __launch_bounds__(kB) void lat_sweep_k(LatArgs a, const float *__restrict__ val,
                                                  const float *__restrict__ inv, const float *__restrict__ x,
                                                  const float *__restrict__ f, float *__restrict__ out, float omega) {
  int c = 0;
  while (c + 1 < a.nclass && (int)blockIdx.x >= a.bstart[c + 1]) c++;
  const long long k = (long long)(blockIdx.x - a.bstart[c]) * kB + threadIdx.x;
  if (k < a.c[c].nodes) lat_class_node<P, MODE>(a, c, k, val, inv, x, f, out, omega);
}

// 1D transfer weights.  Restriction gathers, for a coarse node Xc, the fine
// nodes around 2 Xc; prolongation gathers, for a fine node X, the coarse nodes
// whose basis functions do not vanish at X.
template <int P>
__host__ __device__ inline int restr_1d(int Xc, int *off, float *w) {
  if (P == 1) {
    off[0] = -1; w[0] = 0.5f; off[1] = 0; w[1] = 1.f; off[2] = 1; w[2] = 0.5f;
    return 3;
  }
  if (Xc % 2 == 0) {
    off[0] = -3; w[0] = -0.125f; off[1] = -1; w[1] = 0.375f; off[2] = 0; w[2] = 1.f;
    off[3] = 1; w[3] = 0.375f; off[4] = 3; w[4] = -0.125f;
    return 5;
  }
  off[0] = -1; w[0] = 0.75f; off[1] = 0; w[1] = 1.f; off[2] = 1; w[2] = 0.75f;
  return 3;
}

template <int P>
__host__ __device__ inline int prol_1d(int X, int *idx, float *w) {
  if (P == 1) {
    if ((X & 1) == 0) { idx[0] = X / 2; w[0] = 1.f; return 1; }
    idx[0] = (X - 1) / 2; w[0] = 0.5f; idx[1] = (X + 1) / 2; w[1] = 0.5f;
    return 2;
  }
  const int t = X & 3;
  if (t == 0 || t == 2) { idx[0] = X / 2; w[0] = 1.f; return 1; }
  if (t == 1) {
    idx[0] = (X - 1) / 2; w[0] = 0.375f; idx[1] = (X + 1) / 2; w[1] = 0.75f; idx[2] = (X + 3) / 2; w[2] = -0.125f;
    return 3;
  }
  idx[0] = (X - 3) / 2; w[0] = -0.125f; idx[1] = (X - 1) / 2; w[1] = 0.75f; idx[2] = (X + 1) / 2; w[2] = 0.375f;
  return 3;
}

// coarse interior node t <- R fine (DefectRestriction + CorrectDefect: the
// coarse boundary stays zero)
template <int P>
__host__ __device__ inline void restrict_node(long long t, int nnf, int nnc, const float *__restrict__ fine,
                                              float *__restrict__ coarse) {
  const int m = nnc - 2;
  const int Xc = 1 + (int)(t % m), Yc = 1 + (int)((t / m) % m), Zc = 1 + (int)(t / ((long long)m * m));
  int ox[5], oy[5], oz[5];
  float wx[5], wy[5], wz[5];
  const int nx = restr_1d<P>(Xc, ox, wx), ny = restr_1d<P>(Yc, oy, wy), nz = restr_1d<P>(Zc, oz, wz);
  const long long F = nnf;
  float s = 0.f;
  for (int c = 0; c < nz; c++) {
    const long long zrow = (long long)(2 * Zc + oz[c]) * F * F;
    for (int b = 0; b < ny; b++) {
      const long long yrow = zrow + (long long)(2 * Yc + oy[b]) * F;
      float sy = 0.f;
      for (int a = 0; a < nx; a++) sy += wx[a] * fine[yrow + 2 * Xc + ox[a]];
      s += wz[c] * wy[b] * sy;
    }
  }
  coarse[Xc + (long long)nnc * (Yc + (long long)nnc * Zc)] = s;
}

template <int P>
__global__ __launch_bounds__(kB) void lat_restrict_k(int nnf, int nnc, const float *__restrict__ fine,
                                                     float *__restrict__ coarse) {
  const long long m = nnc - 2, t = (long long)blockIdx.x * kB + threadIdx.x;
  if (t < m * m * m) restrict_node<P>(t, nnf, nnc, fine, coarse);
}

// fine interior node t: x += damp (P e)  (Prolongate + CorrectNodes + Update)
template <int P>
__host__ __device__ inline void prolong_node(long long t, int nnf, int nnc, const float *__restrict__ coarse,
                                             float *__restrict__ x, float damp) {
  const int m = nnf - 2;
  const int X = 1 + (int)(t % m), Y = 1 + (int)((t / m) % m), Z = 1 + (int)(t / ((long long)m * m));
  int ix[3], iy[3], iz[3];
  float wx[3], wy[3], wz[3];
  const int nx = prol_1d<P>(X, ix, wx), ny = prol_1d<P>(Y, iy, wy), nz = prol_1d<P>(Z, iz, wz);
  const long long C = nnc;
  float s = 0.f;
  for (int c = 0; c < nz; c++)
    for (int b = 0; b < ny; b++) {
      float sy = 0.f;
      for (int a = 0; a < nx; a++) sy += wx[a] * coarse[ix[a] + C * (iy[b] + C * iz[c])];
      s += wz[c] * wy[b] * sy;
    }
  x[X + (long long)nnf * (Y + (long long)nnf * Z)] += damp * s;
}

template <int P>
__global__ __launch_bounds__(kB) void lat_prolong_k(int nnf, int nnc, const float *__restrict__ coarse,
                                                    float *__restrict__ x, float damp) {
  const long long m = nnf - 2, t = (long long)blockIdx.x * kB + threadIdx.x;
  if (t < m * m * m) prolong_node<P>(t, nnf, nnc, coarse, x, damp);
}

__global__ void lat_convert_k(long long n, const int *__restrict__ perm, const double *__restrict__ val,
                              float *__restrict__ out) {
  for (long long i = (long long)blockIdx.x * kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB)
    out[i] = (float)val[perm[i]];
}

__global__ void lat_invdiag_k(long long n, const int *__restrict__ node2dof, const int *__restrict__ diag,
                              const double *__restrict__ val, float *__restrict__ inv) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i < n) inv[i] = (float)(1.0 / val[diag[node2dof[i]]]);
}

// DOF-ordered f64 <-> lattice-ordered f32 (the boundary of the lattice vector
// is left untouched: it stays zero)
__global__ void dof_to_lat_k(int n, int nact, const int *__restrict__ dof2node, const double *__restrict__ v,
                             float *__restrict__ l) {
  const int d = blockIdx.x * kB + threadIdx.x;
  if (d < nact) l[dof2node[d]] = (float)v[d];
}

__global__ void lat_to_dof_k(int n, int nact, const int *__restrict__ dof2node, const float *__restrict__ l,
                             double *__restrict__ v) {
  const int d = blockIdx.x * kB + threadIdx.x;
  if (d < n) v[d] = d < nact ? (double)l[dof2node[d]] : 0.0;
}

LatArgs lat_args(const DevLevel &L) {
  LatArgs a{};
  a.nn = L.nn;
  a.nclass = L.nclass;
  int b = 0;
  for (int c = 0; c < L.nclass; c++) {
    const LatClass &C = L.cls[c];
    for (int d = 0; d < 3; d++) {
      a.c[c].px[d] = C.px[d];
      a.c[c].lo[d] = C.lo[d];
      a.c[c].cnt[d] = C.cnt[d];
    }
    a.c[c].step = C.step;
    a.c[c].nsten = C.nsten;
    a.c[c].nodes = C.nodes;
    a.c[c].off = C.off;
    a.c[c].first = C.first;
    a.bstart[c] = b;
    b += (int)((C.nodes + kB - 1) / kB);
  }
  a.bstart[L.nclass] = b;
  return a;
}

template <int MODE>
void lat_sweep(hipStream_t s, const DevLevel &L, int order, const float *x, const float *f, float *out,
               float omega) {
  const LatArgs a = lat_args(L);
  const int grid = a.bstart[L.nclass];
  if (grid <= 0) return;
  if (order == 1) lat_sweep_k<1, MODE><<<grid, kB, 0, s>>>(a, L.lval, L.linv, x, f, out, omega);
  else lat_sweep_k<2, MODE><<<grid, kB, 0, s>>>(a, L.lval, L.linv, x, f, out, omega);
}

}  // namespace

void kl_jacobi(hipStream_t s, const DevLevel &L, int order, const float *x, const float *f, float *out,
               float omega) {
  lat_sweep<kLJacobi>(s, L, order, x, f, out, omega);
}
void kl_jacobi_zero(hipStream_t s, const DevLevel &L, int order, const float *f, float *out, float omega) {
  lat_sweep<kLJacobiZero>(s, L, order, nullptr, f, out, omega);
}
void kl_defect(hipStream_t s, const DevLevel &L, int order, const float *x, const float *f, float *d) {
  lat_sweep<kLDefect>(s, L, order, x, f, d, 0.f);
}
void kl_restrict(hipStream_t s, const DevLevel &F, const DevLevel &C, int order, const float *fine, float *coarse) {
  const long long m = C.nn - 2, n = m * m * m;
  if (n <= 0) return;
  if (order == 1) lat_restrict_k<1><<<(int)((n + kB - 1) / kB), kB, 0, s>>>(F.nn, C.nn, fine, coarse);
  else lat_restrict_k<2><<<(int)((n + kB - 1) / kB), kB, 0, s>>>(F.nn, C.nn, fine, coarse);
}
void kl_prolong_update(hipStream_t s, const DevLevel &F, const DevLevel &C, int order, const float *coarse, float *x,
                       float damp) {
  const long long m = F.nn - 2, n = m * m * m;
  if (n <= 0) return;
  if (order == 1) lat_prolong_k<1><<<(int)((n + kB - 1) / kB), kB, 0, s>>>(F.nn, C.nn, coarse, x, damp);
  else lat_prolong_k<2><<<(int)((n + kB - 1) / kB), kB, 0, s>>>(F.nn, C.nn, coarse, x, damp);
}
void kl_refresh(hipStream_t s, const DevLevel &L) {
  if (L.lat_nval > 0)
    lat_convert_k<<<2048, kB, 0, s>>>(L.lat_nval, L.lperm, L.val, L.lval);
  if (L.lat_nint > 0)
    lat_invdiag_k<<<(int)((L.lat_nint + kB - 1) / kB), kB, 0, s>>>(L.lat_nint, L.lnode2dof, L.diag, L.val, L.linv);
}
void kl_from_dof(hipStream_t s, const DevLevel &L, const double *v, float *l) {
  dof_to_lat_k<<<(L.ndof + kB - 1) / kB, kB, 0, s>>>(L.ndof, L.nact, L.ldof2node, v, l);
}
void kl_to_dof(hipStream_t s, const DevLevel &L, const float *l, double *v) {
  lat_to_dof_k<<<(L.ndof + kB - 1) / kB, kB, 0, s>>>(L.ndof, L.nact, L.ldof2node, l, v);
}

}  // namespace cd3d

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

#include <algorithm>
#include <cstdlib>

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
  int uni[8];     // class stencil the same at every node: compact values
  int pstart[5];  // Q2: first block of each vertex/midpoint pair class (py, pz)
};

// 8-byte vector at a 4-byte aligned address (global_load_dwordx2 needs dword
// alignment only)
typedef float f2u __attribute__((ext_vector_type(2), aligned(4)));

enum LatMode { kLJacobi = 0, kLJacobiZero = 1, kLDefect = 2 };

// (kx, ky, kz) of node k of a class.  A class holds fewer than 2^31 nodes
// (every level's CSR has 32-bit column indices), so 32-bit unsigned division
// is exact and replaces three 64-bit division sequences per thread
__host__ __device__ inline void class_coords(const LatDev &C, long long k, int &kx, int &ky, int &kz) {
  const unsigned u = (unsigned)k, c0 = (unsigned)C.cnt[0], c1 = (unsigned)C.cnt[1], t = u / c0;
  kx = (int)(u - t * c0);
  ky = (int)(t % c1);
  kz = (int)(t / c1);
}

// class of block blk from the classes' first blocks (nondecreasing): seven
// independent comparisons instead of a scan that waits on each entry
__host__ __device__ inline int block_class(const int *bstart, int nclass, int blk) {
  int c = 0;
#pragma unroll
  for (int q = 1; q < 8; q++) c += (q < nclass && blk >= bstart[q]) ? 1 : 0;
  return c;
}

template <int P, int PXS>
__host__ __device__ inline constexpr int lo_off() { return (P == 1 || PXS == 0) ? -P : -1; }
template <int P, int PXS>
__host__ __device__ inline constexpr int hi_off() { return (P == 1 || PXS == 0) ? P : 1; }

// Compact class stencil: kCst floats per class, entry kCst - 1 is 1 / a_ii.
constexpr int kCst = 128;

// The FP64 value of the lattice system matrix alpha (S0 + X SX + Y SY) +
// alpha_m M at entry s of a node of coordinates (X, Y), with the rounding
// spelled out (explicit fma), so that the per-node fill and the sweeps that
// evaluate it on the fly produce the same bits.  s0, sx, sy, sm: the class's
// S0, SX, SY and mass stencils.
__host__ __device__ inline double affine_value(const double *s0, const double *sx, const double *sy,
                                               const double *sm, double alpha, double alpha_m, int s, int X,
                                               int Y) {
  const double e = fma((double)Y, sy[s], fma((double)X, sx[s], s0[s]));
  return alpha_m != 0.0 ? fma(alpha_m, sm[s], alpha * e) : alpha * e;
}

// The same entry for the FP32 preconditioner: the stencils rounded to FP32
// and the expression evaluated with FP32 fma (the FP64 operator keeps
// affine_value and matches the CSR values)
__host__ __device__ inline float affine_value32(const float *s0, const float *sx, const float *sy, const float *sm,
                                                float alpha, float alpha_m, int s, int X, int Y) {
  const float e = fmaf((float)Y, sy[s], fmaf((float)X, sx[s], s0[s]));
  return alpha_m != 0.f ? fmaf(alpha_m, sm[s], alpha * e) : alpha * e;
}

// UNI: the class's stencil (and diagonal) are the same at every node and are
// read from cst (uniform across the block); otherwise per-node values.
template <int P, int PX, int PY, int PZ, int MODE, bool UNI>
__host__ __device__ inline void lat_node(const LatArgs &a, const LatDev &C, long long k, const float *__restrict__ val,
                                         const float *__restrict__ inv, const float *__restrict__ cst,
                                         const float *__restrict__ x, const float *__restrict__ f,
                                         float *__restrict__ out, float omega) {
  int kx, ky, kz;
  class_coords(C, k, kx, ky, kz);
  const int X = C.lo[0] + C.step * kx, Y = C.lo[1] + C.step * ky, Z = C.lo[2] + C.step * kz;
  const long long nn = a.nn, node = X + nn * (Y + nn * Z);
  const float di = UNI ? cst[kCst - 1] : (MODE == kLDefect ? 0.f : inv[C.first + k]);
  if (MODE == kLJacobiZero) {
    out[node] = omega * f[node] * di;
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
        s += (UNI ? cst[si] : v[(long long)si * C.nodes]) * x[node + dx + nn * (dy + nn * dz)];
  if (MODE == kLDefect) out[node] = f[node] - s;
  else out[node] = x[node] + omega * (f[node] - s) * di;
}

// node k of class c (the class fixes the parities, hence the stencil shape)
template <int P, int PX, int PY, int PZ, int MODE>
__host__ __device__ inline void lat_node_any(const LatArgs &a, int c, long long k, const float *__restrict__ val,
                                             const float *__restrict__ inv, const float *__restrict__ cst,
                                             const float *__restrict__ x, const float *__restrict__ f,
                                             float *__restrict__ out, float omega) {
  if (a.uni[c]) lat_node<P, PX, PY, PZ, MODE, true>(a, a.c[c], k, val, inv, cst + c * kCst, x, f, out, omega);
  else lat_node<P, PX, PY, PZ, MODE, false>(a, a.c[c], k, val, inv, cst, x, f, out, omega);
}

// node k of class c (the class fixes the parities, hence the stencil shape)
template <int P, int MODE>
__host__ __device__ inline void lat_class_node(const LatArgs &a, int c, long long k, const float *__restrict__ val,
                                               const float *__restrict__ inv, const float *__restrict__ cst,
                                               const float *__restrict__ x, const float *__restrict__ f,
                                               float *__restrict__ out, float omega) {
  if (P == 1) {
    lat_node_any<1, 0, 0, 0, MODE>(a, c, k, val, inv, cst, x, f, out, omega);
  } else {
    switch (c) {
      case 0: lat_node_any<2, 0, 0, 0, MODE>(a, c, k, val, inv, cst, x, f, out, omega); break;
      case 1: lat_node_any<2, 1, 0, 0, MODE>(a, c, k, val, inv, cst, x, f, out, omega); break;
      case 2: lat_node_any<2, 0, 1, 0, MODE>(a, c, k, val, inv, cst, x, f, out, omega); break;
      case 3: lat_node_any<2, 1, 1, 0, MODE>(a, c, k, val, inv, cst, x, f, out, omega); break;
      case 4: lat_node_any<2, 0, 0, 1, MODE>(a, c, k, val, inv, cst, x, f, out, omega); break;
      case 5: lat_node_any<2, 1, 0, 1, MODE>(a, c, k, val, inv, cst, x, f, out, omega); break;
      case 6: lat_node_any<2, 0, 1, 1, MODE>(a, c, k, val, inv, cst, x, f, out, omega); break;
      default: lat_node_any<2, 1, 1, 1, MODE>(a, c, k, val, inv, cst, x, f, out, omega); break;
    }
  }
}

// Sweep with affine stencils evaluated at each node (rotating convection):
// the class's S0, SX, SY and mass stencils are staged in LDS as FP32 and each
// value is affine_value32, exactly what lat_fill_affine_k stores, so the
// results are bitwise those of lat_node on the filled per-node values,
// without their 4 bytes per nonzero of HBM traffic.
struct LatAff {
  const double *st;   // lst64: S0 at 0, SX at 8 kCst + 8, SY at 2 (8 kCst) + 8
  const double *stm;  // mass stencils (8 kCst) or null
  double alpha, alpha_m;
  int centre[8];      // stencil index of the offset (0, 0, 0) per class
};

template <int P, int PX, int PY, int PZ, int MODE>
__host__ __device__ inline void lat_node_aff(const LatArgs &a, const LatDev &C, long long k, const float *s0,
                                             const float *sx, const float *sy, const float *sm, float alpha,
                                             float alpha_m, int centre, const float *__restrict__ x,
                                             const float *__restrict__ f, float *__restrict__ out, float omega) {
  int kx, ky, kz;
  class_coords(C, k, kx, ky, kz);
  const int X = C.lo[0] + C.step * kx, Y = C.lo[1] + C.step * ky, Z = C.lo[2] + C.step * kz;
  const long long nn = a.nn, node = X + nn * (Y + nn * Z);
  const float di = MODE == kLDefect ? 0.f : 1.0f / affine_value32(s0, sx, sy, sm, alpha, alpha_m, centre, X, Y);
  if (MODE == kLJacobiZero) {
    out[node] = omega * f[node] * di;
    return;
  }
  float s = 0.f;
  int si = 0;
#pragma unroll
  for (int dz = lo_off<P, PZ>(); dz <= hi_off<P, PZ>(); dz++)
#pragma unroll
    for (int dy = lo_off<P, PY>(); dy <= hi_off<P, PY>(); dy++)
#pragma unroll
      for (int dx = lo_off<P, PX>(); dx <= hi_off<P, PX>(); dx++, si++)
        s += affine_value32(s0, sx, sy, sm, alpha, alpha_m, si, X, Y) * x[node + dx + nn * (dy + nn * dz)];
  if (MODE == kLDefect) out[node] = f[node] - s;
  else out[node] = x[node] + omega * (f[node] - s) * di;
}

template <int P, int MODE>
__host__ __device__ inline void lat_class_node_aff(const LatArgs &a, int c, long long k, const float *s0,
                                                   const float *sx, const float *sy, const float *sm,
                                                   float alpha, float alpha_m, int centre, const float *x,
                                                   const float *f, float *out, float omega) {
#define CD3D_AFF(PX, PY, PZ) \
  lat_node_aff<P, PX, PY, PZ, MODE>(a, a.c[c], k, s0, sx, sy, sm, alpha, alpha_m, centre, x, f, out, omega)
  if (P == 1) {
    CD3D_AFF(0, 0, 0);
  } else {
    switch (c) {
      case 0: CD3D_AFF(0, 0, 0); break;
      case 1: CD3D_AFF(1, 0, 0); break;
      case 2: CD3D_AFF(0, 1, 0); break;
      case 3: CD3D_AFF(1, 1, 0); break;
      case 4: CD3D_AFF(0, 0, 1); break;
      case 5: CD3D_AFF(1, 0, 1); break;
      case 6: CD3D_AFF(0, 1, 1); break;
      default: CD3D_AFF(1, 1, 1); break;
    }
  }
#undef CD3D_AFF
}

template <int P, int MODE>
__global__ __launch_bounds__(kB) void lat_aff_sweep_k(LatArgs a, LatAff af, const float *__restrict__ x,
                                                      const float *__restrict__ f, float *__restrict__ out,
                                                      float omega) {
  __shared__ float st[4][kCst];  // S0, SX, SY, mass of the block's class, rounded to FP32
  const int c = block_class(a.bstart, a.nclass, (int)blockIdx.x);
  for (int i = threadIdx.x; i < kCst; i += kB) {
    st[0][i] = (float)af.st[c * kCst + i];
    st[1][i] = (float)af.st[8 * kCst + 8 + c * kCst + i];
    st[2][i] = (float)af.st[2 * 8 * kCst + 8 + c * kCst + i];
    st[3][i] = af.stm ? (float)af.stm[c * kCst + i] : 0.f;
  }
  __syncthreads();
  const long long k = (long long)(blockIdx.x - a.bstart[c]) * kB + threadIdx.x;
  if (k < a.c[c].nodes)
    lat_class_node_aff<P, MODE>(a, c, k, st[0], st[1], st[2], st[3], (float)af.alpha, (float)af.alpha_m,
                                af.centre[c], x, f, out, omega);
}

// Q2 sweeps by vertex/midpoint pairs along X.  A thread takes the nodes
// X0 = 2 ix (cell vertex in X, class (0, py, pz)) and X0 + 1 (cell midpoint,
// class (1, py, pz)).  Both have the same Y/Z stencil extent, and together
// they need the segment X0-2 .. X0+2 of each neighbour line: two 2-element
// and one 1-element load, consecutive lanes on consecutive pairs, instead of
// 5 + 3 loads at stride 2 (half of every cache line touched) for the two
// classes separately.  ix = 0 is a boundary vertex: only its midpoint is
// written.  Each node's sum runs over its stencil in the per-class order, so
// results are bitwise those of the per-class kernels.
template <typename T>
struct Vec2;
template <>
struct Vec2<float> {
  typedef f2u type;
};
template <>
struct Vec2<double> {
  typedef double type __attribute__((ext_vector_type(2), aligned(8)));
};

struct PairPos {
  long long nv;  // lattice index of the vertex node (the midpoint is nv + 1)
  long long kv;  // the vertex's index in its class (when hv)
  bool hv;       // the vertex is interior (ix > 0)
};

// pair k of pair class (PY, PZ); k is also the midpoint's index in its class
template <int PY, int PZ>
__host__ __device__ inline PairPos pair_pos(const LatArgs &a, long long k) {
  const LatDev &M = a.c[2 * PY + 4 * PZ + 1];
  const int n = M.cnt[0];  // pairs per line
  // 32-bit unsigned division: lattice indices stay below 2^31 (class_coords)
  const unsigned u = (unsigned)k, r = u / (unsigned)n;  // ky + cnt[1] kz, the same for both classes
  const unsigned c1 = (unsigned)M.cnt[1], rz = r / c1;
  const int ix = (int)(u - r * (unsigned)n);
  const int Y = M.lo[1] + 2 * (int)(r - rz * c1), Z = M.lo[2] + 2 * (int)rz;
  PairPos q;
  q.nv = 2 * ix + (long long)a.nn * (Y + (long long)a.nn * Z);
  q.kv = (long long)r * (n - 1) + ix - 1;
  q.hv = ix > 0;
  return q;
}

// the stencil sums s0 (vertex) and s1 (midpoint) of a pair and the pair's
// own x values; UNI: compact stencils cs0, cs1, else per-node values val
// ([class][entry][node], FP32)
template <typename T, int PY, int PZ, bool UNI>
__host__ __device__ inline void pair_sums(const LatArgs &a, const PairPos &q, long long k,
                                          const float *__restrict__ val, const T *cs0, const T *cs1,
                                          const T *__restrict__ x, T &s0, T &s1, T &xv, T &xm) {
  typedef typename Vec2<T>::type V2;
  constexpr int c0 = 2 * PY + 4 * PZ, c1 = c0 + 1;
  const LatDev &V = a.c[c0], &M = a.c[c1];
  const long long nn = a.nn;
  const float *__restrict__ v0 = UNI ? nullptr : val + V.off + (q.hv ? q.kv : 0);
  const float *__restrict__ v1 = UNI ? nullptr : val + M.off + k;
  // a boundary vertex (ix = 0) has no sum: its left segment is read from the
  // centre instead (in bounds, unused), so the loads need no branch
  const int lft_off = q.hv ? -2 : 0;
  s0 = s1 = xv = xm = T(0);
  constexpr int ny = hi_off<2, PY>() - lo_off<2, PY>() + 1;
  // planes in a rolled loop: one plane's loads (ny lines) in flight at a time
  // keeps the registers of both stencils' coefficients and x values bounded
#pragma unroll 1
  for (int dz = lo_off<2, PZ>(); dz <= hi_off<2, PZ>(); dz++) {
    const int l0 = (dz - lo_off<2, PZ>()) * ny;
#pragma unroll
    for (int dy = lo_off<2, PY>(); dy <= hi_off<2, PY>(); dy++) {
      const int li = l0 + dy - lo_off<2, PY>();
      const T *xl = x + q.nv + nn * (dy + nn * dz);
      const V2 lft = *(const V2 *)(xl + lft_off);
      const V2 ctr = *(const V2 *)xl;
      const T rgt = xl[2];
      if (dy == 0 && dz == 0) {
        xv = ctr.x;
        xm = ctr.y;
      }
      const T w[5] = {lft.x, lft.y, ctr.x, ctr.y, rgt};
#pragma unroll
      for (int j = 0; j < 5; j++)
        s0 += (UNI ? cs0[li * 5 + j] : T(v0[(long long)(li * 5 + j) * V.nodes])) * w[j];
#pragma unroll
      for (int j = 0; j < 3; j++)
        s1 += (UNI ? cs1[li * 3 + j] : T(v1[(long long)(li * 3 + j) * M.nodes])) * w[j + 2];
    }
  }
}

// FP32 sweep of pair k (modes of lat_node); cs0, cs1: the compact stencils of
// the two classes (UNI)
template <int PY, int PZ, int MODE, bool UNI>
__host__ __device__ inline void lat_pair(const LatArgs &a, long long k, const float *__restrict__ val,
                                         const float *__restrict__ inv, const float *cs0, const float *cs1,
                                         const float *__restrict__ x, const float *__restrict__ f,
                                         float *__restrict__ out, float omega) {
  const PairPos q = pair_pos<PY, PZ>(a, k);
  const long long nv = q.nv;
  const bool hv = q.hv;
  const f2u fp = *(const f2u *)(f + nv);
  float d0 = 0.f, d1 = 0.f;
  if (UNI) {
    d0 = cs0[kCst - 1];
    d1 = cs1[kCst - 1];
  } else if (MODE != kLDefect) {
    d0 = hv ? inv[a.c[2 * PY + 4 * PZ].first + q.kv] : 0.f;
    d1 = inv[a.c[2 * PY + 4 * PZ + 1].first + k];
  }
  float r0, r1;
  if (MODE == kLJacobiZero) {
    r0 = omega * fp.x * d0;
    r1 = omega * fp.y * d1;
  } else {
    float s0, s1, xv, xm;
    pair_sums<float, PY, PZ, UNI>(a, q, k, val, cs0, cs1, x, s0, s1, xv, xm);
    if (MODE == kLDefect) {
      r0 = fp.x - s0;
      r1 = fp.y - s1;
    } else {
      r0 = xv + omega * (fp.x - s0) * d0;
      r1 = xm + omega * (fp.y - s1) * d1;
    }
  }
  // vertex and midpoint are adjacent: one 8-byte store unless the vertex is
  // a boundary node (not written)
  if (hv) *(f2u *)(out + nv) = f2u{r0, r1};
  else out[nv + 1] = r1;
}

// pair k of pair class pc = py + 2 pz; UNI: every class has compact values
// (cs0, cs1 those of the pair class), else every class has per-node values
template <int MODE, bool UNI>
__host__ __device__ inline void lat_pair_node(const LatArgs &a, int pc, long long k, const float *__restrict__ val,
                                              const float *__restrict__ inv, const float *cs0, const float *cs1,
                                              const float *__restrict__ x, const float *__restrict__ f,
                                              float *__restrict__ out, float omega) {
  switch (pc) {
    case 0: lat_pair<0, 0, MODE, UNI>(a, k, val, inv, cs0, cs1, x, f, out, omega); break;
    case 1: lat_pair<1, 0, MODE, UNI>(a, k, val, inv, cs0, cs1, x, f, out, omega); break;
    case 2: lat_pair<0, 1, MODE, UNI>(a, k, val, inv, cs0, cs1, x, f, out, omega); break;
    default: lat_pair<1, 1, MODE, UNI>(a, k, val, inv, cs0, cs1, x, f, out, omega); break;
  }
}

// UNI: the pair class's two compact stencils are staged in LDS (200 values
// do not fit the scalar registers)
template <int MODE, bool UNI>
__global__ __launch_bounds__(kB) void lat_pair_k(LatArgs a, const float *__restrict__ val,
                                                 const float *__restrict__ inv, const float *__restrict__ cst,
                                                 const float *__restrict__ x, const float *__restrict__ f,
                                                 float *__restrict__ out, float omega) {
  __shared__ float cs[2 * kCst];
  // blocks in pair-class order: a CU sweeps one class's code path at a time
  // (measured: interleaving the four classes per tile, also per XCD, saves
  // HBM traffic but is slower)
  int pc = 0;
  while (pc < 3 && (int)blockIdx.x >= a.pstart[pc + 1]) pc++;
  const int tile = blockIdx.x - a.pstart[pc];
  if (UNI) {
    static_assert(kB == 2 * kCst, "one stencil value per thread");
    cs[threadIdx.x] = cst[2 * pc * kCst + threadIdx.x];  // classes 2 pc and 2 pc + 1 are adjacent
    __syncthreads();
  }
  const long long k = (long long)tile * kB + threadIdx.x;
  if (k < a.c[2 * pc + 1].nodes) lat_pair_node<MODE, UNI>(a, pc, k, val, inv, cs, cs + kCst, x, f, out, omega);
}

// 16-byte vector at a 4-byte aligned address
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
// 16 bytes of T (at the alignment of T)
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  typedef f4u type;
  static constexpr int n = 4;
};
template <>
struct Vec16<double> {
  typedef Vec2<double>::type type;
  static constexpr int n = 2;
};

// G vertex/midpoint pairs per thread (compact stencils): the thread's nodes
// X0 .. X0 + 2G - 1 (X0 = 2 G g on its line) need the segment X0-2 .. X0+2G
// of every neighbour line, read as one 2-element load, 16-byte loads and one
// 1-element load (FP32: 2 + G/2 loads for 2G nodes, lat_pair_k needs 2 per
// pair; FP64: 2 + G instead of 3 per pair).  Every node's sum keeps the
// per-class order, so the results are bitwise those of one pair per thread.
// The group's first vertex is a boundary node when g = 0: it is not written,
// and its left segment is read from the centre.
template <int G, int PY, int PZ>
__host__ __device__ inline long long group_pos(const LatArgs &a, long long k, bool &hv) {
  const LatDev &M = a.c[2 * PY + 4 * PZ + 1];
  const int ng = M.cnt[0] / G;  // groups per line
  const unsigned u = (unsigned)k, r = u / (unsigned)ng, c1 = (unsigned)M.cnt[1], rz = r / c1;
  const int g = (int)(u - r * (unsigned)ng);
  const int Y = M.lo[1] + 2 * (int)(r - rz * c1), Z = M.lo[2] + 2 * (int)rz;
  hv = g > 0;
  return 2LL * G * g + (long long)a.nn * (Y + (long long)a.nn * Z);
}

template <typename T, int G, int PY, int PZ>
__host__ __device__ inline void group_sums(const LatArgs &a, long long nv, bool hv, const T *cs0, const T *cs1,
                                           const T *__restrict__ x, T (&s0)[G], T (&s1)[G], T (&xs)[2 * G]) {
  static_assert(G == 2 || G == 4, "two or four pairs per thread");
  typedef typename Vec2<T>::type V2;
  typedef typename Vec16<T>::type V16;
  constexpr int vn = Vec16<T>::n;
  const long long nn = a.nn;
#pragma unroll
  for (int q = 0; q < G; q++) s0[q] = s1[q] = T(0);
  const int lft_off = hv ? -2 : 0;
  constexpr int ny = hi_off<2, PY>() - lo_off<2, PY>() + 1;
#pragma unroll 1
  for (int dz = lo_off<2, PZ>(); dz <= hi_off<2, PZ>(); dz++) {
    const int l0 = (dz - lo_off<2, PZ>()) * ny;
#pragma unroll
    for (int dy = lo_off<2, PY>(); dy <= hi_off<2, PY>(); dy++) {
      const int li = l0 + dy - lo_off<2, PY>();
      const T *xl = x + nv + nn * (dy + nn * dz);
      T w[2 * G + 3];
      const V2 lft = *(const V2 *)(xl + lft_off);
      w[0] = lft.x;
      w[1] = lft.y;
#pragma unroll
      for (int h = 0; h < 2 * G / vn; h++) {
        const V16 m = *(const V16 *)(xl + vn * h);
#pragma unroll
        for (int e = 0; e < vn; e++) w[2 + vn * h + e] = m[e];
      }
      w[2 * G + 2] = xl[2 * G];
      if (dy == 0 && dz == 0) {
#pragma unroll
        for (int j = 0; j < 2 * G; j++) xs[j] = w[2 + j];
      }
#pragma unroll
      for (int q = 0; q < G; q++) {
#pragma unroll
        for (int j = 0; j < 5; j++) s0[q] += cs0[li * 5 + j] * w[2 * q + j];
#pragma unroll
        for (int j = 0; j < 3; j++) s1[q] += cs1[li * 3 + j] * w[2 * q + 2 + j];
      }
    }
  }
}

// FP32 sweep (modes of lat_node) of group k of pair class (PY, PZ)
template <int G, int PY, int PZ, int MODE>
__host__ __device__ inline void lat_pair_group(const LatArgs &a, long long k, const float *cs0, const float *cs1,
                                      const float *__restrict__ x, const float *__restrict__ f,
                                      float *__restrict__ out, float omega) {
  bool hv;
  const long long nv = group_pos<G, PY, PZ>(a, k, hv);
  const float d0 = cs0[kCst - 1], d1 = cs1[kCst - 1];
  float fv[2 * G];
#pragma unroll
  for (int h = 0; h < G / 2; h++) {
    const f4u t = *(const f4u *)(f + nv + 4 * h);
#pragma unroll
    for (int e = 0; e < 4; e++) fv[4 * h + e] = t[e];
  }
  if (MODE == kLJacobiZero) {
#pragma unroll
    for (int q = 0; q < G; q++) {
      if (q > 0 || hv) out[nv + 2 * q] = omega * fv[2 * q] * d0;
      out[nv + 2 * q + 1] = omega * fv[2 * q + 1] * d1;
    }
    return;
  }
  float s0[G], s1[G], xs[2 * G];
  group_sums<float, G, PY, PZ>(a, nv, hv, cs0, cs1, x, s0, s1, xs);
#pragma unroll
  for (int q = 0; q < G; q++) {
    if (MODE == kLDefect) {
      if (q > 0 || hv) out[nv + 2 * q] = fv[2 * q] - s0[q];
      out[nv + 2 * q + 1] = fv[2 * q + 1] - s1[q];
    } else {
      if (q > 0 || hv) out[nv + 2 * q] = xs[2 * q] + omega * (fv[2 * q] - s0[q]) * d0;
      out[nv + 2 * q + 1] = xs[2 * q + 1] + omega * (fv[2 * q + 1] - s1[q]) * d1;
    }
  }
}

template <int G, int MODE>
__host__ __device__ inline void lat_pair_group_node(const LatArgs &a, int pc, long long k, const float *cs0,
                                                    const float *cs1, const float *__restrict__ x,
                                                    const float *__restrict__ f, float *__restrict__ out,
                                                    float omega) {
  switch (pc) {
    case 0: lat_pair_group<G, 0, 0, MODE>(a, k, cs0, cs1, x, f, out, omega); break;
    case 1: lat_pair_group<G, 1, 0, MODE>(a, k, cs0, cs1, x, f, out, omega); break;
    case 2: lat_pair_group<G, 0, 1, MODE>(a, k, cs0, cs1, x, f, out, omega); break;
    default: lat_pair_group<G, 1, 1, MODE>(a, k, cs0, cs1, x, f, out, omega); break;
  }
}

// blocks in pair-class order, a.pstart counted in groups of G pairs
template <int MODE, int G>
__global__ __launch_bounds__(kB) void lat_pair_gk(LatArgs a, const float *__restrict__ cst,
                                                  const float *__restrict__ x, const float *__restrict__ f,
                                                  float *__restrict__ out, float omega) {
  __shared__ float cs[2 * kCst];
  const int bid = blockIdx.x;  // (XCD-contiguous blocks: 25.0 vs 18.6 us, 43.3 vs 26.4 us)
  int pc = 0;
  while (pc < 3 && bid >= a.pstart[pc + 1]) pc++;
  const int tile = bid - a.pstart[pc];
  cs[threadIdx.x] = cst[2 * pc * kCst + threadIdx.x];
  __syncthreads();
  const long long k = (long long)tile * kB + threadIdx.x;
  if (k < a.c[2 * pc + 1].nodes / G) lat_pair_group_node<G, MODE>(a, pc, k, cs, cs + kCst, x, f, out, omega);
}

// Q2 compact-stencil sweep by quads: one thread takes the same X segment (G
// vertex/midpoint pairs) of the four lines (Y0 + oy, Z0 + oz), oy, oz in
// {0, 1}, Y0 = 2 by, Z0 = 2 bz -- one line of each pair class.  The
// neighbour lines of the four together are the 5 x 5 lines Y0 - 2 .. Y0 + 2,
// Z0 - 2 .. Z0 + 2 (a midpoint line's stencil reaches one line either side):
// each is loaded once and feeds every class whose stencil covers it, 4x
// fewer gathers than a thread per pair-class line.  Each node's sum runs
// over its stencil in the order of lat_pair_group, so the results are
// bitwise the same.  The even lines Y0 = 0 / Z0 = 0 are boundary lines: not
// written, and the lines below them are read clamped to the boundary.
// position of quad thread k: segment g (G pairs) of the lines of block (by, bz)
struct QuadPos {
  int g, by, bz;
  long long base;  // node (2 G g, 2 by, 2 bz)
  bool hv;         // the segment's first vertex is interior (g > 0)
};
template <int G>
__host__ __device__ inline bool quad_pos(const LatArgs &a, long long k, QuadPos &p) {
  const int n = a.c[1].cnt[0];  // pairs per line (and blocks per axis)
  const int ng = n / G;
  if (k >= (long long)ng * n * n) return false;
  const unsigned u = (unsigned)k, r = u / (unsigned)ng, rz = r / (unsigned)n;
  p.g = (int)(u - r * (unsigned)ng);
  p.by = (int)(r - rz * (unsigned)n);
  p.bz = (int)rz;
  p.base = 2LL * G * p.g + (long long)a.nn * (2 * p.by + (long long)a.nn * 2 * p.bz);
  p.hv = p.g > 0;
  return true;
}

// the stencil sums of the 4 lines of a quad (cs: the 8 classes' compact
// stencils, class c at cs + c kCst; xs: the lines' own x values).  x is read
// as TX and each value converted to T (FP32 x in an FP64 sum: exact, so the
// sums are those of the FP64 copy of x)
template <typename T, int G, typename TX = T>
__host__ __device__ inline void quad_sums(const LatArgs &a, const QuadPos &p, const T *cs,
                                          const TX *__restrict__ x, T (&s0)[4][G], T (&s1)[4][G],
                                          T (&xs)[4][2 * G]) {
  typedef typename Vec2<TX>::type V2;
  typedef typename Vec16<TX>::type V16;
  constexpr int vn = Vec16<TX>::n;
  const long long nn = a.nn;
  const int Y0 = 2 * p.by, Z0 = 2 * p.bz;
#pragma unroll
  for (int pc = 0; pc < 4; pc++)
#pragma unroll
    for (int q = 0; q < G; q++) s0[pc][q] = s1[pc][q] = T(0);
  const int lft_off = p.hv ? -2 : 0;
#pragma unroll 1
  for (int dz = -2; dz <= 2; dz++) {
    const int zz = Z0 + dz < 0 ? 0 : Z0 + dz;
#pragma unroll
    for (int dy = -2; dy <= 2; dy++) {
      const int yy = Y0 + dy < 0 ? 0 : Y0 + dy;
      const TX *xl = x + 2LL * G * p.g + nn * (yy + nn * zz);
      T w[2 * G + 3];
      const V2 lft = *(const V2 *)(xl + lft_off);
      w[0] = (T)lft.x;
      w[1] = (T)lft.y;
#pragma unroll
      for (int h = 0; h < 2 * G / vn; h++) {
        const V16 m = *(const V16 *)(xl + vn * h);
#pragma unroll
        for (int e = 0; e < vn; e++) w[2 + vn * h + e] = (T)m[e];
      }
      if constexpr (2 * G < vn) {
        // one pair per line in FP32: the pair in one 8-byte load
        const V2 m = *(const V2 *)xl;
        w[2] = (T)m.x;
        w[3] = (T)m.y;
      }
      w[2 * G + 2] = (T)xl[2 * G];
#pragma unroll
      for (int oz = 0; oz < 2; oz++) {
        const int dzr = dz - oz, loz = oz ? -1 : -2, hiz = oz ? 1 : 2;
        if (dzr < loz || dzr > hiz) continue;
#pragma unroll
        for (int oy = 0; oy < 2; oy++) {
          const int dyr = dy - oy, loy = oy ? -1 : -2, hiy = oy ? 1 : 2;
          if (dyr < loy || dyr > hiy) continue;
          const int pc = oy + 2 * oz, ny = hiy - loy + 1;
          const int li = (dzr - loz) * ny + (dyr - loy);
          const T *cs0 = cs + 2 * pc * kCst, *cs1 = cs0 + kCst;
          if (dyr == 0 && dzr == 0) {
#pragma unroll
            for (int j = 0; j < 2 * G; j++) xs[pc][j] = w[2 + j];
          }
#pragma unroll
          for (int q = 0; q < G; q++) {
#pragma unroll
            for (int j = 0; j < 5; j++) s0[pc][q] += cs0[li * 5 + j] * w[2 * q + j];
#pragma unroll
            for (int j = 0; j < 3; j++) s1[pc][q] += cs1[li * 3 + j] * w[2 * q + 2 + j];
          }
        }
      }
    }
  }
}

// quad k of the FP32 sweep: G vertex/midpoint pairs on each of the 4 lines
// (Y0 + oy, Z0 + oz) (cs: the 8 classes' compact stencils)
template <int G, int MODE>
__host__ __device__ __attribute__((always_inline)) inline void lat_quad_node(const LatArgs &a, long long k,
                                                                             const float *cs,
                                              const float *__restrict__ x, const float *__restrict__ f,
                                              float *__restrict__ out, float omega) {
  QuadPos p;
  if (!quad_pos<G>(a, k, p)) return;
  const long long nn = a.nn;
  float s0[4][G], s1[4][G], xs[4][2 * G];
  if (MODE != kLJacobiZero) quad_sums<float, G>(a, p, cs, x, s0, s1, xs);
#pragma unroll
  for (int pc = 0; pc < 4; pc++) {
    const int oy = pc & 1, oz = pc >> 1;
    if ((oy == 0 && p.by == 0) || (oz == 0 && p.bz == 0)) continue;  // a boundary line
    const long long nv = p.base + nn * (oy + nn * oz);
    const float d0 = cs[2 * pc * kCst + kCst - 1], d1 = cs[(2 * pc + 1) * kCst + kCst - 1];
    float fv[2 * G];
    if constexpr (G == 1) {
      const f2u t = *(const f2u *)(f + nv);
      fv[0] = t.x;
      fv[1] = t.y;
    } else {
#pragma unroll
      for (int h = 0; h < G / 2; h++) {
        const f4u t = *(const f4u *)(f + nv + 4 * h);
#pragma unroll
        for (int e = 0; e < 4; e++) fv[4 * h + e] = t[e];
      }
    }
    float r[2 * G];
#pragma unroll
    for (int q = 0; q < G; q++) {
      if (MODE == kLJacobiZero) {
        r[2 * q] = omega * fv[2 * q] * d0;
        r[2 * q + 1] = omega * fv[2 * q + 1] * d1;
      } else if (MODE == kLDefect) {
        r[2 * q] = fv[2 * q] - s0[pc][q];
        r[2 * q + 1] = fv[2 * q + 1] - s1[pc][q];
      } else {
        r[2 * q] = xs[pc][2 * q] + omega * (fv[2 * q] - s0[pc][q]) * d0;
        r[2 * q + 1] = xs[pc][2 * q + 1] + omega * (fv[2 * q + 1] - s1[pc][q]) * d1;
      }
    }
    // the segment's 2G nodes are adjacent: vector stores (4-byte alignment
    // suffices) unless its first vertex is a boundary node (not written)
    if (p.hv) {
      if constexpr (G == 1) {
        *(f2u *)(out + nv) = f2u{r[0], r[1]};
      } else {
#pragma unroll
        for (int h = 0; h < G / 2; h++) *(f4u *)(out + nv + 4 * h) = f4u{r[4 * h], r[4 * h + 1], r[4 * h + 2], r[4 * h + 3]};
      }
    } else {
#pragma unroll
      for (int e = 1; e < 2 * G; e++) out[nv + e] = r[e];
    }
  }
}

template <int G, int MODE, int TB>
__global__ __launch_bounds__(TB) void lat_quad_gk(LatArgs a, const float *__restrict__ cst,
                                                  const float *__restrict__ x, const float *__restrict__ f,
                                                  float *__restrict__ out, float omega, int xcd) {
  __shared__ float cs[8 * kCst];
  for (int i = threadIdx.x; i < 8 * kCst; i += TB) cs[i] = cst[i];
  __syncthreads();
  const int bid = xcd ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  lat_quad_node<G, MODE>(a, (long long)bid * TB + threadIdx.x, cs, x, f, out, omega);
}

template <int P, int MODE>
__global__ __launch_bounds__(kB) void lat_sweep_k(LatArgs a, const float *__restrict__ val,
                                                  const float *__restrict__ inv, const float *__restrict__ cst,
                                                  const float *__restrict__ x, const float *__restrict__ f,
                                                  float *__restrict__ out, float omega) {
  const int c = block_class(a.bstart, a.nclass, (int)blockIdx.x);
  const long long k = (long long)(blockIdx.x - a.bstart[c]) * kB + threadIdx.x;
  if (k < a.c[c].nodes) lat_class_node<P, MODE>(a, c, k, val, inv, cst, x, f, out, omega);
}

// 1D transfer weights (nodal interpolation of the coarse basis at the fine
// nodes: Q1 1, 1/2; Q2 1, 3/4, 3/8, -1/8).
//
// Restriction R = P^T, one thread per coarse interior node: a coarse node Xc
// sits on fine node 2 Xc and gathers, per axis, Q1: offsets -1 0 1 with
// weights 1/2 1 1/2; Q2 even Xc (cell vertex): -3 -1 0 1 3 with -1/8 3/8 1
// 3/8 -1/8; Q2 odd Xc (cell midpoint): -1 0 1 with 3/4 1 3/4.  The parity of
// the coarse node fixes the pattern, so the coarse level's parity classes
// give loops with compile-time bounds (all loads in flight at once).
template <int P, int PXS>
struct Restr1D;
template <int PXS>
struct Restr1D<1, PXS> {
  static constexpr int n = 3;
  __host__ __device__ static constexpr int off(int i) { return i - 1; }
  __host__ __device__ static constexpr float w(int i) { return i == 1 ? 1.0f : 0.5f; }
};
template <>
struct Restr1D<2, 0> {
  static constexpr int n = 5;
  __host__ __device__ static constexpr int off(int i) { return i == 0 ? -3 : i == 1 ? -1 : i == 2 ? 0 : i == 3 ? 1 : 3; }
  __host__ __device__ static constexpr float w(int i) {
    return i == 0 || i == 4 ? -0.125f : i == 2 ? 1.0f : 0.375f;
  }
};
template <>
struct Restr1D<2, 1> {
  static constexpr int n = 3;
  __host__ __device__ static constexpr int off(int i) { return i - 1; }
  __host__ __device__ static constexpr float w(int i) { return i == 1 ? 1.0f : 0.75f; }
};

// The coarse level's first pre-smoothing sweep from a zero start,
// x = omega D^-1 f (lat_node's kLJacobiZero), is pointwise: when every class
// of the coarse level has a uniform stencil the restriction writes it too,
// omega * f * d_c at the nodes it writes (d_c = cst[c kCst + kCst - 1], the
// value and expression the sweep uses, so the bits are the same), and the
// sweep's launch is saved.  x == nullptr: f only.
struct ZeroSweep {
  float *x;
  const float *cst;
  float omega;
};

template <int P, int PX, int PY, int PZ>
__host__ __device__ inline void restrict_node(const LatArgs &ca, const LatDev &C, long long k, int nnf,
                                              const float *__restrict__ fine, float *__restrict__ coarse,
                                              int cls, const ZeroSweep &zs) {
  // the class value first: its load overlaps the gathers
  const float dz = zs.x ? zs.cst[cls * kCst + kCst - 1] : 0.f;
  using RX = Restr1D<P, PX>;
  using RY = Restr1D<P, PY>;
  using RZ = Restr1D<P, PZ>;
  int kx, ky, kz;
  class_coords(C, k, kx, ky, kz);
  const int Xc = C.lo[0] + C.step * kx, Yc = C.lo[1] + C.step * ky, Zc = C.lo[2] + C.step * kz;
  const long long F = nnf, centre = 2 * Xc + F * (2 * Yc + F * 2 * Zc);
  float s = 0.f;
#pragma unroll
  for (int c = 0; c < RZ::n; c++) {
    float sy = 0.f;
#pragma unroll
    for (int b = 0; b < RY::n; b++) {
      float sx = 0.f;
#pragma unroll
      for (int a = 0; a < RX::n; a++) sx += RX::w(a) * fine[centre + RX::off(a) + F * (RY::off(b) + F * RZ::off(c))];
      sy += RY::w(b) * sx;
    }
    s += RZ::w(c) * sy;
  }
  const long long node = Xc + (long long)ca.nn * (Yc + (long long)ca.nn * Zc);
  coarse[node] = s;
  if (zs.x) zs.x[node] = zs.omega * s * dz;
}

template <int P>
__host__ __device__ inline void restrict_class_node(const LatArgs &ca, int c, long long k, int nnf,
                                                    const float *__restrict__ fine, float *__restrict__ coarse,
                                                    const ZeroSweep &zs) {
  const LatDev &C = ca.c[c];
  if (P == 1) {
    restrict_node<1, 0, 0, 0>(ca, C, k, nnf, fine, coarse, c, zs);
  } else {
    switch (c) {
      case 0: restrict_node<2, 0, 0, 0>(ca, C, k, nnf, fine, coarse, c, zs); break;
      case 1: restrict_node<2, 1, 0, 0>(ca, C, k, nnf, fine, coarse, c, zs); break;
      case 2: restrict_node<2, 0, 1, 0>(ca, C, k, nnf, fine, coarse, c, zs); break;
      case 3: restrict_node<2, 1, 1, 0>(ca, C, k, nnf, fine, coarse, c, zs); break;
      case 4: restrict_node<2, 0, 0, 1>(ca, C, k, nnf, fine, coarse, c, zs); break;
      case 5: restrict_node<2, 1, 0, 1>(ca, C, k, nnf, fine, coarse, c, zs); break;
      case 6: restrict_node<2, 0, 1, 1>(ca, C, k, nnf, fine, coarse, c, zs); break;
      default: restrict_node<2, 1, 1, 1>(ca, C, k, nnf, fine, coarse, c, zs); break;
    }
  }
}

// coarse interior <- R fine (DefectRestriction + CorrectDefect: the coarse
// boundary stays zero); ca: the coarse level's classes
template <int P>
__global__ __launch_bounds__(kB) void lat_restrict_k(LatArgs ca, int nnf, const float *__restrict__ fine,
                                                     float *__restrict__ coarse, ZeroSweep zs) {
  const int c = block_class(ca.bstart, ca.nclass, (int)blockIdx.x);
  const long long k = (long long)(blockIdx.x - ca.bstart[c]) * kB + threadIdx.x;
  if (k < ca.c[c].nodes) restrict_class_node<P>(ca, c, k, nnf, fine, coarse, zs);
}

// Q2 restriction by coarse vertex/midpoint pairs along X (the pair classes
// of the coarse level, pair_pos): coarse nodes Xc (vertex) and Xc + 1
// (midpoint) gather fine X offsets -3 -1 0 1 3 around 2 Xc and -1 0 1
// around 2 Xc + 2, together the segment 2 Xc - 3 .. 2 Xc + 3 of each fine
// line (one 1-, one 2- and one 4-element load, consecutive lanes on
// consecutive pairs) instead of 5 + 3 loads at a stride of 4 floats.  Each
// node's sum keeps restrict_node's order.  A boundary vertex (ix = 0) is not
// written; its left values are read from the centre.
template <int PY, int PZ>
__host__ __device__ inline void restrict_pair(const LatArgs &ca, long long k, int nnf, const float *__restrict__ fine,
                                              float *__restrict__ coarse, const ZeroSweep &zs) {
  // vertex class 2 PY + 4 PZ, midpoint class one above; loaded first so the
  // loads overlap the gathers
  constexpr int cv = 2 * PY + 4 * PZ;
  const float dz0 = zs.x ? zs.cst[cv * kCst + kCst - 1] : 0.f, dz1 = zs.x ? zs.cst[(cv + 1) * kCst + kCst - 1] : 0.f;
  using RX0 = Restr1D<2, 0>;
  using RX1 = Restr1D<2, 1>;
  using RY = Restr1D<2, PY>;
  using RZ = Restr1D<2, PZ>;
  const PairPos q = pair_pos<PY, PZ>(ca, k);
  const long long F = nnf;
  const unsigned nnc = (unsigned)ca.nn, u = (unsigned)q.nv, t = u / nnc, tz = t / nnc;
  const long long Xc = u - t * nnc, Yc = t - tz * nnc, Zc = tz;
  const long long centre = 2 * Xc + F * (2 * Yc + F * 2 * Zc);
  const int lo = q.hv ? 0 : 3;
  float s0 = 0.f, s1 = 0.f;
#pragma unroll
  for (int c = 0; c < RZ::n; c++) {
    float sy0 = 0.f, sy1 = 0.f;
#pragma unroll
    for (int b = 0; b < RY::n; b++) {
      const float *fl = fine + centre + F * (RY::off(b) + F * RZ::off(c));
      float w[7];
      w[0] = fl[lo - 3];
      const f2u l2 = *(const f2u *)(fl + lo - 2);
      w[1] = l2.x;
      w[2] = l2.y;
      const f4u m = *(const f4u *)fl;
      w[3] = m.x;
      w[4] = m.y;
      w[5] = m.z;
      w[6] = m.w;
      float sx0 = 0.f, sx1 = 0.f;
#pragma unroll
      for (int a = 0; a < RX0::n; a++) sx0 += RX0::w(a) * w[3 + RX0::off(a)];
#pragma unroll
      for (int a = 0; a < RX1::n; a++) sx1 += RX1::w(a) * w[5 + RX1::off(a)];
      sy0 += RY::w(b) * sx0;
      sy1 += RY::w(b) * sx1;
    }
    s0 += RZ::w(c) * sy0;
    s1 += RZ::w(c) * sy1;
  }
  // vertex and midpoint are adjacent: one 8-byte store (4-byte alignment
  // suffices) unless the vertex is a boundary node
  if (q.hv) {
    *(f2u *)(coarse + q.nv) = f2u{s0, s1};
    if (zs.x) *(f2u *)(zs.x + q.nv) = f2u{zs.omega * s0 * dz0, zs.omega * s1 * dz1};
  } else {
    coarse[q.nv + 1] = s1;
    if (zs.x) zs.x[q.nv + 1] = zs.omega * s1 * dz1;
  }
}

__host__ __device__ inline void restrict_pair_node(const LatArgs &ca, int pc, long long k, int nnf,
                                                   const float *__restrict__ fine, float *__restrict__ coarse,
                                                   const ZeroSweep &zs) {
  switch (pc) {
    case 0: restrict_pair<0, 0>(ca, k, nnf, fine, coarse, zs); break;
    case 1: restrict_pair<1, 0>(ca, k, nnf, fine, coarse, zs); break;
    case 2: restrict_pair<0, 1>(ca, k, nnf, fine, coarse, zs); break;
    default: restrict_pair<1, 1>(ca, k, nnf, fine, coarse, zs); break;
  }
}

// inter: the four pair classes' blocks interleaved (logical block 4 j + pc
// is block j of class pc) and dealt to the XCDs in contiguous ranges, so the
// fine lines that coarse lines of all four parities share are read on one
// XCD at about the same time; otherwise class after class.  Measured on the
// 129^3 -> 65^3-node restriction: interleaved, the L2 fetches 10.3 instead
// of 22.2 MB (algorithmic 9.7 MB) but the kernel takes 6.5 instead of
// 6.0 us (the class-order re-reads come from the Infinity Cache), and the
// mixed solve is unchanged; class order is the default
// (CD3D_RESTRICT_INTER=1 selects the other, bitwise the same)
__global__ __launch_bounds__(kB) void lat_restrict_pair_k(LatArgs ca, int nnf, const float *__restrict__ fine,
                                                          float *__restrict__ coarse, int inter, ZeroSweep zs) {
  int pc = 0, j;
  if (inter) {
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    pc = lb & 3;
    j = lb >> 2;
    if (j >= ca.pstart[pc + 1] - ca.pstart[pc]) return;
  } else {
    while (pc < 3 && (int)blockIdx.x >= ca.pstart[pc + 1]) pc++;
    j = (int)blockIdx.x - ca.pstart[pc];
  }
  const long long k = (long long)j * kB + threadIdx.x;
  if (k < ca.c[2 * pc + 1].nodes) restrict_pair_node(ca, pc, k, nnf, fine, coarse, zs);
}

// Prolongation + CorrectNodes + Update, one thread per coarse cell: the cell's
// (P+1)^3 coarse values interpolated to its fine nodes with local index
// 0..2P-1 per axis (the last plane belongs to the next cell) by three 1D
// passes; fine boundary nodes are left alone.  Local fine position t in the
// cell: Q1 t = 0: (1, 0), t = 1: (1/2, 1/2); Q2 t = 0: (1, 0, 0),
// t = 1: (3/8, 3/4, -1/8), t = 2: (0, 1, 0), t = 3: (-1/8, 3/4, 3/8).
template <int P>
__host__ __device__ constexpr float prol_w(int t, int a) {
  return P == 1 ? (t == 0 ? (a == 0 ? 1.0f : 0.0f) : 0.5f)
                : (t == 0   ? (a == 0 ? 1.0f : 0.0f)
                   : t == 2 ? (a == 1 ? 1.0f : 0.0f)
                   : t == 1 ? (a == 0 ? 0.375f : a == 1 ? 0.75f : -0.125f)
                            : (a == 0 ? -0.125f : a == 1 ? 0.75f : 0.375f));
}

template <int P>
__host__ __device__ inline void prolong_cell(long long t, int ncc, int nnf, int nnc, const float *__restrict__ coarse,
                                             float *__restrict__ x, float damp) {
  constexpr int P1 = P + 1, T = 2 * P;
  const unsigned u = (unsigned)t, n = (unsigned)ncc, r = u / n, rz = r / n;
  const int cx = (int)(u - r * n), cy = (int)(r - rz * n), cz = (int)rz;
  const long long C = nnc, F = nnf;
  float cv[P1][P1][P1];
#pragma unroll
  for (int c = 0; c < P1; c++)
#pragma unroll
    for (int b = 0; b < P1; b++)
#pragma unroll
      for (int a = 0; a < P1; a++) cv[c][b][a] = coarse[(P * cx + a) + C * ((P * cy + b) + C * (P * cz + c))];
#pragma unroll
  for (int tz = 0; tz < T; tz++) {
    const int Z = T * cz + tz;
    if (Z == 0) continue;
    float pz[P1][P1];
#pragma unroll
    for (int b = 0; b < P1; b++)
#pragma unroll
      for (int a = 0; a < P1; a++) {
        float v = 0.f;
#pragma unroll
        for (int c = 0; c < P1; c++) v += prol_w<P>(tz, c) * cv[c][b][a];
        pz[b][a] = v;
      }
#pragma unroll
    for (int ty = 0; ty < T; ty++) {
      const int Y = T * cy + ty;
      if (Y == 0) continue;
      float py[P1];
#pragma unroll
      for (int a = 0; a < P1; a++) {
        float v = 0.f;
#pragma unroll
        for (int b = 0; b < P1; b++) v += prol_w<P>(ty, b) * pz[b][a];
        py[a] = v;
      }
#pragma unroll
      for (int tx = 0; tx < T; tx++) {
        const int X = T * cx + tx;
        if (X == 0) continue;
        float v = 0.f;
#pragma unroll
        for (int a = 0; a < P1; a++) v += prol_w<P>(tx, a) * py[a];
        x[X + F * (Y + F * (long long)Z)] += damp * v;
      }
    }
  }
}

// The same per fine node, one thread per (coarse cell, local fine Z, local
// fine Y): 4 P^2 times the threads of prolong_cell (coarse levels have few
// cells), and consecutive lanes on consecutive cells in X write 2P adjacent
// fine nodes each (coalesced).  Each node's value is formed in prolong_cell's
// order (Z pass, Y pass, X pass); the host check finds the same bits.  On the
// device the weights of prolong_cell are compile-time constants and its
// products may be folded differently, so every device path (the
// launch-per-step V-cycle and the one-workgroup program) uses this form.
template <int P>
__host__ __device__ inline void prolong_row(long long t, int ncc, int nnf, int nnc, const float *__restrict__ coarse,
                                            float *__restrict__ x, float damp) {
  constexpr int P1 = P + 1, T = 2 * P;
  // 32-bit unsigned division: t counts fine-level rows, below 2^31
  const unsigned n = (unsigned)ncc;
  unsigned r = (unsigned)t / n;
  const int cx = (int)((unsigned)t - r * n);
  const int ty = (int)(r % T);
  r /= T;
  const int tz = (int)(r % T);
  r /= T;
  const int cy = (int)(r % n), cz = (int)(r / n);
  const int Y = T * cy + ty, Z = T * cz + tz;
  if (Y == 0 || Z == 0) return;
  const long long C = nnc, F = nnf;
  float pz[P1][P1];
#pragma unroll
  for (int b = 0; b < P1; b++)
#pragma unroll
    for (int a = 0; a < P1; a++) {
      float v = 0.f;
#pragma unroll
      for (int c = 0; c < P1; c++)
        v += prol_w<P>(tz, c) * coarse[(P * cx + a) + C * ((P * cy + b) + C * (P * cz + c))];
      pz[b][a] = v;
    }
  float py[P1];
#pragma unroll
  for (int a = 0; a < P1; a++) {
    float v = 0.f;
#pragma unroll
    for (int b = 0; b < P1; b++) v += prol_w<P>(ty, b) * pz[b][a];
    py[a] = v;
  }
#pragma unroll
  for (int tx = 0; tx < T; tx++) {
    const int X = T * cx + tx;
    if (X == 0) continue;
    float v = 0.f;
#pragma unroll
    for (int a = 0; a < P1; a++) v += prol_w<P>(tx, a) * py[a];
    x[X + F * (Y + F * (long long)Z)] += damp * v;
  }
}

template <int P>
__global__ __launch_bounds__(kB) void lat_prolong_k(int ncc, int nnf, int nnc, const float *__restrict__ coarse,
                                                    float *__restrict__ x, float damp) {
  const long long t = (long long)blockIdx.x * kB + threadIdx.x;
  if (t < (long long)ncc * ncc * ncc * 4 * P * P) prolong_row<P>(t, ncc, nnf, nnc, coarse, x, damp);
}

// ---------------------------------------------------------------- small levels
// The bottom of a mixed V-cycle in one launch: one workgroup runs the
// pre-smoothing, defect and restriction of the small FP32 levels top..1, the
// exact coarse solve x_A = A_AA^-1 f_A (zero iterate, zero Dirichlet data, as
// coarse_lat_k), and the prolongation and post-smoothing back up, with a
// workgroup barrier between steps.  Each step is the per-node code of the
// multi-workgroup kernels (lat_class_node, restrict_class_node,
// prolong_cell) over the same nodes, and the iterates ping-pong between lx
// and lt exactly as the host loop does, so the results are bitwise those of
// the launch-per-step V-cycle.  On small levels every launch is mostly fixed
// cost; one launch replaces 9 per level plus the coarse solve.
constexpr int kSmallMax = 3;  // FP32 levels in one program
// workgroup size: Q2's 125-point stencils need more than 128 registers per
// lane (1024 threads) without spilling
template <int P>
constexpr int small_threads() {
  return P == 1 ? 1024 : 512;
}

struct SmallLevel {
  LatArgs a;
  const float *val, *inv, *cst;
  float *lx, *lt, *la, *lf;
};

struct SmallArgs {
  SmallLevel L[kSmallMax + 1];  // L[0]: the coarsest level (a, lx, lf), L[1..top]: FP32 levels
  int top, pre, post, na;
  float omega, damp;
  const double *Ainv;
  const int *dof2node0;
};
static_assert(sizeof(SmallArgs) <= 4096, "kernel arguments are limited to 4 KiB");

// UNI: every class of the level has compact stencils (else per-node values)
template <int P, int MODE, bool UNI>
__device__ inline void small_sweep(const SmallLevel &L, const float *x, const float *f, float *out, float omega) {
  for (int c = 0; c < L.a.nclass; c++) {
    const float *cst = L.cst + c * kCst;
    for (long long k = threadIdx.x; k < L.a.c[c].nodes; k += small_threads<P>()) {
      if (P == 1) {
        lat_node<1, 0, 0, 0, MODE, UNI>(L.a, L.a.c[c], k, L.val, L.inv, cst, x, f, out, omega);
      } else {
        switch (c) {
          case 0: lat_node<2, 0, 0, 0, MODE, UNI>(L.a, L.a.c[c], k, L.val, L.inv, cst, x, f, out, omega); break;
          case 1: lat_node<2, 1, 0, 0, MODE, UNI>(L.a, L.a.c[c], k, L.val, L.inv, cst, x, f, out, omega); break;
          case 2: lat_node<2, 0, 1, 0, MODE, UNI>(L.a, L.a.c[c], k, L.val, L.inv, cst, x, f, out, omega); break;
          case 3: lat_node<2, 1, 1, 0, MODE, UNI>(L.a, L.a.c[c], k, L.val, L.inv, cst, x, f, out, omega); break;
          case 4: lat_node<2, 0, 0, 1, MODE, UNI>(L.a, L.a.c[c], k, L.val, L.inv, cst, x, f, out, omega); break;
          case 5: lat_node<2, 1, 0, 1, MODE, UNI>(L.a, L.a.c[c], k, L.val, L.inv, cst, x, f, out, omega); break;
          case 6: lat_node<2, 0, 1, 1, MODE, UNI>(L.a, L.a.c[c], k, L.val, L.inv, cst, x, f, out, omega); break;
          default: lat_node<2, 1, 1, 1, MODE, UNI>(L.a, L.a.c[c], k, L.val, L.inv, cst, x, f, out, omega); break;
        }
      }
    }
  }
}

// UNI: every small level has compact stencils in every class
template <int P, bool UNI>
__global__ __launch_bounds__(small_threads<P>()) void lat_small_k(SmallArgs s) {
  constexpr int NT = small_threads<P>();
  float *xs[kSmallMax + 1];
  // down: levels top .. 1
  for (int i = s.top; i >= 1; i--) {
    const SmallLevel &L = s.L[i], &C = s.L[i - 1];
    float *x = L.lx, *t = L.lt;
    for (int k = 0; k < s.pre; k++) {
      if (k == 0) small_sweep<P, kLJacobiZero, UNI>(L, nullptr, L.lf, t, s.omega);
      else small_sweep<P, kLJacobi, UNI>(L, x, L.lf, t, s.omega);
      float *u = x;
      x = t;
      t = u;
      __syncthreads();
    }
    if (s.pre == 0) {
      const long long n3 = (long long)L.a.nn * L.a.nn * L.a.nn;
      for (long long k = threadIdx.x; k < n3; k += NT) x[k] = 0.f;
      __syncthreads();
    }
    small_sweep<P, kLDefect, UNI>(L, x, L.lf, L.la, 0.f);
    __syncthreads();
    for (int c = 0; c < C.a.nclass; c++)
      for (long long k = threadIdx.x; k < C.a.c[c].nodes; k += NT)
        restrict_class_node<P>(C.a, c, k, L.a.nn, L.la, C.lf, ZeroSweep{nullptr, nullptr, 0.f});
    xs[i] = x;
    __syncthreads();
  }
  // coarsest: one wavefront per row, the lanes over the columns
  {
    const int lane = threadIdx.x & 63;
    for (int row = threadIdx.x >> 6; row < s.na; row += NT / 64) {
      double v = 0;
      for (int j = lane; j < s.na; j += 64) v += s.Ainv[(long long)row * s.na + j] * (double)s.L[0].lf[s.dof2node0[j]];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) s.L[0].lx[s.dof2node0[row]] = (float)(0.0 + v);
    }
    __syncthreads();
  }
  // up: levels 1 .. top
  for (int i = 1; i <= s.top; i++) {
    const SmallLevel &L = s.L[i], &C = s.L[i - 1];
    float *x = xs[i], *t = x == L.lx ? L.lt : L.lx;
    const int ncc = (C.a.nn - 1) / P;
    for (long long c = threadIdx.x; c < (long long)ncc * ncc * ncc * 4 * P * P; c += NT)
      prolong_row<P>(c, ncc, L.a.nn, C.a.nn, C.lx, x, s.damp);
    __syncthreads();
    for (int k = 0; k < s.post; k++) {
      small_sweep<P, kLJacobi, UNI>(L, x, L.lf, t, s.omega);
      float *u = x;
      x = t;
      t = u;
      __syncthreads();
    }
    if (x != L.lx) {
      const long long n3 = (long long)L.a.nn * L.a.nn * L.a.nn;
      for (long long k = threadIdx.x; k < n3; k += NT) L.lx[k] = x[k];
      __syncthreads();
    }
  }
}

__global__ void lat_convert_k(long long n, const int *__restrict__ perm, const double *__restrict__ val,
                              float *__restrict__ out) {
  for (long long i = (long long)blockIdx.x * kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB)
    out[i] = (float)val[perm[i]];
}

// Per class: the largest |a| of the first node's stencil (the scale of the
// uniformity test) and the compact stencil of that node; clears the flag.
__global__ void lat_first_node_k(LatArgs a, const int *__restrict__ first, const int *__restrict__ node2dof,
                                 const int *__restrict__ diag, const double *__restrict__ val,
                                 double *__restrict__ scale, int *__restrict__ flag, float *__restrict__ cst) {
  const int c = blockIdx.x;
  const LatDev &C = a.c[c];
  double m = 0.0;
  if (C.nodes > 0)
    for (int s = threadIdx.x; s < C.nsten; s += 64) {
      const double v = val[first[c * kCst + s]];
      cst[c * kCst + s] = (float)v;
      m = fmax(m, fabs(v));
    }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if (threadIdx.x == 0) {
    scale[c] = m;
    flag[c] = C.nodes > 0 ? 0 : 1;
    if (C.nodes > 0) cst[c * kCst + kCst - 1] = (float)(1.0 / val[diag[node2dof[C.first]]]);
  }
}

// flag[c] = 1 when some node of class c has a stencil entry that differs from
// the first node's by more than 1e-12 of the stencil's largest entry
// (active-row CSR entries in CSR order, each with its stencil id)
__global__ void lat_uniform_k(long long n, const uint16_t *__restrict__ sid, const int *__restrict__ first,
                              const double *__restrict__ val, const double *__restrict__ scale,
                              int *__restrict__ flag) {
  for (long long i = (long long)blockIdx.x * kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB) {
    const int id = sid[i], c = id >> 7;
    if (fabs(val[i] - val[first[id]]) > 1e-12 * scale[c]) flag[c] = 1;
  }
}

__global__ void lat_invdiag_k(long long n, const int *__restrict__ node2dof, const int *__restrict__ diag,
                              const double *__restrict__ val, float *__restrict__ inv) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i < n) inv[i] = (float)(1.0 / val[diag[node2dof[i]]]);
}

// ------------------------------------------------------------------------
// Assembly with constant coefficients (TAssembleMat3D::Assemble3D with
// BilinearAssemble / BilinearAssemble_SD, AssembleMat3D.C:230,
// ConvDiff3D.C:61,124).  On the uniform brick mesh every cell has the same
// physical element matrix -- the child frames only renumber its rows and
// columns -- so the element matrix is integrated once (elem_k, the integrand
// of the per-cell kernel in the identity frame), folded into one stencil per
// parity class (the at most 8 cells around an interior node), and written to
// the CSR values of every interior row; Dirichlet rows are set apart.

// E[i][j] (NB x NB, cell-local lexicographic numbering) and the element rhs
// E[NB*NB + i]; one thread per entry.  what: 0 conv-diff, 1 mass (rhs with
// the SUPG test function when supg), 2 the SUPG time-derivative matrix K
template <int P>
__global__ __launch_bounds__(1024) void elem_k(const double *__restrict__ tab, AsmParams ap, double *__restrict__ E) {
  constexpr int NB = (P + 1) * (P + 1) * (P + 1), N1 = P + 2, NQ = N1 * N1 * N1, STRIDE = 4 + 4 * NB;
  // the basis table (values and reference gradients at the quadrature
  // points) is read by every entry: staged in LDS once
  __shared__ double Ts[NQ * STRIDE];
  for (int k = threadIdx.x; k < NQ * STRIDE; k += blockDim.x) Ts[k] = tab[k];
  __syncthreads();
  const int t = threadIdx.x;
  if (t >= NB * NB) return;
  const int i = t / NB, j = t % NB;
  const double detj = (ap.h[0] / 2) * (ap.h[1] / 2) * (ap.h[2] / 2);
  double kap[3], beta[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    kap[a] = ap.eps * 4.0 / (ap.h[a] * ap.h[a]);
    beta[a] = ap.b[a] * 2.0 / ap.h[a];
  }
  const double delta = ap.supg ? ap.delta_const : 0.0;
  double acc = 0.0, racc = 0.0;
  for (int q = 0; q < NQ; q++) {
    const double *T = Ts + q * STRIDE;
    const double w = T[0] * detj;
    const double pi_ = T[4 + 4 * i], gi0 = T[5 + 4 * i], gi1 = T[6 + 4 * i], gi2 = T[7 + 4 * i];
    const double pj = T[4 + 4 * j], gj0 = T[5 + 4 * j], gj1 = T[6 + 4 * j], gj2 = T[7 + 4 * j];
    if (ap.what == 1) {  // mass, rhs with the SUPG test function when supg (TCD3D.C:48)
      acc += w * (pi_ * pj);
      racc += w * ap.f_val * (pi_ + delta * (beta[0] * gi0 + beta[1] * gi1 + beta[2] * gi2));
    } else if (ap.what == 2) {  // SUPG time-derivative matrix K, no rhs (TCD3D.C:102)
      acc += w * (delta * (beta[0] * gi0 + beta[1] * gi1 + beta[2] * gi2) * pj);
    } else {
      const double bgi = beta[0] * gi0 + beta[1] * gi1 + beta[2] * gi2;
      const double bgj = beta[0] * gj0 + beta[1] * gj1 + beta[2] * gj2;
      const double ti = pi_ + delta * bgi;
      const double v = kap[0] * gi0 * gj0 + kap[1] * gi1 * gj1 + kap[2] * gi2 * gj2 + ti * (bgj + ap.c * pj);
      acc += w * v;
      racc += w * ap.f_val * ti;
    }
  }
  E[i * NB + j] = acc;
  if (j == 0) E[NB * NB + i] = racc;
}

// Per axis, the cell-local coordinates of a node of parity px in the cells
// that contain it (two cells on a cell-face plane, one inside a Q2 cell), and
// the stencil offsets of the class.
template <int P>
__host__ __device__ inline void class_axis(int px, int *nl, int *l, int *lo, int *len) {
  if (P == 2 && px == 1) { *nl = 1; l[0] = 1; l[1] = 1; *lo = -1; *len = 3; }
  else { *nl = 2; l[0] = P; l[1] = 0; *lo = -P; *len = 2 * P + 1; }
}

template <int P>
__host__ __device__ inline double stencil_entry(const double *__restrict__ E, const int (&nl)[3], const int (&l)[3][2],
                                       const int (&d)[3]) {
  constexpr int P1 = P + 1, NB = P1 * P1 * P1;
  double v = 0.0;
  for (int az = 0; az < nl[2]; az++)
    for (int ay = 0; ay < nl[1]; ay++)
      for (int ax = 0; ax < nl[0]; ax++) {
        const int lx = l[0][ax], ly = l[1][ay], lz = l[2][az];
        const int mx = lx + d[0], my = ly + d[1], mz = lz + d[2];
        if (mx < 0 || mx > P || my < 0 || my > P || mz < 0 || mz > P) continue;
        v += E[(lx + P1 * (ly + P1 * lz)) * NB + (mx + P1 * (my + P1 * mz))];
      }
  return v;
}

// st[c*kCst + s]: FP64 class stencils; st[8*kCst + c]: class rhs;
// cst: the FP32 compact stencils of the V-cycle (last entry 1/a_ii)
template <int P>
__global__ __launch_bounds__(kCst) void class_stencil_k(LatArgs a, const double *__restrict__ E,
                                                        double *__restrict__ st, float *__restrict__ cst) {
  constexpr int P1 = P + 1, NB = P1 * P1 * P1;
  const int c = blockIdx.x, s = threadIdx.x;
  const LatDev &C = a.c[c];
  int nl[3], l[3][2], lo[3], len[3];
#pragma unroll
  for (int d = 0; d < 3; d++) class_axis<P>(C.px[d], &nl[d], l[d], &lo[d], &len[d]);
  if (s < C.nsten) {
    const int d[3] = {lo[0] + s % len[0], lo[1] + (s / len[0]) % len[1], lo[2] + s / (len[0] * len[1])};
    const double v = stencil_entry<P>(E, nl, l, d);
    st[c * kCst + s] = v;
    cst[c * kCst + s] = (float)v;
  }
  if (s == kCst - 1) {
    const int d0[3] = {0, 0, 0};
    cst[c * kCst + kCst - 1] = (float)(1.0 / stencil_entry<P>(E, nl, l, d0));
    double r = 0.0;
    for (int az = 0; az < nl[2]; az++)
      for (int ay = 0; ay < nl[1]; ay++)
        for (int ax = 0; ax < nl[0]; ax++) r += E[NB * NB + l[0][ax] + P1 * (l[1][ay] + P1 * l[2][az])];
    st[8 * kCst + c] = r;
  }
}

// val at the CSR position of (class z, stencil entry y, node x) = class stencil
// the active-row CSR entries in CSR order (coalesced writes): val = class stencil entry
__global__ void csr_fill_k(long long n, const uint16_t *__restrict__ sid, const double *__restrict__ st,
                           double *__restrict__ val) {
  for (long long i = (long long)blockIdx.x * kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB)
    val[i] = st[sid[i]];
}

// the same, four entries per thread (8-byte id loads, 16-byte value stores;
// measured: non-temporal stores are slower here); sid and val 16-byte
// aligned, entries 4 i .. 4 i + 3 for i < n4
// two entries per lane and store, consecutive lanes on consecutive pairs:
// every store instruction writes one contiguous 1 KB run; two independent
// pairs per iteration (130 M entries: 315 us; four entries per lane in two
// interleaved 16-byte stores: 400 us; 4 or 8 pairs in flight: 327-340 us)
__global__ void csr_fill2_k(long long n2, const uint16_t *__restrict__ sid, const double *__restrict__ st,
                            double *__restrict__ val) {
  typedef unsigned short us2 __attribute__((ext_vector_type(2)));
  typedef double d2 __attribute__((ext_vector_type(2)));
  const long long stride = (long long)gridDim.x * kB;
  long long i = (long long)blockIdx.x * kB + threadIdx.x;
  for (; i + stride < n2; i += 2 * stride) {
    const us2 a = ((const us2 *)sid)[i], b = ((const us2 *)sid)[i + stride];
    const d2 va = {st[a.x], st[a.y]}, vb = {st[b.x], st[b.y]};
    ((d2 *)val)[i] = va;
    ((d2 *)val)[i + stride] = vb;
  }
  if (i < n2) {
    const us2 a = ((const us2 *)sid)[i];
    ((d2 *)val)[i] = d2{st[a.x], st[a.y]};
  }
}

__global__ void rhs_fill_k(LatArgs a, long long n, const int *__restrict__ node2dof, const double *__restrict__ st,
                           double *__restrict__ rhs) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  int c = 0;
  while (c + 1 < a.nclass && i >= a.c[c + 1].first) c++;
  rhs[node2dof[i]] = st[8 * kCst + c];
}

// ------------------------------------------------------------------------
// Assembly with the rotating field b = (-(y - r1), x - r0, b3), Galerkin
// (TCD3D's MatrixARhsAssemble, TCD3D.C:242-295; BilinearAssemble,
// ConvDiff3D.C:61).  b is affine in space and the quadrature integrates the
// element integrand exactly, so a cell with centre (xc, yc) has the matrix
//   A_const + (-(yc - r1)) Dx + (xc - r0) Dy,
//   A_const = eps K + c M + b3 Dz - int (y - yc) d_x phi_j phi_i + int (x - xc) d_y phi_j phi_i,
//   Dx = int d_x phi_j phi_i,  Dy = int d_y phi_j phi_i,
// and the stencil of an interior node (X, Y) is S0 + X SX + Y SY per parity
// class.  The rhs (constant f) is one value per class.

// E: [A_const NB x NB][rhs NB][Dx NB x NB][Dy NB x NB]; one thread per entry
template <int P>
__global__ __launch_bounds__(1024) void elem_affine_k(const double *__restrict__ tab, AsmParams ap,
                                                      double *__restrict__ E) {
  constexpr int NB = (P + 1) * (P + 1) * (P + 1), N1 = P + 2, NQ = N1 * N1 * N1, STRIDE = 4 + 4 * NB;
  // the basis table (values and reference gradients at the quadrature
  // points) is read by every entry: staged in LDS once
  __shared__ double Ts[NQ * STRIDE];
  for (int k = threadIdx.x; k < NQ * STRIDE; k += blockDim.x) Ts[k] = tab[k];
  __syncthreads();
  const int t = threadIdx.x;
  if (t >= NB * NB) return;
  const int i = t / NB, j = t % NB;
  const double detj = (ap.h[0] / 2) * (ap.h[1] / 2) * (ap.h[2] / 2);
  double kap[3], g2[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    kap[a] = ap.eps * 4.0 / (ap.h[a] * ap.h[a]);
    g2[a] = 2.0 / ap.h[a];
  }
  const double b3 = 0.1 * ap.sin2pit;
  double a0 = 0.0, ax = 0.0, ay = 0.0, racc = 0.0;
  for (int q = 0; q < NQ; q++) {
    const double *T = Ts + q * STRIDE;
    const double w = T[0] * detj;
    const double dxq = T[1] * ap.h[0] / 2, dyq = T[2] * ap.h[1] / 2;  // qp - cell centre
    const double pi_ = T[4 + 4 * i], gi0 = T[5 + 4 * i], gi1 = T[6 + 4 * i], gi2 = T[7 + 4 * i];
    const double pj = T[4 + 4 * j], gj0 = T[5 + 4 * j] * g2[0], gj1 = T[6 + 4 * j] * g2[1], gj2 = T[7 + 4 * j] * g2[2];
    const double v = kap[0] * gi0 * T[5 + 4 * j] + kap[1] * gi1 * T[6 + 4 * j] + kap[2] * gi2 * T[7 + 4 * j] +
                     pi_ * (ap.c * pj + b3 * gj2 - dyq * gj0 + dxq * gj1);
    a0 += w * v;
    ax += w * gj0 * pi_;
    ay += w * gj1 * pi_;
    racc += w * ap.f_val * pi_;
  }
  E[i * NB + j] = a0;
  E[NB * NB + NB + i * NB + j] = ax;
  E[2 * NB * NB + NB + i * NB + j] = ay;
  if (j == 0) E[NB * NB + i] = racc;
}

// one stencil entry (offset d) of a node of the class: the sum over the
// cells around the node of A_const + beta1 Dx + beta2 Dy with beta at the
// cell centre, split into the parts constant, linear in X and linear in Y
template <int P>
__host__ __device__ inline void affine_stencil_entry(const double *__restrict__ E, const double *__restrict__ Dx,
                                                     const double *__restrict__ Dy, const int (&nl)[3],
                                                     const int (&l)[3][2], const int (&d)[3], double dnx, double dny,
                                                     double r0, double r1, double *s0, double *sx, double *sy) {
  constexpr int P1 = P + 1, NB = P1 * P1 * P1;
  double a = 0.0, bx = 0.0, by = 0.0;
  for (int az = 0; az < nl[2]; az++)
    for (int ay = 0; ay < nl[1]; ay++)
      for (int ax = 0; ax < nl[0]; ax++) {
        const int lx = l[0][ax], ly = l[1][ay], lz = l[2][az];
        const int mx = lx + d[0], my = ly + d[1], mz = lz + d[2];
        if (mx < 0 || mx > P || my < 0 || my > P || mz < 0 || mz > P) continue;
        const int e = (lx + P1 * (ly + P1 * lz)) * NB + (mx + P1 * (my + P1 * mz));
        // cell centre = node + (P/2 - l) node spacings per axis
        a += E[e] + (r1 - dny * (0.5 * P - ly)) * Dx[e] + (dnx * (0.5 * P - lx) - r0) * Dy[e];
        bx += dnx * Dy[e];
        by -= dny * Dx[e];
      }
  *s0 = a;
  *sx = bx;
  *sy = by;
}

// st: [S0 8 x 128][class rhs 8][SX 8 x 128][SY 8 x 128] (node units: X, Y
// are lattice coordinates, dn the node spacing)
template <int P>
__global__ __launch_bounds__(kCst) void class_stencil_affine_k(LatArgs a, const double *__restrict__ E, double dnx,
                                                               double dny, double r0, double r1,
                                                               double *__restrict__ st) {
  constexpr int P1 = P + 1, NB = P1 * P1 * P1;
  const double *Dx = E + NB * NB + NB, *Dy = E + 2 * NB * NB + NB;
  const int c = blockIdx.x, s = threadIdx.x;
  const LatDev &C = a.c[c];
  int nl[3], l[3][2], lo[3], len[3];
#pragma unroll
  for (int d = 0; d < 3; d++) class_axis<P>(C.px[d], &nl[d], l[d], &lo[d], &len[d]);
  if (s < C.nsten) {
    const int d[3] = {lo[0] + s % len[0], lo[1] + (s / len[0]) % len[1], lo[2] + s / (len[0] * len[1])};
    double s0, sx, sy;
    affine_stencil_entry<P>(E, Dx, Dy, nl, l, d, dnx, dny, r0, r1, &s0, &sx, &sy);
    st[c * kCst + s] = s0;
    st[8 * kCst + 8 + c * kCst + s] = sx;
    st[2 * 8 * kCst + 8 + c * kCst + s] = sy;
  }
  if (s == kCst - 1) {
    double r = 0.0;
    for (int az = 0; az < nl[2]; az++)
      for (int ay = 0; ay < nl[1]; ay++)
        for (int ax = 0; ax < nl[0]; ax++) r += E[NB * NB + l[0][ax] + P1 * (l[1][ay] + P1 * l[2][az])];
    st[8 * kCst + c] = r;
  }
}

// rounding spelled out (explicit fma): the FP64 lattice operator of the
// rotating field evaluates the same expression and gets the CSR values' bits
__host__ __device__ inline double affine_entry(const double *__restrict__ st, int id, int X, int Y) {
  return fma((double)Y, st[2 * 8 * kCst + 8 + id], fma((double)X, st[8 * kCst + 8 + id], st[id]));
}


// CSR values of the active rows, one row block (at most kStreamNnz entries)
// per workgroup: entries in CSR order, each finds its row in the block's row
// pointers (LDS)
__global__ __launch_bounds__(kB) void csr_fill_affine_k(const int *__restrict__ blk, int nact,
                                                        const int *__restrict__ rowptr,
                                                        const int *__restrict__ dof2node, int nn,
                                                        const uint16_t *__restrict__ sid,
                                                        const double *__restrict__ st, double *__restrict__ val) {
  __shared__ int rp[kStreamRows + 1];
  const int r0 = blk[blockIdx.x], r1 = min(blk[blockIdx.x + 1], nact);
  if (r1 <= r0) return;
  const int nr = r1 - r0;
  for (int i = threadIdx.x; i <= nr; i += kB) rp[i] = rowptr[r0 + i];
  __syncthreads();
  const int j0 = rp[0], nz = rp[nr] - j0;
  for (int t = threadIdx.x; t < nz; t += kB) {
    const int k = j0 + t;
    int lo = 0, hi = nr - 1;  // last row with rp <= k
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (rp[mid] <= k) lo = mid;
      else hi = mid - 1;
    }
    const int node = dof2node[r0 + lo];
    val[k] = affine_entry(st, sid[k], node % nn, (node / nn) % nn);
  }
}

// FP32 per-node lattice values and 1/a_ii straight from the affine stencils:
// alpha_m M + alpha (S0 + X SX + Y SY), M the mass class stencils in stm
// (alpha_m = 0: the conv-diff matrix alone), evaluated as affine_value32
__global__ void lat_fill_affine_k(LatArgs a, long long n, const double *__restrict__ st,
                                  const double *__restrict__ stm, double alpha_m, double alpha,
                                  const int *__restrict__ centre, float *__restrict__ lval,
                                  float *__restrict__ linv) {
  for (long long i = (long long)blockIdx.x * kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB) {
    int c = 0;
    while (c + 1 < a.nclass && i >= a.c[c + 1].off) c++;
    const LatDev &C = a.c[c];
    const long long r = i - C.off, s = r / C.nodes, k = r % C.nodes;
    const int X = C.lo[0] + C.step * (int)(k % C.cnt[0]), Y = C.lo[1] + C.step * (int)((k / C.cnt[0]) % C.cnt[1]);
    const int id = c * kCst + (int)s;
    const float s0 = (float)st[id], sx = (float)st[8 * kCst + 8 + id], sy = (float)st[2 * 8 * kCst + 8 + id],
                sm = alpha_m != 0.0 ? (float)stm[id] : 0.f;
    const float v = affine_value32(&s0, &sx, &sy, &sm, (float)alpha, (float)alpha_m, 0, X, Y);
    lval[i] = v;
    if (s == centre[c]) linv[C.first + k] = 1.0f / v;
  }
}

// ------------------------------------------------------------------------
// FP64 operator in lattice order for the outer FGMRES of the mixed form,
// when the class stencils are exactly the assembled CSR values: interior
// rows apply the class stencil, boundary (Dirichlet) rows are identity rows.

enum Lat64Mode { kL64Spmv = 0, kL64Defect = 1 };

template <int P, int PX, int PY, int PZ, int MODE>
__host__ __device__ inline void lat64_node(const LatArgs &a, const LatDev &C, long long k, const double *__restrict__ st,
                                  const double *__restrict__ x, const double *__restrict__ f,
                                  double *__restrict__ out) {
  int kx, ky, kz;
  class_coords(C, k, kx, ky, kz);
  const int X = C.lo[0] + C.step * kx, Y = C.lo[1] + C.step * ky, Z = C.lo[2] + C.step * kz;
  const long long nn = a.nn, node = X + nn * (Y + nn * Z);
  double s = 0.0;
  int si = 0;
#pragma unroll
  for (int dz = lo_off<P, PZ>(); dz <= hi_off<P, PZ>(); dz++)
#pragma unroll
    for (int dy = lo_off<P, PY>(); dy <= hi_off<P, PY>(); dy++)
#pragma unroll
      for (int dx = lo_off<P, PX>(); dx <= hi_off<P, PX>(); dx++, si++)
        s += st[si] * x[node + dx + nn * (dy + nn * dz)];
  out[node] = MODE == kL64Defect ? f[node] - s : s;
}

template <int P, int MODE>
__global__ __launch_bounds__(kB) void lat64_k(LatArgs a, const double *__restrict__ st64, const double *__restrict__ x,
                                              const double *__restrict__ f, double *__restrict__ out) {
  const int c = block_class(a.bstart, a.nclass, (int)blockIdx.x);
  const long long k = (long long)(blockIdx.x - a.bstart[c]) * kB + threadIdx.x;
  if (k >= a.c[c].nodes) return;
  const double *st = st64 + c * kCst;
  if (P == 1) {
    lat64_node<1, 0, 0, 0, MODE>(a, a.c[c], k, st, x, f, out);
  } else {
    switch (c) {
      case 0: lat64_node<2, 0, 0, 0, MODE>(a, a.c[c], k, st, x, f, out); break;
      case 1: lat64_node<2, 1, 0, 0, MODE>(a, a.c[c], k, st, x, f, out); break;
      case 2: lat64_node<2, 0, 1, 0, MODE>(a, a.c[c], k, st, x, f, out); break;
      case 3: lat64_node<2, 1, 1, 0, MODE>(a, a.c[c], k, st, x, f, out); break;
      case 4: lat64_node<2, 0, 0, 1, MODE>(a, a.c[c], k, st, x, f, out); break;
      case 5: lat64_node<2, 1, 0, 1, MODE>(a, a.c[c], k, st, x, f, out); break;
      case 6: lat64_node<2, 0, 1, 1, MODE>(a, a.c[c], k, st, x, f, out); break;
      default: lat64_node<2, 1, 1, 1, MODE>(a, a.c[c], k, st, x, f, out); break;
    }
  }
}

// Q2 FP64 operator by vertex/midpoint pairs (the gathers of lat_pair; the
// class stencils cs0, cs1 of the pair class)
template <int PY, int PZ, int MODE>
__host__ __device__ inline void lat64_pair(const LatArgs &a, long long k, const double *cs0, const double *cs1,
                                           const double *__restrict__ x, const double *__restrict__ f,
                                           double *__restrict__ out) {
  const PairPos q = pair_pos<PY, PZ>(a, k);
  double s0, s1, xv, xm;
  pair_sums<double, PY, PZ, true>(a, q, k, nullptr, cs0, cs1, x, s0, s1, xv, xm);
  if (q.hv) out[q.nv] = MODE == kL64Defect ? f[q.nv] - s0 : s0;
  out[q.nv + 1] = MODE == kL64Defect ? f[q.nv + 1] - s1 : s1;
}

template <int MODE>
__host__ __device__ inline void lat64_pair_node(const LatArgs &a, int pc, long long k, const double *cs0,
                                                const double *cs1, const double *__restrict__ x,
                                                const double *__restrict__ f, double *__restrict__ out) {
  switch (pc) {
    case 0: lat64_pair<0, 0, MODE>(a, k, cs0, cs1, x, f, out); break;
    case 1: lat64_pair<1, 0, MODE>(a, k, cs0, cs1, x, f, out); break;
    case 2: lat64_pair<0, 1, MODE>(a, k, cs0, cs1, x, f, out); break;
    default: lat64_pair<1, 1, MODE>(a, k, cs0, cs1, x, f, out); break;
  }
}

// blocks pstart[4].. take the Dirichlet rows (DOFs nact..ndof-1, identity),
// as lat64_boundary_k
template <int MODE>
__global__ __launch_bounds__(kB) void lat64_pair_k(LatArgs a, const double *__restrict__ st64,
                                                   const double *__restrict__ x, const double *__restrict__ f,
                                                   double *__restrict__ out, int nact, int ndof,
                                                   const int *__restrict__ dof2node) {
  __shared__ double cs[2 * kCst];
  if ((int)blockIdx.x >= a.pstart[4]) {
    const int d = nact + (blockIdx.x - a.pstart[4]) * kB + threadIdx.x;
    if (d < ndof) {
      const int node = dof2node[d];
      out[node] = MODE == kL64Defect ? f[node] - x[node] : x[node];
    }
    return;
  }
  int pc = 0;
  while (pc < 3 && (int)blockIdx.x >= a.pstart[pc + 1]) pc++;
  cs[threadIdx.x] = st64[2 * pc * kCst + threadIdx.x];  // classes 2 pc and 2 pc + 1
  __syncthreads();
  const long long k = (long long)(blockIdx.x - a.pstart[pc]) * kB + threadIdx.x;
  if (k < a.c[2 * pc + 1].nodes) lat64_pair_node<MODE>(a, pc, k, cs, cs + kCst, x, f, out);
}

// FP64 operator by groups of G pairs (group_sums<double>); blocks
// pstart[4].. (counted in groups) take the Dirichlet rows as in lat64_pair_k
template <int G, int PY, int PZ, int MODE>
__host__ __device__ inline void lat64_pair_group(const LatArgs &a, long long k, const double *cs0, const double *cs1,
                                        const double *__restrict__ x, const double *__restrict__ f,
                                        double *__restrict__ out) {
  bool hv;
  const long long nv = group_pos<G, PY, PZ>(a, k, hv);
  double s0[G], s1[G], xs[2 * G];
  group_sums<double, G, PY, PZ>(a, nv, hv, cs0, cs1, x, s0, s1, xs);
#pragma unroll
  for (int q = 0; q < G; q++) {
    const long long v = nv + 2 * q;
    if (q > 0 || hv) out[v] = MODE == kL64Defect ? f[v] - s0[q] : s0[q];
    out[v + 1] = MODE == kL64Defect ? f[v + 1] - s1[q] : s1[q];
  }
}

template <int G, int MODE>
__host__ __device__ inline void lat64_pair_group_node(const LatArgs &a, int pc, long long k, const double *cs0,
                                                      const double *cs1, const double *__restrict__ x,
                                                      const double *__restrict__ f, double *__restrict__ out) {
  switch (pc) {
    case 0: lat64_pair_group<G, 0, 0, MODE>(a, k, cs0, cs1, x, f, out); break;
    case 1: lat64_pair_group<G, 1, 0, MODE>(a, k, cs0, cs1, x, f, out); break;
    case 2: lat64_pair_group<G, 0, 1, MODE>(a, k, cs0, cs1, x, f, out); break;
    default: lat64_pair_group<G, 1, 1, MODE>(a, k, cs0, cs1, x, f, out); break;
  }
}

template <int MODE, int G>
__global__ __launch_bounds__(kB) void lat64_pair_gk(LatArgs a, const double *__restrict__ st64,
                                                    const double *__restrict__ x, const double *__restrict__ f,
                                                    double *__restrict__ out, int nact, int ndof,
                                                    const int *__restrict__ dof2node) {
  __shared__ double cs[2 * kCst];
  const int bid = blockIdx.x;  // (XCD-contiguous blocks: 25.0 vs 18.6 us, 43.3 vs 26.4 us)
  if (bid >= a.pstart[4]) {
    const int d = nact + (bid - a.pstart[4]) * kB + threadIdx.x;
    if (d < ndof) {
      const int node = dof2node[d];
      out[node] = MODE == kL64Defect ? f[node] - x[node] : x[node];
    }
    return;
  }
  int pc = 0;
  while (pc < 3 && bid >= a.pstart[pc + 1]) pc++;
  cs[threadIdx.x] = st64[2 * pc * kCst + threadIdx.x];
  __syncthreads();
  const long long k = (long long)(bid - a.pstart[pc]) * kB + threadIdx.x;
  if (k < a.c[2 * pc + 1].nodes / G) lat64_pair_group_node<G, MODE>(a, pc, k, cs, cs + kCst, x, f, out);
}

// The FP64 lattice operator by quads of lines (quad_sums over the FP64
// compact stencils); blocks from nquad on take the Dirichlet rows (identity).
// Per node the same sums as lat64_pair_group.  TX = float: x is the FP32
// V-cycle output, converted exactly, and xcopy (when given) receives x at
// every node written (every lattice node), the FP32 Krylov column z_i.
template <int G, int MODE, typename TX = double>
__global__ __launch_bounds__(kB) void lat64_quad_gk(LatArgs a, const double *__restrict__ st64,
                                                    const TX *__restrict__ x, const double *__restrict__ f,
                                                    double *__restrict__ out, int nquad, int nact, int ndof,
                                                    const int *__restrict__ dof2node, int xcd,
                                                    TX *__restrict__ xcopy = nullptr) {
  __shared__ double cs[8 * kCst];
  const int bid0 = (int)blockIdx.x;
  if (bid0 >= nquad) {
    const int d = nact + (bid0 - nquad) * kB + threadIdx.x;
    if (d < ndof) {
      const int node = dof2node[d];
      const TX xv = x[node];
      out[node] = MODE == kL64Defect ? f[node] - (double)xv : (double)xv;
      if constexpr (sizeof(TX) == 4) {
        if (xcopy) xcopy[node] = xv;
      }
    }
    return;
  }
  for (int i = threadIdx.x; i < 8 * kCst; i += kB) cs[i] = st64[i];
  __syncthreads();
  const int bid = xcd ? xcd_block(bid0, nquad) : bid0;
  QuadPos p;
  if (!quad_pos<G>(a, (long long)bid * kB + threadIdx.x, p)) return;
  const long long nn = a.nn;
  double s0[4][G], s1[4][G], xs[4][2 * G];
  quad_sums<double, G, TX>(a, p, cs, x, s0, s1, xs);
#pragma unroll
  for (int pc = 0; pc < 4; pc++) {
    const int oy = pc & 1, oz = pc >> 1;
    if ((oy == 0 && p.by == 0) || (oz == 0 && p.bz == 0)) continue;
    const long long nv = p.base + nn * (oy + nn * oz);
#pragma unroll
    for (int q = 0; q < G; q++) {
      const long long v = nv + 2 * q;
      const double r0 = MODE == kL64Defect ? f[v] - s0[pc][q] : s0[pc][q];
      const double r1 = MODE == kL64Defect ? f[v + 1] - s1[pc][q] : s1[pc][q];
      // vertex and midpoint adjacent: paired stores (8-byte alignment of
      // the doubles suffices) unless the vertex is a boundary node
      typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));
      if (q > 0 || p.hv) *(d2u *)(out + v) = d2u{r0, r1};
      else out[v + 1] = r1;
      if constexpr (sizeof(TX) == 4) {
        // the node values again (just read: cache hits), not held in registers
        if (xcopy) {
          if (q > 0 || p.hv) *(f2u *)(xcopy + v) = *(const f2u *)(x + v);
          else xcopy[v + 1] = x[v + 1];
        }
      }
    }
  }
}

// FP64 operator of a rotating-field level: entry s of a node (X, Y) is
// alpha (S0 + X SX + Y SY) (conv-diff alone, alpha 1) or
// fma(alpha, S0 + X SX + Y SY, M) (the TCD3D system matrix M + tau th1 A),
// the exact arithmetic of the CSR values (affine_entry in csr_fill_affine_k,
// then matadd_into_k), so the lattice FGMRES applies the assembled matrix.
// Class-blocked; blocks bstart[nclass].. take the Dirichlet rows (identity).
template <int P, int PX, int PY, int PZ, int MODE>
__host__ __device__ inline void lat64_aff_node(const LatArgs &a, const LatDev &C, long long k, const double *s0,
                                               const double *sx, const double *sy, const double *sm, double alpha,
                                               const double *__restrict__ x, const double *__restrict__ f,
                                               double *__restrict__ out) {
  int kx, ky, kz;
  class_coords(C, k, kx, ky, kz);
  const int X = C.lo[0] + C.step * kx, Y = C.lo[1] + C.step * ky, Z = C.lo[2] + C.step * kz;
  const long long nn = a.nn, node = X + nn * (Y + nn * Z);
  double s = 0.0;
  int si = 0;
#pragma unroll
  for (int dz = lo_off<P, PZ>(); dz <= hi_off<P, PZ>(); dz++)
#pragma unroll
    for (int dy = lo_off<P, PY>(); dy <= hi_off<P, PY>(); dy++)
#pragma unroll
      for (int dx = lo_off<P, PX>(); dx <= hi_off<P, PX>(); dx++, si++) {
        const double e = fma((double)Y, sy[si], fma((double)X, sx[si], s0[si]));
        const double v = sm ? fma(alpha, e, sm[si]) : alpha * e;
        s += v * x[node + dx + nn * (dy + nn * dz)];
      }
  out[node] = MODE == kL64Defect ? f[node] - s : s;
}

template <int P, int MODE>
__global__ __launch_bounds__(kB) void lat64_aff_k(LatArgs a, LatAff af, const double *__restrict__ x,
                                                  const double *__restrict__ f, double *__restrict__ out, int nact,
                                                  int ndof, const int *__restrict__ dof2node) {
  __shared__ double st[4][kCst];
  const int nb = a.bstart[a.nclass];
  if ((int)blockIdx.x >= nb) {  // Dirichlet rows
    const int d = nact + (blockIdx.x - nb) * kB + threadIdx.x;
    if (d < ndof) {
      const int node = dof2node[d];
      out[node] = MODE == kL64Defect ? f[node] - x[node] : x[node];
    }
    return;
  }
  const int c = block_class(a.bstart, a.nclass, (int)blockIdx.x);
  for (int i = threadIdx.x; i < kCst; i += kB) {
    st[0][i] = af.st[c * kCst + i];
    st[1][i] = af.st[8 * kCst + 8 + c * kCst + i];
    st[2][i] = af.st[2 * 8 * kCst + 8 + c * kCst + i];
    st[3][i] = af.stm ? af.stm[c * kCst + i] : 0.0;
  }
  __syncthreads();
  const long long k = (long long)(blockIdx.x - a.bstart[c]) * kB + threadIdx.x;
  if (k >= a.c[c].nodes) return;
  const double *sm = af.stm ? st[3] : nullptr;
#define CD3D_A64(PX, PY, PZ) \
  lat64_aff_node<P, PX, PY, PZ, MODE>(a, a.c[c], k, st[0], st[1], st[2], sm, af.alpha, x, f, out)
  if (P == 1) {
    CD3D_A64(0, 0, 0);
  } else {
    switch (c) {
      case 0: CD3D_A64(0, 0, 0); break;
      case 1: CD3D_A64(1, 0, 0); break;
      case 2: CD3D_A64(0, 1, 0); break;
      case 3: CD3D_A64(1, 1, 0); break;
      case 4: CD3D_A64(0, 0, 1); break;
      case 5: CD3D_A64(1, 0, 1); break;
      case 6: CD3D_A64(0, 1, 1); break;
      default: CD3D_A64(1, 1, 1); break;
    }
  }
#undef CD3D_A64
}

// Dirichlet rows (DOFs nact..ndof-1): identity
__global__ void lat64_boundary_k(int nact, int ndof, const int *__restrict__ dof2node, const double *__restrict__ x,
                                 const double *__restrict__ f, double *__restrict__ out, int defect) {
  const int d = nact + blockIdx.x * kB + threadIdx.x;
  if (d >= ndof) return;
  const int node = dof2node[d];
  out[node] = defect ? f[node] - x[node] : x[node];
}

// DOF order <-> lattice order, f64, every DOF (dir 0: to lattice)
__global__ void lat64_perm_k(int ndof, const int *__restrict__ dof2node, const double *__restrict__ in,
                             double *__restrict__ out, int dir) {
  const int d = blockIdx.x * kB + threadIdx.x;
  if (d >= ndof) return;
  if (dir == 0) out[dof2node[d]] = in[d];
  else out[d] = in[dof2node[d]];
}

__global__ void cast_d2f_k(long long n, const double *__restrict__ a, float *__restrict__ b) {
  for (long long i = (long long)blockIdx.x * kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB)
    b[i] = (float)a[i];
}

__global__ void cast_f2d_k(long long n, const float *__restrict__ a, double *__restrict__ b) {
  for (long long i = (long long)blockIdx.x * kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB)
    b[i] = (double)a[i];
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
    a.uni[c] = L.uni[c];
    b += (int)((C.nodes + kB - 1) / kB);
  }
  a.bstart[L.nclass] = b;
  b = 0;
  if (L.nclass == 8)
    for (int pc = 0; pc < 4; pc++) {
      a.pstart[pc] = b;
      b += (int)((L.cls[2 * pc + 1].nodes + kB - 1) / kB);
    }
  a.pstart[4] = b;
  return a;
}

// small lattices (16^3 cells and below) sweep faster by class: too few
// workgroups for the pair form's longer per-thread chains to pay off
constexpr int kPairMinNN = 40;

// CD3D_LAT_PAIRS=0: Q2 sweeps by class instead of by vertex/midpoint pairs
static bool lat_pairs() {
  static const bool v = [] {
    const char *e = getenv("CD3D_LAT_PAIRS");
    return !(e && atoi(e) == 0);
  }();
  return v;
}

// CD3D_LAT_QUADS=0: the grouped Q2 FP32 sweeps by pair class (lat_pair_gk)
// instead of by quads of lines (lat_quad_gk); read at every launch
static bool lat_quads() {
  const char *e = getenv("CD3D_LAT_QUADS");
  return !(e && atoi(e) == 0);
}

// Pairs per thread of the compact-stencil Q2 kernels: groups of gmax where
// that still leaves kGroupMinThreads threads, else single pairs.  Measured on
// the 129^3-node level of a 64^3-cell Q2 mesh: FP32 sweep 22.8 / 19.0 /
// 18.5 us with 1 / 2 / 4 pairs per thread (gmax 4), FP64 operator 29.7 /
// 26.4 / 37.0 us (gmax 2); the 65^3-node level is slower with groups (too
// few threads).  CD3D_PAIR_GROUP = 1, 2 or 4 forces a size; one that does
// not divide the pairs per line falls back to 1.
constexpr long long kGroupMinThreads = 1 << 17;
static int pair_group(const LatArgs &a, int gmax) {
  long long pairs = 0;
  for (int pc = 0; pc < 4; pc++) pairs += a.c[2 * pc + 1].nodes;
  const char *e = getenv("CD3D_PAIR_GROUP");
  const int G = e ? atoi(e) : pairs / gmax >= kGroupMinThreads ? gmax : 1;
  if (G != 2 && G != 4) return 1;
  for (int pc = 0; pc < 4; pc++)
    if (a.c[2 * pc + 1].cnt[0] % G) return 1;
  return G;
}

// pstart counted in groups of G pairs
static LatArgs group_args(const LatArgs &a, int G) {
  LatArgs ag = a;
  int b = 0;
  for (int pc = 0; pc < 4; pc++) {
    ag.pstart[pc] = b;
    b += (int)((a.c[2 * pc + 1].nodes / G + kB - 1) / kB);
  }
  ag.pstart[4] = b;
  return ag;
}

static LatAff lat_aff(const DevLevel &L) {
  LatAff af{};
  af.st = L.lst64;
  af.stm = L.lat_alpha_m != 0.0 ? L.lst64 + 3 * 8 * kCst + 8 : nullptr;
  af.alpha = L.lat_alpha;
  af.alpha_m = L.lat_alpha_m;
  for (int c = 0; c < 8; c++) af.centre[c] = L.lat_centre[c];
  return af;
}

static bool all_uni(const LatArgs &a) {
  for (int c = 0; c < a.nclass; c++)
    if (!a.uni[c]) return false;
  return true;
}

template <int MODE>
void lat_sweep(hipStream_t s, const DevLevel &L, int order, const float *x, const float *f, float *out,
               float omega) {
  const LatArgs a = lat_args(L);
  const int grid = a.bstart[L.nclass];
  if (grid <= 0) return;
  if (L.lat_affine) {
    const LatAff af = lat_aff(L);
    if (order == 1) lat_aff_sweep_k<1, MODE><<<grid, kB, 0, s>>>(a, af, x, f, out, omega);
    else lat_aff_sweep_k<2, MODE><<<grid, kB, 0, s>>>(a, af, x, f, out, omega);
    return;
  }
  const int G = order == 2 && lat_pairs() && L.nn > kPairMinNN && all_uni(a) ? pair_group(a, 4) : 1;
  if (order == 1) lat_sweep_k<1, MODE><<<grid, kB, 0, s>>>(a, L.lval, L.linv, L.lcst, x, f, out, omega);
  else if (G > 1 && lat_quads()) {
    // measured on the 129^3-node level of a 64^3-cell Q2 mesh (FP32 Jacobi
    // sweep): 2 pairs per thread, 256 threads, XCD-contiguous blocks 12.4 us;
    // without the XCD mapping 13.6; 4 pairs 13.4 / 15.1; 1 pair 14.0 (61
    // instead of 159 VGPRs, but 2.5 instead of 1.75 neighbour-line loads per
    // node); 128 threads 12.6 / 15.0; the pair-class kernel it replaces 18.7
    // (CD3D_QUAD_G, CD3D_QUAD_TB, CD3D_QUAD_XCD select the variants)
    const char *eg = getenv("CD3D_QUAD_G"), *et = getenv("CD3D_QUAD_TB"), *ex = getenv("CD3D_QUAD_XCD");
    const int gq_g = eg ? atoi(eg) : 2, tb = et ? atoi(et) : 256, xcd = ex ? atoi(ex) : 1;
    // G == 4 implies 2 divides the pairs per line too
    const int Gq = gq_g == 1 ? 1 : (gq_g == 4 && G == 4) ? 4 : 2;
    const long long n = a.c[1].cnt[0], threads = n / Gq * n * n;
    const int gq = (int)((threads + tb - 1) / tb);
    if (tb == 128) {
      if (Gq == 1) lat_quad_gk<1, MODE, 128><<<gq, 128, 0, s>>>(a, L.lcst, x, f, out, omega, xcd);
      else if (Gq == 2) lat_quad_gk<2, MODE, 128><<<gq, 128, 0, s>>>(a, L.lcst, x, f, out, omega, xcd);
      else lat_quad_gk<4, MODE, 128><<<gq, 128, 0, s>>>(a, L.lcst, x, f, out, omega, xcd);
    } else {
      if (Gq == 1) lat_quad_gk<1, MODE, 256><<<gq, 256, 0, s>>>(a, L.lcst, x, f, out, omega, xcd);
      else if (Gq == 2) lat_quad_gk<2, MODE, 256><<<gq, 256, 0, s>>>(a, L.lcst, x, f, out, omega, xcd);
      else lat_quad_gk<4, MODE, 256><<<gq, 256, 0, s>>>(a, L.lcst, x, f, out, omega, xcd);
    }
  } else if (G > 1) {
    const LatArgs ag = group_args(a, G);
    if (G == 2) lat_pair_gk<MODE, 2><<<ag.pstart[4], kB, 0, s>>>(ag, L.lcst, x, f, out, omega);
    else lat_pair_gk<MODE, 4><<<ag.pstart[4], kB, 0, s>>>(ag, L.lcst, x, f, out, omega);
  } else if (lat_pairs() && L.nn > kPairMinNN && all_uni(a))
    lat_pair_k<MODE, true><<<a.pstart[4], kB, 0, s>>>(a, L.lval, L.linv, L.lcst, x, f, out, omega);
  else if (lat_pairs() && L.nn > kPairMinNN)
    lat_pair_k<MODE, false><<<a.pstart[4], kB, 0, s>>>(a, L.lval, L.linv, L.lcst, x, f, out, omega);
  else lat_sweep_k<2, MODE><<<grid, kB, 0, s>>>(a, L.lval, L.linv, L.lcst, x, f, out, omega);
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
bool kl_zero_sweep_fusable(const DevLevel &C) {
  if (C.lat_affine || C.nclass <= 0) return false;
  for (int c = 0; c < C.nclass; c++)
    if (!C.uni[c]) return false;
  return true;
}
bool kl_top_zero_sweep(const DevLevel &F, int order, float omega, float *x, TopZeroSweep &zt) {
  // uniform stencils, and classes that are exactly the parity sub-lattices of
  // the interior (class c holds the nodes of parity c, all of them)
  if (!kl_zero_sweep_fusable(F) || F.nclass != (order == 2 ? 8 : 1)) return false;
  const int nn = F.nn;
  long long tot = 0;
  for (int c = 0; c < F.nclass; c++) {
    const LatClass &L = F.cls[c];
    if (L.step != order) return false;
    for (int ax = 0; ax < 3; ax++) {
      if (L.cnt[ax] <= 0 || L.lo[ax] < 1 || L.lo[ax] + L.step * (L.cnt[ax] - 1) > nn - 2) return false;
      if (order == 2 && (L.lo[ax] & 1) != ((c >> ax) & 1)) return false;
    }
    tot += L.nodes;
  }
  if (tot != (long long)(nn - 2) * (nn - 2) * (nn - 2)) return false;
  zt.x = x;
  zt.d = F.lcst + kCst - 1;
  zt.dstride = kCst;
  zt.omega = omega;
  zt.nn = nn;
  zt.order = order;
  return true;
}
void kl_restrict(hipStream_t s, const DevLevel &F, const DevLevel &C, int order, const float *fine, float *coarse,
                 float *zero_sweep, float omega) {
  const LatArgs ca = lat_args(C);
  const int grid = ca.bstart[C.nclass];
  if (grid <= 0) return;
  if (zero_sweep && !kl_zero_sweep_fusable(C))
    throw std::logic_error("the restriction writes the zero-start sweep only for uniform class stencils");
  const ZeroSweep zs{zero_sweep, C.lcst, omega};
  if (order == 1) lat_restrict_k<1><<<grid, kB, 0, s>>>(ca, F.nn, fine, coarse, zs);
  else if (lat_pairs() && C.nn > kPairMinNN) {
    const char *e = getenv("CD3D_RESTRICT_INTER");
    const int inter = e ? atoi(e) : 0;
    int nbmax = 0;
    for (int pc = 0; pc < 4; pc++) nbmax = std::max(nbmax, ca.pstart[pc + 1] - ca.pstart[pc]);
    lat_restrict_pair_k<<<inter ? 4 * nbmax : ca.pstart[4], kB, 0, s>>>(ca, F.nn, fine, coarse, inter, zs);
  }
  else lat_restrict_k<2><<<grid, kB, 0, s>>>(ca, F.nn, fine, coarse, zs);
}
void kl_prolong_update(hipStream_t s, const DevLevel &F, const DevLevel &C, int order, const float *coarse, float *x,
                       float damp) {
  const long long ncc = (C.nn - 1) / order, n = ncc * ncc * ncc * 4 * order * order;  // (cell, fine Z, fine Y)
  if (n <= 0) return;
  if (order == 1) lat_prolong_k<1><<<(int)((n + kB - 1) / kB), kB, 0, s>>>((int)ncc, F.nn, C.nn, coarse, x, damp);
  else lat_prolong_k<2><<<(int)((n + kB - 1) / kB), kB, 0, s>>>((int)ncc, F.nn, C.nn, coarse, x, damp);
}
hipError_t kl_detect_uniform(hipStream_t s, DevLevel &L, int *flag_host, bool *all_uniform) {
  *all_uniform = true;
  if (L.nclass <= 0) return hipSuccess;
  if (L.lat_fresh == 1) {  // kl_assemble_uniform wrote the compact stencils of this matrix
    L.lat_fresh = 0;
    L.lat_exact64 = true;
    for (int c = 0; c < L.nclass; c++) L.uni[c] = 1;
    return hipSuccess;
  }
  L.lat_exact64 = false;
  for (int c = 0; c < L.nclass; c++) L.uni[c] = 0;
  const LatArgs a = lat_args(L);
  lat_first_node_k<<<L.nclass, 64, 0, s>>>(a, L.lfirst, L.lnode2dof, L.diag, L.val, L.lscale, L.lflag, L.lcst);
  lat_uniform_k<<<2048, kB, 0, s>>>(L.lat_nval, L.lsid, L.lfirst, L.val, L.lscale, L.lflag);
  hipError_t e = hipMemcpyAsync(flag_host, L.lflag, sizeof(int) * L.nclass, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    *all_uniform = false;
    return e;
  }
  for (int c = 0; c < L.nclass; c++) {
    L.uni[c] = flag_host[c] == 0;
    *all_uniform = *all_uniform && L.uni[c];
  }
  return hipGetLastError();
}

hipError_t kl_convert(hipStream_t s, const DevLevel &L) {
  lat_convert_k<<<2048, kB, 0, s>>>(L.lat_nval, L.lperm, L.val, L.lval);
  lat_invdiag_k<<<(int)((L.lat_nint + kB - 1) / kB), kB, 0, s>>>(L.lat_nint, L.lnode2dof, L.diag, L.val, L.linv);
  return hipGetLastError();
}

void kl_assemble_uniform(hipStream_t s, DevLevel &L, int order, const double *tab, const AsmParams &ap,
                         double *val, double *rhs) {
  const LatArgs a = lat_args(L);
  if (order == 1) {
    elem_k<1><<<1, 64, 0, s>>>(tab, ap, L.lelem);
    class_stencil_k<1><<<L.nclass, kCst, 0, s>>>(a, L.lelem, L.lst64, L.lcst);
  } else {
    elem_k<2><<<1, 768, 0, s>>>(tab, ap, L.lelem);
    class_stencil_k<2><<<L.nclass, kCst, 0, s>>>(a, L.lelem, L.lst64, L.lcst);
  }
  kl_fill_csr(s, L, val);
  if (L.lat_nint > 0)
    rhs_fill_k<<<(int)((L.lat_nint + kB - 1) / kB), kB, 0, s>>>(a, L.lat_nint, L.lnode2dof, L.lst64, rhs);
}

void kl_assemble_affine(hipStream_t s, DevLevel &L, int order, const double *tab, const AsmParams &ap,
                        double *val, double *rhs, bool csr) {
  const LatArgs a = lat_args(L);
  if (order == 1) {
    elem_affine_k<1><<<1, 64, 0, s>>>(tab, ap, L.lelem);
    class_stencil_affine_k<1><<<L.nclass, kCst, 0, s>>>(a, L.lelem, ap.dn[0], ap.dn[1], ap.rot_c[0], ap.rot_c[1],
                                                        L.lst64);
  } else {
    elem_affine_k<2><<<1, 768, 0, s>>>(tab, ap, L.lelem);
    class_stencil_affine_k<2><<<L.nclass, kCst, 0, s>>>(a, L.lelem, ap.dn[0], ap.dn[1], ap.rot_c[0], ap.rot_c[1],
                                                        L.lst64);
  }
  if (csr && L.nblk > 0)
    csr_fill_affine_k<<<L.nblk, kB, 0, s>>>(L.blk, L.nact, L.rowptr, L.ldof2node, L.nn, L.lsid, L.lst64, val);
  if (L.lat_nint > 0)
    rhs_fill_k<<<(int)((L.lat_nint + kB - 1) / kB), kB, 0, s>>>(a, L.lat_nint, L.lnode2dof, L.lst64, rhs);
}

// stencil index of the offset (0, 0, 0) of every class
static void lat_centres(DevLevel &L) {
  for (int c = 0; c < L.nclass; c++) {
    const LatClass &C = L.cls[c];
    int lo[3], len[3];
    for (int d = 0; d < 3; d++) {
      const int p = C.step == 1 ? 1 : 2;  // Q1 classes step 1, Q2 classes step 2
      const bool mid = p == 2 && C.px[d] == 1;
      lo[d] = mid ? -1 : -p;
      len[d] = mid ? 3 : 2 * p + 1;
    }
    L.lat_centre[c] = -lo[0] + len[0] * (-lo[1] + len[1] * -lo[2]);
  }
}

hipError_t kl_fill_lattice_affine(hipStream_t s, DevLevel &L, double alpha_m, double alpha) {
  if (L.nclass <= 0 || L.lat_nval <= 0) return hipSuccess;
  lat_centres(L);
  int *dcentre = (int *)(L.lst64 + 3 * 8 * kCst + 8 + 8 * kCst);  // spare tail of the stencil buffer
  hipError_t e = hipMemcpyAsync(dcentre, L.lat_centre, sizeof(int) * 8, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  for (int c = 0; c < L.nclass; c++) L.uni[c] = 0;
  L.lat_affine = false;
  lat_fill_affine_k<<<2048, kB, 0, s>>>(lat_args(L), L.lat_nval, L.lst64, L.lst64 + 3 * 8 * kCst + 8, alpha_m, alpha,
                                        dcentre, L.lval, L.linv);
  return hipGetLastError();
}

hipError_t kl_set_lattice_affine(hipStream_t, DevLevel &L, double alpha_m, double alpha) {
  lat_centres(L);
  for (int c = 0; c < L.nclass; c++) L.uni[c] = 0;
  L.lat_alpha_m = alpha_m;
  L.lat_alpha = alpha;
  L.lat_affine = L.nclass > 0;
  return hipSuccess;
}

void kl_fill_csr(hipStream_t s, const DevLevel &L, double *val) {
  // active rows come first in the CSR, and their entries are exactly the stencil entries
  const long long n = L.lat_nval;
  if (n <= 0) return;
  long long done = 0;
  if ((((uintptr_t)val & 15) | ((uintptr_t)L.lsid & 3)) == 0) {
    const long long n2 = n / 2;
    if (n2 > 0)
      csr_fill2_k<<<(int)std::min<long long>((n2 + kB - 1) / kB, 4096), kB, 0, s>>>(n2, L.lsid, L.lst64, val);
    done = 2 * n2;
  }
  if (done < n)
    csr_fill_k<<<(int)std::min<long long>((n - done + kB - 1) / kB, 8192), kB, 0, s>>>(n - done, L.lsid + done,
                                                                                      L.lst64, val + done);
}

template <int MODE>
void lat64_apply(hipStream_t s, const DevLevel &L, int order, const double *x, const double *f, double *out) {
  const LatArgs a = lat_args(L);
  const int grid = a.bstart[L.nclass];
  const int nd = L.ndof - L.nact;
  if (L.lat_aff64) {  // rotating field: interior and Dirichlet rows in one launch
    LatAff af = lat_aff(L);
    af.alpha = L.lat_alpha;
    const int g = grid + (nd + kB - 1) / kB;
    if (g <= 0) return;
    if (order == 1) lat64_aff_k<1, MODE><<<g, kB, 0, s>>>(a, af, x, f, out, L.nact, L.ndof, L.ldof2node);
    else lat64_aff_k<2, MODE><<<g, kB, 0, s>>>(a, af, x, f, out, L.nact, L.ndof, L.ldof2node);
    return;
  }
  if (order == 2 && grid > 0 && lat_pairs() && L.nn > kPairMinNN) {  // interior and Dirichlet rows in one launch
    const int G = pair_group(a, 2);
    if (G > 1 && lat_quads()) {
      // measured on the 129^3-node level (FP64 product): quads of one pair
      // 21.1 us (XCD mapping or not), of two pairs 22.5 (256 VGPRs, one wave
      // per SIMD), the pair-class kernel 27.0
      const char *ex = getenv("CD3D_QUAD_XCD"), *eg = getenv("CD3D_QUAD64_G");
      const int xcd = ex ? atoi(ex) : 1, gq = eg ? atoi(eg) : 1;
      const long long n = a.c[1].cnt[0];
      const int nq = (int)((n / gq * n * n + kB - 1) / kB), ng = nq + (nd + kB - 1) / kB;
      if (gq == 2)
        lat64_quad_gk<2, MODE><<<ng, kB, 0, s>>>(a, L.lst64, x, f, out, nq, L.nact, L.ndof, L.ldof2node, xcd);
      else
        lat64_quad_gk<1, MODE><<<ng, kB, 0, s>>>(a, L.lst64, x, f, out, nq, L.nact, L.ndof, L.ldof2node, xcd);
      return;
    }
    if (G > 1) {
      const LatArgs ag = group_args(a, G);
      const int g = ag.pstart[4] + (nd + kB - 1) / kB;
      if (G == 2) lat64_pair_gk<MODE, 2><<<g, kB, 0, s>>>(ag, L.lst64, x, f, out, L.nact, L.ndof, L.ldof2node);
      else lat64_pair_gk<MODE, 4><<<g, kB, 0, s>>>(ag, L.lst64, x, f, out, L.nact, L.ndof, L.ldof2node);
      return;
    }
    lat64_pair_k<MODE><<<a.pstart[4] + (nd + kB - 1) / kB, kB, 0, s>>>(a, L.lst64, x, f, out, L.nact, L.ndof,
                                                                       L.ldof2node);
    return;
  }
  if (grid > 0) {
    if (order == 1) lat64_k<1, MODE><<<grid, kB, 0, s>>>(a, L.lst64, x, f, out);
    else lat64_k<2, MODE><<<grid, kB, 0, s>>>(a, L.lst64, x, f, out);
  }
  if (nd > 0)
    lat64_boundary_k<<<(nd + kB - 1) / kB, kB, 0, s>>>(L.nact, L.ndof, L.ldof2node, x, f, out,
                                                      MODE == kL64Defect ? 1 : 0);
}

void kl_affine64_apply(hipStream_t s, const DevLevel &L, int order, double alpha, bool with_mass, const double *x,
                       double *y) {
  const LatArgs a = lat_args(L);
  LatAff af = lat_aff(L);
  af.alpha = alpha;
  af.stm = with_mass ? L.lst64 + 3 * 8 * kCst + 8 : nullptr;
  const int g = a.bstart[L.nclass] + (L.ndof - L.nact + kB - 1) / kB;
  if (g <= 0) return;
  if (order == 1) lat64_aff_k<1, kL64Spmv><<<g, kB, 0, s>>>(a, af, x, nullptr, y, L.nact, L.ndof, L.ldof2node);
  else lat64_aff_k<2, kL64Spmv><<<g, kB, 0, s>>>(a, af, x, nullptr, y, L.nact, L.ndof, L.ldof2node);
}
void kl_spmv64(hipStream_t s, const DevLevel &L, int order, const double *x, double *y) {
  lat64_apply<kL64Spmv>(s, L, order, x, nullptr, y);
}
// the quad form of lat64_apply (the one taken for order 2 on uniform levels
// above kPairMinNN nodes per side with quads and pair groups on)
static bool lat64_quad_form(const DevLevel &L, int order, const LatArgs &a) {
  return !L.lat_aff64 && order == 2 && a.bstart[L.nclass] > 0 && lat_pairs() && L.nn > kPairMinNN &&
         pair_group(a, 2) > 1 && lat_quads();
}
bool kl_spmv64_f32_ok(const DevLevel &L, int order) {
  const LatArgs a = lat_args(L);
  const char *eg = getenv("CD3D_QUAD64_G");
  return lat64_quad_form(L, order, a) && !(eg && atoi(eg) == 2);
}
void kl_spmv64_f32(hipStream_t s, const DevLevel &L, int order, const float *x, double *y, float *xcopy) {
  const LatArgs a = lat_args(L);
  if (!kl_spmv64_f32_ok(L, order)) throw std::logic_error("FP32-input lattice operator on a level without quads");
  const char *ex = getenv("CD3D_QUAD_XCD");
  const int xcd = ex ? atoi(ex) : 1, nd = L.ndof - L.nact;
  const long long n = a.c[1].cnt[0];
  const int nq = (int)((n * n * n + kB - 1) / kB), ng = nq + (nd + kB - 1) / kB;
  lat64_quad_gk<1, kL64Spmv, float><<<ng, kB, 0, s>>>(a, L.lst64, x, nullptr, y, nq, L.nact, L.ndof, L.ldof2node,
                                                        xcd, xcopy);
}
void kl_defect64(hipStream_t s, const DevLevel &L, int order, const double *x, const double *f, double *d) {
  lat64_apply<kL64Defect>(s, L, order, x, f, d);
}
void kl_dof_to_lat64(hipStream_t s, const DevLevel &L, const double *v, double *l) {
  lat64_perm_k<<<(L.ndof + kB - 1) / kB, kB, 0, s>>>(L.ndof, L.ldof2node, v, l, 0);
}
void kl_lat64_to_dof(hipStream_t s, const DevLevel &L, const double *l, double *v) {
  lat64_perm_k<<<(L.ndof + kB - 1) / kB, kB, 0, s>>>(L.ndof, L.ldof2node, l, v, 1);
}
void kl_cast_d2f(hipStream_t s, int64_t n, const double *a, float *b) {
  cast_d2f_k<<<(int)std::min<int64_t>((n + kB - 1) / kB, 2048), kB, 0, s>>>(n, a, b);
}
void kl_cast_f2d(hipStream_t s, int64_t n, const float *a, double *b) {
  cast_f2d_k<<<(int)std::min<int64_t>((n + kB - 1) / kB, 2048), kB, 0, s>>>(n, a, b);
}

void kl_small_vcycle(hipStream_t st, const DevLevel *levels, int top, int order, int pre, int post, float omega,
                     float damp, const double *Ainv) {
  if (top < 1 || top > kSmallMax) return;
  SmallArgs s{};
  for (int i = 0; i <= top; i++) {
    const DevLevel &D = levels[i];
    s.L[i].a = lat_args(D);
    s.L[i].val = D.lval;
    s.L[i].inv = D.linv;
    s.L[i].cst = D.lcst;
    s.L[i].lx = D.lx;
    s.L[i].lt = D.lt;
    s.L[i].la = D.la;
    s.L[i].lf = D.lf;
  }
  s.top = top;
  s.pre = pre;
  s.post = post;
  s.na = levels[0].nact;
  s.omega = omega;
  s.damp = damp;
  s.Ainv = Ainv;
  s.dof2node0 = levels[0].ldof2node;
  bool uni = true;
  for (int i = 1; i <= top; i++)
    for (int c = 0; c < s.L[i].a.nclass; c++) uni = uni && s.L[i].a.uni[c];
  if (order == 1 && uni) lat_small_k<1, true><<<1, small_threads<1>(), 0, st>>>(s);
  else if (order == 1) lat_small_k<1, false><<<1, small_threads<1>(), 0, st>>>(s);
  else if (uni) lat_small_k<2, true><<<1, small_threads<2>(), 0, st>>>(s);
  else lat_small_k<2, false><<<1, small_threads<2>(), 0, st>>>(s);
}

void kl_from_dof(hipStream_t s, const DevLevel &L, const double *v, float *l) {
  dof_to_lat_k<<<(L.ndof + kB - 1) / kB, kB, 0, s>>>(L.ndof, L.nact, L.ldof2node, v, l);
}
void kl_to_dof(hipStream_t s, const DevLevel &L, const float *l, double *v) {
  lat_to_dof_k<<<(L.ndof + kB - 1) / kB, kB, 0, s>>>(L.ndof, L.nact, L.ldof2node, l, v);
}

}  // namespace cd3d

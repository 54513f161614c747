// HIP kernels for gfx950 (MI355X): CSR row kernels (SpMV, defect, damped
// Jacobi), grid transfers, deterministic reductions, coarse solves and the
// element assembly.  Everything on this path is HBM-bound sparse work or FP64
// vector arithmetic; nothing is a dense contraction, so no MFMA is used.
//
// CSR row kernels map one row to a group of LANES consecutive lanes of a
// 64-wide wavefront (64/LANES rows per wavefront): the row's column indices
// and values are read as contiguous, coalesced segments and the group reduces
// with DPP/permute shuffles.  LANES is chosen from the mean row length.
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <cstdlib>

#include "cd3d_internal.hpp"
#include "trilinear.hpp"

namespace cd3d {

namespace {

constexpr int kBlock = 256;

enum RowMode { kSpmv = 0, kDefect = 1, kJacobi = 2, kRestrict = 3, kProlongUpd = 4, kActiveSpmv = 5 };

template <int LANES, typename T>
__device__ inline T group_sum(T v) {
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, LANES);
  return v;
}

// One kernel body for every CSR row operation.  MODE:
//   kSpmv       out = A x
//   kDefect     out = f - A x                                   (ScalarDefect)
//   kJacobi     out = (1-w) x + w (f - sum_{j!=i} a_ij x_j)/a_ii on active rows,
//               out = f on Dirichlet rows                       (TMGLevel3D::Jacobi)
//   kRestrict   out = R x on active coarse rows, 0 on the others (DefectRestriction + CorrectDefect)
//   kProlongUpd out += w (P x) on active fine rows               (Prolongate + CorrectNodes + Update)
//   kActiveSpmv out = A x on active rows only                   (MatVectActive)
// UNROLL independent (column, value, x) load chains per lane and iteration
// keep more HBM requests in flight.
template <int LANES, int UNROLL, int MODE, typename T, typename V>
__device__ inline void csr_rows_body(int gid, int n, int nact, const int *__restrict__ rowptr,
                                     const int *__restrict__ col, const V *__restrict__ val, const T *__restrict__ x,
                                     const T *__restrict__ f, T *__restrict__ out, T omega) {
  const int row = gid / LANES, lane = gid % LANES;
  if (row >= n) return;  // whole groups leave together
  if ((MODE == kRestrict || MODE == kProlongUpd || MODE == kActiveSpmv) && row >= nact) {
    if (MODE == kRestrict && lane == 0) out[row] = T(0);
    return;
  }
  const int b = rowptr[row], e = rowptr[row + 1];
  T s = T(0), diag = T(0);
  for (int j0 = b + lane; j0 < e; j0 += LANES * UNROLL) {
    int c[UNROLL];
    T a[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; u++) {
      const int j = j0 + u * LANES;
      c[u] = j < e ? col[j] : row;
      a[u] = j < e ? (T)val[j] : T(0);
    }
#pragma unroll
    for (int u = 0; u < UNROLL; u++) {
      const T xv = x[c[u]];
      if (MODE == kJacobi) {
        if (c[u] == row) diag += a[u];
        else s += a[u] * xv;
      } else {
        s += a[u] * xv;
      }
    }
  }
  s = group_sum<LANES>(s);
  if (MODE == kJacobi) diag = group_sum<LANES>(diag);
  if (lane != 0) return;
  if (MODE == kSpmv || MODE == kRestrict || MODE == kActiveSpmv) out[row] = s;
  else if (MODE == kDefect) out[row] = f[row] - s;
  else if (MODE == kJacobi) out[row] = row < nact ? (T(1) - omega) * x[row] + omega * (f[row] - s) / diag : f[row];
  else if (MODE == kProlongUpd) out[row] += omega * s;
}

template <int LANES, int UNROLL, int MODE, typename T, typename V>
__global__ __launch_bounds__(kBlock) void csr_rows(int n, int nact, const int *__restrict__ rowptr,
                                                   const int *__restrict__ col, const V *__restrict__ val,
                                                   const T *__restrict__ x, const T *__restrict__ f,
                                                   T *__restrict__ out, T omega) {
  csr_rows_body<LANES, UNROLL, MODE, T, V>(blockIdx.x * kBlock + threadIdx.x, n, nact, rowptr, col, val, x, f, out,
                                           omega);
}

// Relaxed Gauss-Seidel over the rows of one colour of the multicolour SSOR
// smoother (rows[0..nrows)): rows of one colour never couple, so every row
// is updated in place from the current x of the other colours,
//   x_i = (1 - w) x_i + w (f_i - sum_{j != i} a_ij x_j) / a_ii.
template <int LANES, int UNROLL>
__global__ __launch_bounds__(kBlock) void colour_sor_k(int nrows, const int *__restrict__ rows,
                                                       const int *__restrict__ rowptr, const int *__restrict__ col,
                                                       const double *__restrict__ val, double *x,
                                                       const double *__restrict__ f, double omega) {
  const int gid = blockIdx.x * kBlock + threadIdx.x;
  const int k = gid / LANES, lane = gid % LANES;
  if (k >= nrows) return;
  const int row = rows[k];
  const int b = rowptr[row], e = rowptr[row + 1];
  double s = 0.0, diag = 0.0;
  for (int j0 = b + lane; j0 < e; j0 += LANES * UNROLL) {
    int c[UNROLL];
    double a[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; u++) {
      const int j = j0 + u * LANES;
      c[u] = j < e ? col[j] : row;
      a[u] = j < e ? val[j] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < UNROLL; u++) {
      if (c[u] == row) diag += a[u];
      else s += a[u] * x[c[u]];
    }
  }
  s = group_sum<LANES>(s);
  diag = group_sum<LANES>(diag);
  if (lane == 0) x[row] = (1.0 - omega) * x[row] + omega * (f[row] - s) / diag;
}

// lanes per row from the mean row length (overridable for tuning through
// CD3D_CSR_LANES / CD3D_CSR_UNROLL)
inline void row_shape(int64_t nnz, int n, int &lanes, int &unroll) {
  const double m = n > 0 ? (double)nnz / n : 1;
  lanes = m <= 12 ? 4 : m <= 40 ? 8 : m <= 96 ? 16 : 32;
  unroll = m <= 12 ? 1 : 4;
  static const int env_l = getenv("CD3D_CSR_LANES") ? atoi(getenv("CD3D_CSR_LANES")) : 0;
  static const int env_u = getenv("CD3D_CSR_UNROLL") ? atoi(getenv("CD3D_CSR_UNROLL")) : 0;
  if (env_l == 4 || env_l == 8 || env_l == 16 || env_l == 32 || env_l == 64) lanes = env_l;
  if (env_u == 1 || env_u == 2 || env_u == 4) unroll = env_u;
}

template <int LANES, int MODE, typename T, typename V>
void launch_lanes(hipStream_t st, int unroll, int grid, int n, int nact, const int *rp, const int *col, const V *val,
                  const T *x, const T *f, T *out, T omega) {
  if (unroll == 4) csr_rows<LANES, 4, MODE, T, V><<<grid, kBlock, 0, st>>>(n, nact, rp, col, val, x, f, out, omega);
  else if (unroll == 2)
    csr_rows<LANES, 2, MODE, T, V><<<grid, kBlock, 0, st>>>(n, nact, rp, col, val, x, f, out, omega);
  else csr_rows<LANES, 1, MODE, T, V><<<grid, kBlock, 0, st>>>(n, nact, rp, col, val, x, f, out, omega);
}

template <int MODE, typename T, typename V>
void launch_rows(hipStream_t st, int64_t nnz, int n, int nact, const int *rp, const int *col, const V *val,
                 const T *x, const T *f, T *out, T omega) {
  if (n <= 0) return;
  int lanes, unroll;
  row_shape(nnz, n, lanes, unroll);
  const int64_t threads = (int64_t)n * lanes;
  const int grid = (int)((threads + kBlock - 1) / kBlock);
  switch (lanes) {
    case 4: launch_lanes<4, MODE, T, V>(st, unroll, grid, n, nact, rp, col, val, x, f, out, omega); break;
    case 8: launch_lanes<8, MODE, T, V>(st, unroll, grid, n, nact, rp, col, val, x, f, out, omega); break;
    case 16: launch_lanes<16, MODE, T, V>(st, unroll, grid, n, nact, rp, col, val, x, f, out, omega); break;
    case 32: launch_lanes<32, MODE, T, V>(st, unroll, grid, n, nact, rp, col, val, x, f, out, omega); break;
    default: launch_lanes<64, MODE, T, V>(st, unroll, grid, n, nact, rp, col, val, x, f, out, omega); break;
  }
}

// ---------------------------------------------------------------- CSR stream
// Row-block kernel for the level matrices.  Workgroup b owns rows
// [blk[b], blk[b+1]) whose nonzeros (at most kStreamNnz) are contiguous in
// the CSR arrays: phase 1 streams them with coalesced, unrolled loads and
// stores a_ij x_j into LDS; phase 2 sums each row from LDS with G lanes per
// row (G = the largest power of two with G * rows <= 256).  Memory access no
// longer depends on the row lengths.  MODE: kSpmv out = A x; kDefect
// out = f - A x; kJacobi out = x + w (f - A x) / a_ii on active rows, f on
// Dirichlet rows (the damped Jacobi step of TMGLevel3D::Jacobi,
// MGLevel3D.C:443-537, written with the full row sum).
constexpr int kStreamPer = kStreamNnz / kStreamBlock;
// matrices above the 256 MB last-level (Infinity) cache are streamed non-temporally
constexpr double kStreamNtBytes = 256.0 * (1 << 20);

template <bool NT, typename P>
__device__ inline P stream_load(const P *p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}

// One row block bi of the stream kernel (the whole workgroup takes part).
// bt[bi] = (first row, end row, first nonzero, nonzero count): the block's
// extent in one scalar load.  Phase 1 streams the nonzeros with coalesced
// loads and stores a_ij x_j into LDS; phase 2 sums each row from LDS with G
// lanes per row (G = the largest power of two with G * rows <= 256, so every
// lane group owns at most one row).  A lane group's row extent and the
// row's operands of the defect and Jacobi modes do not depend on phase 1:
// they are loaded beside the matrix stream, ahead of the barrier, so the
// block's chain of dependent memory trips is the matrix load, the gather of
// x and the store (fine level of the 64^3 Q2 mesh: 346 -> 332 us).  NT:
// non-temporal matrix loads (matrices larger than the caches).
template <int MODE, typename T, typename V, bool NT>
__device__ inline void csr_stream_block(int bi, const int4 *__restrict__ bt, int nact,
                                        const int *__restrict__ rowptr, const int *__restrict__ col,
                                        const V *__restrict__ val, const T *__restrict__ x,
                                        const T *__restrict__ f, const T *__restrict__ invdiag, T *__restrict__ out,
                                        T omega) {
  static_assert(kStreamRows <= kStreamBlock, "one row per lane group");
  __shared__ T prod[kStreamNnz];
  const int4 ext = bt[bi];
  const int r0 = ext.x, nr = ext.y - ext.x, j0 = ext.z, nz = ext.w;
  if (nz > kStreamNnz) {
    // a single row longer than a block: strided partial sums and a block reduction
    __shared__ T red[kStreamBlock / 64];
    T s = T(0);
    for (int t = threadIdx.x; t < nz; t += kStreamBlock) s += (T)val[j0 + t] * x[col[j0 + t]];
    s = group_sum<64>(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      T tot = T(0);
      for (int k = 0; k < kStreamBlock / 64; k++) tot += red[k];
      const int row = r0;
      if (MODE == kSpmv) out[row] = tot;
      else if (MODE == kDefect) out[row] = f[row] - tot;
      else out[row] = row < nact ? x[row] + omega * (f[row] - tot) * invdiag[row] : f[row];
    }
    return;
  }
  // two consecutive nonzeros per thread and load (dword alignment suffices);
  // a matrix larger than the caches is read non-temporally, so that it does
  // not displace the gathered x from L2
  typedef int i2u __attribute__((ext_vector_type(2), aligned(4)));
  typedef V v2u __attribute__((ext_vector_type(2), aligned(sizeof(V))));
  static_assert(kStreamPer % 2 == 0, "pairs of nonzeros per thread");
  int c[kStreamPer];
  T a[kStreamPer];
#pragma unroll
  for (int k = 0; k < kStreamPer / 2; k++) {
    const int t = 2 * threadIdx.x + k * 2 * kStreamBlock;
    if (t + 1 < nz) {
      const i2u cc = stream_load<NT>((const i2u *)(col + j0 + t));
      const v2u vv = stream_load<NT>((const v2u *)(val + j0 + t));
      c[2 * k] = cc.x;
      c[2 * k + 1] = cc.y;
      a[2 * k] = (T)vv.x;
      a[2 * k + 1] = (T)vv.y;
    } else {
      c[2 * k] = t < nz ? col[j0 + t] : 0;
      a[2 * k] = t < nz ? (T)val[j0 + t] : T(0);
      c[2 * k + 1] = 0;
      a[2 * k + 1] = T(0);
    }
  }
  // this lane group's row: extent and operands, loaded ahead of the barrier
  const int lg = nr <= 4 ? 6 : nr <= 8 ? 5 : nr <= 16 ? 4 : nr <= 32 ? 3 : nr <= 64 ? 2 : nr <= 128 ? 1 : 0;
  const int g = threadIdx.x >> lg, l = threadIdx.x & ((1 << lg) - 1), row = r0 + g;
  int ra = 0, re = 0;
  T fr = T(0), xr = T(0), dr = T(0);
  if (g < nr) {
    ra = rowptr[row] - j0;
    re = rowptr[row + 1] - j0;
    if (MODE != kSpmv) fr = f[row];
    if (MODE == kJacobi && row < nact) {
      xr = x[row];
      dr = invdiag[row];
    }
  }
#pragma unroll
  for (int k = 0; k < kStreamPer / 2; k++) {
    const int t = 2 * threadIdx.x + k * 2 * kStreamBlock;
    if (t < nz) prod[t] = a[2 * k] * x[c[2 * k]];
    if (t + 1 < nz) prod[t + 1] = a[2 * k + 1] * x[c[2 * k + 1]];
  }
  __syncthreads();
  T s = T(0);
  for (int t = ra + l; t < re; t += 1 << lg) s += prod[t];
  switch (lg) {
    case 6: s = group_sum<64>(s); break;
    case 5: s = group_sum<32>(s); break;
    case 4: s = group_sum<16>(s); break;
    case 3: s = group_sum<8>(s); break;
    case 2: s = group_sum<4>(s); break;
    case 1: s = group_sum<2>(s); break;
    default: break;
  }
  if (g < nr && l == 0) {
    if (MODE == kSpmv) out[row] = s;
    else if (MODE == kDefect) out[row] = fr - s;
    else out[row] = row < nact ? xr + omega * (fr - s) * dr : fr;
  }
}

template <int MODE, typename T, typename V, bool NT>
__global__ __launch_bounds__(kStreamBlock) void csr_stream(const int4 *__restrict__ bt, int nact,
                                                     const int *__restrict__ rowptr, const int *__restrict__ col,
                                                     const V *__restrict__ val, const T *__restrict__ x,
                                                     const T *__restrict__ f, const T *__restrict__ invdiag,
                                                     T *__restrict__ out, T omega) {
  // workgroups go to the 8 XCDs round-robin: XCD x takes a contiguous range
  // of row blocks, so that its L2 sees about 1/8 of the gathered x (fine
  // level of the 64^3 Q2 mesh: 370.7 -> 352.8 us)
  csr_stream_block<MODE, T, V, NT>(xcd_block(blockIdx.x, gridDim.x), bt, nact, rowptr, col, val, x, f, invdiag, out,
                                   omega);
}

template <typename T>
__global__ void invdiag_k(int n, const int *__restrict__ diag, const double *__restrict__ val, T *__restrict__ inv) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) inv[i] = (T)(1.0 / val[diag[i]]);
}

template <int MODE, typename T, typename V>
void launch_stream(hipStream_t st, const DevLevel &L, const V *val, const T *x, const T *f, const T *invdiag, T *out,
                   T omega) {
  if (L.nblk <= 0) return;
  const bool nt = (double)L.nnz * (sizeof(V) + sizeof(int)) > kStreamNtBytes;
  if (nt)
    csr_stream<MODE, T, V, true><<<L.nblk, kStreamBlock, 0, st>>>(L.bt, L.nact, L.rowptr, L.col, val, x, f, invdiag,
                                                                  out, omega);
  else
    csr_stream<MODE, T, V, false><<<L.nblk, kStreamBlock, 0, st>>>(L.bt, L.nact, L.rowptr, L.col, val, x, f, invdiag,
                                                                   out, omega);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void jacobi_zero_k(int n, int nact, const int *__restrict__ diag,
                                                        const double *__restrict__ val64,
                                                        const float *__restrict__ val32, const T *__restrict__ f,
                                                        T *__restrict__ out, T omega) {
  // first sweep from a zero iterate with zero Dirichlet data: every product
  // with the old iterate vanishes, so out = w f / a_ii exactly as the full sweep
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  if (i < nact) {
    const T d = val32 ? (T)val32[diag[i]] : (T)val64[diag[i]];
    out[i] = omega * f[i] / d;
  } else {
    out[i] = f[i];
  }
}

// ---------------------------------------------------------------- reductions
// Deterministic two-stage sums: each block writes one partial; the consumer
// sums the kRedBlocks partials in a fixed order (every block of the consumer
// does the same sum, so all see the same value).  Products are accumulated
// with explicit fma: under contraction a*b + c*d may fuse either product, so
// written-out sums would leave the rounding to the compiler's choice per
// kernel, and kernels that share a summation order would not agree bitwise.

__device__ inline double block_sum(double v, double *sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += sh[k];
  return t;  // valid in thread 0
}

__device__ inline double sum_partials(const double *__restrict__ p, double *sh) {
  double v = 0;
  for (int k = threadIdx.x; k < kRedBlocks; k += blockDim.x) v += p[k];
  v = block_sum(v, sh);
  __syncthreads();
  if (threadIdx.x == 0) sh[16] = v;
  __syncthreads();
  return sh[16];
}

__global__ __launch_bounds__(kBlock) void dot_k(int n, const double *__restrict__ a, const double *__restrict__ b,
                                                double *__restrict__ partials) {
  __shared__ double sh[17];
  double v = 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) v = fma(a[i], b[i], v);
  v = block_sum(v, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

// The level vectors are streamed as 16-byte pairs (global_load_dwordx4), two
// pairs in flight per thread, when every operand is 16-byte aligned; an odd
// last element goes to thread 0 of block 0.
__global__ __launch_bounds__(kBlock) void dot2_k(int n, const double *__restrict__ a, const double *__restrict__ b,
                                                 double *__restrict__ partials) {
  __shared__ double sh[17];
  const int n2 = n >> 1, gs = gridDim.x * kBlock;
  const double2 *a2 = (const double2 *)a, *b2 = (const double2 *)b;
  double v = 0;
  int i = blockIdx.x * kBlock + threadIdx.x;
  for (; i + gs < n2; i += 2 * gs) {
    const double2 p = a2[i], q = b2[i], r = a2[i + gs], t = b2[i + gs];
    v = fma(p.x, q.x, v);
    v = fma(p.y, q.y, v);
    v = fma(r.x, t.x, v);
    v = fma(r.y, t.y, v);
  }
  if (i < n2) {
    const double2 p = a2[i], q = b2[i];
    v = fma(p.x, q.x, v);
    v = fma(p.y, q.y, v);
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) v = fma(a[n - 1], b[n - 1], v);
  v = block_sum(v, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

__global__ __launch_bounds__(kBlock) void finish_k(const double *__restrict__ partials, double *__restrict__ out) {
  __shared__ double sh[17];
  const double v = sum_partials(partials, sh);
  if (threadIdx.x == 0) *out = v;
}

// y += sign * alpha * x with alpha = *alpha_dev (or the sum of alpha partials
// when alpha_dev is the partial array), then partial dot of the new y with z
__global__ __launch_bounds__(kBlock) void axpy_dot_k(int n, const double *__restrict__ alpha_partials,
                                                     double *__restrict__ alpha_out, double sign,
                                                     const double *__restrict__ x, double *__restrict__ y,
                                                     const double *__restrict__ z, double *__restrict__ partials) {
  __shared__ double sh[17];
  const double alpha = sum_partials(alpha_partials, sh);
  if (blockIdx.x == 0 && threadIdx.x == 0 && alpha_out) *alpha_out = alpha;
  const double a = sign * alpha;
  double v = 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double yi = fma(a, x[i], y[i]);
    y[i] = yi;
    v = fma(yi, z ? z[i] : yi, v);
  }
  v = block_sum(v, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

// axpy_dot_k on 16-byte pairs (Z: a separate z vector, else the dot is <y, y>).
// The loads of a thread's first kAxpyPre pairs do not depend on alpha: they
// are all issued before the reduction of the alpha partials, which hides its
// latency and keeps the whole stream in flight (kRedBlocks x kBlock threads
// take the 2^20 pairs of a 2 M-unknown level in one round).
constexpr int kAxpyPre = 4;
template <bool Z>
__global__ __launch_bounds__(kBlock) void axpy_dot2_k(int n, const double *__restrict__ alpha_partials,
                                                      double *__restrict__ alpha_out, double sign,
                                                      const double *__restrict__ x, double *__restrict__ y,
                                                      const double *__restrict__ z, double *__restrict__ partials) {
  __shared__ double sh[17];
  const int n2 = n >> 1, gs = gridDim.x * kBlock;
  const double2 *x2 = (const double2 *)x, *z2 = (const double2 *)z;
  double2 *y2 = (double2 *)y;
  const int i0 = blockIdx.x * kBlock + threadIdx.x;
  const double2 zero = {0, 0};
  double2 xa[kAxpyPre], ya[kAxpyPre], za[kAxpyPre];
#pragma unroll
  for (int q = 0; q < kAxpyPre; q++) {
    const int i = i0 + q * gs;
    const bool h = i < n2;
    xa[q] = h ? x2[i] : zero;
    ya[q] = h ? y2[i] : zero;
    za[q] = (Z && h) ? z2[i] : zero;
  }
  const double alpha = sum_partials(alpha_partials, sh);
  if (blockIdx.x == 0 && threadIdx.x == 0 && alpha_out) *alpha_out = alpha;
  const double a = sign * alpha;
  double v = 0;
#pragma unroll
  for (int q = 0; q < kAxpyPre; q++) {
    const int i = i0 + q * gs;
    if (i < n2) {
      double2 yi = ya[q];
      yi.x = fma(a, xa[q].x, yi.x);
      yi.y = fma(a, xa[q].y, yi.y);
      y2[i] = yi;
      v = fma(yi.x, Z ? za[q].x : yi.x, v);
      v = fma(yi.y, Z ? za[q].y : yi.y, v);
    }
  }
  for (int i = i0 + kAxpyPre * gs; i < n2; i += gs) {
    double2 yi = y2[i];
    const double2 xi = x2[i], zi = Z ? z2[i] : zero;
    yi.x = fma(a, xi.x, yi.x);
    yi.y = fma(a, xi.y, yi.y);
    y2[i] = yi;
    v = fma(yi.x, Z ? zi.x : yi.x, v);
    v = fma(yi.y, Z ? zi.y : yi.y, v);
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    const double yi = fma(a, x[n - 1], y[n - 1]);
    y[n - 1] = yi;
    v = fma(yi, Z ? z[n - 1] : yi, v);
  }
  v = block_sum(v, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

// y = x / sqrt(sum partials); writes the norm to *norm_out
__global__ __launch_bounds__(kBlock) void scale_copy_k(int n, const double *__restrict__ partials,
                                                       double *__restrict__ norm_out, const double *__restrict__ x,
                                                       double *__restrict__ y) {
  __shared__ double sh[17];
  const double nrm = sqrt(sum_partials(partials, sh));
  if (blockIdx.x == 0 && threadIdx.x == 0) *norm_out = nrm;
  const double s = nrm != 0.0 ? 1.0 / nrm : 1.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) y[i] = x[i] * s;
}

// the finest lattice level's zero-start sweep at lattice node e from the
// FP32 value f just stored there (TopZeroSweep); d: the 8 class values of
// t.d staged in LDS by top_zero_stage
__device__ inline void top_zero_stage(const TopZeroSweep &t, float *d) {
  if (t.x && threadIdx.x < 8) d[threadIdx.x] = threadIdx.x < (t.order == 2 ? 8u : 1u) ? t.d[threadIdx.x * t.dstride] : 0.f;
}
__device__ inline void top_zero_sweep(const TopZeroSweep &t, const float *d, int64_t e, float f) {
  // 32-bit unsigned division: a level has fewer than 2^31 lattice nodes
  const unsigned u = (unsigned)e, nn = (unsigned)t.nn, q = u / nn, Z = q / nn;
  const unsigned X = u - q * nn, Y = q - Z * nn;
  if (X - 1u < nn - 2u && Y - 1u < nn - 2u && Z - 1u < nn - 2u) {
    const int c = t.order == 2 ? (int)((X & 1u) + 2u * (Y & 1u) + 4u * (Z & 1u)) : 0;
    t.x[e] = t.omega * f * d[c];
  }
}

__global__ __launch_bounds__(kBlock) void scale_copy2_k(int n, const double *__restrict__ partials,
                                                        double *__restrict__ norm_out, const double *__restrict__ x,
                                                        double *__restrict__ y, float *__restrict__ y32,
                                                        TopZeroSweep zt) {
  __shared__ double sh[17];
  __shared__ float zd[8];
  const int n2 = n >> 1, gs = gridDim.x * kBlock;
  const double2 *x2 = (const double2 *)x;
  double2 *y2 = (double2 *)y;
  int i = blockIdx.x * kBlock + threadIdx.x;
  const double2 v0 = i < n2 ? x2[i] : double2{0, 0};  // issued before the reduction
  top_zero_stage(zt, zd);                              // visible after its barriers
  const double nrm = sqrt(sum_partials(partials, sh));
  if (blockIdx.x == 0 && threadIdx.x == 0) *norm_out = nrm;
  const double s = nrm != 0.0 ? 1.0 / nrm : 1.0;
  typedef float f2 __attribute__((ext_vector_type(2)));
  if (i < n2) {
    const double2 r = {v0.x * s, v0.y * s};
    y2[i] = r;
    if (y32) {
      const f2 r32 = f2{(float)r.x, (float)r.y};
      ((f2 *)y32)[i] = r32;
      if (zt.x) {
        top_zero_sweep(zt, zd, 2 * (int64_t)i, r32.x);
        top_zero_sweep(zt, zd, 2 * (int64_t)i + 1, r32.y);
      }
    }
  }
  for (i += gs; i < n2; i += gs) {
    const double2 v = x2[i];
    const double2 r = {v.x * s, v.y * s};
    y2[i] = r;
    if (y32) {
      const f2 r32 = f2{(float)r.x, (float)r.y};
      ((f2 *)y32)[i] = r32;
      if (zt.x) {
        top_zero_sweep(zt, zd, 2 * (int64_t)i, r32.x);
        top_zero_sweep(zt, zd, 2 * (int64_t)i + 1, r32.y);
      }
    }
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    y[n - 1] = x[n - 1] * s;
    if (y32) {
      const float v32 = (float)(x[n - 1] * s);
      y32[n - 1] = v32;
      if (zt.x) top_zero_sweep(zt, zd, n - 1, v32);
    }
  }
}

// Low-synchronisation Gram-Schmidt of the mixed form's FGMRES (the
// inverse compact WY form of modified Gram-Schmidt: Swirydowicz, Langou,
// Ananthan, Yang, Thomas, Numer. Linear Algebra Appl. 28 (2021) e2343).
// At Arnoldi step i, with w = A z_i and the basis V_0 .. V_i, one pass
// forms r_k = <w, V_k> (k <= i) and the newest vector's Gram row
// g_k = <V_i, V_k> (k < i); then h = (I + L)^-1 r with L the strictly lower
// part of V^T V (rows kept from the earlier steps) is the MGS column:
// h_k = <w - sum_{m<k} h_m V_m, V_k> = r_k - sum_{m<k} h_m L_km.  A second
// pass forms w - V h and the partial sums of its squared norm.  Two passes
// over the basis per step instead of one pass over w and two basis vectors
// per MGS step.
constexpr int kLsChunk = 2048;  // elements per workgroup of the dot pass
constexpr int kLsMax = 32;      // restart length up to which the form is used

// chunk b of nb: partials r_k at part[k nb + b], g_k at part[(i + 1 + k) nb + b]
// (ws, vs: kLsChunk doubles of LDS each)
__device__ __forceinline__ void ls_dots_chunk(int n, int i, const double *__restrict__ V, int64_t ld,
                                              const double *__restrict__ w, double *__restrict__ part, int b,
                                              int nb, double *ws, double *vs) {
  const int64_t e0 = (int64_t)b * kLsChunk;
  const int ne = (int)min((int64_t)kLsChunk, (int64_t)n - e0);
  const double *Vi = V + (int64_t)i * ld + e0;
  for (int t = threadIdx.x; t < kLsChunk; t += kBlock) {
    ws[t] = t < ne ? w[e0 + t] : 0.0;
    vs[t] = t < ne ? Vi[t] : 0.0;
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  constexpr int kPer = kLsChunk / 64;
  for (int k = wave; k <= i; k += kBlock / 64) {
    double r = 0.0, g = 0.0;
    if (k == i) {
      for (int t = lane; t < ne; t += 64) r = fma(vs[t], ws[t], r);
    } else {
      // all of the lane's loads of V_k first, then the products
      const double *Vk = V + (int64_t)k * ld + e0;
      double v[kPer];
#pragma unroll
      for (int u = 0; u < kPer; u++) {
        const int t = lane + 64 * u;
        v[u] = t < ne ? Vk[t] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kPer; u++) {
        const int t = lane + 64 * u;
        r = fma(v[u], ws[t], r);
        g = fma(v[u], vs[t], g);
      }
    }
    r = group_sum<64>(r);
    g = group_sum<64>(g);
    if (lane == 0) {
      part[(int64_t)k * nb + b] = r;
      if (k < i) part[(int64_t)(i + 1 + k) * nb + b] = g;
    }
  }
}
__global__ __launch_bounds__(kBlock) void ls_dots_k(int n, int i, const double *__restrict__ V, int64_t ld,
                                                    const double *__restrict__ w, double *__restrict__ part) {
  __shared__ double ws[kLsChunk], vs[kLsChunk];
  ls_dots_chunk(n, i, V, ld, w, part, blockIdx.x, gridDim.x, ws, vs);
}

// one workgroup of kLsCoef threads: the sums of the partials (a wave per
// dot, four interleaved accumulators, a fixed order), row i of L (Lm, ld
// kLsMax; rows 0 .. i-1 staged in LDS), and h = (I + L)^-1 r into h[0 .. i]
// by one wave whose lane m holds h_m (one butterfly sum per row)
// (d: 2 kLsMax + 1 doubles of LDS, Ls: kLsMax^2)
constexpr int kLsCoef = 1024;
__device__ __forceinline__ void ls_coeffs_body(int i, int nb, const double *__restrict__ part,
                                               double *__restrict__ Lm, double *__restrict__ h, double *d,
                                               double *Ls) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int t = threadIdx.x; t < i * kLsMax; t += kLsCoef) Ls[t] = Lm[t];
  for (int q = wave; q < 2 * i + 1; q += kLsCoef / 64) {
    const double *pq = part + (int64_t)q * nb;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int t = lane;
    for (; t + 192 < nb; t += 256) {
      a0 += pq[t];
      a1 += pq[t + 64];
      a2 += pq[t + 128];
      a3 += pq[t + 192];
    }
    for (; t < nb; t += 64) a0 += pq[t];
    const double v = group_sum<64>((a0 + a1) + (a2 + a3));
    if (lane == 0) d[q] = v;
  }
  __syncthreads();
  if (threadIdx.x < i) {
    const double g = d[i + 1 + threadIdx.x];
    Ls[i * kLsMax + threadIdx.x] = g;
    Lm[i * kLsMax + threadIdx.x] = g;
  }
  __syncthreads();
  if (wave == 0) {
    double hm = 0.0;  // lane m: h_m once row m is done
    for (int k = 0; k <= i; k++) {
      const double s = group_sum<64>(lane < k ? Ls[k * kLsMax + lane] * hm : 0.0);
      if (lane == k) hm = d[k] - s;
    }
    if (lane <= i) h[lane] = hm;
  }
}
__global__ __launch_bounds__(kLsCoef) void ls_coeffs_k(int i, int nb, const double *__restrict__ part,
                                                       double *__restrict__ Lm, double *__restrict__ h) {
  __shared__ double d[2 * kLsMax + 1], Ls[kLsMax * kLsMax];
  ls_coeffs_body(i, nb, part, Lm, h, d, Ls);
}

// w -= sum_{k <= i} h_k V_k on 16-byte pairs, and the partial sums of |w|^2
// (workgroup blk of nblk writes partials[blk]; sh: 17 doubles of LDS, hs:
// kLsMax)
__device__ __forceinline__ void ls_update_part(int n, int i, const double *__restrict__ h,
                                               const double *__restrict__ V, int64_t ld, double *__restrict__ w,
                                               double *__restrict__ partials, int blk, int nblk, double *sh,
                                               double *hs) {
  if (threadIdx.x <= i) hs[threadIdx.x] = h[threadIdx.x];
  __syncthreads();
  const int n2 = n >> 1;
  double2 *w2 = (double2 *)w;
  double v = 0.0;
  for (int e = blk * kBlock + threadIdx.x; e < n2; e += nblk * kBlock) {
    double2 x = w2[e];
    // the basis vectors four at a time: loads first, then the products
    int k = 0;
    for (; k + 4 <= i + 1; k += 4) {
      double2 z[4];
#pragma unroll
      for (int u = 0; u < 4; u++) z[u] = ((const double2 *)(V + (int64_t)(k + u) * ld))[e];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        x.x = fma(-hs[k + u], z[u].x, x.x);
        x.y = fma(-hs[k + u], z[u].y, x.y);
      }
    }
    for (; k <= i; k++) {
      const double2 z = ((const double2 *)(V + (int64_t)k * ld))[e];
      x.x = fma(-hs[k], z.x, x.x);
      x.y = fma(-hs[k], z.y, x.y);
    }
    w2[e] = x;
    v = fma(x.x, x.x, v);
    v = fma(x.y, x.y, v);
  }
  if ((n & 1) && blk == 0 && threadIdx.x == 0) {
    double x = w[n - 1];
    for (int k = 0; k <= i; k++) x = fma(-hs[k], V[(int64_t)k * ld + n - 1], x);
    w[n - 1] = x;
    v = fma(x, x, v);
  }
  v = block_sum(v, sh);
  if (threadIdx.x == 0) partials[blk] = v;
}
__global__ __launch_bounds__(kBlock) void ls_update2_k(int n, int i, const double *__restrict__ h,
                                                       const double *__restrict__ V, int64_t ld,
                                                       double *__restrict__ w, double *__restrict__ partials) {
  __shared__ double sh[17], hs[kLsMax];
  ls_update_part(n, i, h, V, ld, w, partials, blockIdx.x, gridDim.x, sh, hs);
}

// x *= alpha and, when y32 is given, y32 = (float) x (scalar loop)
__global__ void scale_cast_k(int n, double alpha, double *__restrict__ x, float *__restrict__ y32, TopZeroSweep zt) {
  __shared__ float zd[8];
  top_zero_stage(zt, zd);
  __syncthreads();
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double v = x[i] * alpha;
    x[i] = v;
    if (y32) {
      const float v32 = (float)v;
      y32[i] = v32;
      if (zt.x) top_zero_sweep(zt, zd, i, v32);
    }
  }
}

__global__ void scale_k(int n, double alpha, double *__restrict__ x) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) x[i] *= alpha;
}

__global__ void axpy_k(int n, double alpha, const double *__restrict__ x, double *__restrict__ y) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) y[i] += alpha * x[i];
}

__global__ void axpy_dev_k(int n, const double *__restrict__ alpha, int idx, const double *__restrict__ x,
                           double *__restrict__ y) {
  const double a = alpha[idx];
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) y[i] += a * x[i];
}

__global__ void copy_k(int n, const double *__restrict__ x, double *__restrict__ y) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) y[i] = x[i];
}

__global__ void fill_k(int64_t n, double v, double *__restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) y[i] = v;
}

__global__ void matadd_k(int64_t n, double *__restrict__ C, const double *__restrict__ D, double alpha) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    C[i] += alpha * D[i];
}

// 16-byte aligned (a null pointer counts as aligned)
static inline bool al16(const void *p) { return ((uintptr_t)p & 15) == 0; }

// y = alpha (A x) + beta (B x) on the active rows; A and B share the CSR
// pattern, so one pass reads the pattern and both value arrays once
template <int LANES>
__global__ __launch_bounds__(kBlock) void csr_pair_active_k(int nact, const int *__restrict__ rowptr,
                                                          const int *__restrict__ col, const double *__restrict__ a,
                                                          const double *__restrict__ b, double alpha, double beta,
                                                          const double *__restrict__ x, double *__restrict__ y) {
  const int gid = blockIdx.x * kBlock + threadIdx.x;
  const int row = gid / LANES, lane = gid % LANES;
  if (row >= nact) return;  // whole groups leave together
  double sa = 0, sb = 0;
  for (int j = rowptr[row] + lane; j < rowptr[row + 1]; j += LANES) {
    const double xv = x[col[j]];
    sa += a[j] * xv;
    sb += b[j] * xv;
  }
  sa = group_sum<LANES>(sa);
  sb = group_sum<LANES>(sb);
  if (lane == 0) y[row] = alpha * sa + beta * sb;
}

// C = D + alpha C (C and D of n entries)
__global__ void matadd_into_k(int64_t n, double *__restrict__ C, const double *__restrict__ D, double alpha) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    C[i] = fma(alpha, C[i], D[i]);  // explicit: the FP64 lattice operator of the rotating field repeats it
}

inline int grid_for(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  return (int)(g < 2048 ? (g < 1 ? 1 : g) : 2048);
}

// ---------------------------------------------------------------- coarse level

// x_A += Ainv_AA d_A : one wavefront per row of the explicit inverse of the
// active block (exact coarse solve as one defect correction)
__global__ __launch_bounds__(kBlock) void coarse_update_k(int na, const double *__restrict__ Ainv,
                                                          const double *__restrict__ d, double *__restrict__ x) {
  const int row = (blockIdx.x * kBlock + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (row >= na) return;
  double s = 0;
  for (int j = lane; j < na; j += 64) s += Ainv[(int64_t)row * na + j] * d[j];
  s = group_sum<64>(s);
  if (lane == 0) x[row] += s;
}

// Mixed-form exact coarse solve from a zero iterate with zero Dirichlet data
// (the V-cycle's visit of the coarsest level): x_A = A_AA^-1 f_A, read from
// and written to the FP32 lattice vectors.  The defect f - A x of a zero
// iterate is f itself, so these are the products and the summation order of
// lat_to_dof_k, csr_stream<kDefect> and coarse_update_k into a zeroed x and
// dof_to_lat_k, in one launch instead of six.
__global__ __launch_bounds__(kBlock) void coarse_lat_k(int na, const double *__restrict__ Ainv,
                                                       const int *__restrict__ dof2node, const float *__restrict__ lf,
                                                       float *__restrict__ lx) {
  const int row = (blockIdx.x * kBlock + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (row >= na) return;
  double s = 0;
  for (int j = lane; j < na; j += 64) s += Ainv[(int64_t)row * na + j] * (double)lf[dof2node[j]];
  s = group_sum<64>(s);
  if (lane == 0) lx[dof2node[row]] = (float)(0.0 + s);
}

// Coarse-level smoothing loop of TMultiGrid3D::Smooth(CoarseSmooth)
// (MultiGrid3D.C:197-260) with SOR, SSOR (MGLevel3D.C:254-441) or Jacobi sweeps,
// run by a single wavefront: rows are updated one after another, each row's
// dot product spread over the 64 lanes; the iterate lives in LDS.
__global__ __launch_bounds__(64) void coarse_smooth_k(int n, int nact, const int *__restrict__ rp,
                                                      const int *__restrict__ col, const double *__restrict__ val,
                                                      double *__restrict__ xg, const double *__restrict__ f,
                                                      double *__restrict__ aux, int smoother, double omega,
                                                      int maxit, double red) {
  extern __shared__ double x[];
  const int lane = threadIdx.x;
  for (int i = lane; i < n; i += 64) x[i] = xg[i];
  __syncthreads();
  auto defect_norm = [&]() {
    double s2 = 0;
    for (int i = 0; i < n; i++) {
      double s = 0;
      for (int j = rp[i] + lane; j < rp[i + 1]; j += 64) s += val[j] * x[col[j]];
      s = group_sum<64>(s);
      const double d = f[i] - s;
      s2 += d * d;
    }
    return sqrt(s2);
  };
  auto row_update = [&](int i, const double *xr) {  // (f_i - sum_{j!=i} a_ij xr_j) / a_ii
    double s = 0, diag = 0;
    const int b = rp[i], e = rp[i + 1];
    for (int j = b + lane; j < e; j += 64) {
      const int c = col[j];
      if (c == i) diag += val[j];
      else s += val[j] * xr[c];
    }
    s = group_sum<64>(s);
    diag = group_sum<64>(diag);
    return (f[i] - s) / diag;
  };
  double res = defect_norm();
  const double reduction = red * res;
  int it = 0;
  while (res > reduction && it < maxit) {
    if (smoother == 1) {  // Jacobi: aux = old iterate, then Dirichlet rows = f
      for (int i = lane; i < n; i += 64) aux[i] = x[i];
      __syncthreads();
      for (int i = nact + lane; i < n; i += 64) x[i] = f[i];
      __syncthreads();
      for (int i = 0; i < nact; i++) {
        const double q = row_update(i, aux);
        if (lane == 0) x[i] = (1 - omega) * aux[i] + omega * q;
      }
      __syncthreads();
    } else {  // SOR / SSOR: Dirichlet rows = f, forward (then, SSOR, backward) relaxed Gauss-Seidel
      for (int i = nact + lane; i < n; i += 64) x[i] = f[i];
      __syncthreads();
      for (int i = 0; i < nact; i++) {
        const double q = row_update(i, x);
        __syncthreads();
        if (lane == 0) x[i] = omega * (q - x[i]) + x[i];
        __syncthreads();
      }
      for (int i = smoother == 3 ? nact - 1 : -1; i >= 0; i--) {
        const double q = row_update(i, x);
        __syncthreads();
        if (lane == 0) x[i] = omega * (q - x[i]) + x[i];
        __syncthreads();
      }
    }
    res = defect_norm();
    it++;
  }
  __syncthreads();
  for (int i = lane; i < n; i += 64) xg[i] = x[i];
}

// ---------------------------------------------------------------- assembly

// cell frame code: for local axis a, bits 3a..3a+1 = global axis, bit 3a+2 = negative
__device__ inline void decode_frame(int code, int *dax, int *sgn) {
#pragma unroll
  for (int a = 0; a < 3; a++) {
    dax[a] = (code >> (3 * a)) & 3;
    sgn[a] = ((code >> (3 * a + 2)) & 1) ? -1 : 1;
  }
}

__device__ inline double sdfem_delta(double hconv, double eps, double nb, double nu) {
  // Compute_SDFEM_delta, SDFEM_TYPE 2 (ConvDiff.C:540-556), DELTA0 = 1
  if (fabs(eps) < 1e-20) eps = 1e-20;
  if (!(nb > 0)) return 0.0;
  const double alpha = nu * nb * hconv / (2 * eps);
  return nu * hconv * (1 / tanh(alpha) - 1 / alpha) / (2 * nb);
}

__device__ inline double hconv_brick(const double *h, double b0, double b1, double b2) {
  // Mesh_size_in_convection_direction on an axis-aligned brick (ConvDiff.C:411-469)
  const double nb = sqrt(b0 * b0 + b1 * b1 + b2 * b2);
  double den = 0;
#pragma unroll
  for (int s = 0; s < 8; s++) {
    const double v = ((s & 1) ? -b0 : b0) / h[0] + ((s & 2) ? -b1 : b1) / h[1] + ((s & 4) ? -b2 : b2) / h[2];
    den += fabs(v) / 4;
  }
  if (den < 1e-10) return sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
  return 2 * nb / den;
}

// One thread per (cell, local row i); the NB rows of a cell are consecutive
// threads of a flat index (cells may straddle wavefronts: threads share
// nothing).  Per quadrature point the thread evaluates the coefficients at
// the point and forms, in the cell's reference frame,
//   t_i = phi_i + delta beta.grad phi_i     (test function with the SUPG term)
//   a_i = w (kappa .* grad phi_i + t_i beta),  c_i = w t_i c
// so that row i of the local matrix of BilinearAssemble_SD / BilinearAssemble
// (ConvDiff3D.C:61-190),
//   A_ij += w (eps grad phi_i . grad phi_j + t_i (b.grad phi_j + c phi_j)),
// becomes A_ij += a_i . grad phi_j + c_i phi_j: four FMAs per (q, i, j) whose
// j-operands are the same for every thread (wave-uniform table reads), with
// no exchange between the threads of a cell.  kappa and beta fold in the
// physical gradients J^-T grad phi of THexaAffin (the cell frames are signed
// axis permutations of the brick).  The mass matrix (what 1) is
// A_ij += (w phi_i) phi_j with the rhs w f t_i, and the SUPG time-derivative
// matrix (what 2) A_ij += (w delta beta.grad phi_i) phi_j.  Then the row is added into the CSR values through
// the slot map (cells of one colour share no DOF).
// LDS_TAB: the per-thread operands (phi_i and its gradient at every q) are
// read from an LDS copy of the basis table made at the start of the block;
// otherwise from the table in global memory.  VAR: coefficients that vary
// over the cell (rotating convection, manufactured source); without it the
// kernel carries none of their code (and none of its registers).
template <int P, bool LDS_TAB, bool VAR>
__global__ __launch_bounds__(kBlock) void assemble_k(int ncells, const int *__restrict__ list,
                                                     const int4 *__restrict__ cells, const int *__restrict__ slots,
                                                     const int *__restrict__ gn, int nact,
                                                     const double *__restrict__ tab, AsmParams ap,
                                                     double *__restrict__ val, double *__restrict__ rhs) {
  constexpr int NB = (P + 1) * (P + 1) * (P + 1);
  constexpr int N1 = P + 2;
  constexpr int NQ = N1 * N1 * N1;
  constexpr int STRIDE = 4 + 4 * NB;
  __shared__ double s_tab[LDS_TAB ? NQ * NB * 4 : 1];
  if (LDS_TAB) {
    for (int k = threadIdx.x; k < NQ * NB * 4; k += kBlock) s_tab[k] = tab[(k / (4 * NB)) * STRIDE + 4 + k % (4 * NB)];
    __syncthreads();
  }
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (gid >= (int64_t)ncells * NB) return;
  const int ci = (int)(gid / NB), i = (int)(gid % NB);
  const int c = list ? list[ci] : ci;
  const int4 cg = cells[c];
  int dax[3], sg[3];
  decode_frame(cg.w, dax, sg);
  const int O[3] = {cg.x, cg.y, cg.z};
  const double detj = (ap.h[0] / 2) * (ap.h[1] / 2) * (ap.h[2] / 2);
  // per local axis a: 2 / h along the physical axis dax[a] (J^-T of the brick)
  double ih[3], kap[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const double h = dax[a] == 0 ? ap.h[0] : dax[a] == 1 ? ap.h[1] : ap.h[2];
    ih[a] = 2.0 / h;
    kap[a] = ap.eps * 4.0 / (h * h);
  }
  const double nu = 1.0 / P;
  // constant convection: beta and delta are the same at every point
  double beta[3] = {0, 0, 0};
  if (ap.conv_kind == 0) {
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const double bd = dax[a] == 0 ? ap.b[0] : dax[a] == 1 ? ap.b[1] : ap.b[2];
      beta[a] = sg[a] * bd * ih[a];
    }
  }
  double acc[NB];
#pragma unroll
  for (int j = 0; j < NB; j++) acc[j] = 0;
  double racc = 0;
  double hconv = -1;
  for (int q = 0; q < NQ; q++) {
    const double *T = tab + q * STRIDE;
    const double *Ti = LDS_TAB ? s_tab + (q * NB + i) * 4 : T + 4 + 4 * i;
    const double w = T[0] * detj;
    const double phi_i = Ti[0], g0 = Ti[1], g1 = Ti[2], g2 = Ti[3];
    // physical quadrature point (only variable coefficients need it)
    double X[3] = {0, 0, 0};
    if (VAR && (ap.conv_kind == 1 || ap.f_kind == 1)) {
#pragma unroll
      for (int a = 0; a < 3; a++) {
        const double xa = ap.dn[dax[a]] * (O[dax[a]] + P * sg[a] * (T[1 + a] + 1) * 0.5);
        X[0] = dax[a] == 0 ? xa : X[0];
        X[1] = dax[a] == 1 ? xa : X[1];
        X[2] = dax[a] == 2 ? xa : X[2];
      }
    }
    double b0 = ap.b[0], b1 = ap.b[1], b2 = ap.b[2], cr = ap.c;
    double delta = ap.supg ? ap.delta_const : 0.0;
    double f = ap.f_val;
    if (VAR && ap.qp) {
      // the caller's CoeffFct3D values at this point
      const double *cq = ap.qp + ((int64_t)c * NQ + q) * 6;
      const double eps = cq[0];
      b0 = cq[1];
      b1 = cq[2];
      b2 = cq[3];
      cr = cq[4];
      f = cq[5];
#pragma unroll
      for (int a = 0; a < 3; a++) {
        kap[a] = eps * ih[a] * ih[a];
        beta[a] = sg[a] * (dax[a] == 0 ? b0 : dax[a] == 1 ? b1 : b2) * ih[a];
      }
      if (ap.supg) {
        if (hconv < 0) hconv = hconv_brick(ap.h, b0, b1, b2);
        delta = sdfem_delta(hconv, eps, sqrt(b0 * b0 + b1 * b1 + b2 * b2), nu);
      }
    } else if (VAR && ap.conv_kind == 1) {
      b0 = -(X[1] - ap.rot_c[1]);
      b1 = X[0] - ap.rot_c[0];
      b2 = 0.1 * ap.sin2pit;
#pragma unroll
      for (int a = 0; a < 3; a++) beta[a] = sg[a] * (dax[a] == 0 ? b0 : dax[a] == 1 ? b1 : b2) * ih[a];
      if (ap.supg) {
        if (hconv < 0) hconv = hconv_brick(ap.h, b0, b1, b2);
        delta = sdfem_delta(hconv, ap.eps, sqrt(b0 * b0 + b1 * b1 + b2 * b2), nu);
      }
    }
    if (VAR && !ap.qp && ap.f_kind == 1) {
      const double sx = sin(M_PI * X[0]), sy = sin(M_PI * X[1]), sz = sin(M_PI * X[2]);
      const double cx = cos(M_PI * X[0]), cy = cos(M_PI * X[1]), cz = cos(M_PI * X[2]), u = sx * sy * sz;
      f = ap.eps * 3 * M_PI * M_PI * u + b0 * M_PI * cx * sy * sz + b1 * M_PI * sx * cy * sz +
          b2 * M_PI * sx * sy * cz + ap.c * u;
    }
    const double bg = beta[0] * g0 + beta[1] * g1 + beta[2] * g2;
    const double ti = phi_i + delta * bg;
    if (ap.what == 1) {  // mass; rhs with the SUPG test function (MatrixMRhsAssemble_SUPG, TCD3D.C:48)
      const double bw = w * phi_i;
#pragma unroll
      for (int j = 0; j < NB; j++) acc[j] = fma(bw, T[4 + 4 * j], acc[j]);
      racc += w * f * ti;
      continue;
    }
    if (ap.what == 2) {  // SUPG time-derivative matrix K (MatricesAKRhsAssemble_SUPG, TCD3D.C:102)
      const double kw = w * (delta * bg);
#pragma unroll
      for (int j = 0; j < NB; j++) acc[j] = fma(kw, T[4 + 4 * j], acc[j]);
      continue;
    }
    const double a0 = w * (kap[0] * g0 + ti * beta[0]);
    const double a1 = w * (kap[1] * g1 + ti * beta[1]);
    const double a2 = w * (kap[2] * g2 + ti * beta[2]);
    const double cc = w * (ti * cr);
#pragma unroll
    for (int j = 0; j < NB; j++)
      acc[j] = fma(a0, T[5 + 4 * j], fma(a1, T[6 + 4 * j], fma(a2, T[7 + 4 * j], fma(cc, T[4 + 4 * j], acc[j]))));
    racc += w * f * ti;
  }
  if (ap.E) {  // two-pass assembly: the element row, contiguously
    const int64_t row = (int64_t)c * NB + i;
#pragma unroll
    for (int j = 0; j < NB; j++) ap.E[row * NB + j] = acc[j];
    ap.Er[row] = racc;
    return;
  }
  const int r = gn[c * NB + i];
  if (r >= nact) return;
  const int *sl = slots + ((int64_t)c * NB + i) * NB;
#pragma unroll
  for (int j = 0; j < NB; j++) val[sl[j]] += acc[j];
  rhs[r] += racc;
}

// The element kernel on general hexahedra (mapped hierarchies): the same
// thread layout and j-loop as assemble_k, with the geometry of THexaTrilinear
// at every quadrature point.  With K = J^-T = C / det (C the cofactors) and
// the reference gradients g_i, the physical gradient is K g_i, and
//   eps (K g_i) . (K g_j) + t_i b . K g_j = (eps K^T K g_i + t_i K^T b) . g_j,
// so a_i = w |det| (eps K^T (K g_i) + t_i beta), beta = K^T b,
// t_i = phi_i + delta beta . g_i keep the brick kernel's form
// A_ij += a_i . g_j + c_i phi_j.  The SUPG delta uses the cell's mesh size in
// convection direction from hconv[c] (hconv_geo_k, run before).
template <int P, bool VAR>
__global__ __launch_bounds__(kBlock) void assemble_geo_k(int ncells, const int *__restrict__ list,
                                                         const int4 *__restrict__ cells, const int *__restrict__ slots,
                                                         const int *__restrict__ gn, int nact,
                                                         const double *__restrict__ tab, AsmParams ap,
                                                         double *__restrict__ val, double *__restrict__ rhs) {
  constexpr int NB = (P + 1) * (P + 1) * (P + 1);
  constexpr int N1 = P + 2;
  constexpr int NQ = N1 * N1 * N1;
  constexpr int STRIDE = 4 + 4 * NB;
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (gid >= (int64_t)ncells * NB) return;
  const int ci = (int)(gid / NB), i = (int)(gid % NB);
  const int c = list ? list[ci] : ci;
  TriMap M;
  tri_load(ap.vx, ap.nv, cells[c], P, M);
  const double nu = 1.0 / P;
  const double hc = ap.supg ? ap.hconv[c] : 0.0;
  // constant coefficients: one delta per cell
  double delta_c = 0.0;
  if (!VAR && ap.supg)
    delta_c = sdfem_delta(hc, ap.eps, sqrt(ap.b[0] * ap.b[0] + ap.b[1] * ap.b[1] + ap.b[2] * ap.b[2]), nu);
  double acc[NB];
#pragma unroll
  for (int j = 0; j < NB; j++) acc[j] = 0;
  double racc = 0;
  for (int q = 0; q < NQ; q++) {
    const double *T = tab + q * STRIDE;
    const double *Ti = T + 4 + 4 * i;
    const double phi_i = Ti[0], g0 = Ti[1], g1 = Ti[2], g2 = Ti[3];
    double eps = ap.eps, b0 = ap.b[0], b1 = ap.b[1], b2 = ap.b[2], cr = ap.c, f = ap.f_val;
    double delta = delta_c;
    if (VAR) {
      double X[3] = {0, 0, 0};
      if (ap.conv_kind == 1 || ap.f_kind == 1) tri_point(M, T[1], T[2], T[3], X);
      if (ap.qp) {
        const double *cq = ap.qp + ((int64_t)c * NQ + q) * 6;
        eps = cq[0];
        b0 = cq[1];
        b1 = cq[2];
        b2 = cq[3];
        cr = cq[4];
        f = cq[5];
      } else if (ap.conv_kind == 1) {
        b0 = -(X[1] - ap.rot_c[1]);
        b1 = X[0] - ap.rot_c[0];
        b2 = 0.1 * ap.sin2pit;
      }
      if (!ap.qp && ap.f_kind == 1) {
        const double sx = sin(M_PI * X[0]), sy = sin(M_PI * X[1]), sz = sin(M_PI * X[2]);
        const double cx = cos(M_PI * X[0]), cy = cos(M_PI * X[1]), cz = cos(M_PI * X[2]), u = sx * sy * sz;
        f = ap.eps * 3 * M_PI * M_PI * u + b0 * M_PI * cx * sy * sz + b1 * M_PI * sx * cy * sz +
            b2 * M_PI * sx * sy * cz + ap.c * u;
      }
      if (ap.supg) delta = sdfem_delta(hc, eps, sqrt(b0 * b0 + b1 * b1 + b2 * b2), nu);
    }
    double K[3][3];
    const double det = tri_jac(M, T[1], T[2], T[3], K);
    const double rd = 1.0 / det;
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int a = 0; a < 3; a++) K[d][a] *= rd;
    const double w = T[0] * fabs(det);
    double beta[3];
#pragma unroll
    for (int a = 0; a < 3; a++) beta[a] = b0 * K[0][a] + b1 * K[1][a] + b2 * K[2][a];
    const double bg = beta[0] * g0 + beta[1] * g1 + beta[2] * g2;
    const double ti = phi_i + delta * bg;
    if (ap.what == 1) {  // mass; rhs with the SUPG test function
      const double bw = w * phi_i;
#pragma unroll
      for (int j = 0; j < NB; j++) acc[j] = fma(bw, T[4 + 4 * j], acc[j]);
      racc += w * f * ti;
      continue;
    }
    if (ap.what == 2) {  // SUPG time-derivative matrix K
      const double kw = w * (delta * bg);
#pragma unroll
      for (int j = 0; j < NB; j++) acc[j] = fma(kw, T[4 + 4 * j], acc[j]);
      continue;
    }
    double G[3];  // physical gradient of phi_i
#pragma unroll
    for (int d = 0; d < 3; d++) G[d] = K[d][0] * g0 + K[d][1] * g1 + K[d][2] * g2;
    double a[3];
#pragma unroll
    for (int ax = 0; ax < 3; ax++) a[ax] = w * (eps * (K[0][ax] * G[0] + K[1][ax] * G[1] + K[2][ax] * G[2]) + ti * beta[ax]);
    const double cc = w * (ti * cr);
#pragma unroll
    for (int j = 0; j < NB; j++)
      acc[j] = fma(a[0], T[5 + 4 * j], fma(a[1], T[6 + 4 * j], fma(a[2], T[7 + 4 * j], fma(cc, T[4 + 4 * j], acc[j]))));
    racc += w * f * ti;
  }
  if (ap.E) {  // two-pass assembly: the element row, contiguously
    const int64_t row = (int64_t)c * NB + i;
#pragma unroll
    for (int j = 0; j < NB; j++) ap.E[row * NB + j] = acc[j];
    ap.Er[row] = racc;
    return;
  }
  const int r = gn[c * NB + i];
  if (r >= nact) return;
  const int *sl = slots + ((int64_t)c * NB + i) * NB;
#pragma unroll
  for (int j = 0; j < NB; j++) val[sl[j]] += acc[j];
  rhs[r] += racc;
}

// per cell of a mapped level: the SUPG mesh size in convection direction,
// with b at the cell's first quadrature point (the reference evaluates it
// once per cell, at the first point, and caches it: AssembleMat3D.C:457)
template <int P>
__global__ __launch_bounds__(kBlock) void hconv_geo_k(int ncell, const int4 *__restrict__ cells,
                                                      const double *__restrict__ tab, AsmParams ap,
                                                      double *__restrict__ hconv) {
  constexpr int NQ = (P + 2) * (P + 2) * (P + 2);
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= ncell) return;
  TriMap M;
  tri_load(ap.vx, ap.nv, cells[c], P, M);
  double b0 = ap.b[0], b1 = ap.b[1], b2 = ap.b[2];
  if (ap.qp) {
    const double *cq = ap.qp + (int64_t)c * NQ * 6;
    b0 = cq[1];
    b1 = cq[2];
    b2 = cq[3];
  } else if (ap.conv_kind == 1) {
    double X[3];
    tri_point(M, tab[1], tab[2], tab[3], X);
    b0 = -(X[1] - ap.rot_c[1]);
    b1 = X[0] - ap.rot_c[0];
    b2 = 0.1 * ap.sin2pit;
  }
  hconv[c] = tri_hconv(M, b0, b1, b2);
}

// Explicit inverse of the active block of the coarsest matrix by Gauss-Jordan
// elimination with partial pivoting, spread over many workgroups: one launch
// per pivot column.  Rows are never swapped or scaled during the sweep: the
// pivot row of column k (the unused row with the largest |M[r][k]|, ties to
// the lower index) is recorded in perm[k]; every other row subtracts its
// multiple of the (unmodified) pivot row.  At the end row perm[k] holds
// d_k e_k in the left half, and row k of the inverse is its right half / d_k.
// M is na x 2na ([A_AA | I]); one wavefront per row, lanes over columns.
constexpr int kGjWaves = kBlock / 64;

__device__ inline void gj_better(double &bv, int &bi, double v, int r) {
  if (v > bv || (v == bv && r < bi)) { bv = v; bi = r; }
}
// |x| as a pivot score; a NaN scores below every number, so that some unused
// row is always chosen and the pivot index stays in range
__device__ inline double gj_score(double x) {
  const double a = fabs(x);
  return a >= 0.0 ? a : -0.5;
}

__global__ __launch_bounds__(kBlock) void gj_init_k(int na, const int *__restrict__ rp, const int *__restrict__ col,
                                                    const double *__restrict__ val, double *__restrict__ M,
                                                    int *__restrict__ used, double *__restrict__ pv,
                                                    int *__restrict__ pi) {
  const int w2 = 2 * na, wave = blockIdx.x * kGjWaves + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  __shared__ double sv[kGjWaves];
  __shared__ int si[kGjWaves];
  double best = -1.0;
  int bi = na;
  for (int r = wave; r < na; r += gridDim.x * kGjWaves) {
    for (int j = lane; j < w2; j += 64) M[(int64_t)r * w2 + j] = (j == na + r) ? 1.0 : 0.0;
    __threadfence_block();  // the zeros land before the entries below
    double s0 = -3.0;       // score of column 0 (-3: no entry, i.e. 0)
    for (int t = rp[r] + lane; t < rp[r + 1]; t += 64) {
      const int c = col[t];
      if (c < na) M[(int64_t)r * w2 + c] = val[t];
      if (c == 0) s0 = gj_score(val[t]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s0 = fmax(s0, __shfl_xor(s0, o));
    if (lane == 0) used[r] = 0;
    gj_better(best, bi, s0 == -3.0 ? 0.0 : s0, r);
  }
  if (lane == 0) { sv[threadIdx.x >> 6] = best; si[threadIdx.x >> 6] = bi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kGjWaves; k++)
      if (sv[k] > sv[0] || (sv[k] == sv[0] && si[k] < si[0])) { sv[0] = sv[k]; si[0] = si[k]; }
    pv[blockIdx.x] = sv[0];
    pi[blockIdx.x] = si[0];
  }
}

// arg-max over the block partials (largest value, ties to the lower row),
// computed by every wavefront: lanes stride the partials, then a butterfly
__device__ inline int gj_pivot(const double *__restrict__ pv, const int *__restrict__ pi, int nparts) {
  double best = -1.0;
  int bi = 1 << 30;
  for (int t = threadIdx.x & 63; t < nparts; t += 64) {
    const double v = pv[t];
    const int r = pi[t];
    if (v > best || (v == best && r < bi)) { best = v; bi = r; }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double v = __shfl_xor(best, o);
    const int r = __shfl_xor(bi, o);
    if (v > best || (v == best && r < bi)) { best = v; bi = r; }
  }
  return bi;
}

__global__ __launch_bounds__(kBlock) void gj_step_k(int na, int k, double *__restrict__ M, int *__restrict__ used,
                                                    int *__restrict__ perm, const double *__restrict__ pv_in,
                                                    const int *__restrict__ pi_in, double *__restrict__ pv_out,
                                                    int *__restrict__ pi_out) {
  const int w2 = 2 * na, wave = blockIdx.x * kGjWaves + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int p = gj_pivot(pv_in, pi_in, gridDim.x);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    perm[k] = p;
    used[p] = 1;
  }
  __shared__ double sv[kGjWaves];
  __shared__ int si[kGjWaves];
  const double *__restrict__ prow = M + (int64_t)p * w2;
  const double pk = prow[k];
  double best = -1.0;
  int bi = na;
  for (int r = wave; r < na; r += gridDim.x * kGjWaves) {
    if (r == p) continue;
    double *__restrict__ row = M + (int64_t)r * w2;
    const double f = row[k] / pk;
    if (f != 0.0)
      for (int j = k + 1 + lane; j < w2; j += 64) row[j] -= f * prow[j];
    __syncwarp();
    if (lane == 0) {
      row[k] = 0.0;
      if (k + 1 < na && !used[r]) {
        const double v = gj_score(row[k + 1]);
        if (v > best) { best = v; bi = r; }
      }
    }
  }
  if (lane == 0) { sv[threadIdx.x >> 6] = best; si[threadIdx.x >> 6] = bi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int t = 1; t < kGjWaves; t++)
      if (sv[t] > sv[0] || (sv[t] == sv[0] && si[t] < si[0])) { sv[0] = sv[t]; si[0] = si[t]; }
    pv_out[blockIdx.x] = sv[0];
    pi_out[blockIdx.x] = si[0];
  }
}

// dval: the pivot values (blocked form), else M[perm[k]][k]
__global__ __launch_bounds__(kBlock) void gj_final_k(int na, const double *__restrict__ M,
                                                     const int *__restrict__ perm, const double *__restrict__ dval,
                                                     double *__restrict__ inv) {
  const int w2 = 2 * na, wave = blockIdx.x * kGjWaves + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  for (int k = wave; k < na; k += gridDim.x * kGjWaves) {
    const double *row = M + (int64_t)perm[k] * w2;
    const double d = 1.0 / (dval ? dval[k] : row[k]);
    for (int j = lane; j < na; j += 64) inv[(int64_t)k * na + j] = row[na + j] * d;
  }
}

// Small coarse blocks (na <= kGjLdsMax): the whole in-place Gauss-Jordan
// inversion in one workgroup with the matrix in LDS.  Partial pivoting with
// row interchanges (the largest |a_ik| for i >= k, ties to the lower row);
// the interchanges are undone as column interchanges in reverse order at the
// end.  One launch instead of na + 2.
constexpr int kGjLdsMax = 128;

__global__ __launch_bounds__(1024) void gj_lds_k(int na, const int *__restrict__ rp, const int *__restrict__ col,
                                                 const double *__restrict__ val, double *__restrict__ inv) {
  extern __shared__ double A[];  // na x na, row-major
  __shared__ double fcol[kGjLdsMax];
  __shared__ int ipiv[kGjLdsMax];
  __shared__ int pivrow;
  const int t = threadIdx.x, nt = blockDim.x, nn = na * na;
  for (int i = t; i < nn; i += nt) A[i] = 0.0;
  __syncthreads();
  for (int r = t; r < na; r += nt)
    for (int q = rp[r]; q < rp[r + 1]; q++)
      if (col[q] < na) A[r * na + col[q]] = val[q];
  __syncthreads();
  for (int k = 0; k < na; k++) {
    if (t < 64) {  // pivot search by the first wavefront
      double best = -1.0;
      int bi = k;
      for (int i = k + t; i < na; i += 64) {
        const double v = gj_score(A[i * na + k]);
        if (v > best || (v == best && i < bi)) { best = v; bi = i; }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double v = __shfl_xor(best, o);
        const int r = __shfl_xor(bi, o);
        if (v > best || (v == best && r < bi)) { best = v; bi = r; }
      }
      if (t == 0) {
        pivrow = bi;
        ipiv[k] = bi;
      }
    }
    __syncthreads();
    const int p = pivrow;
    if (p != k)
      for (int j = t; j < na; j += nt) {
        const double a = A[k * na + j];
        A[k * na + j] = A[p * na + j];
        A[p * na + j] = a;
      }
    __syncthreads();
    const double dinv = 1.0 / A[k * na + k];
    for (int j = t; j < na; j += nt) {
      if (j != k) A[k * na + j] *= dinv;
      fcol[j] = A[j * na + k];  // column k before the elimination
    }
    __syncthreads();
    if (t == 0) A[k * na + k] = dinv;
    for (int e = t; e < nn; e += nt) {
      const int i = e / na, j = e % na;
      if (i == k) continue;
      if (j == k) A[e] = -fcol[i] * dinv;
      else A[e] -= fcol[i] * A[k * na + j];
    }
    __syncthreads();
  }
  for (int k = na - 1; k >= 0; k--) {
    const int p = ipiv[k];
    if (p != k)
      for (int i = t; i < na; i += nt) {
        const double a = A[i * na + k];
        A[i * na + k] = A[i * na + p];
        A[i * na + p] = a;
      }
    __syncthreads();
  }
  for (int i = t; i < nn; i += nt) inv[i] = A[i];
}

// Blocked form of the same elimination for kGjLdsMax < na <= kGjBlockedMax:
// the same pivots (largest unused |m_rk|, ties to the lower row), the same
// multipliers and the same extraction, reorganised by panels of NB pivot
// columns.  The panel kernel (one workgroup; each thread holds RPT rows of the
// panel in registers) finds the NB pivots of the panel and every row's
// multipliers W[r][j] (0 for the pivot row of step j).  The update kernel then
// applies their combined effect to the columns right of the panel as one
// rank-NB update,
//   U_j  = M[p_j] - sum_{i<j} W[p_j][i] U_i   (pivot row p_j as it was at step j)
//   M'[r] = M[r] - sum_j W[r][j] U_j           (every row r),
// reading M and writing M' (two buffers, so that the U rows are formed from
// the old values while M' is written).  The pivot values go to dval[k].  Two
// launches per panel instead of NB.
constexpr int kGjPanel = 16;        // stride of the multiplier rows W
constexpr int kGjBlockedMax = 4096;  // 4 rows per thread of the panel kernel
constexpr int kGjUpdRows = 16;      // rows per workgroup of the update kernel

template <int NB, int RPT>
__global__ __launch_bounds__(1024) void gj_panel_k(int na, int k0, int b, const double *__restrict__ M,
                                                   int *__restrict__ used, int *__restrict__ perm,
                                                   double *__restrict__ dval, double *__restrict__ W) {
  __shared__ double sv[2][16];
  __shared__ int si[2][16];
  __shared__ double srow[2][16][NB];  // each wave's candidate row, double-buffered by step
  const int w2 = 2 * na, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  double v[RPT][NB];
  bool un[RPT];
  int pj[RPT];
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const int r = t + q * 1024;
    pj[q] = -1;
    un[q] = r < na && !used[r];
#pragma unroll
    for (int j = 0; j < NB; j++) v[q][j] = (r < na && j < b) ? M[(int64_t)r * w2 + k0 + j] : 0.0;
  }
  // the wave's best unused row for column j; its owner lane stages the row
  auto propose = [&](const int j, const int buf) {
    double bv = -1.0;
    int bi = 1 << 30;
#pragma unroll
    for (int q = 0; q < RPT; q++)
      if (un[q]) gj_better(bv, bi, gj_score(v[q][j]), t + q * 1024);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(bv, o);
      const int oi = __shfl_xor(bi, o);
      gj_better(bv, bi, ov, oi);
    }
#pragma unroll
    for (int q = 0; q < RPT; q++)
      if (un[q] && t + q * 1024 == bi) {
#pragma unroll
        for (int jj = j; jj < NB; jj++) srow[buf][wv][jj] = v[q][jj];
      }
    if (lane == 0) {
      sv[buf][wv] = bv;
      si[buf][wv] = bi;
    }
  };
  const int nw = blockDim.x >> 6;
  propose(0, 0);
#pragma unroll
  for (int j = 0; j < NB; j++) {
    if (j < b) {
      const int buf = j & 1;
      __syncthreads();
      double gv = lane < nw ? sv[buf][lane] : -2.0;
      int gi = lane < nw ? si[buf][lane] : (1 << 30), gs = lane;
#pragma unroll
      for (int o = 8; o > 0; o >>= 1) {
        const double ov = __shfl_xor(gv, o);
        const int oi = __shfl_xor(gi, o), os = __shfl_xor(gs, o);
        if (ov > gv || (ov == gv && oi < gi)) { gv = ov; gi = oi; gs = os; }
      }
      const int p = __shfl(gi, 0), ps = __shfl(gs, 0);
      const double d = srow[buf][ps][j];
      if (t == 0) {
        perm[k0 + j] = p;
        used[p] = 1;
        dval[k0 + j] = d;
      }
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        const int r = t + q * 1024;
        if (r == p) {
          un[q] = false;
          pj[q] = j;
        } else if (r < na) {
          const double f = v[q][j] / d;
          if (f != 0.0)
#pragma unroll
            for (int jj = j + 1; jj < NB; jj++) v[q][jj] -= f * srow[buf][ps][jj];
          v[q][j] = f;
        }
      }
      if (j + 1 < b) propose(j + 1, buf ^ 1);
    }
  }
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const int r = t + q * 1024;
    if (r < na)
#pragma unroll
      for (int j = 0; j < NB; j++) W[(int64_t)r * kGjPanel + j] = (j < b && j != pj[q]) ? v[q][j] : 0.0;
  }
}

// rank-NB update of the columns right of the panel; a workgroup forms U for
// its 64 columns (each wave its own copy, in registers) and updates
// kGjUpdRows rows of them
template <int NB>
__global__ __launch_bounds__(256) void gj_update_k(int na, int k0, int b, const double *__restrict__ Min,
                                                   double *__restrict__ Mout, const int *__restrict__ perm,
                                                   const double *__restrict__ W) {
  const int w2 = 2 * na, c = k0 + b + blockIdx.x * 64 + (threadIdx.x & 63), wv = threadIdx.x >> 6;
  if (c >= w2) return;
  double u[NB];
#pragma unroll
  for (int j = 0; j < NB; j++) {
    u[j] = 0.0;
    if (j < b) {
      const int p = perm[k0 + j];
      double s = Min[(int64_t)p * w2 + c];
#pragma unroll
      for (int i = 0; i < j; i++) s -= W[(int64_t)p * kGjPanel + i] * u[i];
      u[j] = s;
    }
  }
  const int r1 = min(na, (int)(blockIdx.y + 1) * kGjUpdRows);
  for (int r = blockIdx.y * kGjUpdRows + wv; r < r1; r += 4) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < NB; j++) s += W[(int64_t)r * kGjPanel + j] * u[j];
    Mout[(int64_t)r * w2 + c] = Min[(int64_t)r * w2 + c] - s;
  }
}

// x += sum_k coef[k] Z_k  (FGMRES update x += Z y, UpdateGmresIterate)
// multi_axpy_k on 16-byte pairs (Z, x and ldz 16-byte aligned)
__global__ void multi_axpy2_k(int n, int m, const double *__restrict__ coef, const double *__restrict__ Z,
                              int64_t ldz, double *__restrict__ x) {
  const int n2 = n >> 1;
  double2 *x2 = (double2 *)x;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n2; i += gridDim.x * kBlock) {
    double2 t = {0, 0};
    for (int k = 0; k < m; k++) {
      const double2 z = ((const double2 *)(Z + k * ldz))[i];
      t.x += coef[k] * z.x;
      t.y += coef[k] * z.y;
    }
    double2 v = x2[i];
    v.x += t.x;
    v.y += t.y;
    x2[i] = v;
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    double t = 0;
    for (int k = 0; k < m; k++) t += coef[k] * Z[k * ldz + n - 1];
    x[n - 1] += t;
  }
}

// x += sum_k coef[k] Z_k with FP32 columns (the mixed form's Krylov columns
// z_k, FP32 V-cycle outputs): each value converted exactly, so the sums of
// multi_axpy_k over their FP64 copies
__global__ void multi_axpy32_k(int n, int m, const double *__restrict__ coef, const float *__restrict__ Z,
                               int64_t ldz, double *__restrict__ x) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double s = 0;
    for (int k = 0; k < m; k++) s += coef[k] * (double)Z[k * ldz + i];
    x[i] += s;
  }
}

__global__ void multi_axpy_k(int n, int m, const double *__restrict__ coef, const double *__restrict__ Z,
                             int64_t ldz, double *__restrict__ x) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double s = 0;
    for (int k = 0; k < m; k++) s += coef[k] * Z[k * ldz + i];
    x[i] += s;
  }
}

// x_i = v for i in [b, e)  (Dirichlet part of an iterate)
__global__ void copy_range_k(int b, int e, const double *__restrict__ src, double *__restrict__ dst) {
  const int i = b + blockIdx.x * kBlock + threadIdx.x;
  if (i < e) dst[i] = src[i];
}

__global__ void dirichlet_rows_k(int n, int nact, const int *__restrict__ rp, double *__restrict__ val,
                                 double *__restrict__ rhs) {
  const int r = nact + blockIdx.x * kBlock + threadIdx.x;
  if (r >= n) return;
  val[rp[r]] = 1.0;  // AddLocalSqMatToGlobal: Dirichlet row = identity row
  rhs[r] = 0.0;      // AddLocalRhsToGlobal: Dirichlet value g = 0
}

// Smoothers 2 and 3 in the reference's row order (TMGLevel3D::SOR / SSOR,
// MGLevel3D.C:254-441).  The rows come grouped by wavefront level
// (build_gs_schedule): every row of a level reads only rows of lower levels
// (new values) or of higher levels (old values), exactly what the sequential
// sweep reads, so a level's rows update at once.  One workgroup walks the
// levels with a barrier between them; each wave takes one row at a time,
// its lanes over the row's nonzeros.  Every wave reaches every barrier (the
// level loop is uniform).  The sweep is as sequential as the reference's:
// levels hold a handful of rows (about 8 on the 32^3-cell Q2 level).  Only
// the gather of x depends on the previous level, so each wave fetches its
// row of the next level (index, extent, the first 128 entries, f_i, x_i)
// while it works on the current one.
constexpr int kGsBlock = 1024;
struct GsRow {
  int i, b, e, c0, c1;
  double v0, v1, fi, xi;
};
__device__ inline void gs_fetch(int r, int end, const int *__restrict__ rows, const int *__restrict__ rp,
                                const int *__restrict__ col, const double *__restrict__ val,
                                const double *x, const double *__restrict__ f, int lane, GsRow &q) {
  q.i = -1;
  if (r >= end) return;
  q.i = rows[r];
  q.b = rp[q.i];
  q.e = rp[q.i + 1];
  const int j0 = q.b + lane, j1 = j0 + 64;
  q.c0 = j0 < q.e ? col[j0] : -1;
  q.v0 = j0 < q.e ? val[j0] : 0.0;
  q.c1 = j1 < q.e ? col[j1] : -1;
  q.v1 = j1 < q.e ? val[j1] : 0.0;
  q.fi = f[q.i];
  q.xi = x[q.i];  // written only by this row's own update
}
__device__ inline void gs_update(const GsRow &q, const int *__restrict__ col, const double *__restrict__ val,
                                 double *x, double omega, int lane) {
  double s = 0.0, diag = 0.0;
  if (q.c0 >= 0) {
    if (q.c0 == q.i) diag = q.v0;
    else s += q.v0 * x[q.c0];
  }
  if (q.c1 >= 0) {
    if (q.c1 == q.i) diag = q.v1;
    else s += q.v1 * x[q.c1];
  }
  for (int j = q.b + 128 + lane; j < q.e; j += 64) {  // rows longer than 128 entries
    const int c = col[j];
    if (c == q.i) diag = val[j];
    else s += val[j] * x[c];
  }
  s = group_sum<64>(s);
  diag = group_sum<64>(diag);
  if (lane == 0) x[q.i] = omega * ((q.fi - s) / diag - q.xi) + q.xi;
}
__global__ __launch_bounds__(kGsBlock) void gs_sweep_k(int nlev, const int *__restrict__ start,
                                                       const int *__restrict__ rows, const int *__restrict__ rp,
                                                       const int *__restrict__ col, const double *__restrict__ val,
                                                       double *x, const double *__restrict__ f, double omega) {
  constexpr int W = kGsBlock / 64;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (nlev <= 0) return;
  int sb = start[0], se = start[1];
  GsRow cur, nxt;
  gs_fetch(sb + w, se, rows, rp, col, val, x, f, lane, cur);
  for (int k = 0; k < nlev; k++) {
    const int ne = k + 2 <= nlev ? start[k + 2] : se;
    // the next level's row: x of it is old until its own update
    gs_fetch(se + w, ne, rows, rp, col, val, x, f, lane, nxt);
    if (cur.i >= 0) gs_update(cur, col, val, x, omega, lane);
    for (int r = sb + w + W; r < se; r += W) {  // levels with more rows than waves
      GsRow q;
      gs_fetch(r, se, rows, rp, col, val, x, f, lane, q);
      gs_update(q, col, val, x, omega, lane);
    }
    __syncthreads();
    cur = nxt;
    sb = se;
    se = ne;
  }
}

}  // namespace

// ---------------------------------------------------------------- launchers

void k_spmv(hipStream_t s, const DevLevel &L, const double *x, double *y) {
  launch_stream<kSpmv, double, double>(s, L, L.val, x, nullptr, nullptr, y, 0.0);
}
void k_defect(hipStream_t s, const DevLevel &L, const double *x, const double *f, double *d) {
  launch_stream<kDefect, double, double>(s, L, L.val, x, f, nullptr, d, 0.0);
}
void k_jacobi(hipStream_t s, const DevLevel &L, const double *x, const double *f, double *out, double omega) {
  launch_stream<kJacobi, double, double>(s, L, L.val, x, f, L.invdiag, out, omega);
}
void k_invdiag(hipStream_t s, const DevLevel &L) {
  invdiag_k<double><<<(L.ndof + kBlock - 1) / kBlock, kBlock, 0, s>>>(L.ndof, L.diag, L.val, L.invdiag);
}
void k_jacobi_zero(hipStream_t s, const DevLevel &L, const double *f, double *out, double omega) {
  jacobi_zero_k<double><<<(L.ndof + kBlock - 1) / kBlock, kBlock, 0, s>>>(L.ndof, L.nact, L.diag, L.val, nullptr,
                                                                         f, out, omega);
}
void k_matvec_pair_active(hipStream_t s, const DevLevel &L, const double *a, const double *b, double alpha,
                          double beta, const double *x, double *y) {
  if (L.nact <= 0) return;
  const double m = (double)L.nnz / L.ndof;
  const int lanes = m <= 12 ? 4 : m <= 40 ? 8 : 16;
  const int grid = (int)(((int64_t)L.nact * lanes + kBlock - 1) / kBlock);
  if (lanes == 4) csr_pair_active_k<4><<<grid, kBlock, 0, s>>>(L.nact, L.rowptr, L.col, a, b, alpha, beta, x, y);
  else if (lanes == 8) csr_pair_active_k<8><<<grid, kBlock, 0, s>>>(L.nact, L.rowptr, L.col, a, b, alpha, beta, x, y);
  else csr_pair_active_k<16><<<grid, kBlock, 0, s>>>(L.nact, L.rowptr, L.col, a, b, alpha, beta, x, y);
}
void k_matadd_into(hipStream_t s, int64_t n, double *C, const double *D, double alpha) {
  matadd_into_k<<<grid_for(n), kBlock, 0, s>>>(n, C, D, alpha);
}
void k_matvec_active(hipStream_t s, const DevLevel &L, const double *x, double *y) {
  launch_rows<kActiveSpmv, double, double>(s, L.nnz, L.ndof, L.nact, L.rowptr, L.col, L.val, x,
                                           nullptr, y, 0.0);
}
void k_mcssor(hipStream_t s, const DevLevel &L, double *x, const double *f, double omega) {
  // x_D = f_D, then the colours forward and backward
  k_copy_range(s, L.nact, L.ndof, f, x);
  const int ncol = (int)L.mc_start.size() - 1;
  int lanes, unroll;
  row_shape(L.nnz, L.ndof, lanes, unroll);
  for (int t = 0; t < 2 * ncol; t++) {
    const int c = t < ncol ? t : 2 * ncol - 1 - t;
    const int n = L.mc_start[c + 1] - L.mc_start[c];
    if (n <= 0) continue;
    const int *rows = L.mc_rows + L.mc_start[c];
    const int grid = (int)(((int64_t)n * lanes + kBlock - 1) / kBlock);
    switch (lanes) {
      case 4: colour_sor_k<4, 1><<<grid, kBlock, 0, s>>>(n, rows, L.rowptr, L.col, L.val, x, f, omega); break;
      case 8: colour_sor_k<8, 4><<<grid, kBlock, 0, s>>>(n, rows, L.rowptr, L.col, L.val, x, f, omega); break;
      case 16: colour_sor_k<16, 4><<<grid, kBlock, 0, s>>>(n, rows, L.rowptr, L.col, L.val, x, f, omega); break;
      default: colour_sor_k<32, 4><<<grid, kBlock, 0, s>>>(n, rows, L.rowptr, L.col, L.val, x, f, omega); break;
    }
  }
}
void k_gs_sweep(hipStream_t s, const DevLevel &L, double *x, const double *f, double omega, int kind) {
  k_copy_range(s, L.nact, L.ndof, f, x);
  gs_sweep_k<<<1, kGsBlock, 0, s>>>(L.gs_nlev[0], L.gs_start[0], L.gs_rows[0], L.rowptr, L.col, L.val, x, f, omega);
  if (kind == 3)
    gs_sweep_k<<<1, kGsBlock, 0, s>>>(L.gs_nlev[1], L.gs_start[1], L.gs_rows[1], L.rowptr, L.col, L.val, x, f,
                                      omega);
}
void k_restrict(hipStream_t s, const DevLevel &F, const DevLevel &C, const double *fine, double *coarse) {
  launch_rows<kRestrict, double, double>(s, F.r_nnz, C.ndof, C.nact, F.r_rowptr, F.r_col,
                                         F.r_val, fine, nullptr, coarse, 0.0);
}
void k_prolong_update(hipStream_t s, const DevLevel &F, const DevLevel &C, const double *coarse, double *x,
                      double damp) {
  launch_rows<kProlongUpd, double, double>(s, F.p_nnz, F.ndof, F.nact, F.p_rowptr, F.p_col,
                                           F.p_val, coarse, nullptr, x, damp);
}
void k_prolong(hipStream_t s, const DevLevel &F, const DevLevel &C, const double *coarse, double *fine) {
  (void)C;
  launch_rows<kSpmv, double, double>(s, F.p_nnz, F.ndof, F.nact, F.p_rowptr, F.p_col, F.p_val,
                                     coarse, nullptr, fine, 0.0);
}
void k_dot(hipStream_t s, int n, const double *a, const double *b, double *partials, double *out) {
  if (al16(a) && al16(b)) dot2_k<<<kRedBlocks, kBlock, 0, s>>>(n, a, b, partials);
  else dot_k<<<kRedBlocks, kBlock, 0, s>>>(n, a, b, partials);
  if (out) finish_k<<<1, kBlock, 0, s>>>(partials, out);
}
void k_axpy_dot(hipStream_t s, int n, const double *alpha_partials, double sign, const double *x, double *y,
                const double *z, double *partials, double *alpha_out) {
  if (al16(x) && al16(y) && al16(z)) {
    if (z) axpy_dot2_k<true><<<kRedBlocks, kBlock, 0, s>>>(n, alpha_partials, alpha_out, sign, x, y, z, partials);
    else axpy_dot2_k<false><<<kRedBlocks, kBlock, 0, s>>>(n, alpha_partials, alpha_out, sign, x, y, z, partials);
  } else {
    axpy_dot_k<<<kRedBlocks, kBlock, 0, s>>>(n, alpha_partials, alpha_out, sign, x, y, z, partials);
  }
}
void k_scale_copy(hipStream_t s, int n, const double *partials, double *norm_out, const double *x, double *y,
                  float *y32, const TopZeroSweep *zt) {
  const TopZeroSweep z = zt && y32 ? *zt : TopZeroSweep{};
  if (al16(x) && al16(y) && ((uintptr_t)y32 & 7) == 0) {
    scale_copy2_k<<<kRedBlocks, kBlock, 0, s>>>(n, partials, norm_out, x, y, y32, z);
  } else {
    scale_copy_k<<<kRedBlocks, kBlock, 0, s>>>(n, partials, norm_out, x, y);
    if (y32) scale_cast_k<<<grid_for(n), kBlock, 0, s>>>(n, 1.0, y, y32, z);
  }
}
int k_ls_blocks(int n) { return (n + kLsChunk - 1) / kLsChunk; }
int k_ls_max() { return kLsMax; }
void k_ls_step(hipStream_t s, int n, int i, const double *V, int64_t ld, double *w, double *part, double *Lm,
               double *h, double *partials) {
  const int nb = k_ls_blocks(n);
  ls_dots_k<<<nb, kBlock, 0, s>>>(n, i, V, ld, w, part);
  ls_coeffs_k<<<1, kLsCoef, 0, s>>>(i, nb, part, Lm, h);
  ls_update2_k<<<kRedBlocks, kBlock, 0, s>>>(n, i, h, V, ld, w, partials);
}
void k_scale_cast(hipStream_t s, int n, double alpha, double *x, float *y32, const TopZeroSweep *zt) {
  scale_cast_k<<<grid_for(n), kBlock, 0, s>>>(n, alpha, x, y32, zt && y32 ? *zt : TopZeroSweep{});
}
void k_scale(hipStream_t s, int n, double alpha, double *x) { scale_k<<<grid_for(n), kBlock, 0, s>>>(n, alpha, x); }
void k_axpy(hipStream_t s, int n, double alpha, const double *x, double *y) {
  axpy_k<<<grid_for(n), kBlock, 0, s>>>(n, alpha, x, y);
}
void k_axpy_dev(hipStream_t s, int n, const double *alpha_dev, int idx, const double *x, double *y) {
  axpy_dev_k<<<grid_for(n), kBlock, 0, s>>>(n, alpha_dev, idx, x, y);
}
void k_copy(hipStream_t s, int n, const double *x, double *y) { copy_k<<<grid_for(n), kBlock, 0, s>>>(n, x, y); }
void k_fill(hipStream_t s, int64_t n, double v, double *y) { fill_k<<<grid_for(n), kBlock, 0, s>>>(n, v, y); }
void k_matadd(hipStream_t s, int64_t n, double *C, const double *D, double alpha) {
  matadd_k<<<grid_for(n), kBlock, 0, s>>>(n, C, D, alpha);
}
void k_coarse_from_lattice(hipStream_t s, const DevLevel &C, const double *Ainv, const float *lf, float *lx) {
  if (C.nact > 0)
    coarse_lat_k<<<(C.nact * 64 + kBlock - 1) / kBlock, kBlock, 0, s>>>(C.nact, Ainv, C.ldof2node, lf, lx);
}
void k_gemv_dense(hipStream_t s, int n, const double *A, const double *x, double *y) {
  if (n > 0) coarse_update_k<<<(n * 64 + kBlock - 1) / kBlock, kBlock, 0, s>>>(n, A, x, y);
}
void k_coarse_smooth(hipStream_t s, const DevLevel &L, double *x, const double *f, double *d, int smoother,
                     double omega, int maxit, double red) {
  coarse_smooth_k<<<1, 64, sizeof(double) * L.ndof, s>>>(L.ndof, L.nact, L.rowptr, L.col, L.val, x, f, d, smoother,
                                                         omega, maxit, red);
}
void k_assemble(hipStream_t s, const DevLevel &L, const double *tab, const AsmParams &ap, double *val, double *rhs,
                int colour_begin, int colour_count) {
  if (colour_count <= 0) return;
  const int *list = L.colour_cells + colour_begin;
  // CD3D_ASM_LDS_TAB=1: per-thread table operands from an LDS copy of the
  // table (measured on the 64^3 Q2 level, constant SUPG coefficients: 9.3 ms
  // per assembly against 6.7 ms from global memory: the 55 KB copy limits a
  // CU to two workgroups and costs a staging pass per workgroup)
  static const int lds_tab = [] {
    const char *e = getenv("CD3D_ASM_LDS_TAB");
    return e ? atoi(e) : 0;
  }();
  const bool var = ap.conv_kind != 0 || ap.f_kind != 0 || ap.qp != nullptr;
  const int64_t threads = (int64_t)colour_count * (ap.order == 1 ? 8 : 27);
  const int grid = (int)((threads + kBlock - 1) / kBlock);
  if (ap.vx) {  // general hexahedra
    if (ap.order == 1) {
      if (var) assemble_geo_k<1, true><<<grid, kBlock, 0, s>>>(colour_count, list, L.cells, L.slots, L.gn, L.nact, tab, ap, val, rhs);
      else assemble_geo_k<1, false><<<grid, kBlock, 0, s>>>(colour_count, list, L.cells, L.slots, L.gn, L.nact, tab, ap, val, rhs);
    } else {
      if (var) assemble_geo_k<2, true><<<grid, kBlock, 0, s>>>(colour_count, list, L.cells, L.slots, L.gn, L.nact, tab, ap, val, rhs);
      else assemble_geo_k<2, false><<<grid, kBlock, 0, s>>>(colour_count, list, L.cells, L.slots, L.gn, L.nact, tab, ap, val, rhs);
    }
    return;
  }
#define CD3D_ASM(P_, LDS_, VAR_) \
  assemble_k<P_, LDS_, VAR_><<<grid, kBlock, 0, s>>>(colour_count, list, L.cells, L.slots, L.gn, L.nact, tab, ap, val, rhs)
  if (ap.order == 1) {
    if (lds_tab) { if (var) CD3D_ASM(1, true, true); else CD3D_ASM(1, true, false); }
    else { if (var) CD3D_ASM(1, false, true); else CD3D_ASM(1, false, false); }
  } else {
    if (lds_tab) { if (var) CD3D_ASM(2, true, true); else CD3D_ASM(2, true, false); }
    else { if (var) CD3D_ASM(2, false, true); else CD3D_ASM(2, false, false); }
  }
#undef CD3D_ASM
}
// CSR entry e of an active row: the sum of its element-row contributions in
// colour order (gather lists of build_gather); active rhs entries likewise
__global__ __launch_bounds__(kBlock) void gather_entries_k(int64_t n, const int *__restrict__ gptr,
                                                           const int *__restrict__ gidx,
                                                           const double *__restrict__ E, double *__restrict__ val) {
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += (int64_t)gridDim.x * kBlock) {
    double s = 0.0;
    for (int k = gptr[e]; k < gptr[e + 1]; k++) s += E[gidx[k]];
    val[e] = s;
  }
}

void k_assemble_rows(hipStream_t s, const DevLevel &L, const double *tab, const AsmParams &ap) {
  const int nb = ap.order == 1 ? 8 : 27;
  const int64_t threads = (int64_t)L.ncell * nb;
  const int grid = (int)((threads + kBlock - 1) / kBlock);
  if (ap.vx) {
    if (ap.order == 1) {
      if (ap.conv_kind != 0 || ap.f_kind != 0 || ap.qp)
        assemble_geo_k<1, true><<<grid, kBlock, 0, s>>>(L.ncell, nullptr, L.cells, nullptr, L.gn, L.nact, tab, ap, nullptr, nullptr);
      else
        assemble_geo_k<1, false><<<grid, kBlock, 0, s>>>(L.ncell, nullptr, L.cells, nullptr, L.gn, L.nact, tab, ap, nullptr, nullptr);
    } else {
      if (ap.conv_kind != 0 || ap.f_kind != 0 || ap.qp)
        assemble_geo_k<2, true><<<grid, kBlock, 0, s>>>(L.ncell, nullptr, L.cells, nullptr, L.gn, L.nact, tab, ap, nullptr, nullptr);
      else
        assemble_geo_k<2, false><<<grid, kBlock, 0, s>>>(L.ncell, nullptr, L.cells, nullptr, L.gn, L.nact, tab, ap, nullptr, nullptr);
    }
    return;
  }
  const bool var = ap.conv_kind != 0 || ap.f_kind != 0 || ap.qp != nullptr;
  if (ap.order == 1) {
    if (var) assemble_k<1, false, true><<<grid, kBlock, 0, s>>>(L.ncell, nullptr, L.cells, nullptr, L.gn, L.nact, tab, ap, nullptr, nullptr);
    else assemble_k<1, false, false><<<grid, kBlock, 0, s>>>(L.ncell, nullptr, L.cells, nullptr, L.gn, L.nact, tab, ap, nullptr, nullptr);
  } else {
    if (var) assemble_k<2, false, true><<<grid, kBlock, 0, s>>>(L.ncell, nullptr, L.cells, nullptr, L.gn, L.nact, tab, ap, nullptr, nullptr);
    else assemble_k<2, false, false><<<grid, kBlock, 0, s>>>(L.ncell, nullptr, L.cells, nullptr, L.gn, L.nact, tab, ap, nullptr, nullptr);
  }
}

void k_gather_assembly(hipStream_t s, const DevLevel &L, int64_t nnz_act, double *val, double *rhs) {
  if (nnz_act > 0)
    gather_entries_k<<<(int)std::min<int64_t>((nnz_act + kBlock - 1) / kBlock, 1 << 16), kBlock, 0, s>>>(
        nnz_act, L.gptr, L.gidx, L.E, val);
  if (L.nact > 0)
    gather_entries_k<<<(L.nact + kBlock - 1) / kBlock, kBlock, 0, s>>>(L.nact, L.rptr, L.ridx, L.Er, rhs);
}

void k_hconv_geo(hipStream_t s, const DevLevel &L, const double *tab, const AsmParams &ap, double *hconv) {
  if (L.ncell <= 0) return;
  const int grid = (L.ncell + kBlock - 1) / kBlock;
  if (ap.order == 1) hconv_geo_k<1><<<grid, kBlock, 0, s>>>(L.ncell, L.cells, tab, ap, hconv);
  else hconv_geo_k<2><<<grid, kBlock, 0, s>>>(L.ncell, L.cells, tab, ap, hconv);
}

// blocked elimination (panels of NB pivots); returns the buffer holding the result
template <int NB, int RPT>
const double *gj_blocked(hipStream_t s, int na, double *M, double *M1, int *used, int *perm, double *dval,
                         double *W) {
  double *Mi = M, *Mo = M1;
  for (int k0 = 0; k0 < na; k0 += NB) {
    const int b = std::min(NB, na - k0), ncol = 2 * na - k0 - b;  // ncol >= na > 0
    gj_panel_k<NB, RPT><<<1, RPT == 1 ? (na + 63) / 64 * 64 : 1024, 0, s>>>(na, k0, b, Mi, used, perm, dval, W);
    const dim3 grid((ncol + 63) / 64, (na + kGjUpdRows - 1) / kGjUpdRows);
    gj_update_k<NB><<<grid, 256, 0, s>>>(na, k0, b, Mi, Mo, perm, W);
    std::swap(Mi, Mo);
  }
  return Mi;
}

// CD3D_GJ_UNBLOCKED=1: one launch per pivot also for large blocks (comparison runs)
static bool gj_unblocked() {
  static const bool v = [] {
    const char *e = getenv("CD3D_GJ_UNBLOCKED");
    return e && atoi(e) != 0;
  }();
  return v;
}

void k_invert_active(hipStream_t s, const DevLevel &L, double *scratch, double *inv) {
  // scratch layout: M, M' (2 na^2 doubles each), W (kGjPanel na), dval[na],
  // two partial (value, index) buffers, used[na], perm[na]
  const int na = L.nact;
  if (na <= 0) return;
  if (na <= kGjLdsMax) {
    // the larger-than-64 KiB LDS opt-in, once per device; the serving loop's
    // worker threads reach this concurrently
    static std::mutex lds_mu;
    static int lds_ok[64];
    static bool lds_init = false;
    int dev = 0;
    (void)hipGetDevice(&dev);
    bool ok;
    {
      std::lock_guard<std::mutex> g(lds_mu);
      if (!lds_init) {
        for (int &v : lds_ok) v = -1;
        lds_init = true;
      }
      int &st = lds_ok[dev & 63];
      if (st < 0)
        st = hipFuncSetAttribute((const void *)gj_lds_k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)(sizeof(double) * kGjLdsMax * kGjLdsMax)) == hipSuccess;
      ok = st != 0;
    }
    const size_t bytes = sizeof(double) * na * na;
    if (ok || bytes <= 65536) {
      gj_lds_k<<<1, 1024, bytes, s>>>(na, L.rowptr, L.col, L.val, inv);
      return;
    }
    (void)hipGetLastError();
  }
  const int nblk = std::min((na + kGjWaves - 1) / kGjWaves, 256);
  double *M = scratch, *M1 = M + (int64_t)2 * na * na, *W = M1 + (int64_t)2 * na * na, *dval = W + (int64_t)kGjPanel * na;
  double *pv0 = dval + na, *pv1 = pv0 + 256;
  int *pi0 = (int *)(pv1 + 256), *pi1 = pi0 + 256, *used = pi1 + 256, *perm = used + na;
  gj_init_k<<<nblk, kBlock, 0, s>>>(na, L.rowptr, L.col, L.val, M, used, pv0, pi0);
  if (na <= kGjBlockedMax && !gj_unblocked()) {
    const int rpt = (na + 1023) / 1024;
    const double *Mf = rpt == 1   ? gj_blocked<16, 1>(s, na, M, M1, used, perm, dval, W)
                       : rpt == 2 ? gj_blocked<8, 2>(s, na, M, M1, used, perm, dval, W)
                       : rpt == 3 ? gj_blocked<8, 3>(s, na, M, M1, used, perm, dval, W)
                                  : gj_blocked<8, 4>(s, na, M, M1, used, perm, dval, W);
    gj_final_k<<<nblk, kBlock, 0, s>>>(na, Mf, perm, dval, inv);
    return;
  }
  for (int k = 0; k < na; k++) {
    const bool even = (k % 2) == 0;
    gj_step_k<<<nblk, kBlock, 0, s>>>(na, k, M, used, perm, even ? pv0 : pv1, even ? pi0 : pi1, even ? pv1 : pv0,
                                      even ? pi1 : pi0);
  }
  gj_final_k<<<nblk, kBlock, 0, s>>>(na, M, perm, nullptr, inv);
}
bool k_invert_multi_launch(int na) { return na > kGjLdsMax; }

int64_t gj_scratch_doubles(int na) {
  // M and M' (na x 2na each), W (na x kGjPanel), dval, partials, used + perm (ints)
  return (int64_t)4 * na * na + (int64_t)(kGjPanel + 1) * na + 512 + (512 + 2 * (int64_t)na + 2) / 2 + 1;
}
void k_multi_axpy32(hipStream_t s, int n, int m, const double *coef, const float *Z, int64_t ldz, double *x) {
  multi_axpy32_k<<<grid_for(n), kBlock, 0, s>>>(n, m, coef, Z, ldz, x);
}
void k_multi_axpy(hipStream_t s, int n, int m, const double *coef, const double *Z, int64_t ldz, double *x) {
  if (al16(Z) && al16(x) && ldz % 2 == 0) multi_axpy2_k<<<grid_for((n + 1) / 2), kBlock, 0, s>>>(n, m, coef, Z, ldz, x);
  else multi_axpy_k<<<grid_for(n), kBlock, 0, s>>>(n, m, coef, Z, ldz, x);
}
void k_copy_range(hipStream_t s, int b, int e, const double *src, double *dst) {
  if (e > b) copy_range_k<<<(e - b + kBlock - 1) / kBlock, kBlock, 0, s>>>(b, e, src, dst);
}
void k_dirichlet_rows(hipStream_t s, const DevLevel &L, double *val, double *rhs) {
  const int nd = L.ndof - L.nact;
  if (nd <= 0) return;
  dirichlet_rows_k<<<(nd + kBlock - 1) / kBlock, kBlock, 0, s>>>(L.ndof, L.nact, L.rowptr, val, rhs);
}

// ---------------------------------------------------------------- same-level batches
// blockIdx.y is the problem; each kernel takes its arrays from the tables and
// returns at once for a finished problem (the test is uniform per workgroup).

__device__ inline double *bvec(const BatchLevel &L, const BatchProblem &P, int id) {
  if (id < 4) return L.vec[id];
  if (id == kBXs) return P.xs;
  if (id == kBBs) return P.bs;
  if (id == kBW) return P.w;
  if (id >= kBZ0) return P.Z + (int64_t)(id - kBZ0) * P.ld;
  return P.V + (int64_t)(id - kBV0) * P.ld;
}

// the level vectors of the V-cycle in the batch's working precision: FP64,
// or FP32 for the mixed form (vec32, 1/a_ii in invdiag32); the matrices
// stay FP64 and are rounded to FP32 as they are read
template <typename T>
__device__ inline T *bvec_t(const BatchLevel &L, const BatchProblem &P, int id);
template <>
__device__ inline double *bvec_t<double>(const BatchLevel &L, const BatchProblem &P, int id) {
  return bvec(L, P, id);
}
template <>
__device__ inline float *bvec_t<float>(const BatchLevel &L, const BatchProblem &, int id) {
  return L.vec32[id];
}
template <typename T>
__device__ inline const T *binv_t(const BatchLevel &L);
template <>
__device__ inline const double *binv_t<double>(const BatchLevel &L) {
  return L.invdiag;
}
template <>
__device__ inline const float *binv_t<float>(const BatchLevel &L) {
  return L.invdiag32;
}

template <int MODE, typename T>
__global__ __launch_bounds__(kStreamBlock) void bstream_k(const BatchLevel *__restrict__ lev,
                                                    const BatchProblem *__restrict__ prob,
                                                    const int *__restrict__ active, int xid, int fid, int oid,
                                                    double omega) {
  const int b = blockIdx.y;
  if (!active[b]) return;
  const BatchLevel &L = lev[b];
  if ((int)blockIdx.x >= L.nblk) return;
  const BatchProblem &P = prob[b];
  csr_stream_block<MODE, T, double, false>(blockIdx.x, L.bt, L.nact, L.rowptr, L.col, L.val, bvec_t<T>(L, P, xid),
                                           fid >= 0 ? bvec_t<T>(L, P, fid) : nullptr, binv_t<T>(L),
                                           bvec_t<T>(L, P, oid), (T)omega);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void bjacobi_zero_k(const BatchLevel *__restrict__ lev,
                                                         const BatchProblem *__restrict__ prob,
                                                         const int *__restrict__ active, int fid, int oid,
                                                         double omega) {
  const int b = blockIdx.y;
  if (!active[b]) return;
  const BatchLevel &L = lev[b];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= L.ndof) return;
  const T *f = bvec_t<T>(L, prob[b], fid);
  T *out = bvec_t<T>(L, prob[b], oid);
  out[i] = i < L.nact ? (T)omega * f[i] / (T)L.val[L.diag[i]] : f[i];
}

// restriction onto the coarser level's rows: 8 lanes per row
template <typename T>
__global__ __launch_bounds__(kBlock) void brestrict_k(const BatchLevel *__restrict__ fine,
                                                      const BatchProblem *__restrict__ prob,
                                                      const int *__restrict__ active,
                                                      const BatchLevel *__restrict__ coarse, int xid, int oid) {
  const int b = blockIdx.y;
  if (!active[b]) return;
  const BatchLevel &F = fine[b], &Cl = coarse[b];
  csr_rows_body<8, 1, kRestrict, T, double>(blockIdx.x * kBlock + threadIdx.x, F.ndof_c, F.nact_c, F.r_rowptr,
                                            F.r_col, F.r_val, bvec_t<T>(F, prob[b], xid), nullptr,
                                            bvec_t<T>(Cl, prob[b], oid), T(0));
}

template <typename T>
__global__ __launch_bounds__(kBlock) void bprolong_k(const BatchLevel *__restrict__ fine,
                                                     const BatchProblem *__restrict__ prob,
                                                     const int *__restrict__ active,
                                                     const BatchLevel *__restrict__ coarse, int xid, int oid,
                                                     double damp) {
  const int b = blockIdx.y;
  if (!active[b]) return;
  const BatchLevel &F = fine[b], &Cl = coarse[b];
  csr_rows_body<8, 1, kProlongUpd, T, double>(blockIdx.x * kBlock + threadIdx.x, F.ndof, F.nact, F.p_rowptr,
                                              F.p_col, F.p_val, bvec_t<T>(Cl, prob[b], xid), nullptr,
                                              bvec_t<T>(F, prob[b], oid), (T)damp);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void bcopy_k(const BatchLevel *__restrict__ lev,
                                                  const BatchProblem *__restrict__ prob,
                                                  const int *__restrict__ active, int xid, int oid,
                                                  int dirichlet_only) {
  const int b = blockIdx.y;
  if (!active[b]) return;
  const BatchLevel &L = lev[b];
  const int i = (dirichlet_only ? L.nact : 0) + blockIdx.x * kBlock + threadIdx.x;
  if (i >= L.ndof) return;
  T *out = bvec_t<T>(L, prob[b], oid);
  out[i] = xid >= 0 ? bvec_t<T>(L, prob[b], xid)[i] : T(0);
}

// coarsest level, accumulated in FP64 in both precisions
template <typename T>
__global__ __launch_bounds__(kBlock) void bgemv_k(const BatchLevel *__restrict__ lev,
                                                  const BatchProblem *__restrict__ prob,
                                                  const int *__restrict__ active, int did, int oid) {
  const int b = blockIdx.y;
  if (!active[b]) return;
  const BatchProblem &P = prob[b];
  const int na = P.na0, row = (blockIdx.x * kBlock + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (row >= na) return;
  const T *d = bvec_t<T>(lev[b], P, did);
  double s = 0;
  for (int j = lane; j < na; j += 64) s += P.cinv[(int64_t)row * na + j] * (double)d[j];
  s = group_sum<64>(s);
  if (lane == 0) bvec_t<T>(lev[b], P, oid)[row] += (T)s;
}

// the mixed form's precision boundary: FP64 vector sid -> FP32 level vector
// did (to32) or back (from32), and 1/a_ii rounded to FP32
__global__ __launch_bounds__(kBlock) void bto32_k(const BatchLevel *__restrict__ lev,
                                                  const BatchProblem *__restrict__ prob,
                                                  const int *__restrict__ active, int sid, int did) {
  const int b = blockIdx.y;
  if (!active[b]) return;
  const BatchLevel &L = lev[b];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < L.ndof) L.vec32[did][i] = (float)bvec(L, prob[b], sid)[i];
}
__global__ __launch_bounds__(kBlock) void bfrom32_k(const BatchLevel *__restrict__ lev,
                                                    const BatchProblem *__restrict__ prob,
                                                    const int *__restrict__ active, int sid, int did) {
  const int b = blockIdx.y;
  if (!active[b]) return;
  const BatchLevel &L = lev[b];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < L.ndof) bvec(L, prob[b], did)[i] = (double)L.vec32[sid][i];
}
__global__ __launch_bounds__(kBlock) void binv32_k(const BatchLevel *__restrict__ lev, int B) {
  const int b = blockIdx.y;
  if (b >= B) return;
  const BatchLevel &L = lev[b];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < L.ndof) L.invdiag32[i] = (float)L.invdiag[i];
}

__device__ inline double bsum_partials(const double *__restrict__ p, double *sh) {
  double v = threadIdx.x < kBatchRed ? p[threadIdx.x] : 0.0;
  v = block_sum(v, sh);
  __syncthreads();
  if (threadIdx.x == 0) sh[16] = v;
  __syncthreads();
  return sh[16];
}

__global__ __launch_bounds__(kBlock) void bdot_k(const BatchLevel *__restrict__ fine,
                                                 const BatchProblem *__restrict__ prob,
                                                 const int *__restrict__ active, int xid, int yid, int slot) {
  __shared__ double sh[17];
  const int b = blockIdx.y;
  if (!active[b]) return;
  const BatchProblem &P = prob[b];
  const double *x = bvec(fine[b], P, xid), *y = bvec(fine[b], P, yid);
  double v = 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < P.n; i += kBatchRed * kBlock) v = fma(x[i], y[i], v);
  v = block_sum(v, sh);
  if (threadIdx.x == 0) P.red[slot * kBatchRed + blockIdx.x] = v;
}

// scal[idx] = sum of the partials (root: its square root)
__global__ __launch_bounds__(kBlock) void bfinish_k(const BatchProblem *__restrict__ prob,
                                                    const int *__restrict__ active, int slot, int idx, int root) {
  __shared__ double sh[17];
  const int b = blockIdx.y;
  if (!active[b]) return;
  const BatchProblem &P = prob[b];
  const double v = bsum_partials(P.red + slot * kBatchRed, sh);
  if (threadIdx.x == 0) P.scal[idx] = root ? sqrt(v) : v;
}

// y += sign alpha x with alpha the sum of the partials in slot_in (stored at
// scal[alpha_idx]), then partials of <y, z> (z < 0: <y, y>) into slot_out
__global__ __launch_bounds__(kBlock) void baxpy_dot_k(const BatchLevel *__restrict__ fine,
                                                      const BatchProblem *__restrict__ prob,
                                                      const int *__restrict__ active, int slot_in, int alpha_idx,
                                                      double sign, int xid, int yid, int zid, int slot_out) {
  __shared__ double sh[17];
  const int b = blockIdx.y;
  if (!active[b]) return;
  const BatchProblem &P = prob[b];
  const double alpha = bsum_partials(P.red + slot_in * kBatchRed, sh);
  if (blockIdx.x == 0 && threadIdx.x == 0) P.scal[alpha_idx] = alpha;
  const double a = sign * alpha;
  const double *x = bvec(fine[b], P, xid), *z = zid >= 0 ? bvec(fine[b], P, zid) : nullptr;
  double *y = bvec(fine[b], P, yid);
  double v = 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < P.n; i += kBatchRed * kBlock) {
    const double yi = fma(a, x[i], y[i]);
    y[i] = yi;
    v = fma(yi, z ? z[i] : yi, v);
  }
  v = block_sum(v, sh);
  if (threadIdx.x == 0) P.red[slot_out * kBatchRed + blockIdx.x] = v;
}

// y = x / sqrt(sum of the partials in slot_in); the norm to scal[norm_idx]
__global__ __launch_bounds__(kBlock) void bscale_copy_k(const BatchLevel *__restrict__ fine,
                                                        const BatchProblem *__restrict__ prob,
                                                        const int *__restrict__ active, int slot_in, int norm_idx,
                                                        int xid, int yid) {
  __shared__ double sh[17];
  const int b = blockIdx.y;
  if (!active[b]) return;
  const BatchProblem &P = prob[b];
  const double nrm = sqrt(bsum_partials(P.red + slot_in * kBatchRed, sh));
  if (blockIdx.x == 0 && threadIdx.x == 0) P.scal[norm_idx] = nrm;
  const double sc = nrm != 0.0 ? 1.0 / nrm : 1.0;
  const double *x = bvec(fine[b], P, xid);
  double *y = bvec(fine[b], P, yid);
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < P.n; i += kBatchRed * kBlock) y[i] = x[i] * sc;
}

// the low-synchronisation Gram-Schmidt step of the single-problem kernels
// above for every active problem of a batch (blockIdx.y = problem): dots over
// the problem's own ceil(n / kLsChunk) chunks into P.ls_part, the MGS column
// h into P.scal[0 .. i] (row i of L into P.ls_L), then w -= V h with the
// partial sums of |w|^2 in reduction slot `slot`
__global__ __launch_bounds__(kBlock) void bls_dots_k(const BatchProblem *__restrict__ prob,
                                                     const int *__restrict__ active, int i) {
  __shared__ double ws[kLsChunk], vs[kLsChunk];
  const int b = blockIdx.y;
  if (!active[b]) return;
  const BatchProblem &P = prob[b];
  const int nb = (P.n + kLsChunk - 1) / kLsChunk;
  if ((int)blockIdx.x >= nb) return;
  ls_dots_chunk(P.n, i, P.V, P.ld, P.w, P.ls_part, blockIdx.x, nb, ws, vs);
}
__global__ __launch_bounds__(kLsCoef) void bls_coeffs_k(const BatchProblem *__restrict__ prob,
                                                        const int *__restrict__ active, int i) {
  __shared__ double d[2 * kLsMax + 1], Ls[kLsMax * kLsMax];
  const int b = blockIdx.y;
  if (!active[b]) return;
  const BatchProblem &P = prob[b];
  ls_coeffs_body(i, (P.n + kLsChunk - 1) / kLsChunk, P.ls_part, P.ls_L, P.scal, d, Ls);
}
__global__ __launch_bounds__(kBlock) void bls_update_k(const BatchProblem *__restrict__ prob,
                                                       const int *__restrict__ active, int i, int slot) {
  __shared__ double sh[17], hs[kLsMax];
  const int b = blockIdx.y;
  if (!active[b]) return;
  const BatchProblem &P = prob[b];
  ls_update_part(P.n, i, P.scal, P.V, P.ld, P.w, P.red + slot * kBatchRed, blockIdx.x, kBatchRed, sh, hs);
}

__global__ __launch_bounds__(kBlock) void bscale_k(const BatchLevel *__restrict__ fine,
                                                   const BatchProblem *__restrict__ prob,
                                                   const int *__restrict__ active, int xid) {
  const int b = blockIdx.y;
  const BatchProblem &P = prob[b];
  if (!active[b] || P.coef[0] == 0.0) return;
  const double a = P.coef[1];
  double *x = bvec(fine[b], P, xid);
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < P.n; i += kBatchRed * kBlock) x[i] *= a;
}

__global__ __launch_bounds__(kBlock) void bupdate_k(const BatchProblem *__restrict__ prob) {
  const int b = blockIdx.y;
  const BatchProblem &P = prob[b];
  const int m = (int)P.coef[0];
  if (m <= 0) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < P.n; i += gridDim.x * kBlock) {
    double t = 0;
    for (int k = 0; k < m; k++) t += P.coef[1 + k] * P.Z[(int64_t)k * P.ld + i];
    P.xs[i] += t;
  }
}

template <typename T>
static void kb_stream_t(hipStream_t s, int mode, int B, int maxblk, const BatchLevel *lev, const BatchProblem *prob,
                        const int *active, int x, int f, int out, double omega) {
  const dim3 grid(maxblk, B);
  if (mode == kSpmv) bstream_k<kSpmv, T><<<grid, kStreamBlock, 0, s>>>(lev, prob, active, x, f, out, omega);
  else if (mode == kDefect) bstream_k<kDefect, T><<<grid, kStreamBlock, 0, s>>>(lev, prob, active, x, f, out, omega);
  else bstream_k<kJacobi, T><<<grid, kStreamBlock, 0, s>>>(lev, prob, active, x, f, out, omega);
}
void kb_stream(hipStream_t s, int mode, int B, int maxblk, const BatchLevel *lev, const BatchProblem *prob,
               const int *active, int x, int f, int out, double omega, bool f32) {
  if (maxblk <= 0 || B <= 0) return;
  if (f32) kb_stream_t<float>(s, mode, B, maxblk, lev, prob, active, x, f, out, omega);
  else kb_stream_t<double>(s, mode, B, maxblk, lev, prob, active, x, f, out, omega);
}
void kb_jacobi_zero(hipStream_t s, int B, int maxn, const BatchLevel *lev, const BatchProblem *prob,
                    const int *active, int f, int out, double omega, bool f32) {
  if (maxn <= 0) return;
  const dim3 grid((maxn + kBlock - 1) / kBlock, B);
  if (f32) bjacobi_zero_k<float><<<grid, kBlock, 0, s>>>(lev, prob, active, f, out, omega);
  else bjacobi_zero_k<double><<<grid, kBlock, 0, s>>>(lev, prob, active, f, out, omega);
}
void kb_restrict(hipStream_t s, int B, int maxn_c, const BatchLevel *fine, const BatchProblem *prob,
                 const int *active, int x, int out_coarse, const BatchLevel *coarse, bool f32) {
  if (maxn_c <= 0) return;
  const dim3 grid((maxn_c * 8 + kBlock - 1) / kBlock, B);
  if (f32) brestrict_k<float><<<grid, kBlock, 0, s>>>(fine, prob, active, coarse, x, out_coarse);
  else brestrict_k<double><<<grid, kBlock, 0, s>>>(fine, prob, active, coarse, x, out_coarse);
}
void kb_prolong_update(hipStream_t s, int B, int maxn, const BatchLevel *fine, const BatchProblem *prob,
                       const int *active, const BatchLevel *coarse, int x_coarse, int out, double damp, bool f32) {
  if (maxn <= 0) return;
  const dim3 grid((maxn * 8 + kBlock - 1) / kBlock, B);
  if (f32) bprolong_k<float><<<grid, kBlock, 0, s>>>(fine, prob, active, coarse, x_coarse, out, damp);
  else bprolong_k<double><<<grid, kBlock, 0, s>>>(fine, prob, active, coarse, x_coarse, out, damp);
}
void kb_copy(hipStream_t s, int B, int maxn, const BatchLevel *lev, const BatchProblem *prob, const int *active,
             int x, int out, bool dirichlet_only, bool f32) {
  if (maxn <= 0) return;
  const dim3 grid((maxn + kBlock - 1) / kBlock, B);
  if (f32) bcopy_k<float><<<grid, kBlock, 0, s>>>(lev, prob, active, x, out, dirichlet_only ? 1 : 0);
  else bcopy_k<double><<<grid, kBlock, 0, s>>>(lev, prob, active, x, out, dirichlet_only ? 1 : 0);
}
void kb_fill_zero(hipStream_t s, int B, int maxn, const BatchLevel *lev, const BatchProblem *prob, const int *active,
                  int out, bool f32) {
  kb_copy(s, B, maxn, lev, prob, active, -1, out, false, f32);
}
void kb_gemv_coarse(hipStream_t s, int B, int maxna, const BatchLevel *lev, const BatchProblem *prob,
                    const int *active, int d, int out, bool f32) {
  if (maxna <= 0) return;
  const dim3 grid((maxna * 64 + kBlock - 1) / kBlock, B);
  if (f32) bgemv_k<float><<<grid, kBlock, 0, s>>>(lev, prob, active, d, out);
  else bgemv_k<double><<<grid, kBlock, 0, s>>>(lev, prob, active, d, out);
}
void kb_to32(hipStream_t s, int B, int maxn, const BatchLevel *lev, const BatchProblem *prob, const int *active,
             int x, int out32) {
  if (maxn > 0) bto32_k<<<dim3((maxn + kBlock - 1) / kBlock, B), kBlock, 0, s>>>(lev, prob, active, x, out32);
}
void kb_from32(hipStream_t s, int B, int maxn, const BatchLevel *lev, const BatchProblem *prob, const int *active,
               int x32, int out) {
  if (maxn > 0) bfrom32_k<<<dim3((maxn + kBlock - 1) / kBlock, B), kBlock, 0, s>>>(lev, prob, active, x32, out);
}
void kb_invdiag32(hipStream_t s, int B, int maxn, const BatchLevel *lev) {
  if (maxn > 0 && B > 0) binv32_k<<<dim3((maxn + kBlock - 1) / kBlock, B), kBlock, 0, s>>>(lev, B);
}
void kb_dot(hipStream_t s, int B, const BatchLevel *fine, const BatchProblem *prob, const int *active, int x, int y,
            int slot) {
  bdot_k<<<dim3(kBatchRed, B), kBlock, 0, s>>>(fine, prob, active, x, y, slot);
}
void kb_finish(hipStream_t s, int B, const BatchProblem *prob, const int *active, int slot, int idx, bool root) {
  bfinish_k<<<dim3(1, B), kBlock, 0, s>>>(prob, active, slot, idx, root ? 1 : 0);
}
void kb_axpy_dot(hipStream_t s, int B, const BatchLevel *fine, const BatchProblem *prob, const int *active,
                 int slot_in, int alpha_idx, double sign, int x, int y, int z, int slot_out) {
  baxpy_dot_k<<<dim3(kBatchRed, B), kBlock, 0, s>>>(fine, prob, active, slot_in, alpha_idx, sign, x, y, z, slot_out);
}
void kb_ls_step(hipStream_t s, int B, int maxn, const BatchProblem *prob, const int *active, int i, int slot) {
  bls_dots_k<<<dim3(k_ls_blocks(maxn), B), kBlock, 0, s>>>(prob, active, i);
  bls_coeffs_k<<<dim3(1, B), kLsCoef, 0, s>>>(prob, active, i);
  bls_update_k<<<dim3(kBatchRed, B), kBlock, 0, s>>>(prob, active, i, slot);
}
void kb_scale_copy(hipStream_t s, int B, const BatchLevel *fine, const BatchProblem *prob, const int *active,
                   int slot_in, int norm_idx, int x, int y) {
  bscale_copy_k<<<dim3(kBatchRed, B), kBlock, 0, s>>>(fine, prob, active, slot_in, norm_idx, x, y);
}
void kb_scale(hipStream_t s, int B, const BatchLevel *fine, const BatchProblem *prob, const int *active, int x) {
  bscale_k<<<dim3(kBatchRed, B), kBlock, 0, s>>>(fine, prob, active, x);
}
void kb_update(hipStream_t s, int B, const BatchLevel *fine, const BatchProblem *prob) {
  (void)fine;
  bupdate_k<<<dim3(kBatchRed, B), kBlock, 0, s>>>(prob);
}

}  // namespace cd3d

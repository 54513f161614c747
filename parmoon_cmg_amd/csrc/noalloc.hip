// y = A x on a level with host vectors, staged through a device array the
// library holds for the life of the process: cd3d_matvec without any device
// allocation, for callers that cannot allocate (tight serving loops,
// harnesses that forbid hipMalloc).
//
// `scratch` holds x at offset 0 and y at offset ld (N rounded up to 32
// doubles), so a level may have at most kScratchDoubles / 2 unknowns.  Its
// device address is looked up once per device (hipGetSymbolAddress, with that
// device current) and only that address is handed to kernels and copies.
// Every call on a device uses the same array: calls are serialised by a
// process-wide mutex, so concurrent callers in one process wait for each
// other; nothing in the array needs clearing between calls (x is copied in
// whole and every row of y is written).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>

#include "../../include/cd3d.h"

// capi.cpp: y = A x on `level` staged through the device buffer `base` of
// `capacity` doubles (x at 0, y at N rounded up to 32; CD3D_ERR_ARG when the
// level does not fit); the handle's device must be current
cd3d_status cd3d_internal_matvec_staged(cd3d_problem *p, int level, const double *x_host, double *y_host,
                                        double *base, size_t capacity);
int cd3d_internal_device(const cd3d_problem *p);

namespace {

constexpr size_t kScratchDoubles = size_t(1) << 23;  // 64 MiB per device
constexpr int kMaxDevices = 64;

std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

// device address of `scratch` on device dev (current); g_mu held
static hipError_t scratch_address(int dev, double **out) {
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = (double *)ptr;
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

extern "C" cd3d_status cd3d_matvec_noalloc(cd3d_problem *p, int level, const double *x_host, double *y_host) {
  if (!p || !x_host || !y_host) return CD3D_ERR_ARG;
  const int dev = cd3d_internal_device(p);
  if (dev < 0 || dev >= kMaxDevices) return CD3D_ERR_ARG;
  std::lock_guard<std::mutex> g(g_mu);
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) return CD3D_ERR_HIP;
  if (prev != dev && hipSetDevice(dev) != hipSuccess) return CD3D_ERR_HIP;
  double *base = nullptr;
  cd3d_status st = CD3D_ERR_HIP;
  const hipError_t e = scratch_address(dev, &base);
  if (e == hipSuccess) {
    st = cd3d_internal_matvec_staged(p, level, x_host, y_host, base, kScratchDoubles);
  } else {
    fprintf(stderr, "cd3d_matvec_noalloc: hipGetSymbolAddress failed: %s\n", hipGetErrorString(e));
  }
  if (prev != dev) (void)hipSetDevice(prev);
  return st;
}

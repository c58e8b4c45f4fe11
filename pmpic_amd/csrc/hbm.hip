// HBM bandwidth probe: a STREAM-style triad a = b + s*c over three device arrays, timed with HIP
// events.  It is the practical ceiling the parcel kernels' achieved bandwidth is compared with
// (beside the runtime's device-to-device copy); it has no counterpart in the reference.
#include "common.h"

#include "../../include/pmpic_hip.h"

namespace pmpic {
namespace {

// 16-byte elements, a grid-stride loop over the whole arrays, several loads in flight per
// thread; the launch covers every CU many times over
__global__ void __launch_bounds__(256) k_triad(double2 *__restrict__ a,
                                               const double2 *__restrict__ b,
                                               const double2 *__restrict__ c, double s,
                                               int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    double2 x[4], y[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      x[u] = b[i + u * stride];
      y[u] = c[i + u * stride];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) a[i + u * stride] = make_double2(x[u].x + s * y[u].x, x[u].y + s * y[u].y);
  }
  for (; i < n; i += stride) {
    const double2 x = b[i], y = c[i];
    a[i] = make_double2(x.x + s * y.x, x.y + s * y.y);
  }
}

__global__ void k_fill2(double2 *a, int64_t n, double v) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    a[i] = make_double2(v, v);
}

}  // namespace
}  // namespace pmpic

using namespace pmpic;

extern "C" int pmpic_hbm_triad(int32_t device, int64_t bytes_per_array, int32_t reps,
                               double *best_ms, double *gbps) {
  if (bytes_per_array < 16 || reps < 1 || !best_ms || !gbps) return PMPIC_ERR_ARG;
  PMPIC_CHECK(hipSetDevice(device));
  const int64_t n = bytes_per_array / 16;
  double2 *a = nullptr, *b = nullptr, *c = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = PMPIC_OK;
  if (hipMalloc((void **)&a, n * 16) != hipSuccess || hipMalloc((void **)&b, n * 16) != hipSuccess ||
      hipMalloc((void **)&c, n * 16) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess)
    st = PMPIC_ERR_HIP;
  float best = 0.f;
  if (st == PMPIC_OK) {
    const dim3 tb(256), gb(256 * 32);
    hipLaunchKernelGGL(k_fill2, gb, tb, 0, 0, a, n, 0.0);
    hipLaunchKernelGGL(k_fill2, gb, tb, 0, 0, b, n, 1.0);
    hipLaunchKernelGGL(k_fill2, gb, tb, 0, 0, c, n, 2.0);
    for (int r = 0; r <= reps && st == PMPIC_OK; r++) {  // the first launch warms up
      float ms = 0.f;
      if (hipEventRecord(e0, 0) != hipSuccess) st = PMPIC_ERR_HIP;
      hipLaunchKernelGGL(k_triad, gb, tb, 0, 0, a, b, c, 0.5, n);
      if (st == PMPIC_OK && (hipEventRecord(e1, 0) != hipSuccess ||
                             hipEventSynchronize(e1) != hipSuccess ||
                             hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        st = PMPIC_ERR_HIP;
      if (r > 0 && (best == 0.f || ms < best)) best = ms;
    }
    if (st == PMPIC_OK && hipGetLastError() != hipSuccess) st = PMPIC_ERR_HIP;
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipFree(a);
  hipFree(b);
  hipFree(c);
  if (st != PMPIC_OK) return st;
  *best_ms = best;
  *gbps = 3.0 * (double)n * 16.0 / ((double)best * 1e-3) / 1e9;
  return PMPIC_OK;
}

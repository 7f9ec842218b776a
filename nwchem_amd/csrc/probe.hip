// probe.hip -- calibration kernel for HBM-byte counters, built as its own
// measurement library (libtce_t_probe.so, loaded by profiling.py), not part
// of libtce_t.so.  A plain grid-stride copy of n doubles moves exactly 8n
// bytes in and 8n bytes out; profiling it with the same FETCH_SIZE /
// WRITE_SIZE counters as the (T) kernel gives the factor that turns counter
// readings into bytes.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace {
__global__ void __launch_bounds__(256) stream_copy(const double *__restrict__ src,
                                                   double *__restrict__ dst, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace

extern "C" int probe_stream_copy(long n, int repeats) {
  if (n <= 0 || repeats <= 0) return TCE_T_ERR_ARG;
  double *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * sizeof(double)) != hipSuccess) return TCE_T_ERR_NOMEM;
  if (hipMalloc(&b, n * sizeof(double)) != hipSuccess) {
    (void)hipFree(a);
    return TCE_T_ERR_NOMEM;
  }
  int rc = TCE_T_OK;
  if (hipMemset(a, 0, n * sizeof(double)) != hipSuccess) rc = TCE_T_ERR_HIP;
  int cus = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0);
  for (int r = 0; r < repeats && rc == TCE_T_OK; ++r) {
    hipLaunchKernelGGL(stream_copy, dim3(std::max(1, cus) * 8), dim3(256), 0, 0, a, b, n);
    if (hipGetLastError() != hipSuccess) rc = TCE_T_ERR_HIP;
  }
  if (hipDeviceSynchronize() != hipSuccess) rc = TCE_T_ERR_HIP;
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

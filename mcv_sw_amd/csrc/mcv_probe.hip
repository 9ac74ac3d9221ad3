// Plain device-to-device copy kernel of known size: the yardstick for the
// HBM traffic counters (FETCH_SIZE / WRITE_SIZE) and for the bandwidth a
// simple streaming kernel reaches on this GPU.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mcv_hip.h"

namespace {

__global__ __launch_bounds__(256) void copy_probe_kernel(const double2 *__restrict__ src,
                                                         double2 *__restrict__ dst, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace

extern "C" int mcv_hip_copy_probe(int32_t device, int64_t bytes, int32_t reps,
                                  double *ms_per_copy) {
  if (bytes < 16 || reps < 1 || !ms_per_copy) return MCV_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return MCV_EDEVICE;
  const long long n = bytes / 16;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * 16) != hipSuccess) return MCV_ENOMEM;
  if (hipMalloc(&b, n * 16) != hipSuccess) {
    (void)hipFree(a);
    return MCV_ENOMEM;
  }
  int rc = MCV_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const dim3 grid(8192), block(256);
  if (hipMemset(a, 0, n * 16) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = MCV_EDEVICE;
  } else {
    hipLaunchKernelGGL(copy_probe_kernel, grid, block, 0, 0, (const double2 *)a, (double2 *)b, n);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(copy_probe_kernel, grid, block, 0, 0, (const double2 *)a, (double2 *)b,
                         n);
    (void)hipEventRecord(e1, 0);
    float ms = 0.f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = MCV_EDEVICE;
    *ms_per_copy = ms / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

// Device-to-device copy kernels of known size: the plain copy is the
// yardstick for the HBM traffic counters (FETCH_SIZE / WRITE_SIZE); the
// streaming variants measure the bandwidth a copy reaches on this GPU with
// several loads in flight per thread and non-temporal accesses.
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

typedef double d2v __attribute__((vector_size(16)));

// Streaming copy: U independent 16-byte loads in flight per thread per round,
// then U stores, both marked non-temporal (streamed past the caches); a grid
// of a few workgroups per CU walks the array in rounds of gridDim*blockDim*U
// elements.
template <int U>
__global__ __launch_bounds__(256) void stream_probe_kernel(const d2v *__restrict__ src,
                                                           d2v *__restrict__ dst, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < n; i += U * stride) {
    d2v r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) r[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < U; ++u) __builtin_nontemporal_store(r[u], dst + i + u * stride);
  }
  for (; i < n; i += stride) dst[i] = src[i];
}

// Launches copy variant v: 0 the plain copy on 8192 x 256 threads; 1..3 the
// streaming copy with U = 2, 4, 8 on `per_cu` workgroups per CU.
void launch_variant(int v, int cus, int per_cu, const void *a, void *b, long long n) {
  const dim3 block(256);
  const dim3 grid(v == 0 ? 8192 : cus * per_cu);
  switch (v) {
    case 0:
      hipLaunchKernelGGL(copy_probe_kernel, grid, block, 0, 0, (const double2 *)a, (double2 *)b,
                         n);
      break;
    case 1:
      hipLaunchKernelGGL(stream_probe_kernel<2>, grid, block, 0, 0, (const d2v *)a, (d2v *)b, n);
      break;
    case 2:
      hipLaunchKernelGGL(stream_probe_kernel<4>, grid, block, 0, 0, (const d2v *)a, (d2v *)b, n);
      break;
    default:
      hipLaunchKernelGGL(stream_probe_kernel<8>, grid, block, 0, 0, (const d2v *)a, (d2v *)b, n);
      break;
  }
}

// Mean milliseconds per copy of variant v, reps launches after one warm-up.
int time_variant(int v, int cus, int per_cu, const void *a, void *b, long long n, int reps,
                 hipEvent_t e0, hipEvent_t e1, double *ms_out) {
  launch_variant(v, cus, per_cu, a, b, n);
  if (hipGetLastError() != hipSuccess) return MCV_EDEVICE;
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < reps; ++r) launch_variant(v, cus, per_cu, a, b, n);
  (void)hipEventRecord(e1, 0);
  float ms = 0.f;
  if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
    return MCV_EDEVICE;
  *ms_out = ms / reps;
  return MCV_OK;
}

}  // namespace

// Allocates two buffers of `bytes`, runs `fn(a, b, n, e0, e1)` and releases
// everything.
template <typename F>
int with_buffers(int32_t device, int64_t bytes, F fn) {
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
  if (hipMemset(a, 0, n * 16) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess)
    rc = MCV_EDEVICE;
  else
    rc = fn(a, b, n, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

extern "C" int mcv_hip_copy_probe(int32_t device, int64_t bytes, int32_t reps,
                                  double *ms_per_copy) {
  if (bytes < 16 || reps < 1 || !ms_per_copy) return MCV_EINVAL;
  return with_buffers(device, bytes, [&](void *a, void *b, long long n, hipEvent_t e0,
                                         hipEvent_t e1) {
    return time_variant(0, 0, 0, a, b, n, reps, e0, e1, ms_per_copy);
  });
}

extern "C" int mcv_hip_stream_probe(int32_t device, int64_t bytes, int32_t reps, double *ms,
                                    int32_t *variant, int32_t *per_cu) {
  if (bytes < 16 || reps < 1 || !ms) return MCV_EINVAL;
  int cus = 0;
  if (hipSetDevice(device) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return MCV_EDEVICE;
  return with_buffers(device, bytes, [&](void *a, void *b, long long n, hipEvent_t e0,
                                         hipEvent_t e1) {
    double best = 1e30;
    int bv = 0, bp = 0;
    static const int kPerCu[3] = {2, 4, 8};
    for (int v = 1; v <= 3; ++v)
      for (int p : kPerCu) {
        double t = 0.0;
        int rc = time_variant(v, cus, p, a, b, n, reps, e0, e1, &t);
        if (rc) return rc;
        if (t < best) { best = t; bv = v; bp = p; }
      }
    *ms = best;
    if (variant) *variant = bv;
    if (per_cu) *per_cu = bp;
    return (int)MCV_OK;
  });
}

// Calibration kernel: a plain device-to-device copy of known size.  Its measured time gives
// the achievable streaming bandwidth, and its rocprofv3 FETCH_SIZE / WRITE_SIZE readings
// scale the byte counters of the step kernels.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/pyfds_amd.h"

namespace {

__global__ __launch_bounds__(256) void k_stream_copy(const float4 *__restrict__ src,
                                                     float4 *__restrict__ dst, int64_t n4) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
    dst[i] = src[i];
}

}  // namespace

extern "C" int fds_stream_copy_bench(int64_t bytes, int reps, double *gbs, double *ms_each) {
  if (bytes < 16 || reps < 1 || !gbs || !ms_each) return FDS_E_ARG;
  const int64_t n4 = bytes / 16;
  void *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = FDS_E_HIP;
  if (hipMalloc(&a, n4 * 16) != hipSuccess || hipMalloc(&b, n4 * 16) != hipSuccess) goto done;
  if (hipMemset(a, 0, n4 * 16) != hipSuccess) goto done;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) goto done;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto done;
  {
    const int grid = 256 * 16;
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, s, (const float4 *)a,
                       (float4 *)b, n4);
    if (hipEventRecord(e0, s) != hipSuccess) goto done;
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, s, (const float4 *)a,
                         (float4 *)b, n4);
    if (hipEventRecord(e1, s) != hipSuccess) goto done;
    if (hipStreamSynchronize(s) != hipSuccess) goto done;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) goto done;
    *ms_each = ms / reps;
    *gbs = 2.0 * (double)(n4 * 16) / (*ms_each * 1e-3) / 1e9;
    rc = 0;
  }
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

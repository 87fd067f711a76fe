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

typedef float v4f __attribute__((ext_vector_type(4)));

// Four independent 16-byte loads in flight per thread before the stores, non-temporal stores.
__global__ __launch_bounds__(256) void k_stream_copy4(const v4f *__restrict__ src,
                                                      v4f *__restrict__ dst, int64_t n4) {
  const int64_t base = ((int64_t)blockIdx.x * 256 * 4) + threadIdx.x;
  v4f v[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int64_t i = base + u * 256;
    if (i < n4) v[u] = __builtin_nontemporal_load(src + i);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int64_t i = base + u * 256;
    if (i < n4) __builtin_nontemporal_store(v[u], dst + i);
  }
}

}  // namespace

extern "C" int fds_stream_copy_bench(int64_t bytes, int reps, double *gbs, double *ms_each) {
  if (bytes < 16 || reps == 0 || !gbs || !ms_each) return FDS_E_ARG;
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
    const bool unrolled = reps < 0;
    const int nrep = unrolled ? -reps : reps;
    const int grid4 = (int)((n4 + 1023) / 1024);
    auto launch = [&]() {
      if (unrolled)
        hipLaunchKernelGGL(k_stream_copy4, dim3(grid4), dim3(256), 0, s, (const v4f *)a,
                           (v4f *)b, n4);
      else
        hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, s, (const float4 *)a,
                           (float4 *)b, n4);
    };
    launch();
    if (hipEventRecord(e0, s) != hipSuccess) goto done;
    for (int r = 0; r < nrep; ++r) launch();
    reps = nrep;
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

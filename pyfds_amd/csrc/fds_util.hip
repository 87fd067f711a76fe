// Calibration kernels: plain device-to-device copies of known size.  Their measured time gives
// the achievable streaming bandwidth, and their rocprofv3 FETCH_SIZE / WRITE_SIZE readings
// scale the byte counters of the step kernels.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/pyfds_amd.h"

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

// Variant 0: grid-stride loop, one 16-byte load and store per iteration.
__global__ __launch_bounds__(256) void k_stream_copy(const v4f *__restrict__ src,
                                                     v4f *__restrict__ dst, int64_t n4) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
    dst[i] = src[i];
}

// Variants 1-4: one pass, a workgroup copies U * 256 16-byte elements; all U loads of a thread
// are issued before the first store (element u*256 + t of the workgroup's block, so each
// load and store instruction is coalesced over the wave).  NT: non-temporal loads and stores.
template <int U, bool NT>
__global__ __launch_bounds__(256) void k_stream_copy_u(const v4f *__restrict__ src,
                                                       v4f *__restrict__ dst, int64_t n4) {
  const int64_t base = (int64_t)blockIdx.x * (256 * U) + threadIdx.x;
  v4f v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t i = base + u * 256;
    if (i < n4) v[u] = NT ? __builtin_nontemporal_load(src + i) : src[i];
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t i = base + u * 256;
    if (i < n4) {
      if (NT)
        __builtin_nontemporal_store(v[u], dst + i);
      else
        dst[i] = v[u];
    }
  }
}

constexpr int kCopyVariants = 5;

void launch_copy(int variant, const v4f *a, v4f *b, int64_t n4, hipStream_t s) {
  auto one_pass = [&](int u) { return dim3((unsigned)((n4 + 256 * u - 1) / (256 * u))); };
  switch (variant) {
    case 0:
      hipLaunchKernelGGL(k_stream_copy, dim3(256 * 16), dim3(256), 0, s, a, b, n4);
      break;
    case 1:
      hipLaunchKernelGGL((k_stream_copy_u<8, false>), one_pass(8), dim3(256), 0, s, a, b, n4);
      break;
    case 2:
      hipLaunchKernelGGL((k_stream_copy_u<8, true>), one_pass(8), dim3(256), 0, s, a, b, n4);
      break;
    case 3:
      hipLaunchKernelGGL((k_stream_copy_u<16, true>), one_pass(16), dim3(256), 0, s, a, b, n4);
      break;
    default:
      hipLaunchKernelGGL((k_stream_copy_u<4, true>), one_pass(4), dim3(256), 0, s, a, b, n4);
      break;
  }
}

}  // namespace

extern "C" int fds_stream_copy_bench(int64_t bytes, int reps, int variant, double *gbs,
                                     double *ms_each) {
  if (bytes < 16 || reps < 1 || variant < 0 || variant >= kCopyVariants || !gbs || !ms_each)
    return FDS_E_ARG;
  const int64_t n4 = bytes / 16;
  if ((n4 + 1023) / 1024 > 0x7fffffff) return FDS_E_ARG;   // one-pass grid size
  void *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = FDS_E_HIP;
  float ms = 0.f;
  if (hipMalloc(&a, n4 * 16) != hipSuccess || hipMalloc(&b, n4 * 16) != hipSuccess) goto done;
  if (hipMemset(a, 0, n4 * 16) != hipSuccess) goto done;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) goto done;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto done;
  launch_copy(variant, (const v4f *)a, (v4f *)b, n4, s);   // untimed first copy
  if (hipEventRecord(e0, s) != hipSuccess) goto done;
  for (int r = 0; r < reps; ++r) launch_copy(variant, (const v4f *)a, (v4f *)b, n4, s);
  if (hipEventRecord(e1, s) != hipSuccess) goto done;
  if (hipStreamSynchronize(s) != hipSuccess) goto done;
  if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) goto done;
  *ms_each = ms / reps;
  *gbs = 2.0 * (double)(n4 * 16) / (*ms_each * 1e-3) / 1e9;
  rc = 0;
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

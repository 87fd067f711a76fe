// Calibration: the LDS-ring streaming kernels' memory access pattern without their arithmetic.
// Its time per launch is the floor that the pattern itself sets on those step kernels (what
// they would take if their arithmetic were free), measured at the same tile geometry, tile
// order, ring loads and LDS footprint.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fds_stream2d.h"

namespace fds {
namespace {

template <typename T, int NS, int NC>
struct PatternArgs {
  const T *src[NS + NC];   // NS state arrays, then NC coefficient arrays
  T *dst[NS];
  int nx, ny, G, H, tiles_x, ntiles;
};

// One one-wave workgroup per tile of 64 columns (64 - 2G owned) and H rows, tiles in the
// step kernels' XCD order.  Rows -G-1 .. H+G of the tile stream through the two rings of an
// S-stage step kernel, one row ahead (state rows into the ring of NS arrays, coefficient rows
// into the ring of NC arrays, 16-byte global-to-LDS loads of the aligned window of columns
// -1 .. 64); every row is read back from LDS, and rows 0 .. H-1 store the NS state arrays'
// owned columns.
template <typename T, int NS, int NC, int S>
__global__ __launch_bounds__(64) void k_stream_pattern(PatternArgs<T, NS, NC> a) {
  using RG = RingGeom<T, 64, kRingX4>;
  constexpr int PF = 1, RS = PF + 1, RC = PF + 2 * S - 1, NA = NS + NC;
  __shared__ T ring[ring_elems<T, 64, PF, S, NS, NC, kRingX4>()];
  const int tile = xcd_tile(blockIdx.x, gridDim.x, a.ntiles);
  if (tile < 0) return;
  const int t = threadIdx.x, G = a.G, H = a.H, nx = a.nx, n = a.nx * a.ny;
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const int x0 = tx * (64 - 2 * G) - G, y0 = ty * H;
  const int lead = ((x0 - 1) % RG::E + RG::E) % RG::E;
  const bool owned = t >= G && t < 64 - G && x0 + t < nx;
  T *const cring = ring + RS * NS * RG::CW;
  const lds_ptr<T> lring = ring_lds(&ring[0]), lcring = lring + RS * NS * RG::CW;
  const T *ssrc[NS], *csrc[NC];
#pragma unroll
  for (int q = 0; q < NS; ++q) ssrc[q] = a.src[q];
#pragma unroll
  for (int q = 0; q < NC; ++q) csrc[q] = a.src[NS + q];
  // row rho (ordinal r = rho + G + 1) into state slot r % RS and coefficient slot r % RC
  auto issue = [&](int rho, int r) __attribute__((always_inline)) {
    const int fs = (y0 + rho) * nx + x0 - 1 - lead;
    ring_issue<T, 64, NS, kRingX4>(lring + (r % RS) * NS * RG::CW, ssrc, fs, n);
    ring_issue<T, 64, NC, kRingX4>(lcring + (r % RC) * NC * RG::CW, csrc, fs, n);
  };
  const int rows = H + 2 * G + 2;
  issue(-G - 1, 0);
  T acc = T(0);
  for (int r = 0; r < rows; ++r) {
    const int rho = r - G - 1;
    compiler_fence();
    if (r + 1 < rows) {
      issue(rho + 1, r + 1);
      vm_wait<NA>();
    } else {
      vm_wait<0>();
    }
    compiler_fence();
    const T *sb = ring + (r % RS) * NS * RG::CW + lead + 1 + t;
    const T *cb = cring + (r % RC) * NC * RG::CW + lead + 1 + t;
    T v = T(0);
#pragma unroll
    for (int q = 0; q < NS; ++q) v += sb[q * RG::CW];
#pragma unroll
    for (int q = 0; q < NC; ++q) v += cb[q * RG::CW];
    acc += v;
    if (owned && rho >= 0 && rho < H && y0 + rho < a.ny) {
      const int o = (y0 + rho) * nx + x0 + t;
#pragma unroll
      for (int q = 0; q < NS; ++q) a.dst[q][o] = v + T(q);
    }
  }
  if (acc == T(-1.25e-7)) a.dst[0][0] = acc;   // keeps the LDS reads
}

template <typename T, int NS, int NC, int S>
int run_pattern(int nx, int ny, int G, int H, int reps, double *ms_each, double *bytes) {
  using RG = RingGeom<T, 64, kRingX4>;
  const int64_t n = (int64_t)nx * ny;
  PatternArgs<T, NS, NC> a = {};
  a.nx = nx;
  a.ny = ny;
  a.G = G;
  a.H = H;
  a.tiles_x = (nx + 64 - 2 * G - 1) / (64 - 2 * G);
  a.ntiles = a.tiles_x * ((ny + H - 1) / H);
  void *buf[2 * NS + NC] = {};
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = FDS_E_HIP;
  float ms = 0.f;
  const dim3 grid((unsigned)xcd_grid(a.ntiles));
  for (int k = 0; k < 2 * NS + NC; ++k) {
    if (hipMalloc(&buf[k], n * sizeof(T)) != hipSuccess) goto done;
    if (hipMemset(buf[k], 0, n * sizeof(T)) != hipSuccess) goto done;
  }
  for (int q = 0; q < NS + NC; ++q) a.src[q] = (const T *)buf[q];
  for (int q = 0; q < NS; ++q) a.dst[q] = (T *)buf[NS + NC + q];
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) goto done;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto done;
  hipLaunchKernelGGL((k_stream_pattern<T, NS, NC, S>), grid, dim3(64), 0, s, a);   // untimed
  if (hipGetLastError() != hipSuccess) goto done;
  if (hipEventRecord(e0, s) != hipSuccess) goto done;
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL((k_stream_pattern<T, NS, NC, S>), grid, dim3(64), 0, s, a);
  if (hipEventRecord(e1, s) != hipSuccess) goto done;
  if (hipStreamSynchronize(s) != hipSuccess) goto done;
  if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) goto done;
  *ms_each = ms / reps;
  // requested bytes: every tile loads H + 2G + 2 rows of NS + NC windows of CW elements;
  // every cell is stored once per state array
  *bytes = ((double)a.ntiles * (H + 2 * G + 2) * (NS + NC) * RG::CW + (double)n * NS) *
           sizeof(T);
  rc = 0;
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  for (int k = 0; k < 2 * NS + NC; ++k)
    if (buf[k]) (void)hipFree(buf[k]);
  return rc;
}

template <typename T, int NS, int NC>
int run_pattern_s(int stages, int nx, int ny, int G, int H, int reps, double *ms, double *b) {
  if (stages == 2) return run_pattern<T, NS, NC, 2>(nx, ny, G, H, reps, ms, b);
  return run_pattern<T, NS, NC, 3>(nx, ny, G, H, reps, ms, b);
}

template <typename T>
int run_pattern_t(int stages, int ns, int nx, int ny, int G, int H, int reps, double *ms,
                  double *b) {
  if (ns == 3) return run_pattern_s<T, 3, 2>(stages, nx, ny, G, H, reps, ms, b);
  return run_pattern_s<T, 1, 2>(stages, nx, ny, G, H, reps, ms, b);
}

}  // namespace
}  // namespace fds

extern "C" int fds_stream_pattern_bench(int nx, int ny, int elem, int stages, int ghost,
                                        int strip_h, int n_state, int reps, double *ms_each,
                                        double *bytes) {
  if (!ms_each || !bytes || reps < 1 || (elem != 8 && elem != 4) ||
      (stages != 2 && stages != 3) || (n_state != 1 && n_state != 3) || ghost < 0 ||
      ghost > 16 || strip_h < 1 || nx < 1 || ny < 1 || nx % (16 / elem) != 0 ||
      (int64_t)nx * ny > (int64_t)1 << 28)
    return FDS_E_ARG;
  if (elem == 8)
    return fds::run_pattern_t<double>(stages, n_state, nx, ny, ghost, strip_h, reps, ms_each,
                                      bytes);
  return fds::run_pattern_t<float>(stages, n_state, nx, ny, ghost, strip_h, reps, ms_each,
                                   bytes);
}

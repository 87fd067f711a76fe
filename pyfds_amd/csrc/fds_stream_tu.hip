// Streaming step kernels of one numeric form and one shape of the streaming table.  The build
// compiles this file once per form (FDS_FORM 0: float64, 1: float32, 2: float32 + Kahan) and
// per shape index (FDS_SHAPE), so the shapes compile in parallel.
#include "fds_kernels.h"
#include "fds_stream2d.h"
#include "fds_stream_lossy.h"

#if !defined(FDS_FORM) || !defined(FDS_SHAPE)
#error "compile with -DFDS_FORM=<0|1|2> -DFDS_SHAPE=<index>"
#endif

#if FDS_FORM == 0
#define FDS_FORM_NS f64
#define FDS_T double
#define FDS_KAHAN false
#elif FDS_FORM == 1
#define FDS_FORM_NS f32
#define FDS_T float
#define FDS_KAHAN false
#else
#define FDS_FORM_NS k32
#define FDS_T float
#define FDS_KAHAN true
#endif

#define FDS_CAT2(a, b) a##b
#define FDS_CAT(a, b) FDS_CAT2(a, b)

namespace fds {
namespace FDS_FORM_NS {

using T = FDS_T;
constexpr bool K = FDS_KAHAN;

template <int TX, bool AXI, int S, bool NT, bool ISO>
__global__ __launch_bounds__(TX) void k_stream_acoustic(StepArgs<T> a) {
  __shared__ StreamLds<T, TX, S> L;
  const int tile = xcd_tile(blockIdx.x, gridDim.x, a.ntiles);
  if (tile < 0) return;
  acoustic_stream<T, K, TX, AXI, S, NT, ISO>(a, tile, a.dyn[0] + a.k, a.dyn[1] + a.k, L);
}

template <int TX, bool AXI, int S, bool NT, bool ISO>
__global__ __launch_bounds__(TX) void k_stream_thermal(StepArgs<T> a) {
  __shared__ StreamLds<T, TX, S> L;
  const int tile = xcd_tile(blockIdx.x, gridDim.x, a.ntiles);
  if (tile < 0) return;
  thermal_stream<T, K, TX, AXI, S, NT, ISO>(a, tile, a.dyn[0] + a.k, a.dyn[1] + a.k, L);
}

template <int TX, bool AXI, bool NT, bool ISO>
__global__ __launch_bounds__(TX) void k_stream_lossy(StepArgs<T> a) {
  __shared__ LossyLds<T, TX, AXI> L;
  const int tile = xcd_tile(blockIdx.x, gridDim.x, a.ntiles);
  if (tile < 0) return;
  lossy_stream<T, K, TX, AXI, NT, ISO>(a, tile, a.dyn[0] + a.k, a.dyn[1] + a.k, L);
}

template <int TX, int S, bool NT>
__global__ __launch_bounds__(TX) void k_stream_batch(const StepArgs<T> *__restrict__ descs,
                                                     const int32_t *__restrict__ tile_inst,
                                                     const int32_t *__restrict__ inst_base,
                                                     int ntiles, int k) {
  __shared__ StreamLds<T, TX, S> L;
  const int g = xcd_tile(blockIdx.x, gridDim.x, ntiles);
  if (g < 0) return;
  const int inst = tile_inst[g];
  const StepArgs<T> &a = descs[inst];
  any_stream<T, K, TX, S, NT>(a, g - inst_base[inst], a.dyn[0] + k, a.dyn[1] + k, L);
}

template <int TX, bool AXI, int S, bool NT, bool ISO>
void launch_stream_acoustic(const void *args, int nblocks, size_t, hipStream_t s) {
  hipLaunchKernelGGL((k_stream_acoustic<TX, AXI, S, NT, ISO>), dim3(nblocks), dim3(TX), 0, s,
                     *static_cast<const StepArgs<T> *>(args));
}

template <int TX, bool AXI, int S, bool NT, bool ISO>
void launch_stream_thermal(const void *args, int nblocks, size_t, hipStream_t s) {
  hipLaunchKernelGGL((k_stream_thermal<TX, AXI, S, NT, ISO>), dim3(nblocks), dim3(TX), 0, s,
                     *static_cast<const StepArgs<T> *>(args));
}

template <int TX, bool AXI, bool NT, bool ISO>
void launch_stream_lossy(const void *args, int nblocks, size_t, hipStream_t s) {
  hipLaunchKernelGGL((k_stream_lossy<TX, AXI, NT, ISO>), dim3(nblocks), dim3(TX), 0, s,
                     *static_cast<const StepArgs<T> *>(args));
}

template <int TX, int S, bool NT>
void launch_stream_batch(const BatchLaunch &b, int nblocks, size_t, hipStream_t s) {
  hipLaunchKernelGGL((k_stream_batch<TX, S, NT>), dim3(nblocks), dim3(TX), 0, s,
                     static_cast<const StepArgs<T> *>(b.descs), b.tile_inst, b.inst_base,
                     b.ntiles, b.k);
}

// Variants [v]: 0 plain, 1 axial-symmetric, 2 isotropic coefficients (y arrays = x arrays, not
// axial-symmetric).  Fused [.][1]: the shape's stage count; [.][0]: one stage (remainder steps).
template <int TX, int S, bool NT>
void fill_variants(KernelSet &ks, int I, int f) {
  ks.stream_acoustic[I][0][f] = launch_stream_acoustic<TX, false, S, NT, false>;
  ks.stream_acoustic[I][1][f] = launch_stream_acoustic<TX, true, S, NT, false>;
  ks.stream_acoustic[I][2][f] = launch_stream_acoustic<TX, false, S, NT, true>;
  ks.stream_thermal[I][0][f] = launch_stream_thermal<TX, false, S, NT, false>;
  ks.stream_thermal[I][1][f] = launch_stream_thermal<TX, true, S, NT, false>;
  ks.stream_thermal[I][2][f] = launch_stream_thermal<TX, false, S, NT, true>;
  ks.stream_batch[I][f] = launch_stream_batch<TX, S, NT>;
  if (f == 0) {   // lossy acoustics: one step per launch on every shape
    ks.stream_lossy[I][0] = launch_stream_lossy<TX, false, NT, false>;
    ks.stream_lossy[I][1] = launch_stream_lossy<TX, true, NT, false>;
    ks.stream_lossy[I][2] = launch_stream_lossy<TX, false, NT, true>;
  }
}

void FDS_CAT(fill_stream_, FDS_SHAPE)(KernelSet &ks) {
  constexpr int I = FDS_SHAPE;
  static_assert(I < kNumStream, "shape index out of range");
  constexpr int TX = kStreamShapes[I][0];
  constexpr int S = kStreamShapes[I][2];
  constexpr bool NT = kStreamShapes[I][3] != 0;
  fill_variants<TX, S, NT>(ks, I, 1);
  fill_variants<TX, 1, NT>(ks, I, 0);
}

}  // namespace FDS_FORM_NS
}  // namespace fds

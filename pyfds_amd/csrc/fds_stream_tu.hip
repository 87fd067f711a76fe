// Streaming step kernels of one numeric form and one shape of the streaming table.  The build
// compiles this file once per form (FDS_FORM 0: float64, 1: float32, 2: float32 + Kahan) and
// per shape index (FDS_SHAPE), so the shapes compile in parallel.
#include <algorithm>

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
// PF: rows of input prefetched ahead by stage 0 (the shape's own depth for its full stage
// count; remainder launches of fewer stages use one row), into registers or (RING, one-wave
// workgroups) into LDS rings filled by direct global-to-LDS loads.
template <int TX, bool AXI, int S, bool NT, bool ISO, int PF, int RING>
__global__ __launch_bounds__(TX) void k_stream_acoustic(StepArgs<T> a) {
  static_assert(!RING || TX == 64, "the LDS ring serves one-wave workgroups");
  __shared__ StreamLds<T, TX, S> L;
  __shared__ T ring[RING ? ring_elems<T, TX, PF, S, K ? 6 : 3, ISO ? 2 : 4, RING>() : 1];
  const int tile = xcd_tile(blockIdx.x, gridDim.x, a.ntiles);
  if (tile < 0) return;
  acoustic_stream<T, K, TX, AXI, S, NT, ISO, PF, RING>(a, tile, a.dyn[0] + a.k, a.dyn[1] + a.k,
                                                       L, ring);
}

template <int TX, bool AXI, int S, bool NT, bool ISO, int PF, int RING>
__global__ __launch_bounds__(TX) void k_stream_thermal(StepArgs<T> a) {
  static_assert(!RING || TX == 64, "the LDS ring serves one-wave workgroups");
  __shared__ StreamLds<T, TX, S> L;
  __shared__ T ring[RING ? ring_elems<T, TX, PF, S, K ? 2 : 1, ISO ? 2 : 4, RING>() : 1];
  const int tile = xcd_tile(blockIdx.x, gridDim.x, a.ntiles);
  if (tile < 0) return;
  thermal_stream<T, K, TX, AXI, S, NT, ISO, PF, RING>(a, tile, a.dyn[0] + a.k, a.dyn[1] + a.k,
                                                      L, ring);
}

template <int TX, bool AXI, bool NT, bool ISO, int RING>
__global__ __launch_bounds__(TX) void k_stream_lossy(StepArgs<T> a) {
  static_assert(!RING || TX == 64, "the LDS ring serves one-wave workgroups");
  __shared__ LossyLds<T, TX, AXI> L;
  __shared__ T ring[RING ? lossy_ring_elems<T, TX, ISO, AXI, K>() : 1];
  const int tile = xcd_tile(blockIdx.x, gridDim.x, a.ntiles);
  if (tile < 0) return;
  lossy_stream<T, K, TX, AXI, NT, ISO, RING>(a, tile, a.dyn[0] + a.k, a.dyn[1] + a.k, L, ring);
}

// Batch kernels with RING stream rows into LDS rings in kRingDword mode (members of any
// width); sized for the largest member model (acoustic, separate y coefficients).
template <int TX, int S, int PF, int RING>
constexpr int batch_ring_elems() {
  return RING ? ring_elems<T, TX, PF, S, K ? 6 : 3, 4, RING>() : 1;
}

template <int TX, int S, bool NT, int PF, int RING>
__global__ __launch_bounds__(TX) void k_stream_batch(const StepArgs<T> *__restrict__ descs,
                                                     const int32_t *__restrict__ tile_inst,
                                                     const int32_t *__restrict__ inst_base,
                                                     int ntiles, int k) {
  static_assert(!RING || (TX == 64 && RING == kRingDword), "batch rings: one-wave dword mode");
  __shared__ StreamLds<T, TX, S> L;
  __shared__ T ring[batch_ring_elems<TX, S, PF, RING>()];
  const int g = xcd_tile(blockIdx.x, gridDim.x, ntiles);
  if (g < 0) return;
  const int inst = tile_inst[g];
  const StepArgs<T> &a = descs[inst];
  any_stream<T, K, TX, S, NT, PF, RING>(a, g - inst_base[inst], a.dyn[0] + k, a.dyn[1] + k, L,
                                        ring);
}

// Persistent batch launch with a tile work queue: a resident grid of workgroups claims tiles
// from a counter until the launch's tiles are gone, in the order `order` (largest tiles
// first).  Launch L of a run (L = its first stage's step offset in the run / spl) has its own
// counter queue[L], zeroed at the run's start, so a claim c is tile order[c] for c < ntiles
// and every workgroup exits at its first claim past the end.  L < nq is guaranteed by the
// host (the counters cover the run); a launch outside them does no work.
template <int TX, int S, bool NT, int PF, int RING>
__global__ __launch_bounds__(TX) void k_stream_batch_wq(const StepArgs<T> *__restrict__ descs,
                                                        const int32_t *__restrict__ tile_inst,
                                                        const int32_t *__restrict__ inst_base,
                                                        const int32_t *__restrict__ order,
                                                        int ntiles, int k, int spl,
                                                        unsigned *queue, int nq) {
  static_assert(!RING || (TX == 64 && RING == kRingDword), "batch rings: one-wave dword mode");
  __shared__ StreamLds<T, TX, S> L;
  __shared__ T ring[batch_ring_elems<TX, S, PF, RING>()];
  __shared__ unsigned claim;
  const int64_t launch = (descs[0].dyn[1] + k) / spl;
  if (launch < 0 || launch >= nq) return;
  unsigned *q = queue + launch;
  for (;;) {
    if (threadIdx.x == 0) claim = atomicAdd(q, 1u);
    __syncthreads();
    const unsigned t = claim;
    __syncthreads();   // every thread has its claim (and left the previous tile's LDS rows)
    if (t >= (unsigned)ntiles) break;
    const int g = order[t];
    const int inst = tile_inst[g];
    const StepArgs<T> &a = descs[inst];
    any_stream<T, K, TX, S, NT, PF, RING>(a, g - inst_base[inst], a.dyn[0] + k, a.dyn[1] + k,
                                          L, ring);
  }
}

template <int TX, bool AXI, int S, bool NT, bool ISO, int PF, int RING>
void launch_stream_acoustic(const void *args, int nblocks, size_t, hipStream_t s) {
  hipLaunchKernelGGL((k_stream_acoustic<TX, AXI, S, NT, ISO, PF, RING>), dim3(nblocks), dim3(TX),
                     0, s, *static_cast<const StepArgs<T> *>(args));
}

template <int TX, bool AXI, int S, bool NT, bool ISO, int PF, int RING>
void launch_stream_thermal(const void *args, int nblocks, size_t, hipStream_t s) {
  hipLaunchKernelGGL((k_stream_thermal<TX, AXI, S, NT, ISO, PF, RING>), dim3(nblocks), dim3(TX),
                     0, s, *static_cast<const StepArgs<T> *>(args));
}

template <int TX, bool AXI, bool NT, bool ISO, int RING>
void launch_stream_lossy(const void *args, int nblocks, size_t, hipStream_t s) {
  hipLaunchKernelGGL((k_stream_lossy<TX, AXI, NT, ISO, RING>), dim3(nblocks), dim3(TX), 0, s,
                     *static_cast<const StepArgs<T> *>(args));
}

template <int TX, int S, bool NT, int PF, int RING>
void launch_stream_batch(const BatchLaunch &b, int nblocks, size_t, hipStream_t s) {
  hipLaunchKernelGGL((k_stream_batch<TX, S, NT, PF, RING>), dim3(nblocks), dim3(TX), 0, s,
                     static_cast<const StepArgs<T> *>(b.descs), b.tile_inst, b.inst_base,
                     b.ntiles, b.k);
}

// Resident grid: as many workgroups as the device holds at once, never more than the tiles.
template <int TX, int S, bool NT, int PF, int RING>
void launch_stream_batch_wq(const BatchLaunch &b, int, size_t, hipStream_t s) {
  static int per_cu = 0;
  if (per_cu == 0 &&
      hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, k_stream_batch_wq<TX, S, NT, PF, RING>, TX, 0) != hipSuccess)
    per_cu = 1;
  const int grid = std::max(1, std::min(b.ntiles, std::max(1, per_cu) * b.num_cu));
  hipLaunchKernelGGL((k_stream_batch_wq<TX, S, NT, PF, RING>), dim3(grid), dim3(TX), 0, s,
                     static_cast<const StepArgs<T> *>(b.descs), b.tile_inst, b.inst_base,
                     b.order, b.ntiles, b.k, b.spl, b.queue, b.nq);
}

// Variants [v]: 0 plain, 1 axial-symmetric, 2 isotropic coefficients (y arrays = x arrays, not
// axial-symmetric).  [.][f]: f + 1 stages per launch, for f + 1 from 1 (remainder steps) to the
// shape's own stage count; fewer stages than the tiling's ghost depth allows are exact too.
// Remainder variants (fewer stages than the shape's) prefetch one row into registers.  Batches
// (BATCH) run their full launches on the shape's stage count and remainder steps one stage at
// a time; the work-queue variant (WQ) exists for the full launches (remainder launches of a
// work-queue batch run on the dispatcher kernel: they only end runs, and a run starts the
// queue afresh).  RING: single fields stream through LDS rings in kRingX4 mode, batches in
// kRingDword mode.
template <int TX, int S, bool NT, int PF, int RING, bool BATCH, bool WQ>
void fill_variants(KernelSet &ks, int I, int f) {
  ks.stream_acoustic[I][0][f] = launch_stream_acoustic<TX, false, S, NT, false, PF, RING>;
  ks.stream_acoustic[I][1][f] = launch_stream_acoustic<TX, true, S, NT, false, PF, RING>;
  ks.stream_acoustic[I][2][f] = launch_stream_acoustic<TX, false, S, NT, true, PF, RING>;
  ks.stream_thermal[I][0][f] = launch_stream_thermal<TX, false, S, NT, false, PF, RING>;
  ks.stream_thermal[I][1][f] = launch_stream_thermal<TX, true, S, NT, false, PF, RING>;
  ks.stream_thermal[I][2][f] = launch_stream_thermal<TX, false, S, NT, true, PF, RING>;
  if constexpr (BATCH) {
    constexpr int BR = RING ? kRingDword : 0;
    ks.stream_batch[I][f] = launch_stream_batch<TX, S, NT, PF, BR>;
    if constexpr (WQ)
      ks.stream_batch_wq[I][f] = launch_stream_batch_wq<TX, S, NT, PF, BR>;
    else
      ks.stream_batch_wq[I][f] = launch_stream_batch<TX, S, NT, PF, BR>;
  }
}

void FDS_CAT(fill_stream_, FDS_SHAPE)(KernelSet &ks) {
  constexpr int I = FDS_SHAPE;
  static_assert(I < kNumStream, "shape index out of range");
  constexpr int TX = kStreamShapes[I][0];
  constexpr int S = kStreamShapes[I][2];
  constexpr bool NT = kStreamShapes[I][3] != 0;
  constexpr int PF = kStreamShapes[I][4];
  constexpr int RING = kStreamShapes[I][5] != 0 ? kRingX4 : 0;
  static_assert(S >= 1 && S <= kMaxStages, "stage count out of range");
  static_assert(PF >= 1, "prefetch depth out of range");
  // shapes with deeper prefetch serve single fields (no batch, no lossy kernels); a shape's
  // own depth and ring apply to its full stage count
  constexpr bool B = !stream_single_only(I);
  fill_variants<TX, 1, NT, S == 1 ? PF : 1, S == 1 ? RING : 0, B, B && S == 1>(ks, I, 0);
  if constexpr (S >= 2)
    fill_variants<TX, 2, NT, S == 2 ? PF : 1, S == 2 ? RING : 0, B && S == 2, B>(ks, I, 1);
  if constexpr (S >= 3) fill_variants<TX, 3, NT, PF, RING, B, B>(ks, I, 2);
  // lossy acoustics: one step per launch on every shape (rows in registers; the LDS-ring shape
  // streams them into LDS rings)
  if constexpr (B) {
    ks.stream_lossy[I][0] = launch_stream_lossy<TX, false, NT, false, RING>;
    ks.stream_lossy[I][1] = launch_stream_lossy<TX, true, NT, false, RING>;
    ks.stream_lossy[I][2] = launch_stream_lossy<TX, false, NT, true, RING>;
  }
}

}  // namespace FDS_FORM_NS
}  // namespace fds

// Launch tables of the step kernels, one per numeric form.
#pragma once

#include <hip/hip_runtime.h>

namespace fds {

// Tile shapes (TX cells along x, TY rows) of the 2D step kernels.  The last entry serves
// one-row grids (1D models run on the 2D kernels with nx = n, ny = 1).
constexpr int kNumTiles = 6;
constexpr int kTileShapes[kNumTiles][2] = {
    {64, 16}, {64, 8}, {128, 8}, {256, 4}, {128, 4}, {256, 1}};
constexpr int kDefaultTile2D = 1;       // default of the tile engine (lossy models)
constexpr int kDefaultLossless2D = kNumTiles + 8;   // lossless 2D: streaming 64 x 16, 2 steps
constexpr int kDefaultLossy2D = kNumTiles + 7;      // lossy 2D: streaming 64 x 16, one step
constexpr int kTileRow = 5;

// Row-streaming engine (2D models; lossy acoustics one step per launch): TX threads = columns,
// strips of TY rows, S time steps fused per launch, NT = non-temporal state reads and writes,
// PF = input rows loaded ahead (loads in flight per wave), RING = how: 0 into registers, 1 by
// direct global-to-LDS loads into a ring of PF + 1 rows (no registers held by loads in
// flight); TX = 64 runs one wave per workgroup and exchanges neighbours by lane shuffles.
// Shape indices continue the tile table: index kNumTiles + i is
// kStreamShapes[i] = {TX, TY, S, NT, PF, RING}.
constexpr int kNumStream = 15;
constexpr int kStreamShapes[kNumStream][6] = {
    {256, 16, 1, 0, 1, 0}, {256, 16, 2, 0, 1, 0}, {128, 32, 2, 0, 1, 0}, {256, 32, 3, 0, 1, 0},
    {256, 16, 1, 1, 1, 0}, {256, 16, 2, 1, 1, 0}, {256, 32, 3, 1, 1, 0}, {64, 16, 1, 0, 1, 0},
    {64, 16, 2, 0, 1, 0},  {64, 32, 3, 0, 1, 0},  {64, 16, 1, 1, 1, 0},  {64, 16, 2, 1, 1, 0},
    {64, 32, 3, 0, 2, 1},  {64, 16, 2, 0, 2, 1},  {64, 32, 3, 0, 1, 1}};
// Shapes that serve single lossless fields only (no batch or lossy kernels): deeper prefetch.
// (Shapes with the LDS ring and one row of prefetch also run batches, streaming their rows in
// dword mode, and lossy fields, on the lossy kernel.)
constexpr bool stream_single_only(int i) { return kStreamShapes[i][4] > 1; }
// The register-path shape with an LDS-ring shape's geometry (width, height, stages, cache
// policy; prefetch depth aside): it computes the same bits and runs a single field whose rows
// do not start on 16-byte boundaries.  Shapes without the ring map to themselves.
constexpr int ring_fallback(int i) {
  if (!kStreamShapes[i][5]) return i;
  int best = i;
  for (int j = 0; j < kNumStream; ++j)
    if (!kStreamShapes[j][5] && kStreamShapes[j][0] == kStreamShapes[i][0] &&
        kStreamShapes[j][1] == kStreamShapes[i][1] && kStreamShapes[j][2] == kStreamShapes[i][2] &&
        kStreamShapes[j][3] == kStreamShapes[i][3] && kStreamShapes[j][4] == 1)
      best = j;
  return best;
}
static_assert(ring_fallback(12) == 9 && ring_fallback(13) == 8 && ring_fallback(14) == 9,
              "every LDS-ring shape needs a register-path shape of its geometry");
constexpr int kMaxStages = 3;   // largest S in kStreamShapes
// X-macro over the streaming shape indices (each shape is its own translation unit).
#define FDS_STREAM_SHAPES(X) \
  X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14)

using LaunchFn = void (*)(const void *args, int nblocks, size_t lds_bytes, hipStream_t s);
// resident 1D launches: `grid` workgroups (segments) of `threads` threads; the coupled one
// takes acoustic and thermal step arguments and the coupling arguments
using LineFn = void (*)(const void *args, int grid, int threads, hipStream_t s);
using CoupleFn = void (*)(const void *ac, const void *th, const void *cp, int grid,
                          int threads, hipStream_t s);

// Batched launch: one workgroup per tile of every instance; descs[i] are the instances'
// step arguments, tile_inst maps a global tile to its instance, inst_base[i] is the first
// global tile of instance i.
struct BatchLaunch {
  const void *descs;
  const int32_t *tile_inst;
  const int32_t *inst_base;
  int ntiles;
  int k;
  // tile work queue (persistent launches): hand-out order of the global tiles, one claim
  // counter per launch of the run (nq of them), the steps of a full launch and the device's
  // compute units
  const int32_t *order;
  unsigned *queue;
  int nq;
  int spl;
  int num_cu;
};
using BatchFn = void (*)(const BatchLaunch &b, int nblocks, size_t lds_bytes, hipStream_t s);

struct KernelSet {
  LaunchFn acoustic[kNumTiles][2][2];   // [tile][axi][lossy]
  LaunchFn thermal[kNumTiles][2];       // [tile][axi]
  LaunchFn nonlinear;
  BatchFn batch[kNumTiles];
  // [shape][plain, axi, isotropic][stages - 1]: every stage count up to the shape's own
  LaunchFn stream_acoustic[kNumStream][3][kMaxStages];
  LaunchFn stream_thermal[kNumStream][3][kMaxStages];
  BatchFn stream_batch[kNumStream][kMaxStages];          // [shape][stages - 1]
  // persistent variant: a resident grid takes tiles from a work queue (grid chosen inside)
  BatchFn stream_batch_wq[kNumStream][kMaxStages];
  LaunchFn stream_lossy[kNumStream][3];         // lossy acoustics, one step: [shape][variant]
  // resident 1D engine: [cells per thread][acoustic, lossy acoustic, thermal, nonlinear]
  LineFn line[4][4];
  // ThermoAcoustic1D on the resident engine: [cells per thread][lossless, lossy acoustics]
  CoupleFn couple[4][2];
};

const KernelSet &kernels_f64();
const KernelSet &kernels_f32();
const KernelSet &kernels_k32();

}  // namespace fds

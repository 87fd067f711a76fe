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
constexpr int kDefaultLossless2D = kNumTiles;   // lossless 2D: streaming 256 x 16
constexpr int kTileRow = 5;

// Row-streaming engine (lossless 2D models): TX threads = columns, strips of TY rows, S time
// steps fused per launch.  Shape indices continue the tile table: index kNumTiles + i is
// kStreamShapes[i] = {TX, TY, S}.
constexpr int kNumStream = 8;
constexpr int kStreamShapes[kNumStream][3] = {{256, 16, 1}, {256, 32, 1}, {128, 16, 1},
                                              {256, 8, 1},  {256, 16, 2}, {256, 32, 2},
                                              {128, 32, 2}, {256, 32, 3}};

using LaunchFn = void (*)(const void *args, int nblocks, size_t lds_bytes, hipStream_t s);

// Batched launch: one workgroup per tile of every instance; descs[i] are the instances'
// step arguments, tile_inst maps a global tile to its instance, inst_base[i] is the first
// global tile of instance i.
struct BatchLaunch {
  const void *descs;
  const int32_t *tile_inst;
  const int32_t *inst_base;
  int ntiles;
  int k;
};
using BatchFn = void (*)(const BatchLaunch &b, int nblocks, size_t lds_bytes, hipStream_t s);

struct KernelSet {
  LaunchFn acoustic[kNumTiles][2][2];   // [tile][axi][lossy]
  LaunchFn thermal[kNumTiles][2];       // [tile][axi]
  LaunchFn nonlinear;
  BatchFn batch[kNumTiles];
  LaunchFn stream_acoustic[kNumStream][2][2];   // [shape][axi][fused]
  LaunchFn stream_thermal[kNumStream][2][2];
  BatchFn stream_batch[kNumStream][2];          // [shape][fused]
};

const KernelSet &kernels_f64();
const KernelSet &kernels_f32();
const KernelSet &kernels_k32();

}  // namespace fds

// Learned absorbing boundary: shared definitions for the step loop (fds_api.hip) and the
// MFMA kernels (fds_lbc.hip).
//
// Geometry.  A square G x G grid (G = interior + 2L) whose outer frame of depth L is the
// boundary layer.  Side s in {0 top (iy small), 1 bottom, 2 left (ix small), 3 right} has
// "depth" d (0 = outermost cell) and "along" coordinate a in [0, G):
//   top: (ix, iy) = (a, d); bottom: (a, G-1-d); left: (d, a); right: (G-1-d, a).
// The network of side s reads the band rows d in [L, L+NB) at the current and the previous
// time level (zero for a outside [L, G-L), i.e. where the band cell lies in another side's
// layer) and predicts the L layer rows of p, vx, vy.  Top/bottom write a in [0, G), left/right
// a in [L, G-L), so no cell is written twice and no written cell is read in the same launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fds {

constexpr int kLbcBand = 8;                  // band rows read per side
constexpr int kLbcDepth = 16;                // layer rows predicted per side
constexpr int kLbcCin = 2 * 3 * kLbcBand;    // 48: {t, t-1} x {p, vx, vy} x band rows
constexpr int kLbcHidden = 32;
constexpr int kLbcCout = 3 * kLbcDepth;      // 48: {p, vx, vy} x layer rows
constexpr int kLbcTaps = 5;                  // conv kernel width along the boundary
constexpr int kLbcSeg = 16;                  // layer positions per workgroup (4 x 66 groups at 1056)

// Map (side, depth, along) to a flat index of the G x G grid.
__host__ __device__ inline int lbc_cell(int side, int d, int a, int g) {
  switch (side) {
    case 0: return d * g + a;
    case 1: return (g - 1 - d) * g + a;
    case 2: return a * g + d;
    default: return a * g + (g - 1 - d);
  }
}

// Input channel layout: ci = h*24 + c*8 + j  (h: 0 current step, 1 previous; c: p, vx, vy;
// j: band row).  Output channel layout: co = c*16 + d.
struct LbcStep {
  const float *prev[3];        // state before this step
  float *cur[3];               // state after the stencil of this step (layer overwritten)
  float *cur_comp[3];          // Kahan compensation of cur (nullptr if none): zeroed in layer
  const uint16_t *w1;          // bf16 [4][kLbcHidden][kLbcTaps][kLbcCin]
  const uint16_t *w2;          // bf16 [4][kLbcCout][kLbcTaps][kLbcHidden]
  int g, l;                    // grid side, layer depth (must equal kLbcDepth)
  float cs[3];                 // network units: p * cs[0], vx * cs[1], vy * cs[2]
};

// Launch the fused inference kernel for one step on stream s.
void launch_lbc_step(const LbcStep &a, hipStream_t s);

}  // namespace fds

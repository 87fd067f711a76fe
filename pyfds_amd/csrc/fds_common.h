// Shared device-side views and helpers for the pyfds_amd step kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fds {

constexpr int kThreads = 256;   // workgroup size of every step kernel (4 waves of 64)
constexpr int kNumXcd = 8;      // MI355X: 8 XCDs, workgroups dispatched round-robin

// Per-tile event lists (CSR over tiles): events ptr[t]..ptr[t+1]-1 of tile t touch LDS
// position pos[e] and refer to id[e] (a unique boundary cell, or an output slot).
struct EvView {
  const int32_t *ptr;
  const int32_t *pos;
  const int32_t *id;
};

// Compiled boundary list of one component: unique cell u runs entries ptr[u]..ptr[u+1]-1 in
// list order, x = a*x + table[off + step*stride].
struct ProgView {
  const int32_t *ptr;
  const double *a;
  const int64_t *off;
  const int64_t *stride;
  const double *table;
};

struct OutView {
  EvView ev;
  double *buf;   // [rows][m]
  int32_t m;
};

template <typename T>
struct StepArgs {
  int32_t nx, ny;
  int64_t n;
  int32_t tiles_x, ntiles;
  const T *in[3];
  T *out[3];
  const T *cin[3];   // Kahan compensation (accumulated components only)
  T *cout[3];
  const T *coef[8];
  double s[8];
  EvView bev[3];
  ProgView prog[3];
  OutView outp[3];
  const int64_t *dyn;   // dyn[0] = pyfds step of launch k = 0, dyn[1] = output row of k = 0
  int32_t k;            // step index within the launch sequence
  int32_t flags;        // bit0: store assigned components (thermal flux) this step
  int32_t steps;        // resident 1D kernels: steps to run in this launch
  int32_t strip_h;      // row-streaming kernels: rows per strip
  int32_t ghost;        // row-streaming kernels: ghost columns / rows per side (fused steps - 1)
  const void *lhead[3]; // resident 1D kernels: per-thread first boundary cell (LineEvHead)
  const void *ohead[3]; // resident 1D kernels: per-thread first recorded point (LineOutHead)
};

// Resident 1D engine: each thread's first boundary cell of a component, precomputed on the
// host so that a launch starts with one load instead of a chain through the event and
// program lists.
struct LineEvHead {
  int32_t info;         // bits 0-7: register slot (0xff: none cached); bit 8: general path
  int32_t stride;       // signal table stride per step
  int64_t off;          // signal table index of step 0
  double a;             // x = a*x + table[off + step*stride]
};

struct LineOutHead {
  int32_t info;         // bits 0-7: register slot of the first point (0xff: none); bit 8: more
  int32_t id;           // its output slot
};

// Map workgroup b to a tile so that each XCD (b % 8) gets one contiguous band of tiles:
// vertically and horizontally adjacent tiles then share halo rows in that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nblocks, int ntiles) {
  int per = nblocks / kNumXcd;
  int t = (b % kNumXcd) * per + b / kNumXcd;
  return t < ntiles ? t : -1;
}

__host__ __device__ inline int xcd_grid(int ntiles) {
  return ((ntiles + kNumXcd - 1) / kNumXcd) * kNumXcd;
}

// Apply the boundary program of unique cell u at `step` to the value held as (x, c).
// FP64: x = a*x + v per entry, exactly the reference's float64 expression
// (pyfds/regions.py:136-145).  FP32: evaluated in float64 and rounded; with Kahan the
// pair (x, c) is re-split so that x - c carries the float64 result.
template <typename T, bool KAHAN>
__device__ __forceinline__ void apply_prog(const ProgView &pg, int u, int64_t step, T &x, T &c) {
  const int e0 = pg.ptr[u], e1 = pg.ptr[u + 1];
  double v = KAHAN ? (double)x - (double)c : (double)x;
  for (int e = e0; e < e1; ++e) {
    const double val = pg.table[pg.off[e] + step * pg.stride[e]];
    const double ax = pg.a[e] * v;
    v = ax + val;
  }
  x = (T)v;
  if (KAHAN) c = (T)((double)x - v);
}

// x <- x - d with Kahan compensation c (x - c carries the running value).
template <typename T>
__device__ __forceinline__ void kahan_sub(T &x, T &c, T d) {
  const T y = -d - c;
  const T t = x + y;
  c = (t - x) - y;
  x = t;
}

template <typename T, bool KAHAN>
__device__ __forceinline__ double out_value(T x, T c) {
  return KAHAN ? (double)x - (double)c : (double)x;
}

// One-wave workgroups (64 threads) exchange neighbours by DPP wave shifts (one VALU move per
// 32 bits, no LDS round trip, no barrier): lane i receives lane i-1's value (lane 0 gets zero)
// or lane i+1's value (lane 63 gets zero).  Called with all 64 lanes active.
__device__ __forceinline__ int dpp_from_left(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false);   // wave_shr:1
}
__device__ __forceinline__ int dpp_from_right(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false);   // wave_shl:1
}
__device__ __forceinline__ float wave_from_left(float v) {
  return __int_as_float(dpp_from_left(__float_as_int(v)));
}
__device__ __forceinline__ float wave_from_right(float v) {
  return __int_as_float(dpp_from_right(__float_as_int(v)));
}
__device__ __forceinline__ double wave_from_left(double v) {
  return __hiloint2double(dpp_from_left(__double2hiint(v)), dpp_from_left(__double2loint(v)));
}
__device__ __forceinline__ double wave_from_right(double v) {
  return __hiloint2double(dpp_from_right(__double2hiint(v)), dpp_from_right(__double2loint(v)));
}

}  // namespace fds

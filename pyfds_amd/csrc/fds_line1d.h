// Resident 1D engine: one workgroup holds a whole 1D line (Acoustic1D, Thermal1D) in registers
// and advances it up to kLineSteps time steps per launch.
//
// Thread t owns the CPT consecutive cells [t*CPT, t*CPT + CPT) and keeps their state,
// Kahan compensation and coefficients in registers for the whole launch.  A step needs one
// neighbour value from each side of a thread's block; those pass through small LDS arrays
// behind one barrier each, so a step costs two barriers and touches global memory only for
// boundary signal values (fetched one step ahead) and recorded outputs.
//
// Acoustic1D step (pyfds/acoustics.py:40-52), in the reference's FP64 order:
//   p boundaries, record p
//   v <- v - ((Dbwd(B) p) - (D2(M) v))        lossless fields: v <- v - Dbwd(B) p
//   v boundaries, record v
//   p <- p - Dfwd(K) v
// Thermal1D step (pyfds/thermal.py:40-51):
//   T boundaries, record T;  q <- -(Dbwd(Q) T);  q boundaries, record q;  T <- T - Dfwd(G) q
// The difference operators are the scipy DIA products of pyfds/fields.py:157-202 with
// column-indexed factors (dbwd / dfwd / the D2 sum below); only i < 0 and i >= n are zero
// ghosts.  The arithmetic is the LDS-tile engine's for one-row grids, term for term, so both
// engines give the same bits in every form.
#pragma once

#include "fds_common.h"
#include "fds_step2d.h"
#include "../../include/pyfds_amd.h"

namespace fds {

constexpr int kLineThreads = 1024;                 // largest workgroup of the resident kernel
constexpr int kNumLineCpt = 4;
constexpr int kLineCpt[kNumLineCpt] = {1, 2, 4, 8};   // cells per thread of the variants
// workgroup size limit of each variant: the 4- and 8-cell variants need more than the 128
// VGPRs a 1024-thread workgroup leaves each lane
constexpr int kLineMaxThreads[kNumLineCpt] = {1024, 1024, 512, 512};
constexpr int kLineMaxCells = 4096;
constexpr int kLineSteps = 2048;                   // time steps per launch at most

enum LineFamily { kLineAcoustic = 0, kLineAcousticLossy = 1, kLineThermal = 2 };

// A thread's boundary cells of one component.  Lists are CSR over threads: entries
// ev.ptr[t]..ev.ptr[t+1]-1, ev.pos = register slot, ev.id = unique boundary cell.  The first
// cell, when its program has a single entry (the usual case), is cached in registers with its
// signal value fetched one step ahead; any further cell takes the general path.
struct LineBound {
  int e0, e1;        // this thread's cells
  int slot;          // cached cell's register slot, -1 if none
  int32_t idx;       // table index of the value for the next step to apply
  int32_t stride;
  double a, val;     // x = a*x + val

  template <typename T>
  __device__ __forceinline__ void init(const EvView &ev, const ProgView &pg, int t,
                                       int64_t step0) {
    e0 = ev.ptr[t];
    e1 = ev.ptr[t + 1];
    slot = -1;
    idx = stride = 0;
    a = val = 0.0;
    if (e0 < e1) {
      const int u = ev.id[e0];
      const int q0 = pg.ptr[u];
      if (pg.ptr[u + 1] - q0 == 1) {
        slot = ev.pos[e0];
        a = pg.a[q0];
        stride = (int32_t)pg.stride[q0];
        idx = (int32_t)(pg.off[q0] + step0 * pg.stride[q0]);
        val = pg.table[idx];
      }
    }
  }

  // Apply this step's programs to the thread's cells (x, c); `more`: a next step follows in
  // this launch, so its value may be fetched.
  template <typename T, bool KAHAN, int CPT>
  __device__ __forceinline__ void apply(const EvView &ev, const ProgView &pg, int64_t step,
                                        bool more, T (&x)[CPT], T (&c)[CPT]) {
    if (slot >= 0) {
      // the slot is tested against each register position (compile-time index) so that the
      // arrays stay in registers
#pragma unroll
      for (int m = 0; m < CPT; ++m) {
        if (slot != m) continue;
        double v = KAHAN ? (double)x[m] - (double)c[m] : (double)x[m];
        const double ax = a * v;
        v = ax + val;
        x[m] = (T)v;
        if (KAHAN) c[m] = (T)((double)x[m] - v);
      }
      idx += stride;
      if (more) val = pg.table[idx];
    }
    for (int e = e0 + (slot >= 0 ? 1 : 0); e < e1; ++e) {
      const int s = ev.pos[e], u = ev.id[e];
#pragma unroll
      for (int m = 0; m < CPT; ++m)
        if (s == m) apply_prog<T, KAHAN>(pg, u, step, x[m], c[m]);
    }
  }
};

// A thread's recorded points of one component (CSR over threads like LineBound; ev.id = the
// output slot).  The first point is cached in registers.
struct LineRecord {
  int e0, e1, slot, id;

  __device__ __forceinline__ void init(const OutView &o, int t) {
    e0 = o.ev.ptr[t];
    e1 = o.ev.ptr[t + 1];
    slot = e0 < e1 ? o.ev.pos[e0] : -1;
    id = e0 < e1 ? o.ev.id[e0] : 0;
  }

  template <typename T, bool KAHAN, int CPT>
  __device__ __forceinline__ void write(const OutView &o, int64_t row, const T (&x)[CPT],
                                        const T (&c)[CPT]) const {
    if (slot < 0) return;
    double *dst = o.buf + row * o.m;
#pragma unroll
    for (int m = 0; m < CPT; ++m)
      if (slot == m) dst[id] = out_value<T, KAHAN>(x[m], c[m]);
    for (int e = e0 + 1; e < e1; ++e) {
      const int s = o.ev.pos[e], i = o.ev.id[e];
#pragma unroll
      for (int m = 0; m < CPT; ++m)
        if (s == m) dst[i] = out_value<T, KAHAN>(x[m], c[m]);
    }
  }
};

// The line kernel's body for `a.steps` steps starting at step a.dyn[0] + a.k (output row
// a.dyn[1] + a.k).  Reads state from a.in / a.cin, writes a.out / a.cout once at the end.
template <typename T, bool KAHAN, int FAM, int CPT>
__device__ void line1d(const StepArgs<T> &a) {
  constexpr bool THERMAL = FAM == kLineThermal;
  constexpr bool LOSSY = FAM == kLineAcousticLossy;
  constexpr bool KV = KAHAN && !THERMAL;   // the second component is accumulated (velocity)
  __shared__ T xa[kLineThreads], xb[LOSSY ? kLineThreads : 1], xc[LOSSY ? kLineThreads : 1];
  __shared__ T xd[kLineThreads];

  const int t = threadIdx.x, nt = blockDim.x;
  const int n = (int)a.n;
  const int j0 = t * CPT;
  // coefficient slots: backward factor (B or Q), forward factor (K or G), viscous factor M
  const T *Fb = a.coef[THERMAL ? FDS_C_QX : FDS_C_BX];
  const T *Ff = a.coef[THERMAL ? FDS_C_GX : FDS_C_KX];
  const T *Fm = a.coef[FDS_C_MX];

  T x0[CPT], c0[CPT], x1[CPT], c1[CPT];
  T fb[CPT + 1], ff[CPT + 1], fm[LOSSY ? CPT + 2 : 1];
  unsigned valid = 0;
#pragma unroll
  for (int m = 0; m < CPT; ++m) {
    const int j = j0 + m;
    const bool in = j < n;
    if (in) valid |= 1u << m;
    x0[m] = in ? a.in[0][j] : T(0);
    c0[m] = (KAHAN && in) ? a.cin[0][j] : T(0);
    x1[m] = (!THERMAL && in) ? a.in[1][j] : T(0);
    c1[m] = (KV && in) ? a.cin[1][j] : T(0);
  }
#pragma unroll
  for (int m = 0; m <= CPT; ++m) {
    const int jb = j0 + m - 1, jf = j0 + m;
    fb[m] = (jb >= 0 && jb < n) ? Fb[jb] : T(0);
    ff[m] = jf < n ? Ff[jf] : T(0);
  }
  if constexpr (LOSSY) {
#pragma unroll
    for (int m = 0; m < CPT + 2; ++m) {
      const int j = j0 + m - 1;
      fm[m] = (j >= 0 && j < n) ? Fm[j] : T(0);
    }
  }

  const int64_t step0 = a.dyn[0] + a.k, row0 = a.dyn[1] + a.k;
  LineBound b0, b1;
  b0.init<T>(a.bev[0], a.prog[0], t, step0);
  b1.init<T>(a.bev[1], a.prog[1], t, step0);
  LineRecord r0, r1;
  r0.init(a.outp[0], t);
  r1.init(a.outp[1], t);

  for (int s = 0; s < a.steps; ++s) {
    const int64_t step = step0 + s, row = row0 + s;
    const bool more = s + 1 < a.steps;

    // ---- first component: boundaries, record, publish the block's last cell
    b0.apply<T, KAHAN, CPT>(a.bev[0], a.prog[0], step, more, x0, c0);
    r0.write<T, KAHAN, CPT>(a.outp[0], row, x0, c0);
    xa[t] = x0[CPT - 1];
    if constexpr (LOSSY) {
      xb[t] = x1[0];
      xc[t] = x1[CPT - 1];
    }
    __syncthreads();

    // ---- second component from the first (velocity update / heat flux assignment)
    const T pm = t > 0 ? xa[t - 1] : T(0);
    T vo[LOSSY ? CPT + 2 : 1];
    if constexpr (LOSSY) {
      vo[0] = t > 0 ? xc[t - 1] : T(0);
      vo[CPT + 1] = t + 1 < nt ? xb[t + 1] : T(0);
#pragma unroll
      for (int m = 0; m < CPT; ++m) vo[m + 1] = x1[m];
    }
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const T xm = m > 0 ? x0[m - 1] : pm;
      T d = dbwd(fb[m], xm, fb[m + 1], x0[m]);
      if constexpr (THERMAL) {
        x1[m] = ((valid >> m) & 1) ? -d : T(0);
      } else {
        if constexpr (LOSSY) {
          // D2(M) v: offsets -1, 0, +1 in that order (dia_matvec), out-of-line terms absent
          const int j = j0 + m;
          T y = T(0);
          if (j >= 1) { const T q = fm[m] * vo[m]; y = y + q; }
          { const T q = (T(-2) * fm[m + 1]) * vo[m + 1]; y = y + q; }
          if (j + 1 < n) { const T q = fm[m + 2] * vo[m + 2]; y = y + q; }
          d = d - y;
        }
        if ((valid >> m) & 1) {
          if (KAHAN) kahan_sub(x1[m], c1[m], d); else x1[m] = x1[m] - d;
        }
      }
    }

    // ---- second component: boundaries, record, publish the block's first cell
    b1.apply<T, KV, CPT>(a.bev[1], a.prog[1], step, more, x1, c1);
    r1.write<T, KV, CPT>(a.outp[1], row, x1, c1);
    xd[t] = x1[0];
    __syncthreads();

    // ---- first component from the second
    const T vn = t + 1 < nt ? xd[t + 1] : T(0);
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const T xp = m + 1 < CPT ? x1[m + 1] : vn;
      const T d = dfwd(ff[m], x1[m], ff[m + 1], xp);
      if ((valid >> m) & 1) {
        if (KAHAN) kahan_sub(x0[m], c0[m], d); else x0[m] = x0[m] - d;
      }
    }
  }

#pragma unroll
  for (int m = 0; m < CPT; ++m) {
    const int j = j0 + m;
    if (j >= n) continue;
    a.out[0][j] = x0[m];
    a.out[1][j] = x1[m];
    if (KAHAN) a.cout[0][j] = c0[m];
    if (KV) a.cout[1][j] = c1[m];
  }
}

}  // namespace fds

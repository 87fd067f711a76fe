// Resident 1D engine: workgroups hold a 1D line (Acoustic1D, Thermal1D, IdealGas1D,
// Acoustic2ndOrder1D, and the ThermoAcoustic1D pair) in registers and advance it many time
// steps per launch.
//
// Thread t owns the CPT consecutive cells [t*CPT, t*CPT + CPT) of its workgroup's segment and
// keeps their state, Kahan compensation and coefficients in registers for the whole launch.
// One workgroup may hold the whole line (segment = line, up to kLineSteps steps per launch),
// or W workgroups split it: workgroup w owns cells [w*L, w*L + L) (L = a.strip_h) and also
// steps G = a.ghost ghost cells on each side, whose values go wrong from the segment's outer
// edges inwards by one cell per step (two in the first step on the right where a component
// reads its right neighbour's old value), so a launch runs at most G - 1 steps and stores only
// the owned cells (temporal blocking).  A step needs one neighbour value from each side of a
// thread's block: one-wave workgroups move them by DPP lane shifts with no barrier, larger
// ones through small LDS arrays behind one barrier each.  Global memory is touched only for
// boundary signal values (fetched ahead) and recorded outputs.
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
#include "fds_step_nl.h"
#include "../../include/pyfds_amd.h"

namespace fds {

constexpr int kLineThreads = 1024;                 // largest workgroup of the resident kernel
constexpr int kNumLineCpt = 4;
constexpr int kLineCpt[kNumLineCpt] = {1, 2, 4, 8};   // cells per thread of the variants
// workgroup size limit of each variant: the 4- and 8-cell variants need more than the 128
// VGPRs a 1024-thread workgroup leaves each lane
constexpr int kLineMaxThreads[kNumLineCpt] = {1024, 1024, 512, 512};
// the coupled kernel holds two fields: half the workgroup from 2 cells per thread on
constexpr int kCoupleMaxThreads[kNumLineCpt] = {1024, 512, 512, 256};
// the nonlinear models hold three components and call pow_cr / expm1 / log1p
constexpr int kNlLineMaxThreads[kNumLineCpt] = {512, 512, 512, 256};
constexpr int kLineSteps = 2048;                   // time steps per launch at most
// default layout: one workgroup up to kLineSegmentAbove cells; longer lines in one-wave
// segments (DPP neighbour exchange, no barriers) of 64 threads x 2 cells with 48 ghost cells
// per side: 32 owned cells, 47 steps per launch.  Measured on Acoustic1D, 2000 cells: 0.30 us
// per step, against 0.42 for 128-thread segments with LDS exchange and 0.64 for one
// workgroup (one CU's issue rate bounds a single workgroup).  A step of a one-wave segment
// costs ~0.27 us of dependent latency and a launch ~2.8 us, so more ghost cells (more steps
// per launch) pay until the redundant ghost work matters; lines past kLineSegLongAbove cells,
// which fill the GPU, use 4 cells per thread and 64 ghost cells (half the redundancy).
constexpr int kLineSegmentAbove = 512;
constexpr int kLineSegCpt = 2, kLineSegThreads = 64, kLineSegGhost = 48;
constexpr int kLineSegLongAbove = 16384;
constexpr int kLineSegLongCpt = 4, kLineSegLongGhost = 64;
// longest line on the resident engine: its per-thread event tables hold one entry per thread
// of every segment (half a line's cells); longer lines step on the LDS-tile kernels
constexpr int64_t kLineMaxResident = int64_t(1) << 24;

enum LineFamily { kLineAcoustic = 0, kLineAcousticLossy = 1, kLineThermal = 2, kLineNonlinear = 3 };

// workgroup size limit of variant i for a family
__host__ __device__ constexpr int line_max_threads(int fam, int i) {
  return fam == kLineNonlinear ? kNlLineMaxThreads[i] : kLineMaxThreads[i];
}

// Index of this thread in the launch (event lists are CSR over all threads of all segments).
__device__ __forceinline__ int line_thread() { return blockIdx.x * blockDim.x + threadIdx.x; }

// First cell of this workgroup's segment (ghost cells included) and its owned range.
template <typename T>
__device__ __forceinline__ int line_base(const StepArgs<T> &a) {
  return blockIdx.x * a.strip_h - a.ghost;
}

// A thread's boundary cells of one component.  Lists are CSR over threads: entries
// ev.ptr[t]..ev.ptr[t+1]-1, ev.pos = register slot, ev.id = unique boundary cell.  The first
// cell, when its program has a single entry (the usual case), is cached in registers; a
// constant value (stride 0) is loaded once, a signal's values two steps ahead (a step of a
// segment is shorter than a cache round trip).  Any other cell of the thread takes the
// general path, which re-reads the thread's list.
struct LineBound {
  int info;          // bits 0-7: cached cell's register slot (0xff: none); bit 8: general path
  int32_t idx;       // table index of the value of the step after the next to apply
  int32_t stride;
  double a, val, val2;   // x = a*x + val this step; val2 the next step's value

  __device__ __forceinline__ int slot() const { return info & 0xff; }

  // steps: steps of this launch (the signal table covers them all, and no more)
  template <typename T>
  __device__ __forceinline__ void init(const void *heads, const ProgView &pg, int t,
                                       int64_t step0, int steps) {
    const LineEvHead h = static_cast<const LineEvHead *>(heads)[t];
    info = h.info;
    stride = h.stride;
    a = h.a;
    idx = (int32_t)(h.off + step0 * (int64_t)h.stride);
    const bool on = (info & 0xff) != 0xff;
    val = on ? pg.table[idx] : 0.0;
    val2 = (on && stride != 0 && steps > 1) ? pg.table[idx + stride] : val;
  }

  // Apply this step's programs to the thread's cells (x, c); `rem`: steps of this launch
  // after this one (values up to that step may be fetched).
  template <typename T, bool KAHAN, int CPT>
  __device__ __forceinline__ void apply(const EvView &ev, const ProgView &pg, int64_t step,
                                        int rem, T (&x)[CPT], T (&c)[CPT]) {
    const int sl = slot();
    if (sl != 0xff) {
      // the slot is tested against each register position (compile-time index) so that the
      // arrays stay in registers
#pragma unroll
      for (int m = 0; m < CPT; ++m) {
        if (sl != m) continue;
        double v = KAHAN ? (double)x[m] - (double)c[m] : (double)x[m];
        const double ax = a * v;
        v = ax + val;
        x[m] = (T)v;
        if (KAHAN) c[m] = (T)((double)x[m] - v);
      }
      if (stride != 0) {
        idx += stride;                     // now the next step's index
        val = val2;
        if (rem >= 2) val2 = pg.table[idx + stride];
      }
    }
    if (info & 0x100) {
      const int t = line_thread();
      const int e1 = ev.ptr[t + 1];
      for (int e = ev.ptr[t] + (sl != 0xff ? 1 : 0); e < e1; ++e) {
        const int s = ev.pos[e], u = ev.id[e];
#pragma unroll
        for (int m = 0; m < CPT; ++m)
          if (s == m) apply_prog<T, KAHAN>(pg, u, step, x[m], c[m]);
      }
    }
  }
};

// A thread's recorded points of one component (CSR over threads like LineBound; ev.id = the
// output slot).  The first point is cached in registers; further points re-read the list.
struct LineRecord {
  int info, id;      // info: bits 0-7 first point's register slot (0xff: none), bit 8: more

  __device__ __forceinline__ void init(const void *heads, int t) {
    const LineOutHead h = static_cast<const LineOutHead *>(heads)[t];
    info = h.info;
    id = h.id;
  }

  template <typename T, bool KAHAN, int CPT>
  __device__ __forceinline__ void write(const OutView &o, int64_t row, const T (&x)[CPT],
                                        const T (&c)[CPT]) const {
    const int sl = info & 0xff;
    if (sl == 0xff) return;
    double *dst = o.buf + row * o.m;
#pragma unroll
    for (int m = 0; m < CPT; ++m)
      if (sl == m) dst[id] = out_value<T, KAHAN>(x[m], c[m]);
    if (info & 0x100) {
      const int t = line_thread();
      const int e1 = o.ev.ptr[t + 1];
      for (int e = o.ev.ptr[t] + 1; e < e1; ++e) {
        const int s = o.ev.pos[e], i = o.ev.id[e];
#pragma unroll
        for (int m = 0; m < CPT; ++m)
          if (s == m) dst[i] = out_value<T, KAHAN>(x[m], c[m]);
      }
    }
  }
};

// LDS exchange arrays of one field: a = first component's last cell (before the first
// barrier), b / c = lossy velocity's first / last cell (before the first barrier), d / e = the
// second component's first / last cell (before the second barrier).
template <typename T>
struct LineXchg {
  T *a, *b, *c, *d, *e;
};

// One field of the resident 1D engine: a thread's block of CPT cells in registers and the
// three parts of a step, separated by the caller's barriers.
template <typename T, bool KAHAN, int FAM, int CPT>
struct LineField {
  static constexpr bool THERMAL = FAM == kLineThermal;
  static constexpr bool LOSSY = FAM == kLineAcousticLossy;
  static constexpr bool KV = KAHAN && !THERMAL;   // second component accumulated (velocity)

  T x0[CPT], c0[CPT], x1[CPT], c1[CPT];
  T fb[CPT + 1], ff[CPT + 1], fm[LOSSY ? CPT + 2 : 1];
  unsigned valid;
  int n, j0;
  LineBound b0, b1;
  LineRecord r0, r1;

  __device__ __forceinline__ void load(const StepArgs<T> &a, int t, int64_t step0) {
    n = (int)a.n;
    j0 = line_base(a) + t * CPT;
    // coefficient slots: backward factor (B or Q), forward factor (K or G), viscous factor M
    const T *Fb = a.coef[THERMAL ? FDS_C_QX : FDS_C_BX];
    const T *Ff = a.coef[THERMAL ? FDS_C_GX : FDS_C_KX];
    const T *Fm = a.coef[FDS_C_MX];
    valid = 0;
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const int j = j0 + m;
      const bool in = j >= 0 && j < n;
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
      ff[m] = (jf >= 0 && jf < n) ? Ff[jf] : T(0);
    }
    if constexpr (LOSSY) {
#pragma unroll
      for (int m = 0; m < CPT + 2; ++m) {
        const int j = j0 + m - 1;
        fm[m] = (j >= 0 && j < n) ? Fm[j] : T(0);
      }
    }
    const int g = line_thread();
    b0.init<T>(a.lhead[0], a.prog[0], g, step0, a.steps);
    b1.init<T>(a.lhead[1], a.prog[1], g, step0, a.steps);
    r0.init(a.ohead[0], g);
    r1.init(a.ohead[1], g);
  }

  __device__ __forceinline__ bool ok(int m) const { return (valid >> m) & 1; }

  // first component: boundaries, record, publish the block's last cell (WAVE: neighbours come
  // by DPP in the next part, nothing to publish)
  template <bool WAVE>
  __device__ __forceinline__ void part1(const StepArgs<T> &a, int t, int64_t step, int64_t row,
                                        int rem, const LineXchg<T> &x) {
    b0.apply<T, KAHAN, CPT>(a.bev[0], a.prog[0], step, rem, x0, c0);
    r0.write<T, KAHAN, CPT>(a.outp[0], row, x0, c0);
    if constexpr (!WAVE) {
      x.a[t] = x0[CPT - 1];
      if constexpr (LOSSY) {
        x.b[t] = x1[0];
        x.c[t] = x1[CPT - 1];
      }
    }
  }

  // second component from the first (velocity update / heat flux assignment), its
  // boundaries and record; publish its first cell (and, with LAST, its last cell)
  template <bool WAVE, bool LAST = false>
  __device__ __forceinline__ void part2(const StepArgs<T> &a, int t, int nt, int64_t step,
                                        int64_t row, int rem, const LineXchg<T> &x) {
    const T pm = WAVE ? wave_from_left(x0[CPT - 1]) : (t > 0 ? x.a[t - 1] : T(0));
    T vo[LOSSY ? CPT + 2 : 1];
    if constexpr (LOSSY) {
      vo[0] = WAVE ? wave_from_left(x1[CPT - 1]) : (t > 0 ? x.c[t - 1] : T(0));
      vo[CPT + 1] = WAVE ? wave_from_right(x1[0]) : (t + 1 < nt ? x.b[t + 1] : T(0));
#pragma unroll
      for (int m = 0; m < CPT; ++m) vo[m + 1] = x1[m];
    }
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const T xm = m > 0 ? x0[m - 1] : pm;
      T d = dbwd(fb[m], xm, fb[m + 1], x0[m]);
      if constexpr (THERMAL) {
        x1[m] = ok(m) ? -d : T(0);
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
        if (ok(m)) {
          if (KAHAN) kahan_sub(x1[m], c1[m], d); else x1[m] = x1[m] - d;
        }
      }
    }
    b1.apply<T, KV, CPT>(a.bev[1], a.prog[1], step, rem, x1, c1);
    r1.write<T, KV, CPT>(a.outp[1], row, x1, c1);
    if constexpr (!WAVE) {
      x.d[t] = x1[0];
      if constexpr (LAST) x.e[t] = x1[CPT - 1];
    }
  }

  // first component from the second
  template <bool WAVE>
  __device__ __forceinline__ void part3(int t, int nt, const LineXchg<T> &x) {
    const T vn = WAVE ? wave_from_right(x1[0]) : (t + 1 < nt ? x.d[t + 1] : T(0));
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const T xp = m + 1 < CPT ? x1[m + 1] : vn;
      const T d = dfwd(ff[m], x1[m], ff[m + 1], xp);
      if (ok(m)) {
        if (KAHAN) kahan_sub(x0[m], c0[m], d); else x0[m] = x0[m] - d;
      }
    }
  }

  __device__ __forceinline__ void store(const StepArgs<T> &a) const {
    const int lo = blockIdx.x * a.strip_h, hi = min(n, lo + a.strip_h);
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const int j = j0 + m;
      if (j < lo || j >= hi) continue;
      a.out[0][j] = x0[m];
      a.out[1][j] = x1[m];
      if (KAHAN) a.cout[0][j] = c0[m];
      if (KV) a.cout[1][j] = c1[m];
    }
  }
};

// One field for `a.steps` steps starting at step a.dyn[0] + a.k (output row a.dyn[1] + a.k).
// Reads state from a.in / a.cin, writes a.out / a.cout once at the end.  WAVE: one-wave
// workgroups, neighbours by DPP, no barriers.
template <typename T, bool KAHAN, int FAM, int CPT, bool WAVE>
__device__ void line1d(const StepArgs<T> &a) {
  constexpr bool LOSSY = FAM == kLineAcousticLossy;
  constexpr int NX = WAVE ? 1 : kLineThreads;
  __shared__ T xa[NX], xb[LOSSY ? NX : 1], xc[LOSSY ? NX : 1];
  __shared__ T xd[NX];
  const LineXchg<T> x{xa, xb, xc, xd, nullptr};
  const int t = threadIdx.x, nt = blockDim.x;
  const int64_t step0 = a.dyn[0] + a.k, row0 = a.dyn[1] + a.k;
  LineField<T, KAHAN, FAM, CPT> f;
  f.load(a, t, step0);
  for (int s = 0; s < a.steps; ++s) {
    const int64_t step = step0 + s, row = row0 + s;
    const int rem = a.steps - s - 1;
    f.template part1<WAVE>(a, t, step, row, rem, x);
    if (!WAVE) __syncthreads();
    f.template part2<WAVE>(a, t, nt, step, row, rem, x);
    if (!WAVE) __syncthreads();
    f.template part3<WAVE>(t, nt, x);
  }
  f.store(a);
}

// IdealGas1D / Acoustic2ndOrder1D (pyfds/nonlinear_acoustics.py:38-89, 133-171) for `a.steps`
// steps, the LDS-tile kernel's arithmetic (fds_step_nl.h) term for term:
//   p boundaries, record p
//   v -= (Dbwd(f) p / (rho0 + rho') [+ Dcen(h) v] - D2(M) v / (rho0 + rho'))
//   v boundaries, record v
//   rho' -= Dfwd(f) ((rho' + rho0) v);  rho' boundaries, record rho'
//   p = equation of state(rho')
// f = dt/dx and h = dt/dx/2/2 are scalars; out-of-line neighbours are zeros, and the three D2
// terms are always summed (zero factors outside), like the tile kernel.  Two barriers per
// step: p's last cell and v's first and last cells before the first, the mass flux's first
// cell before the second.
template <typename T, bool KAHAN, int CPT, bool WAVE>
__device__ void line1d_nonlinear(const StepArgs<T> &a) {
  constexpr int NX = WAVE ? 1 : kLineThreads;
  __shared__ T xp[NX], xv0[NX], xv1[NX], xw[NX];
  const int t = threadIdx.x, nt = blockDim.x;
  const int n = (int)a.n, j0 = line_base(a) + t * CPT;
  const T f = (T)a.s[FDS_S_DTDX], h = (T)a.s[FDS_S_DTDX_4];
  const int mode = (int)a.s[FDS_S_FLAGS];
  const T *R0 = a.coef[FDS_C_RHO0], *Mv = a.coef[FDS_C_M];
  const T *E1 = a.coef[FDS_C_EOS1], *E2 = a.coef[FDS_C_EOS2], *CS = a.coef[FDS_C_CSQ];

  T p[CPT], cp0[CPT], v[CPT], cv[CPT], r[CPT], cr[CPT];
  T r0[CPT], mm[CPT + 2], e1[CPT], e2[CPT], cs[CPT];
  unsigned valid = 0;
#pragma unroll
  for (int m = 0; m < CPT; ++m) {
    const int j = j0 + m;
    const bool in = j >= 0 && j < n;
    if (in) valid |= 1u << m;
    p[m] = in ? a.in[0][j] : T(0);
    cp0[m] = T(0);
    v[m] = in ? a.in[1][j] : T(0);
    cv[m] = (KAHAN && in) ? a.cin[1][j] : T(0);
    r[m] = in ? a.in[2][j] : T(0);
    cr[m] = (KAHAN && in) ? a.cin[2][j] : T(0);
    r0[m] = in ? R0[j] : T(0);
    e1[m] = in ? E1[j] : T(0);
    e2[m] = in ? E2[j] : T(0);
    cs[m] = in ? CS[j] : T(0);
  }
#pragma unroll
  for (int m = 0; m < CPT + 2; ++m) {
    const int j = j0 + m - 1;
    mm[m] = (j >= 0 && j < n) ? Mv[j] : T(0);
  }
  const int64_t step0 = a.dyn[0] + a.k, row0 = a.dyn[1] + a.k;
  LineBound b0, b1, b2;
  const int g = line_thread();
  b0.init<T>(a.lhead[0], a.prog[0], g, step0, a.steps);
  b1.init<T>(a.lhead[1], a.prog[1], g, step0, a.steps);
  b2.init<T>(a.lhead[2], a.prog[2], g, step0, a.steps);
  LineRecord o0, o1, o2;
  o0.init(a.ohead[0], g);
  o1.init(a.ohead[1], g);
  o2.init(a.ohead[2], g);

  for (int s = 0; s < a.steps; ++s) {
    const int64_t step = step0 + s, row = row0 + s;
    const int rem = a.steps - s - 1;
    b0.apply<T, false, CPT>(a.bev[0], a.prog[0], step, rem, p, cp0);
    o0.write<T, false, CPT>(a.outp[0], row, p, cp0);
    if constexpr (!WAVE) {
      xp[t] = p[CPT - 1];
      xv0[t] = v[0];
      xv1[t] = v[CPT - 1];
      __syncthreads();
    }

    // velocity
    T vo[CPT + 2];
    vo[0] = WAVE ? wave_from_left(v[CPT - 1]) : (t > 0 ? xv1[t - 1] : T(0));
    vo[CPT + 1] = WAVE ? wave_from_right(v[0]) : (t + 1 < nt ? xv0[t + 1] : T(0));
#pragma unroll
    for (int m = 0; m < CPT; ++m) vo[m + 1] = v[m];
    const T pm = WAVE ? wave_from_left(p[CPT - 1]) : (t > 0 ? xp[t - 1] : T(0));
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const T pa = (-f) * (m > 0 ? p[m - 1] : pm);
      const T pb = f * p[m];
      const T A = pa + pb;                               // a_v_p . p
      const T sr = r0[m] + r[m];                         // rho0 + rho'
      T D = T(0);                                        // a_v_v . v, offsets -1, 0, 1
      { const T q = mm[m] * vo[m]; D = D + q; }
      { const T q = (T(-2) * mm[m + 1]) * vo[m + 1]; D = D + q; }
      { const T q = mm[m + 2] * vo[m + 2]; D = D + q; }
      T inc;
      if (mode & kNlConvective) {
        const T ca = (-h) * vo[m];
        const T cb = h * vo[m + 2];
        const T C = ca + cb;                             // a_v_v2 . v (central)
        inc = ((A / sr) + C) - (D / sr);
      } else {
        inc = (A / sr) - (D / sr);
      }
      if ((valid >> m) & 1) {
        if (KAHAN) kahan_sub(v[m], cv[m], inc); else v[m] = v[m] - inc;
      }
    }
    b1.apply<T, KAHAN, CPT>(a.bev[1], a.prog[1], step, rem, v, cv);
    o1.write<T, KAHAN, CPT>(a.outp[1], row, v, cv);

    // mass flux (rho' + rho0) v, then density
    T w[CPT];
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const T sr = r[m] + r0[m];
      w[m] = ((valid >> m) & 1) ? sr * v[m] : T(0);
    }
    T wn;
    if constexpr (WAVE) {
      wn = wave_from_right(w[0]);
    } else {
      xw[t] = w[0];
      __syncthreads();
      wn = t + 1 < nt ? xw[t + 1] : T(0);
    }
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const T ta = (-f) * w[m];
      const T tb = f * (m + 1 < CPT ? w[m + 1] : wn);
      const T d = ta + tb;
      if ((valid >> m) & 1) {
        if (KAHAN) kahan_sub(r[m], cr[m], d); else r[m] = r[m] - d;
      }
    }
    b2.apply<T, KAHAN, CPT>(a.bev[2], a.prog[2], step, rem, r, cr);
    o2.write<T, KAHAN, CPT>(a.outp[2], row, r, cr);

    // equation of state
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      const T dr = KAHAN ? (T)((double)r[m] - (double)cr[m]) : r[m];
      T pn;
      if (mode & kNlSecondOrder) {
        const T sq = dr * dr;
        const T pa = e1[m] * dr;
        const T pb = e2[m] * sq;
        pn = pa + pb;
      } else if (mode & kNlState) {
        if (sizeof(T) == 8) {
          const T ratio = (r0[m] + dr) / r0[m];
          const T pw = (T)crm::pow_cr((double)ratio, (double)e2[m]);  // correctly rounded
          pn = e1[m] * (pw - T(1));
        } else {
          const double xr = (double)dr / (double)r0[m];
          pn = (T)((double)e1[m] * expm1((double)e2[m] * log1p(xr)));
        }
      } else {
        pn = cs[m] * dr;
      }
      p[m] = ((valid >> m) & 1) ? pn : T(0);
    }
  }

  const int lo = blockIdx.x * a.strip_h, hi = min(n, lo + a.strip_h);
#pragma unroll
  for (int m = 0; m < CPT; ++m) {
    const int j = j0 + m;
    if (j < lo || j >= hi) continue;
    a.out[0][j] = p[m];
    a.out[1][j] = v[m];
    a.out[2][j] = r[m];
    if (KAHAN) {
      a.cout[1][j] = cv[m];
      a.cout[2][j] = cr[m];
    }
  }
}

// Acoustic-loss coupling of ThermoAcoustic1D (pyfds/coupled_fields.py:46-65 through
// pyfds/coupling.py:107-116): per-cell inputs and the accumulated transfer (in/out).
template <typename T>
struct LineCouple {
  const T *rho;      // acoustic density
  const T *loss;     // mu / rho_thermal / c_p
  double *acc;       // accumulated transfer, n values
  T dt;
  int64_t stepping;
  int accumulate;
};

// ThermoAcoustic1D for `ac.steps` steps: per step the acoustic field (a step of
// pyfds/acoustics.py:40-52), then the thermal field (pyfds/thermal.py:40-51), then the
// coupling (SynchronizedFields.sim_step, pyfds/coupling.py:78-85): with v the velocity just
// computed, d = a_v_p . (v * rho) = Dbwd(B)(v * rho) and transfer = ((loss * d**2) * dt);
// accumulated every step when `accumulate`, added to the temperature when step % stepping == 0
// (then the accumulator restarts from zero).  Four barriers per step.  In segments the
// transfer reads the velocity's left neighbour, so wrong values reach one cell further into
// the temperature: a launch runs at most ghost - 2 steps.
template <typename T, bool KAHAN, int AFAM, int CPT, bool WAVE>
__device__ void line1d_coupled(const StepArgs<T> &ac, const StepArgs<T> &th,
                               const LineCouple<T> &cp) {
  constexpr bool LOSSY = AFAM == kLineAcousticLossy;
  constexpr int NX = WAVE ? 1 : kLineThreads;
  __shared__ T xa[NX], xb[LOSSY ? NX : 1], xc[LOSSY ? NX : 1];
  __shared__ T xd[NX], xe[NX], ya[NX], yd[NX];
  const LineXchg<T> xA{xa, xb, xc, xd, xe}, xT{ya, nullptr, nullptr, yd, nullptr};
  const int t = threadIdx.x, nt = blockDim.x;
  const int64_t step0 = ac.dyn[0] + ac.k, row0 = ac.dyn[1] + ac.k;
  LineField<T, KAHAN, AFAM, CPT> A;
  LineField<T, KAHAN, kLineThermal, CPT> H;
  A.load(ac, t, step0);
  H.load(th, t, step0);
  const int n = A.n, j0 = A.j0;
  T rho[CPT + 1], loss[CPT];
  double acc[CPT];
#pragma unroll
  for (int m = 0; m <= CPT; ++m) {
    const int j = j0 + m - 1;
    rho[m] = (j >= 0 && j < n) ? cp.rho[j] : T(0);
  }
#pragma unroll
  for (int m = 0; m < CPT; ++m) {
    const int j = j0 + m;
    const bool in = j >= 0 && j < n;
    loss[m] = in ? cp.loss[j] : T(0);
    acc[m] = (cp.accumulate && in) ? cp.acc[j] : 0.0;
  }
  for (int s = 0; s < ac.steps; ++s) {
    const int64_t step = step0 + s, row = row0 + s;
    const int rem = ac.steps - s - 1;
    A.template part1<WAVE>(ac, t, step, row, rem, xA);
    if (!WAVE) __syncthreads();
    A.template part2<WAVE, true>(ac, t, nt, step, row, rem, xA);
    if (!WAVE) __syncthreads();
    A.template part3<WAVE>(t, nt, xA);
    // loss transfer from the step's final velocity
    T tr[CPT];
    {
      const T vm = WAVE ? wave_from_left(A.x1[CPT - 1]) : (t > 0 ? xA.e[t - 1] : T(0));
#pragma unroll
      for (int m = 0; m < CPT; ++m) {
        const T wm = (m > 0 ? A.x1[m - 1] : vm) * rho[m];
        const T wc = A.x1[m] * rho[m + 1];
        const T d = dbwd(A.fb[m], wm, A.fb[m + 1], wc);
        const T d2 = d * d;
        const T ld = loss[m] * d2;
        tr[m] = ld * cp.dt;
      }
    }
    H.template part1<WAVE>(th, t, step, row, rem, xT);
    if (!WAVE) __syncthreads();
    H.template part2<WAVE>(th, t, nt, step, row, rem, xT);
    if (!WAVE) __syncthreads();
    H.template part3<WAVE>(t, nt, xT);
    const bool deliver = step % cp.stepping == 0;
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      if (!H.ok(m)) continue;
      if (cp.accumulate) acc[m] = acc[m] + (double)tr[m];
      if (!deliver) continue;
      const T add = cp.accumulate ? (T)acc[m] : tr[m];
      if (KAHAN) kahan_sub(H.x0[m], H.c0[m], -add); else H.x0[m] = H.x0[m] + add;
      if (cp.accumulate) acc[m] = 0.0;
    }
  }
  A.store(ac);
  H.store(th);
  if (cp.accumulate) {
    const int lo = blockIdx.x * ac.strip_h, hi = min(n, lo + ac.strip_h);
#pragma unroll
    for (int m = 0; m < CPT; ++m)
      if (j0 + m >= lo && j0 + m < hi) cp.acc[j0 + m] = acc[m];
  }
}

}  // namespace fds

// One time step of pyfds' 2D staggered leapfrog solvers as a single LDS-tiled kernel.
//
// The grid is the reference's flat array i = ix + nx*iy: the x neighbour of i is i+-1
// (so the x stencil wraps from the end of one row to the start of the next) and the y
// neighbour is i+-nx; only i < 0 and i >= n are (zero) ghosts (pyfds/fields.py:273-365 build
// exactly these DIA offsets).  A workgroup owns a TX x TY tile and stages a box of
// (TY+2) x (TX+2) flat positions: box (r, c) holds flat (y0+r)*nx + x0 + c for r in [-1, TY],
// c in [-1, TX].  Because every box position is addressed by its flat index, the row wrap and
// partial tiles need no special cases.
//
// Per step (Acoustic2D, pyfds/acoustics.py:111-128; Axi :205-225):
//   p box <- p (+ p boundaries of this step), record p
//   vx on rows [0,TY) x cols [0,TX], vy on rows [0,TY] x cols [0,TX)   (tile + 1 halo)
//   vx/vy boundaries, record vx/vy
//   p_new on the tile
// Thermal2D (pyfds/thermal.py:92-107; Axi :161-176) is the same with q assigned, not updated.
// Every multiply/add appears in the reference's order (scipy dia_matvec: y = 0, then one
// stored diagonal after another; numpy element-wise expressions left to right) and the
// library is compiled with -ffp-contract=off, so the FP64 form reproduces the reference's
// float64 results bit for bit.
//
// State is double-buffered (in -> out) because neighbouring tiles read this tile's halo.
#pragma once

#include "fds_common.h"
#include "../../include/pyfds_amd.h"

namespace fds {

// LDS layout shared by host (launch size) and device.
struct Lds2D {
  int nb, nk, nvx, nvy;       // element counts
  int o_p, o_b, o_by, o_k, o_ky, o_vx, o_vy, o_cp, o_cvx, o_cvy;   // element offsets
  int total;
};

__host__ __device__ inline Lds2D lds2d_layout(int tx, int ty, bool sep_b, bool sep_k, bool kahan) {
  Lds2D L;
  L.nb = (tx + 2) * (ty + 2);
  L.nk = (tx + 1) * (ty + 1);
  L.nvx = ty * (tx + 1);
  L.nvy = (ty + 1) * tx;
  int o = 0;
  L.o_p = o; o += L.nb;
  L.o_b = o; o += L.nb;
  L.o_by = sep_b ? o : L.o_b; if (sep_b) o += L.nb;
  L.o_k = o; o += L.nk;
  L.o_ky = sep_k ? o : L.o_k; if (sep_k) o += L.nk;
  L.o_vx = o; o += L.nvx;
  L.o_vy = o; o += L.nvy;
  L.o_cp = o; if (kahan) o += L.nb;
  L.o_cvx = o; if (kahan) o += L.nvx;
  L.o_cvy = o; if (kahan) o += L.nvy;
  L.total = o;
  return L;
}

// flag bits of StepArgs::flags
constexpr int kFlagStore = 1;    // write assigned components (thermal heat flux)
constexpr int kFlagIsoB = 2;     // y backward coefficients alias the x ones
constexpr int kFlagIsoK = 4;     // y forward coefficients alias the x ones
constexpr int kFlagAxi = 8;      // batched launches: axial-symmetric instance
constexpr int kFlagLossy = 16;   // batched launches: lossy acoustic instance
constexpr int kFlagThermal = 32; // batched launches: thermal instance

// Radius of the cells in column ix (pyfds/acoustics.py:176: np.tile(x.vector, ny) + dx/2 with
// x.vector = arange(nx)*dx), rounded like the reference: (ix*dx) + dx/2 in float64.
template <typename T>
__device__ __forceinline__ T radius(int ix, double dx, double half_dx) {
  const double xi = (double)ix * dx;
  return (T)(xi + half_dx);
}

// Column of box column c of a tile starting at x0: the flat x +-1 neighbour wraps into the
// previous / next row, so column x0 + c is taken modulo nx (c >= -1).
__device__ __forceinline__ int wrap_col(int v, int nx) {
  if (v < 0) v += nx;
  if (v >= nx) v -= nx;
  if (v >= nx) v %= nx;
  return v;
}

template <typename T, bool KAHAN>
__device__ __forceinline__ void run_events(const EvView &ev, const ProgView &pg, int tile,
                                           int64_t step, T *vals, T *comp) {
  const int e0 = ev.ptr[tile], e1 = ev.ptr[tile + 1];
  for (int e = e0 + (int)threadIdx.x; e < e1; e += kThreads) {
    const int pos = ev.pos[e];
    T x = vals[pos];
    T c = KAHAN ? comp[pos] : T(0);
    apply_prog<T, KAHAN>(pg, ev.id[e], step, x, c);
    vals[pos] = x;
    if (KAHAN) comp[pos] = c;
  }
}

template <typename T, bool KAHAN>
__device__ __forceinline__ void record(const OutView &o, int tile, int64_t row, const T *vals,
                                       const T *comp) {
  const int e0 = o.ev.ptr[tile], e1 = o.ev.ptr[tile + 1];
  for (int e = e0 + (int)threadIdx.x; e < e1; e += kThreads) {
    const int pos = o.ev.pos[e];
    o.buf[row * o.m + o.ev.id[e]] = out_value<T, KAHAN>(vals[pos], KAHAN ? comp[pos] : T(0));
  }
}

// Backward difference with column-indexed factors at one point:
// (-f[j-s]) * x[j-s] + f[j] * x[j]  (dia_matvec order for offsets [-s, 0]).
template <typename T>
__device__ __forceinline__ T dbwd(T fm, T xm, T fc, T xc) {
  const T t0 = (-fm) * xm;
  const T t1 = fc * xc;
  return t0 + t1;
}

// Forward difference: (-f[j]) * x[j] + f[j+s] * x[j+s]  (offsets [0, s]).
template <typename T>
__device__ __forceinline__ T dfwd(T fc, T xc, T fp, T xp) {
  const T t0 = (-fc) * xc;
  const T t1 = fp * xp;
  return t0 + t1;
}

// ------------------------------------------------------------------------------------------
// Acoustic2D / Acoustic3DAxi (and Acoustic1D as nx = n, ny = 1 with zero y coefficients)
// ------------------------------------------------------------------------------------------
template <typename T, bool KAHAN, int TX, int TY, bool AXI, bool LOSSY>
__device__ void acoustic_tile(const StepArgs<T> &a, int tile, int64_t step, int64_t row, T *lds) {
  constexpr int BW = TX + 2, KW = TX + 1;
  const bool iso_b = (a.flags & kFlagIsoB) != 0;
  const bool iso_k = !AXI && (a.flags & kFlagIsoK) != 0;
  const Lds2D L = lds2d_layout(TX, TY, !iso_b, !iso_k, KAHAN);
  T *sp = lds + L.o_p, *sb = lds + L.o_b, *sby = lds + L.o_by;
  T *sk = lds + L.o_k, *sky = lds + L.o_ky;
  T *svx = lds + L.o_vx, *svy = lds + L.o_vy;
  T *scp = lds + L.o_cp, *scvx = lds + L.o_cvx, *scvy = lds + L.o_cvy;

  const int tid = threadIdx.x;
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const int x0 = tx * TX, y0 = ty * TY;
  const int nx = a.nx, n = (int)a.n;
  const int base = y0 * nx + x0;
  const double dx = a.s[FDS_S_DX], hdx = a.s[FDS_S_HALF_DX];
  const T *Kx = a.coef[FDS_C_KX], *Ky = a.coef[FDS_C_KY];
  const T *Bx = a.coef[FDS_C_BX], *By = a.coef[FDS_C_BY];
  const T *Mx = a.coef[FDS_C_MX], *My = a.coef[FDS_C_MY];
  const T *Mc = a.coef[FDS_C_MC], *Mt = a.coef[FDS_C_MT];
  const T *vx_in = a.in[1], *vy_in = a.in[2];

  // ---- stage p (+compensation) and the backward coefficients over the box
  for (int e = tid; e < L.nb; e += kThreads) {
    const int rr = e / BW;
    const int r = rr - 1, c = e - rr * BW - 1;
    const int j = base + r * nx + c;
    const bool in = (j >= 0) && (j < n);
    sp[e] = in ? a.in[0][j] : T(0);
    sb[e] = in ? Bx[j] : T(0);
    if (!iso_b) sby[e] = in ? By[j] : T(0);
    if (KAHAN) scp[e] = in ? a.cin[0][j] : T(0);
  }
  // forward coefficients over rows [0, TY] x cols [0, TX]; Axi: x factor is K/r
  for (int e = tid; e < L.nk; e += kThreads) {
    const int r = e / KW, c = e - r * KW;
    const int j = base + r * nx + c;
    const bool in = j < n;
    if (AXI) {
      sk[e] = in ? Kx[j] / radius<T>(wrap_col(x0 + c, nx), dx, hdx) : T(0);
      sky[e] = in ? Ky[j] : T(0);
    } else {
      sk[e] = in ? Kx[j] : T(0);
      if (!iso_k) sky[e] = in ? Ky[j] : T(0);
    }
  }
  __syncthreads();

  // ---- pressure boundaries of this step on every staged position, then record
  if (a.bev[0].ptr[tile + 1] > a.bev[0].ptr[tile]) {
    run_events<T, KAHAN>(a.bev[0], a.prog[0], tile, step, sp, scp);
    __syncthreads();
  }
  record<T, KAHAN>(a.outp[0], tile, row, sp, scp);

  // viscous operator V (a_vx_vx == a_vy_vy) at flat j applied to field f:
  // offsets [-nx, -1, 0, 1, nx] in that order; Axi adds the central term at +-1.
  auto visc = [&](const T *f, int j, int ix) -> T {
    T y = T(0);
    if (j - nx >= 0) { const T t = My[j - nx] * f[j - nx]; y = y + t; }
    if (j - 1 >= 0) {
      T cm = Mx[j - 1];
      if (AXI) {
        const T fc = Mc[j - 1] / radius<T>(wrap_col(ix - 1, nx), dx, hdx);
        cm = cm + (-fc) / T(2);
      }
      const T t = cm * f[j - 1];
      y = y + t;
    }
    {
      const T c0 = (T(-2) * Mx[j]) + (T(-2) * My[j]);
      const T t = c0 * f[j];
      y = y + t;
    }
    if (j + 1 < n) {
      T cp = Mx[j + 1];
      if (AXI) {
        const T fc = Mc[j + 1] / radius<T>(wrap_col(ix + 1, nx), dx, hdx);
        cp = cp + fc / T(2);
      }
      const T t = cp * f[j + 1];
      y = y + t;
    }
    if (j + nx < n) { const T t = My[j + nx] * f[j + nx]; y = y + t; }
    return y;
  };

  // ---- velocities on tile + one halo column / row
  for (int e = tid; e < L.nvx; e += kThreads) {
    const int r = e / KW, c = e - r * KW;
    const int j = base + r * nx + c;
    T x = T(0), cc = T(0);
    if (j < n) {
      const int bm = (r + 1) * BW + c, bc = bm + 1;   // box (r, c-1), (r, c)
      T d = dbwd(sb[bm], sp[bm], sb[bc], sp[bc]);
      if (LOSSY) {
        const int ix = wrap_col(x0 + c, nx);
        d = d - visc(vx_in, j, ix);
        if (AXI) {
          const T rj = radius<T>(ix, dx, hdx);
          const T term = (Mt[j] * vx_in[j]) / (rj * rj);
          d = d + term;
        }
      }
      x = vx_in[j];
      if (KAHAN) { cc = a.cin[1][j]; kahan_sub(x, cc, d); } else { x = x - d; }
    }
    svx[e] = x;
    if (KAHAN) scvx[e] = cc;
  }
  for (int e = tid; e < L.nvy; e += kThreads) {
    const int r = e / TX, c = e - r * TX;
    const int j = base + r * nx + c;
    T x = T(0), cc = T(0);
    if (j < n) {
      const int bm = r * BW + c + 1, bc = bm + BW;    // box (r-1, c), (r, c)
      T d = dbwd(sby[bm], sp[bm], sby[bc], sp[bc]);
      if (LOSSY) d = d - visc(vy_in, j, wrap_col(x0 + c, nx));
      x = vy_in[j];
      if (KAHAN) { cc = a.cin[2][j]; kahan_sub(x, cc, d); } else { x = x - d; }
    }
    svy[e] = x;
    if (KAHAN) scvy[e] = cc;
  }
  __syncthreads();

  const bool ev1 = a.bev[1].ptr[tile + 1] > a.bev[1].ptr[tile];
  const bool ev2 = a.bev[2].ptr[tile + 1] > a.bev[2].ptr[tile];
  if (ev1 || ev2) {
    run_events<T, KAHAN>(a.bev[1], a.prog[1], tile, step, svx, scvx);
    run_events<T, KAHAN>(a.bev[2], a.prog[2], tile, step, svy, scvy);
    __syncthreads();
  }
  record<T, KAHAN>(a.outp[1], tile, row, svx, scvx);
  record<T, KAHAN>(a.outp[2], tile, row, svy, scvy);

  // ---- pressure update on the tile, write everything back
  for (int e = tid; e < TX * TY; e += kThreads) {
    const int r = e / TX, c = e - r * TX;
    if (x0 + c >= a.nx || y0 + r >= a.ny) continue;
    const int j = base + r * nx + c;
    const int ki = r * KW + c;
    T wc = svx[ki], wp = svx[ki + 1];
    if (AXI) {
      wc = wc * radius<T>(x0 + c, dx, hdx);
      wp = wp * radius<T>(wrap_col(x0 + c + 1, nx), dx, hdx);
    }
    const T dxt = dfwd(sk[ki], wc, sk[ki + 1], wp);
    const int vi = r * TX + c;
    const T dyt = dfwd(sky[ki], svy[vi], sky[ki + KW], svy[vi + TX]);
    const T d = dxt + dyt;
    const int bi = (r + 1) * BW + c + 1;
    T p = sp[bi];
    if (KAHAN) {
      T cp = scp[bi];
      kahan_sub(p, cp, d);
      a.cout[0][j] = cp;
      a.cout[1][j] = scvx[ki];
      a.cout[2][j] = scvy[vi];
    } else {
      p = p - d;
    }
    a.out[0][j] = p;
    a.out[1][j] = svx[ki];
    a.out[2][j] = svy[vi];
  }
}

// ------------------------------------------------------------------------------------------
// Thermal2D / Thermal3DAxi (and Thermal1D as nx = n, ny = 1 with zero y coefficients)
// ------------------------------------------------------------------------------------------
template <typename T, bool KAHAN, int TX, int TY, bool AXI>
__device__ void thermal_tile(const StepArgs<T> &a, int tile, int64_t step, int64_t row, T *lds) {
  constexpr int BW = TX + 2, KW = TX + 1;
  const bool iso_b = (a.flags & kFlagIsoB) != 0;      // conduction factors (a_q*_t)
  const bool iso_k = !AXI && (a.flags & kFlagIsoK) != 0;   // a_t_q* factors
  const Lds2D L = lds2d_layout(TX, TY, !iso_b, !iso_k, KAHAN);
  T *st = lds + L.o_p, *sq = lds + L.o_b, *sqy = lds + L.o_by;
  T *sg = lds + L.o_k, *sgy = lds + L.o_ky;
  T *sqx_new = lds + L.o_vx, *sqy_new = lds + L.o_vy;
  T *sct = lds + L.o_cp;

  const int tid = threadIdx.x;
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const int x0 = tx * TX, y0 = ty * TY;
  const int nx = a.nx, n = (int)a.n;
  const int base = y0 * nx + x0;
  const double dx = a.s[FDS_S_DX], hdx = a.s[FDS_S_HALF_DX];
  const T *Gx = a.coef[FDS_C_GX], *Gy = a.coef[FDS_C_GY];
  const T *Qx = a.coef[FDS_C_QX], *Qy = a.coef[FDS_C_QY];

  for (int e = tid; e < L.nb; e += kThreads) {
    const int rr = e / BW;
    const int r = rr - 1, c = e - rr * BW - 1;
    const int j = base + r * nx + c;
    const bool in = (j >= 0) && (j < n);
    st[e] = in ? a.in[0][j] : T(0);
    sq[e] = in ? Qx[j] : T(0);
    if (!iso_b) sqy[e] = in ? Qy[j] : T(0);
    if (KAHAN) sct[e] = in ? a.cin[0][j] : T(0);
  }
  for (int e = tid; e < L.nk; e += kThreads) {
    const int r = e / KW, c = e - r * KW;
    const int j = base + r * nx + c;
    const bool in = j < n;
    if (AXI) {
      sg[e] = in ? Gx[j] / radius<T>(wrap_col(x0 + c, nx), dx, hdx) : T(0);
      sgy[e] = in ? Gy[j] : T(0);
    } else {
      sg[e] = in ? Gx[j] : T(0);
      if (!iso_k) sgy[e] = in ? Gy[j] : T(0);
    }
  }
  __syncthreads();

  if (a.bev[0].ptr[tile + 1] > a.bev[0].ptr[tile]) {
    run_events<T, KAHAN>(a.bev[0], a.prog[0], tile, step, st, sct);
    __syncthreads();
  }
  record<T, KAHAN>(a.outp[0], tile, row, st, sct);

  // heat flux q = -(Dbwd(k/d) T), assigned every step (pyfds/thermal.py:98-99)
  for (int e = tid; e < L.nvx; e += kThreads) {
    const int r = e / KW, c = e - r * KW;
    const int j = base + r * nx + c;
    T q = T(0);
    if (j < n) {
      const int bm = (r + 1) * BW + c, bc = bm + 1;
      q = -dbwd(sq[bm], st[bm], sq[bc], st[bc]);
    }
    sqx_new[e] = q;
  }
  for (int e = tid; e < L.nvy; e += kThreads) {
    const int r = e / TX, c = e - r * TX;
    const int j = base + r * nx + c;
    T q = T(0);
    if (j < n) {
      const int bm = r * BW + c + 1, bc = bm + BW;
      q = -dbwd(sqy[bm], st[bm], sqy[bc], st[bc]);
    }
    sqy_new[e] = q;
  }
  __syncthreads();

  const bool ev1 = a.bev[1].ptr[tile + 1] > a.bev[1].ptr[tile];
  const bool ev2 = a.bev[2].ptr[tile + 1] > a.bev[2].ptr[tile];
  if (ev1 || ev2) {
    run_events<T, false>(a.bev[1], a.prog[1], tile, step, sqx_new, nullptr);
    run_events<T, false>(a.bev[2], a.prog[2], tile, step, sqy_new, nullptr);
    __syncthreads();
  }
  record<T, false>(a.outp[1], tile, row, sqx_new, nullptr);
  record<T, false>(a.outp[2], tile, row, sqy_new, nullptr);

  const bool store = (a.flags & kFlagStore) != 0;
  for (int e = tid; e < TX * TY; e += kThreads) {
    const int r = e / TX, c = e - r * TX;
    if (x0 + c >= a.nx || y0 + r >= a.ny) continue;
    const int j = base + r * nx + c;
    const int ki = r * KW + c;
    T wc = sqx_new[ki], wp = sqx_new[ki + 1];
    if (AXI) {
      wc = wc * radius<T>(x0 + c, dx, hdx);
      wp = wp * radius<T>(wrap_col(x0 + c + 1, nx), dx, hdx);
    }
    const T dxt = dfwd(sg[ki], wc, sg[ki + 1], wp);
    const int vi = r * TX + c;
    const T dyt = dfwd(sgy[ki], sqy_new[vi], sgy[ki + KW], sqy_new[vi + TX]);
    const T d = dxt + dyt;
    const int bi = (r + 1) * BW + c + 1;
    T t = st[bi];
    if (KAHAN) {
      T ct = sct[bi];
      kahan_sub(t, ct, d);
      a.cout[0][j] = ct;
    } else {
      t = t - d;
    }
    a.out[0][j] = t;
    if (store) {
      a.out[1][j] = sqx_new[ki];
      a.out[2][j] = sqy_new[vi];
    }
  }
}

// One tile of instance `a` in a batched launch: the model variant comes from a.flags.
template <typename T, bool KAHAN, int TX, int TY>
__device__ __forceinline__ void any_tile(const StepArgs<T> &a, int tile, int64_t step, int64_t row,
                                         T *lds) {
  const int f = a.flags;
  if (f & kFlagThermal) {
    if (f & kFlagAxi) thermal_tile<T, KAHAN, TX, TY, true>(a, tile, step, row, lds);
    else thermal_tile<T, KAHAN, TX, TY, false>(a, tile, step, row, lds);
  } else if (f & kFlagAxi) {
    if (f & kFlagLossy) acoustic_tile<T, KAHAN, TX, TY, true, true>(a, tile, step, row, lds);
    else acoustic_tile<T, KAHAN, TX, TY, true, false>(a, tile, step, row, lds);
  } else {
    if (f & kFlagLossy) acoustic_tile<T, KAHAN, TX, TY, false, true>(a, tile, step, row, lds);
    else acoustic_tile<T, KAHAN, TX, TY, false, false>(a, tile, step, row, lds);
  }
}

}  // namespace fds

// One time step of pyfds' 1D nonlinear acoustics (IdealGas1D, Acoustic2ndOrder1D) as an
// LDS-tiled kernel.  Tile = TX cells [x0, x0+TX); staged halo: p on [-1, TX], v on
// [-1, TX+1], density and static density on [0, TX].
//
// Step (pyfds/nonlinear_acoustics.py:266-300 and :360-382):
//   p boundaries, record p
//   v -= (Dbwd(dt/dx) p / (rho0 + rho') [+ Dcen(dt/dx/2) v] - D2(dt/dx**2 mu) v / (rho0 + rho'))
//   v boundaries, record v
//   rho' -= Dfwd(dt/dx) ((rho' + rho0) v)
//   rho' boundaries, record rho'
//   p = equation of state(rho')
#pragma once

#include "fds_common.h"
#include "fds_crmath.h"
#include "../../include/pyfds_amd.h"

namespace fds {

constexpr int kNlTile = 256;   // cells per tile (one per thread)

constexpr int kNlConvective = 1;   // FDS_S_FLAGS bits
constexpr int kNlState = 2;
constexpr int kNlSecondOrder = 4;

template <typename T, bool KAHAN>
__device__ void nonlinear1d_tile(const StepArgs<T> &a, int tile, int64_t step, int64_t row) {
  constexpr int TX = kNlTile;
  __shared__ T sp[TX + 2], sv[TX + 3], scv[KAHAN ? TX + 3 : 1];
  __shared__ T sd[TX + 1], s0[TX + 1], sm[TX + 3];
  __shared__ T svn[TX + 1], scvn[KAHAN ? TX + 1 : 1], sw[TX + 1];
  __shared__ T sdn[TX], scdn[KAHAN ? TX : 1];

  const int tid = threadIdx.x;
  const int64_t n = a.n;
  const int64_t x0 = (int64_t)tile * TX;
  const T h = (T)a.s[FDS_S_DTDX_4];
  const T f = (T)a.s[FDS_S_DTDX];
  const int mode = (int)a.s[FDS_S_FLAGS];
  const T *rho0 = a.coef[FDS_C_RHO0], *M = a.coef[FDS_C_M];

  // p: box index c+1 for c in [-1, TX]; v, M: c+1 for c in [-1, TX+1]; rho', rho0: c in [0, TX]
  for (int e = tid; e < TX + 3; e += kThreads) {
    const int64_t j = x0 + e - 1;
    const bool in = (j >= 0) && (j < n);
    if (e < TX + 2) sp[e] = in ? a.in[0][j] : T(0);
    sv[e] = in ? a.in[1][j] : T(0);
    if (KAHAN) scv[e] = in ? a.cin[1][j] : T(0);
    sm[e] = in ? M[j] : T(0);
    if (e >= 1 && e < TX + 2) {
      sd[e - 1] = in ? a.in[2][j] : T(0);
      s0[e - 1] = in ? rho0[j] : T(0);
    }
  }
  __syncthreads();
  if (a.bev[0].ptr[tile + 1] > a.bev[0].ptr[tile]) {
    for (int e = a.bev[0].ptr[tile] + tid; e < a.bev[0].ptr[tile + 1]; e += kThreads) {
      const int pos = a.bev[0].pos[e];
      T x = sp[pos], c = T(0);
      apply_prog<T, false>(a.prog[0], a.bev[0].id[e], step, x, c);
      sp[pos] = x;
    }
    __syncthreads();
  }
  for (int e = a.outp[0].ev.ptr[tile] + tid; e < a.outp[0].ev.ptr[tile + 1]; e += kThreads)
    a.outp[0].buf[row * a.outp[0].m + a.outp[0].ev.id[e]] = (double)sp[a.outp[0].ev.pos[e]];

  // velocity on [0, TX]
  for (int c = tid; c <= TX; c += kThreads) {
    const int64_t j = x0 + c;
    T x = T(0), cc = T(0);
    if (j < n) {
      const T pa = (-f) * sp[c];
      const T pb = f * sp[c + 1];
      const T A = pa + pb;                                 // a_v_p . p
      const T s = s0[c] + sd[c];                           // rho0 + rho'
      T D = T(0);                                          // a_v_v . v, offsets -1, 0, 1
      { const T t = sm[c] * sv[c]; D = D + t; }
      { const T t = (T(-2) * sm[c + 1]) * sv[c + 1]; D = D + t; }
      { const T t = sm[c + 2] * sv[c + 2]; D = D + t; }
      T inc;
      if (mode & kNlConvective) {
        const T ca = (-h) * sv[c];
        const T cb = h * sv[c + 2];
        const T C = ca + cb;                               // a_v_v2 . v (central)
        inc = ((A / s) + C) - (D / s);
      } else {
        inc = (A / s) - (D / s);
      }
      x = sv[c + 1];
      if (KAHAN) { cc = scv[c + 1]; kahan_sub(x, cc, inc); } else { x = x - inc; }
    }
    svn[c] = x;
    if (KAHAN) scvn[c] = cc;
  }
  __syncthreads();
  if (a.bev[1].ptr[tile + 1] > a.bev[1].ptr[tile]) {
    for (int e = a.bev[1].ptr[tile] + tid; e < a.bev[1].ptr[tile + 1]; e += kThreads) {
      const int pos = a.bev[1].pos[e];
      T x = svn[pos], c = KAHAN ? scvn[pos] : T(0);
      apply_prog<T, KAHAN>(a.prog[1], a.bev[1].id[e], step, x, c);
      svn[pos] = x;
      if (KAHAN) scvn[pos] = c;
    }
    __syncthreads();
  }
  for (int e = a.outp[1].ev.ptr[tile] + tid; e < a.outp[1].ev.ptr[tile + 1]; e += kThreads) {
    const int pos = a.outp[1].ev.pos[e];
    a.outp[1].buf[row * a.outp[1].m + a.outp[1].ev.id[e]] =
        out_value<T, KAHAN>(svn[pos], KAHAN ? scvn[pos] : T(0));
  }

  // mass flux w = (rho' + rho0) * v on [0, TX]; density on the tile
  for (int c = tid; c <= TX; c += kThreads) {
    const int64_t j = x0 + c;
    T w = T(0);
    if (j < n) { const T s = sd[c] + s0[c]; w = s * svn[c]; }
    sw[c] = w;
  }
  __syncthreads();
  for (int c = tid; c < TX; c += kThreads) {
    const int64_t j = x0 + c;
    T x = T(0), cc = T(0);
    if (j < n) {
      const T ta = (-f) * sw[c];
      const T tb = f * sw[c + 1];
      const T d = ta + tb;
      x = sd[c];
      if (KAHAN) { cc = a.cin[2][j]; kahan_sub(x, cc, d); } else { x = x - d; }
    }
    sdn[c] = x;
    if (KAHAN) scdn[c] = cc;
  }
  __syncthreads();
  if (a.bev[2].ptr[tile + 1] > a.bev[2].ptr[tile]) {
    for (int e = a.bev[2].ptr[tile] + tid; e < a.bev[2].ptr[tile + 1]; e += kThreads) {
      const int pos = a.bev[2].pos[e];
      T x = sdn[pos], c = KAHAN ? scdn[pos] : T(0);
      apply_prog<T, KAHAN>(a.prog[2], a.bev[2].id[e], step, x, c);
      sdn[pos] = x;
      if (KAHAN) scdn[pos] = c;
    }
    __syncthreads();
  }
  for (int e = a.outp[2].ev.ptr[tile] + tid; e < a.outp[2].ev.ptr[tile + 1]; e += kThreads) {
    const int pos = a.outp[2].ev.pos[e];
    a.outp[2].buf[row * a.outp[2].m + a.outp[2].ev.id[e]] =
        out_value<T, KAHAN>(sdn[pos], KAHAN ? scdn[pos] : T(0));
  }

  // equation of state and write-back
  const T *e1 = a.coef[FDS_C_EOS1], *e2 = a.coef[FDS_C_EOS2], *csq = a.coef[FDS_C_CSQ];
  for (int c = tid; c < TX; c += kThreads) {
    const int64_t j = x0 + c;
    if (j >= n) continue;
    const T dr = KAHAN ? (T)((double)sdn[c] - (double)scdn[c]) : sdn[c];
    T p;
    if (mode & kNlSecondOrder) {
      const T sq = dr * dr;
      const T pa = e1[j] * dr;
      const T pb = e2[j] * sq;
      p = pa + pb;
    } else if (mode & kNlState) {
      const T r0 = s0[c];
      if (sizeof(T) == 8) {
        // float64: the reference's expression, term by term, with the correctly rounded power
        const T ratio = (r0 + dr) / r0;
        const T pw = (T)crm::pow_cr((double)ratio, (double)e2[j]);
        p = e1[j] * (pw - T(1));
      } else {
        // float32: ((rho0 + rho')/rho0)**gamma - 1 cancels catastrophically for small rho';
        // evaluate it as expm1(gamma * log1p(rho'/rho0)) in float64 instead
        const double xr = (double)dr / (double)r0;
        const double pw1 = expm1((double)e2[j] * log1p(xr));
        p = (T)((double)e1[j] * pw1);
      }
    } else {
      p = csq[j] * dr;
    }
    a.out[0][j] = p;
    a.out[1][j] = svn[c];
    a.out[2][j] = sdn[c];
    if (KAHAN) {
      a.cout[1][j] = scvn[c];
      a.cout[2][j] = scdn[c];
    }
  }
}

}  // namespace fds

// Row-streaming step kernel for lossy acoustics (Acoustic2D / Acoustic3DAxi with absorption),
// one time step per launch.
//
// Same strip walk as the lossless kernels (a workgroup of TX threads owns a block of columns and
// marches down a strip of H rows, one barrier per row, or DPP wave shifts and no barrier for
// TX = 64), plus the viscous operator V (pyfds/acoustics.py:104-107, :192-201) on both
// velocities:
//   vx <- vx - (((Dbwd(Bx) p) - V vx) [+ (Mt vx) / r^2])      vy <- vy - ((Dbwd(By) p) - V vy)
//   V f at j = (((My f)[j-nx] + (cm f)[j-1]) + (c0 f)[j]) + (cp f)[j+1]) + (My f)[j+nx]
// with c0 = -2 Mx - 2 My and the x factors cm = Mx (+ (-Mc/r)/2), cp = Mx (+ (Mc/r)/2) taken at
// the neighbour's column (the DIA sum's stored diagonals in ascending offset order, the LDS-tile
// kernel's expression term for term, so FP64 stays bit-identical to the reference).
//
// V reads the velocities' old values one row above and below, so rows are prefetched two
// iterations ahead (a three-row ring: the row below is in registers before it is needed), and
// the x neighbours' products (cm f and cp f) pass through LDS like Bx*p.  Tiles carry one ghost
// column per side (x stride TX - 2): the last thread needs its right neighbour's old vx, which
// it loads as a halo, and the first and last threads' velocities and pressures are not stored.
// No ghost rows: the strip reads rows -1 .. H+1 and writes rows 0 .. H-1.
#pragma once

#include "fds_stream2d.h"

namespace fds {

template <typename T>
struct LsRow {   // one row of one column: state, coefficients, viscous factors
  T p, cp, vx, cvx, vy, cvy, bx, by, kx, ky, mx, my, mc, mt;
};

template <typename T>
struct LsHalo {  // last thread: column TX's old vx and its viscous x factors
  T vx, mx, mc;
};

template <typename T, int TX, bool AXI>
struct LossyLds {                     // (one-wave workgroups exchange by DPP wave shifts)
  static constexpr int W = TX == 64 ? 1 : TX + 2;
  T bwd[2][W];                        // Bx*p, columns -1..TX
  T fwd[2][W];                        // forward x products of the new vx, columns 0..TX
  T vis[2][AXI ? 4 : 2][W];           // viscous x products of the old vx, vy, columns 0..TX
};

// Whether every read of a lossy streaming tile (rows y0-1 .. y0+H+1, columns x0 .. x0+TX) falls
// inside the arrays.
__device__ __forceinline__ bool lossy_strip_inside(int x0, int y0, int H, int TX, int nx,
                                                   int64_t n) {
  const int64_t lo = (int64_t)(y0 - 1) * nx + x0 - 1;
  const int64_t hi = (int64_t)(y0 + H + 1) * nx + x0 + TX;
  return lo >= 0 && hi < n;
}

// LDS rings of the lossy kernel (RING, one-wave workgroups): row ordinal r = rho + 1.  The
// neighbour ring holds vx, vy and My of rows r-1 .. r+1+PF (PF + 3 slots: the viscous term reads
// the old velocities of the rows above and below); the row ring holds p, Bx, Kx[, Mx, By, Ky]
// [, Mc, Mt][, cp, cvx, cvy] of rows r .. r+1+PF (PF + 2 slots).  Every input row arrives by
// direct global-to-LDS loads PF rows ahead of the row below the one being updated.
constexpr int kLossyRingPF = 1;

template <bool ISO, bool AXI, bool KAHAN>
constexpr int lossy_row_arrays() {
  return 3 + (ISO ? 0 : 3) + (AXI ? 2 : 0) + (KAHAN ? 3 : 0);
}

template <typename T, int TX, bool ISO, bool AXI, bool KAHAN>
constexpr int lossy_ring_elems() {
  return ((kLossyRingPF + 3) * 3 + (kLossyRingPF + 2) * lossy_row_arrays<ISO, AXI, KAHAN>()) *
         RingGeom<T, TX, kRingX4>::CW;
}

template <typename T, bool KAHAN, int TX, bool AXI, bool NT, bool EV, bool CHK, bool ISO,
          int RING = 0>
__device__ __forceinline__ void lossy_strip(const StepArgs<T> &a, int tile, int64_t step,
                                            int64_t row, LossyLds<T, TX, AXI> &L,
                                            T *ring = nullptr) {
  constexpr bool WAVE = TX == 64;
  constexpr int NV = AXI ? 4 : 2;     // viscous products per row: (left, right) x (vx, vy)
  const int H = a.strip_h;
  const int t = threadIdx.x;
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const int x0 = tx * (TX - 2) - 1, y0 = ty * H;
  const int nx = a.nx, n = (int)a.n;
  const bool iso_b = (a.flags & kFlagIsoB) != 0;
  const T *Kx = a.coef[FDS_C_KX], *Ky = a.coef[FDS_C_KY];
  const T *Bx = a.coef[FDS_C_BX], *By = iso_b ? Bx : a.coef[FDS_C_BY];
  const T *Mx = a.coef[FDS_C_MX], *My = a.coef[FDS_C_MY];
  const T *Mc = a.coef[FDS_C_MC], *Mt = a.coef[FDS_C_MT];
  const T *pin = a.in[0], *vxin = a.in[1], *vyin = a.in[2];
  const double dx = a.s[FDS_S_DX], hdx = a.s[FDS_S_HALF_DX];
  const bool rh = (t == TX - 1);
  const bool owned = t >= 1 && t < TX - 1 && x0 + t < nx;
  T rad = T(1), radh = T(1);
  if (AXI) {
    rad = radius<T>(wrap_any(x0 + t, nx), dx, hdx);
    radh = radius<T>(wrap_any(x0 + TX, nx), dx, hdx);
  }
  if constexpr (!WAVE) {
    for (int b = 0; b < 2; ++b) {
      L.bwd[b][t + 1] = T(0);
      L.fwd[b][t] = T(0);
      for (int k = 0; k < NV; ++k) L.vis[b][k][t] = T(0);
      if (t == 0) {
        L.bwd[b][0] = T(0);
        L.fwd[b][TX] = T(0);
        for (int k = 0; k < NV; ++k) L.vis[b][k][TX] = L.vis[b][k][TX + 1] = T(0);
      }
    }
  }
  StageCursors cur;
  if constexpr (EV) cur = stage_cursors(a, tile);

  constexpr int RS = 3;       // rows rho-1, rho, rho+1 in registers; rho+2 in flight
  LsRow<T> ring_regs[RS] = {};
  LsHalo<T> halo[RS] = {};
  T bpy[2] = {}, kvy[2] = {}, kvd[2] = {};
  T pp[2] = {}, cpp[2] = {};  // p of the previous row after its events

  auto load_row = [&](int rho, LsRow<T> &d, LsHalo<T> &h) __attribute__((always_inline)) {
    const int rb = (y0 + rho) * nx + x0;
    d.p = ldn<NT, CHK>(pin, rb, t, n);
    d.vx = ldn<NT, CHK>(vxin, rb, t, n);
    d.vy = ldn<NT, CHK>(vyin, rb, t, n);
    d.bx = ldr<CHK>(Bx, rb, t, n);
    d.by = ISO ? d.bx : ldr<CHK>(By, rb, t, n);
    d.kx = ldr<CHK>(Kx, rb, t, n);
    d.ky = ISO ? d.kx : ldr<CHK>(Ky, rb, t, n);
    d.mx = ldr<CHK>(Mx, rb, t, n);
    d.my = ISO ? d.mx : ldr<CHK>(My, rb, t, n);
    if (AXI) {
      d.mc = ldr<CHK>(Mc, rb, t, n);
      d.mt = ldr<CHK>(Mt, rb, t, n);
    }
    if (KAHAN) {
      d.cp = ldn<NT, CHK>(a.cin[0], rb, t, n);
      d.cvx = ldn<NT, CHK>(a.cin[1], rb, t, n);
      d.cvy = ldn<NT, CHK>(a.cin[2], rb, t, n);
    }
    // column TX: every lane loads it (one cache line), so the row needs no branch; only the
    // last thread uses the values
    h.vx = ldr<CHK>(vxin, rb, TX, n);
    h.mx = ldr<CHK>(Mx, rb, TX, n);
    if (AXI) h.mc = ldr<CHK>(Mc, rb, TX, n);
  };

  // ---- LDS rings (RING)
  constexpr int PF = kLossyRingPF, RN = PF + 3, RR = PF + 2;
  using RG = RingGeom<T, TX, kRingX4>;
  constexpr int NRW = lossy_row_arrays<ISO, AXI, KAHAN>();
  constexpr int NA = 3 + NRW;                 // ring loads per row
  constexpr int QP = 0, QBX = 1, QKX = 2, QMX = 3, QBY = 4, QKY = 5;
  constexpr int QMC = ISO ? 3 : 6, QCP = QMC + (AXI ? 2 : 0);
  constexpr int NVX = 0, NVY = 1, NMY = 2;    // neighbour ring arrays
  const int lead = ((x0 - 1) % RG::E + RG::E) % RG::E;
  T *const rring = ring + RN * 3 * RG::CW;
  const lds_ptr<T> lring = ring_lds(ring), lrring = lring + RN * 3 * RG::CW;
  const T *nsrc[3], *rsrc[NRW];
  if constexpr (RING != 0) {
    nsrc[NVX] = vxin;
    nsrc[NVY] = vyin;
    nsrc[NMY] = ISO ? Mx : My;
    rsrc[QP] = pin;
    rsrc[QBX] = Bx;
    rsrc[QKX] = Kx;
    if constexpr (!ISO) {
      rsrc[QMX] = Mx;
      rsrc[QBY] = By;
      rsrc[QKY] = Ky;
    }
    if constexpr (AXI) {
      rsrc[QMC] = Mc;
      rsrc[QMC + 1] = Mt;
    }
    if constexpr (KAHAN) {
      rsrc[QCP] = a.cin[0];
      rsrc[QCP + 1] = a.cin[1];
      rsrc[QCP + 2] = a.cin[2];
    }
  }
  auto nbase = [&](int r) __attribute__((always_inline)) {
    return ring + ring_slot(r, 0, RN) * 3 * RG::CW;
  };
  auto rbase = [&](int r) __attribute__((always_inline)) {
    return rring + ring_slot(r, 0, RR) * NRW * RG::CW;
  };
  auto issue = [&](int r) __attribute__((always_inline)) {   // row rho = r - 1
    const int fs = (y0 + r - 1) * nx + x0 - 1 - lead;
    ring_issue<T, TX, 3, kRingX4>(lring + ring_slot(r, 0, RN) * 3 * RG::CW, nsrc, fs, n);
    ring_issue<T, TX, NRW, kRingX4>(lrring + ring_slot(r, 0, RR) * NRW * RG::CW, rsrc, fs, n);
  };
  // value of ring array q (base: the row's slot) at column c of row ordinal r
  auto rd = [&](const T *base, int q, int r, int c) __attribute__((always_inline)) -> T {
    const T v = base[q * RG::CW + lead + 1 + c];
    if (!CHK) return v;
    const int j = (y0 + r - 1) * nx + x0 + c;
    return (j >= 0 && j < n) ? v : T(0);
  };
  // the row of ordinal r into (d, h) as load_row would (the halo: column TX)
  auto read_row = [&](int r, LsRow<T> &d, LsHalo<T> &h) __attribute__((always_inline)) {
    const T *nb = nbase(r), *rb = rbase(r);
    d.p = rd(rb, QP, r, t);
    d.vx = rd(nb, NVX, r, t);
    d.vy = rd(nb, NVY, r, t);
    d.bx = rd(rb, QBX, r, t);
    d.kx = rd(rb, QKX, r, t);
    d.my = rd(nb, NMY, r, t);
    if constexpr (ISO) {
      d.by = d.bx;
      d.ky = d.kx;
      d.mx = d.my;
    } else {
      d.mx = rd(rb, QMX, r, t);
      d.by = rd(rb, QBY, r, t);
      d.ky = rd(rb, QKY, r, t);
    }
    if constexpr (AXI) {
      d.mc = rd(rb, QMC, r, t);
      d.mt = rd(rb, QMC + 1, r, t);
    }
    if constexpr (KAHAN) {
      d.cp = rd(rb, QCP, r, t);
      d.cvx = rd(rb, QCP + 1, r, t);
      d.cvy = rd(rb, QCP + 2, r, t);
    }
    // column TX: every lane reads it (an LDS broadcast); only the last thread uses it
    h.vx = rd(nb, NVX, r, TX);
    h.mx = ISO ? rd(nb, NMY, r, TX) : rd(rb, QMX, r, TX);
    if (AXI) h.mc = rd(rb, QMC, r, TX);
  };

  if constexpr (RING != 0) {
#pragma unroll
    for (int r = 0; r <= PF; ++r) issue(r);   // rows -1 .. PF-1
  } else {
    load_row(-1, ring_regs[0], halo[0]);
    load_row(0, ring_regs[1], halo[1]);
  }
  const int iters = H + 2;    // rows -1 .. H
  run_loop<RING ? 1 : RS>(iters, [&](auto Qc, auto Wc, int i) __attribute__((always_inline)) {
    constexpr int Q = decltype(Qc)::value, W = decltype(Wc)::value;
    constexpr int WU = (W + 2) % RS, WD = (W + 1) % RS;   // rows rho-1 and rho+1
    const int rho = i - 1;
    LsRow<T> rowP = {}, rowD = {};
    LsHalo<T> hP = {}, hD = {};
    T vxu, vyu, myu;
    if constexpr (RING != 0) {
      // row ordinal i is rho; issue row i + 1 + PF, then wait for row i + 1 (the rows after
      // it that are in flight: min(PF, H + 2 - (i + 1)))
      compiler_fence();
      if (i + 1 + PF <= H + 2) issue(i + 1 + PF);
      ring_wait<NA, PF>(H + 1 - i);
      compiler_fence();
      read_row(i, rowP, hP);
      const T *nd = nbase(i + 1), *nu = nbase(i - 1);
      rowD.vx = rd(nd, NVX, i + 1, t);
      rowD.vy = rd(nd, NVY, i + 1, t);
      rowD.my = rd(nd, NMY, i + 1, t);
      // row rho-1 (ordinal i-1): read only where the updates use it (i >= 1)
      vxu = rd(nu, NVX, i - 1, t);
      vyu = rd(nu, NVY, i - 1, t);
      myu = rd(nu, NMY, i - 1, t);
    } else {
      // row rho-1's old velocities and y factor, read before its slot takes row rho+2
      vxu = ring_regs[WU].vx;
      vyu = ring_regs[WU].vy;
      myu = ring_regs[WU].my;
      if (rho + 2 <= H + 1) load_row(rho + 2, ring_regs[WU], halo[WU]);
    }
    LsRow<T> &P = RING ? rowP : ring_regs[RING ? 0 : W];
    const LsRow<T> &D = RING ? rowD : ring_regs[RING ? 0 : WD];
    const LsHalo<T> &h = RING ? hP : halo[RING ? 0 : W];
    (void)hD;

    // ---- before the barrier: p boundaries and record, row products
    T pcur = P.p, cpcur = P.cp;
    if constexpr (EV) {
      T dummy = T(0), dummyc = T(0);
      stream_events<T, KAHAN>(a.bev[0], a.prog[0], cur.b[0], TX + 2, -1, -1, rho, t, step,
                              pcur, cpcur, false, 0, dummy, dummyc);
      stream_record<T, KAHAN>(a.outp[0], cur.o[0], TX + 2, -1, -1, rho, t, row, pcur, cpcur);
    }
    const T bpx = P.bx * pcur;
    bpy[Q] = P.by * pcur;
    // viscous x products of the old velocities: (left factor, right factor) x (vx, vy)
    T q[4];
    T qh = T(0);   // last thread: column TX's right product of vx
    if (AXI) {
      const T fc = P.mc / rad;
      const T cm = P.mx + (-fc) / T(2), cpf = P.mx + fc / T(2);
      q[0] = cm * P.vx;
      q[1] = cpf * P.vx;
      q[2] = cm * P.vy;
      q[3] = cpf * P.vy;
      if (rh) {
        const T fch = h.mc / radh;
        qh = (h.mx + fch / T(2)) * h.vx;
      }
    } else {
      q[0] = q[1] = P.mx * P.vx;
      q[2] = q[3] = P.mx * P.vy;
      if (rh) qh = h.mx * h.vx;
    }
    if constexpr (!WAVE) {
      L.bwd[Q][t + 1] = bpx;
      if (AXI) {
#pragma unroll
        for (int k = 0; k < 4; ++k) L.vis[Q][k][t] = q[k];
        if (rh) L.vis[Q][1][TX] = qh;
      } else {
        L.vis[Q][0][t] = q[0];
        L.vis[Q][1][t] = q[2];
        if (rh) L.vis[Q][0][TX] = qh;
      }
    }
    if (!WAVE) __syncthreads();

    // ---- after the barrier
    T vxn = T(0), cvxn = T(0), vyn = T(0), cvyn = T(0);
    const T c0 = (T(-2) * P.mx) + (T(-2) * P.my);
    // neighbour products: left's left-factor product, right's right-factor product
    T qlx, qrx, qly, qry;
    if constexpr (WAVE) {
      qlx = wave_from_left(q[0]);
      qrx = wave_from_right(q[1]);
      qly = wave_from_left(q[2]);
      qry = wave_from_right(q[3]);
      if (rh) qrx = qh;
    } else if (AXI) {
      qlx = t > 0 ? L.vis[Q][0][t - 1] : T(0);
      qrx = L.vis[Q][1][t + 1];
      qly = t > 0 ? L.vis[Q][2][t - 1] : T(0);
      qry = L.vis[Q][3][t + 1];
    } else {
      qlx = t > 0 ? L.vis[Q][0][t - 1] : T(0);
      qrx = L.vis[Q][0][t + 1];
      qly = t > 0 ? L.vis[Q][1][t - 1] : T(0);
      qry = L.vis[Q][1][t + 1];
    }
    if (rho >= 0 && rho <= H - 1) {
      T left;
      if constexpr (WAVE) left = wave_from_left(bpx);
      else left = L.bwd[Q][t];
      T d = (-left) + bpx;
      T y = T(0);
      { const T u = myu * vxu; y = y + u; }
      y = y + qlx;
      { const T u = c0 * P.vx; y = y + u; }
      y = y + qrx;
      { const T u = D.my * D.vx; y = y + u; }
      d = d - y;
      if (AXI) {
        const T term = (P.mt * P.vx) / (rad * rad);
        d = d + term;
      }
      vxn = P.vx;
      cvxn = P.cvx;
      if (KAHAN) kahan_sub(vxn, cvxn, d); else vxn = vxn - d;
      if constexpr (EV) {
        T dummy = T(0), dummyc = T(0);
        stream_events<T, KAHAN>(a.bev[1], a.prog[1], cur.b[1], TX + 1, 0, 0, rho, t, step,
                                vxn, cvxn, false, 0, dummy, dummyc);
        stream_record<T, KAHAN>(a.outp[1], cur.o[1], TX + 1, 0, 0, rho, t, row, vxn, cvxn);
      }
      T kvx;
      if (AXI) kvx = (P.kx / rad) * (vxn * rad);
      else kvx = P.kx * vxn;
      if constexpr (WAVE) {
        const T right = wave_from_right(kvx);
        kvd[Q] = (-kvx) + right;   // x difference of row rho, used by the next iteration
      } else {
        L.fwd[Q][t] = kvx;
      }
    }
    kvy[Q] = T(0);
    if (rho >= 0) {
      T d = (-bpy[Q ^ 1]) + bpy[Q];
      T y = T(0);
      { const T u = myu * vyu; y = y + u; }
      y = y + qly;
      { const T u = c0 * P.vy; y = y + u; }
      y = y + qry;
      { const T u = D.my * D.vy; y = y + u; }
      d = d - y;
      vyn = P.vy;
      cvyn = P.cvy;
      if (KAHAN) kahan_sub(vyn, cvyn, d); else vyn = vyn - d;
      if constexpr (EV) {
        T dummy = T(0), dummyc = T(0);
        stream_events<T, KAHAN>(a.bev[2], a.prog[2], cur.b[2], TX, 0, 0, rho, t, step, vyn,
                                cvyn, false, 0, dummy, dummyc);
        stream_record<T, KAHAN>(a.outp[2], cur.o[2], TX, 0, 0, rho, t, row, vyn, cvyn);
      }
      kvy[Q] = P.ky * vyn;
    }
    T pnew = T(0), cpnew = T(0);
    if (rho > 0) {    // p of row rho-1 (its x products were published an iteration ago)
      T dxt;
      if constexpr (WAVE) dxt = kvd[Q ^ 1];
      else dxt = (-L.fwd[Q ^ 1][t]) + L.fwd[Q ^ 1][t + 1];
      const T dyt = (-kvy[Q ^ 1]) + kvy[Q];
      const T d = dxt + dyt;
      pnew = pp[Q ^ 1];
      cpnew = cpp[Q ^ 1];
      if (KAHAN) kahan_sub(pnew, cpnew, d); else pnew = pnew - d;
    }
    pp[Q] = pcur;
    cpp[Q] = cpcur;
    if (owned && rho >= 1 && rho - 1 < H && y0 + rho - 1 < a.ny) {
      const int rb = (y0 + rho - 1) * nx + x0;
      stn<NT>(a.out[0], rb, t, pnew);
      if (KAHAN) stn<NT>(a.cout[0], rb, t, cpnew);
    }
    if (owned && rho >= 0 && rho < H && y0 + rho < a.ny) {
      const int rb = (y0 + rho) * nx + x0;
      stn<NT>(a.out[1], rb, t, vxn);
      stn<NT>(a.out[2], rb, t, vyn);
      if (KAHAN) {
        stn<NT>(a.cout[1], rb, t, cvxn);
        stn<NT>(a.cout[2], rb, t, cvyn);
      }
    }
  });
}

template <typename T, int TX>
__device__ __forceinline__ int lossy_path(const StepArgs<T> &a, int tile) {
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const bool inside = lossy_strip_inside(tx * (TX - 2) - 1, ty * a.strip_h, a.strip_h, TX, a.nx,
                                         a.n);
  return __builtin_amdgcn_readfirstlane((tile_has_events(a, tile) ? 2 : 0) | (inside ? 1 : 0));
}

// ISO kernels serve fields whose y coefficient arrays (B, K, M) are the x ones.
template <typename T, bool KAHAN, int TX, bool AXI, bool NT, bool ISO, int RING = 0>
__device__ __forceinline__ void lossy_stream(const StepArgs<T> &a, int tile, int64_t step,
                                             int64_t row, LossyLds<T, TX, AXI> &L,
                                             T *ring = nullptr) {
  switch (lossy_path<T, TX>(a, tile)) {
    case 0:
      lossy_strip<T, KAHAN, TX, AXI, NT, false, true, ISO, RING>(a, tile, step, row, L, ring);
      break;
    case 1:
      lossy_strip<T, KAHAN, TX, AXI, NT, false, false, ISO, RING>(a, tile, step, row, L, ring);
      break;
    case 2:
      lossy_strip<T, KAHAN, TX, AXI, NT, true, true, ISO, RING>(a, tile, step, row, L, ring);
      break;
    case 3:
      lossy_strip<T, KAHAN, TX, AXI, NT, true, false, ISO, RING>(a, tile, step, row, L, ring);
      break;
    default: break;
  }
}

}  // namespace fds

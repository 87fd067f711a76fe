// Row-streaming step kernels for the lossless 2D models (Acoustic2D / Acoustic3DAxi and
// Thermal2D / Thermal3DAxi), with S time steps fused per launch.
//
// A workgroup of TX threads owns a block of columns and marches down a strip of H rows: thread
// t keeps its column's values in registers, prefetches the next row while it computes the
// current one, and exchanges x-neighbour products through double-buffered LDS rows, so one row
// costs one barrier.
//
// Temporal fusion: the launch runs S stages, stage s advancing the field from step n+s-1 to
// n+s.  Stage 1 reads rows from memory; stage s+1 consumes the rows stage s has finished, two
// rows behind it, so the intermediate time levels never leave the registers.  Each extra stage
// needs one more ghost row above and below the strip and one more ghost column on each side
// (the stencil reaches one cell per step): with G >= S-1 ghosts, workgroups overlap by 2G
// columns and write only their TX-2G owned columns and H owned rows.  Tiles are laid out with
// x stride TX-2G; thread t of tile (tx, ty) holds column tx*(TX-2G) - G + t of rows ty*H + r.
//
// Arithmetic is the tile kernel's, term by term (x - d with d formed in the same order; the
// column-indexed differences from products of both columns via (-a)*b == -(a*b)), so FP64
// results stay bit-identical to the reference whatever S, G, TX and H are.
//
// Events (boundaries, outputs) are per-tile lists sorted by row then column over the union of
// the stages' rows; every stage walks them with its own cursor and its own step index.
#pragma once

#include "fds_common.h"
#include "fds_step2d.h"
#include "../../include/pyfds_amd.h"

namespace fds {

struct Cursor {
  int cur, end;
};

// LDS of one streaming workgroup: per stage, double-buffered rows of the backward-difference
// products (columns -1..TX) and of the forward-difference products (columns 0..TX).
template <typename T, int TX, int S>
struct StreamLds {
  T bwd[S][2][TX + 2];
  T fwd[S][2][TX + 1];
};

__device__ __forceinline__ Cursor cursor(const EvView &ev, int tile) {
  return Cursor{ev.ptr[tile], ev.ptr[tile + 1]};
}

// Apply the events of row r of a region (row width w, origin (rmin, cmin)) whose column is
// `col` (or `hcol` when has_halo) to (x, c) / (hx, hc); events of earlier rows are skipped.
template <typename T, bool KAHAN>
__device__ __forceinline__ void stream_events(const EvView &ev, const ProgView &pg, Cursor &q,
                                              int w, int rmin, int cmin, int r, int col,
                                              int64_t step, T &x, T &c, bool has_halo, int hcol,
                                              T &hx, T &hc) {
  int e = q.cur;
  while (e < q.end) {
    const int pos = ev.pos[e];
    const int er = pos / w + rmin;
    if (er > r) break;
    if (er == r) {
      const int ec = pos - (er - rmin) * w + cmin;
      if (ec == col) apply_prog<T, KAHAN>(pg, ev.id[e], step, x, c);
      if (has_halo && ec == hcol) apply_prog<T, KAHAN>(pg, ev.id[e], step, hx, hc);
    }
    ++e;
  }
  q.cur = e;
}

template <typename T, bool KAHAN>
__device__ __forceinline__ void stream_record(const OutView &o, Cursor &q, int w, int rmin,
                                              int cmin, int r, int col, int64_t row, T x, T c) {
  int e = q.cur;
  while (e < q.end) {
    const int pos = o.ev.pos[e];
    const int er = pos / w + rmin;
    if (er > r) break;
    if (er == r && pos - (er - rmin) * w + cmin == col)
      o.buf[row * o.m + o.ev.id[e]] = out_value<T, KAHAN>(x, c);
    ++e;
  }
  q.cur = e;
}

template <typename T>
__device__ __forceinline__ T ld(const T *p, int j, int n) {
  return (j >= 0 && j < n) ? p[j] : T(0);
}

// x modulo nx for any x (ghost columns of the first tile are negative, those of the last may
// pass nx): the flat x +-1 neighbour wraps into the previous / next row.
__device__ __forceinline__ int wrap_any(int x, int nx) {
  if (x >= 0 && x < nx) return x;
  const int v = x % nx;
  return v < 0 ? v + nx : v;
}

// Rows of stage sg (0-based) for ghost depth G and strip height H: the stage updates rows
// [lo, hi] (vy also hi+1) and reads input rows lo-1 .. hi+1.  At loop iteration i it holds
// input row i - G - 1 - 2*sg.
struct StageRows {
  int lo, hi;
  __device__ __forceinline__ StageRows(int G, int H, int sg) : lo(sg - G), hi(H - 1 + G - sg) {}
};

// ------------------------------------------------------------------------------------------
// Acoustic (lossless): v -= Dbwd(B) p, then p -= Dfwd(K) v
// ------------------------------------------------------------------------------------------
template <typename T>
struct AcIn {          // one input row of one column
  T p, cp, vx, cvx, vy, cvy, bx, by, kx, ky;
};

template <typename T>
struct AcHalo {        // left thread: p, B of column -1; right thread: p, B, vx, K of column TX
  T p, cp, b, vx, cvx, k;
};

template <typename T>
struct AcStage {
  AcIn<T> in;          // row being processed
  AcIn<T> hold;        // v_new and coefficients of the previous row (handed to the next stage)
  T bpx, bpy, bpy_prev, p_prev, cp_prev, kvy_prev;
  Cursor pc, vc1, vc2, oc0, oc1, oc2;
};

template <typename T>
__device__ __forceinline__ void ac_stage_init(AcStage<T> &st, const StepArgs<T> &a, int tile) {
  st.pc = cursor(a.bev[0], tile);
  st.vc1 = cursor(a.bev[1], tile);
  st.vc2 = cursor(a.bev[2], tile);
  st.oc0 = cursor(a.outp[0].ev, tile);
  st.oc1 = cursor(a.outp[1].ev, tile);
  st.oc2 = cursor(a.outp[2].ev, tile);
  st.bpx = st.bpy = st.bpy_prev = st.p_prev = st.cp_prev = st.kvy_prev = T(0);
  st.in = AcIn<T>{};
  st.hold = AcIn<T>{};
}

// Before the barrier: boundaries of p in row rho, record p, publish Bx*p.
template <typename T, bool KAHAN, int TX, bool HALO>
__device__ __forceinline__ void ac_pre(AcStage<T> &st, AcHalo<T> &h, const StepArgs<T> &a,
                                       int G, int rho, const StageRows &R, int64_t step,
                                       int64_t orow, T *sbp) {
  const int t = threadIdx.x;
  const bool lh = (t == 0), rh = (t == TX - 1);
  stream_events<T, KAHAN>(a.bev[0], a.prog[0], st.pc, TX + 2, -G - 1, -1, rho, t, step, st.in.p,
                          st.in.cp, HALO && (lh || rh), lh ? -1 : TX, h.p, h.cp);
  stream_record<T, KAHAN>(a.outp[0], st.oc0, TX + 2, -G - 1, -1, rho, t, orow, st.in.p,
                          st.in.cp);
  st.bpx = st.in.bx * st.in.p;
  st.bpy = st.in.by * st.in.p;
  if (rho <= R.hi) {
    sbp[t + 1] = st.bpx;
    if (HALO && lh) sbp[0] = h.b * h.p;
  }
}

// After the barrier: v of row rho, x products, p of row rho-1.  Returns the new p of row rho-1
// in (pn, cpn) and the new v of row rho in (vxn, cvxn, vyn, cvyn).
template <typename T, bool KAHAN, int TX, bool AXI, bool HALO>
__device__ __forceinline__ void ac_post(AcStage<T> &st, const AcHalo<T> &h, const StepArgs<T> &a,
                                        int G, int rho, const StageRows &R, int64_t step,
                                        int64_t orow, const T *sbp, T *skv_cur,
                                        const T *skv_prev, T rad, T radh, T &pn, T &cpn,
                                        T &vxn, T &cvxn, T &vyn, T &cvyn) {
  const int t = threadIdx.x;
  const bool rh = (t == TX - 1);
  vxn = cvxn = vyn = cvyn = T(0);
  T hvxn = T(0), hcvxn = T(0);
  if (rho >= R.lo && rho <= R.hi) {
    const T d = (-sbp[t]) + st.bpx;
    vxn = st.in.vx;
    cvxn = st.in.cvx;
    if (KAHAN) kahan_sub(vxn, cvxn, d); else vxn = vxn - d;
    if (HALO && rh) {
      const T dh = (-st.bpx) + h.b * h.p;
      hvxn = h.vx;
      hcvxn = h.cvx;
      if (KAHAN) kahan_sub(hvxn, hcvxn, dh); else hvxn = hvxn - dh;
    }
    stream_events<T, KAHAN>(a.bev[1], a.prog[1], st.vc1, TX + 1, -G, 0, rho, t, step, vxn, cvxn,
                            HALO && rh, TX, hvxn, hcvxn);
    stream_record<T, KAHAN>(a.outp[1], st.oc1, TX + 1, -G, 0, rho, t, orow, vxn, cvxn);
    T kvx;
    if (AXI) kvx = (st.in.kx / rad) * (vxn * rad);
    else kvx = st.in.kx * vxn;
    skv_cur[t] = kvx;
    if (HALO && rh) {
      if (AXI) skv_cur[TX] = (h.k / radh) * (hvxn * radh);
      else skv_cur[TX] = h.k * hvxn;
    }
  }
  T kvy = T(0);
  if (rho >= R.lo) {
    const T d = (-st.bpy_prev) + st.bpy;
    vyn = st.in.vy;
    cvyn = st.in.cvy;
    if (KAHAN) kahan_sub(vyn, cvyn, d); else vyn = vyn - d;
    T dummy = T(0), dummyc = T(0);
    stream_events<T, KAHAN>(a.bev[2], a.prog[2], st.vc2, TX, -G, 0, rho, t, step, vyn, cvyn,
                            false, 0, dummy, dummyc);
    stream_record<T, KAHAN>(a.outp[2], st.oc2, TX, -G, 0, rho, t, orow, vyn, cvyn);
    kvy = st.in.ky * vyn;
  }
  pn = cpn = T(0);
  if (rho > R.lo) {    // p of row rho-1 (its x products were published an iteration ago)
    const T dxt = (-skv_prev[t]) + skv_prev[t + 1];
    const T dyt = (-st.kvy_prev) + kvy;
    const T d = dxt + dyt;
    pn = st.p_prev;
    cpn = st.cp_prev;
    if (KAHAN) kahan_sub(pn, cpn, d); else pn = pn - d;
  }
  st.p_prev = st.in.p;
  st.cp_prev = st.in.cp;
  st.kvy_prev = kvy;
  st.bpy_prev = st.bpy;
}

// S stages over one strip (S <= G + 1).
template <typename T, bool KAHAN, int TX, bool AXI, int S>
__device__ void acoustic_stream(const StepArgs<T> &a, int tile, int64_t step, int64_t row,
                                StreamLds<T, TX, S> &L) {
  auto &sbp = L.bwd;    // Bx*p of the row, columns -1..TX
  auto &skv = L.fwd;    // forward x products of the row, columns 0..TX
  const int H = a.strip_h, G = a.ghost;
  const int t = threadIdx.x;
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const int x0 = tx * (TX - 2 * G) - G, y0 = ty * H;
  const int nx = a.nx, n = (int)a.n;
  const bool iso_b = (a.flags & kFlagIsoB) != 0;
  const T *Kx = a.coef[FDS_C_KX], *Ky = a.coef[FDS_C_KY];
  const T *Bx = a.coef[FDS_C_BX], *By = iso_b ? Bx : a.coef[FDS_C_BY];
  const T *pin = a.in[0], *vxin = a.in[1], *vyin = a.in[2];
  const double dx = a.s[FDS_S_DX], hdx = a.s[FDS_S_HALF_DX];
  const bool lh = (t == 0), rh = (t == TX - 1);
  const bool owned = t >= G && t < TX - G && x0 + t < nx;
  T rad = T(1), radh = T(1);
  if (AXI) {
    rad = radius<T>(wrap_any(x0 + t, nx), dx, hdx);
    radh = radius<T>(wrap_any(x0 + TX, nx), dx, hdx);
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    for (int b = 0; b < 2; ++b) {
      sbp[s][b][t + 1] = T(0);
      skv[s][b][t] = T(0);
      if (lh) {
        sbp[s][b][0] = T(0);
        skv[s][b][TX] = T(0);
      }
    }
  }

  AcStage<T> st[S];
#pragma unroll
  for (int s = 0; s < S; ++s) ac_stage_init(st[s], a, tile);
  AcHalo<T> h{}, nh{};
  AcIn<T> nin{};

  // stage 1 input row rho into (nin, nh)
  auto load_row = [&](int rho) {
    const int j = (y0 + rho) * nx + x0 + t;
    nin.p = ld(pin, j, n);
    nin.vx = ld(vxin, j, n);
    nin.vy = ld(vyin, j, n);
    nin.bx = ld(Bx, j, n);
    nin.by = ld(By, j, n);
    nin.kx = ld(Kx, j, n);
    nin.ky = ld(Ky, j, n);
    if (KAHAN) {
      nin.cp = ld(a.cin[0], j, n);
      nin.cvx = ld(a.cin[1], j, n);
      nin.cvy = ld(a.cin[2], j, n);
    }
    if (lh) {
      nh.p = ld(pin, j - 1, n);
      nh.b = ld(Bx, j - 1, n);
      if (KAHAN) nh.cp = ld(a.cin[0], j - 1, n);
    }
    if (rh) {
      nh.p = ld(pin, j + 1, n);
      nh.b = ld(Bx, j + 1, n);
      nh.vx = ld(vxin, j + 1, n);
      nh.k = ld(Kx, j + 1, n);
      if (KAHAN) {
        nh.cp = ld(a.cin[0], j + 1, n);
        nh.cvx = ld(a.cin[1], j + 1, n);
      }
    }
  };

  load_row(-G - 1);
  const int iters = H + 2 * G + S + 1;
  for (int i = 0; i < iters; ++i) {
    const int buf = i & 1;
    // ---- before the barrier (stage 1 takes the prefetched row and prefetches the next)
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int rho = i - G - 1 - 2 * s;
      const StageRows R(G, H, s);
      if (rho < R.lo - 1 || rho > R.hi + 1) continue;
      if (s == 0) {
        st[0].in = nin;
        h = nh;
        if (rho < R.hi + 1) load_row(rho + 1);
        ac_pre<T, KAHAN, TX, true>(st[0], h, a, G, rho, R, step, row, sbp[0][buf]);
      } else {
        ac_pre<T, KAHAN, TX, false>(st[s], h, a, G, rho, R, step + s, row + s, sbp[s][buf]);
      }
    }
    __syncthreads();
    // ---- after the barrier: consumers first, so a stage hands its finished row to the next
    // stage once that stage is done with its current one
#pragma unroll
    for (int s = S - 1; s >= 0; --s) {
      const int rho = i - G - 1 - 2 * s;
      const StageRows R(G, H, s);
      if (rho < R.lo - 1 || rho > R.hi + 1) continue;
      T pn, cpn, vxn, cvxn, vyn, cvyn;
      if (s == 0)
        ac_post<T, KAHAN, TX, AXI, true>(st[0], h, a, G, rho, R, step, row, sbp[0][buf],
                                         skv[0][buf], skv[0][buf ^ 1], rad, radh, pn, cpn, vxn,
                                         cvxn, vyn, cvyn);
      else
        ac_post<T, KAHAN, TX, AXI, false>(st[s], h, a, G, rho, R, step + s, row + s,
                                          sbp[s][buf], skv[s][buf], skv[s][buf ^ 1], rad, radh,
                                          pn, cpn, vxn, cvxn, vyn, cvyn);
      if (s == S - 1) {
        // last stage: write the owned cells
        if (owned && rho >= 1 && rho - 1 < H && y0 + rho - 1 < a.ny) {
          const int jp = (y0 + rho - 1) * nx + x0 + t;
          a.out[0][jp] = pn;
          if (KAHAN) a.cout[0][jp] = cpn;
        }
        if (owned && rho >= 0 && rho < H && y0 + rho < a.ny) {
          const int jv = (y0 + rho) * nx + x0 + t;
          a.out[1][jv] = vxn;
          a.out[2][jv] = vyn;
          if (KAHAN) {
            a.cout[1][jv] = cvxn;
            a.cout[2][jv] = cvyn;
          }
        }
      } else {
        // hand row rho-1 (p just updated; v and coefficients held) to the next stage
        AcIn<T> &nxt = st[s + 1 < S ? s + 1 : s].in;
        nxt = st[s].hold;
        nxt.p = pn;
        nxt.cp = cpn;
      }
      AcIn<T> &hd = st[s].hold;
      hd.vx = vxn;
      hd.cvx = cvxn;
      hd.vy = vyn;
      hd.cvy = cvyn;
      hd.bx = st[s].in.bx;
      hd.by = st[s].in.by;
      hd.kx = st[s].in.kx;
      hd.ky = st[s].in.ky;
    }
  }
}

// ------------------------------------------------------------------------------------------
// Thermal: q = -(Dbwd(Q) T) assigned, T -= Dfwd(G) q
// ------------------------------------------------------------------------------------------
template <typename T>
struct ThIn {
  T t, ct, qx, qy, gx, gy;
};

template <typename T>
struct ThHalo {        // left thread: T, Q of column -1; right thread: T, Q, G of column TX
  T t, ct, q, g;
};

template <typename T>
struct ThStage {
  ThIn<T> in, hold;
  T qtx, qty, qty_prev, t_prev, ct_prev, gqy_prev;
  Cursor pc, vc1, vc2, oc0, oc1, oc2;
};

template <typename T>
__device__ __forceinline__ void th_stage_init(ThStage<T> &st, const StepArgs<T> &a, int tile) {
  st.pc = cursor(a.bev[0], tile);
  st.vc1 = cursor(a.bev[1], tile);
  st.vc2 = cursor(a.bev[2], tile);
  st.oc0 = cursor(a.outp[0].ev, tile);
  st.oc1 = cursor(a.outp[1].ev, tile);
  st.oc2 = cursor(a.outp[2].ev, tile);
  st.qtx = st.qty = st.qty_prev = st.t_prev = st.ct_prev = st.gqy_prev = T(0);
  st.in = ThIn<T>{};
  st.hold = ThIn<T>{};
}

template <typename T, bool KAHAN, int TX, bool HALO>
__device__ __forceinline__ void th_pre(ThStage<T> &st, ThHalo<T> &h, const StepArgs<T> &a,
                                       int G, int rho, const StageRows &R, int64_t step,
                                       int64_t orow, T *sqt) {
  const int t = threadIdx.x;
  const bool lh = (t == 0), rh = (t == TX - 1);
  stream_events<T, KAHAN>(a.bev[0], a.prog[0], st.pc, TX + 2, -G - 1, -1, rho, t, step, st.in.t,
                          st.in.ct, HALO && (lh || rh), lh ? -1 : TX, h.t, h.ct);
  stream_record<T, KAHAN>(a.outp[0], st.oc0, TX + 2, -G - 1, -1, rho, t, orow, st.in.t,
                          st.in.ct);
  st.qtx = st.in.qx * st.in.t;
  st.qty = st.in.qy * st.in.t;
  if (rho <= R.hi) {
    sqt[t + 1] = st.qtx;
    if (HALO && lh) sqt[0] = h.q * h.t;
  }
}

template <typename T, bool KAHAN, int TX, bool AXI, bool HALO>
__device__ __forceinline__ void th_post(ThStage<T> &st, const ThHalo<T> &h, const StepArgs<T> &a,
                                        int G, int rho, const StageRows &R, int64_t step,
                                        int64_t orow, const T *sqt, T *sgq_cur,
                                        const T *sgq_prev, T rad, T radh, T &tn, T &ctn, T &fx,
                                        T &fy) {
  const int t = threadIdx.x;
  const bool rh = (t == TX - 1);
  T zero = T(0), zeroc = T(0), dummy = T(0), dummyc = T(0);
  fx = fy = T(0);
  T hfx = T(0);
  if (rho >= R.lo && rho <= R.hi) {
    fx = -((-sqt[t]) + st.qtx);
    if (HALO && rh) hfx = -((-st.qtx) + h.q * h.t);
    stream_events<T, false>(a.bev[1], a.prog[1], st.vc1, TX + 1, -G, 0, rho, t, step, fx, zeroc,
                            HALO && rh, TX, hfx, zero);
    stream_record<T, false>(a.outp[1], st.oc1, TX + 1, -G, 0, rho, t, orow, fx, zeroc);
    T gqx;
    if (AXI) gqx = (st.in.gx / rad) * (fx * rad);
    else gqx = st.in.gx * fx;
    sgq_cur[t] = gqx;
    if (HALO && rh) {
      if (AXI) sgq_cur[TX] = (h.g / radh) * (hfx * radh);
      else sgq_cur[TX] = h.g * hfx;
    }
  }
  T gqy = T(0);
  if (rho >= R.lo) {
    fy = -((-st.qty_prev) + st.qty);
    stream_events<T, false>(a.bev[2], a.prog[2], st.vc2, TX, -G, 0, rho, t, step, fy, zeroc,
                            false, 0, dummy, dummyc);
    stream_record<T, false>(a.outp[2], st.oc2, TX, -G, 0, rho, t, orow, fy, zeroc);
    gqy = st.in.gy * fy;
  }
  tn = ctn = T(0);
  if (rho > R.lo) {
    const T dxt = (-sgq_prev[t]) + sgq_prev[t + 1];
    const T dyt = (-st.gqy_prev) + gqy;
    const T d = dxt + dyt;
    tn = st.t_prev;
    ctn = st.ct_prev;
    if (KAHAN) kahan_sub(tn, ctn, d); else tn = tn - d;
  }
  st.t_prev = st.in.t;
  st.ct_prev = st.in.ct;
  st.gqy_prev = gqy;
  st.qty_prev = st.qty;
}

template <typename T, bool KAHAN, int TX, bool AXI, int S>
__device__ void thermal_stream(const StepArgs<T> &a, int tile, int64_t step, int64_t row,
                               StreamLds<T, TX, S> &L) {
  auto &sqt = L.bwd;    // Qx*T of the row
  auto &sgq = L.fwd;    // forward x products of the row
  const int H = a.strip_h, G = a.ghost;
  const int t = threadIdx.x;
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const int x0 = tx * (TX - 2 * G) - G, y0 = ty * H;
  const int nx = a.nx, n = (int)a.n;
  const bool iso_b = (a.flags & kFlagIsoB) != 0;
  const bool store = (a.flags & kFlagStore) != 0;
  const T *Gx = a.coef[FDS_C_GX], *Gy = a.coef[FDS_C_GY];
  const T *Qx = a.coef[FDS_C_QX], *Qy = iso_b ? Qx : a.coef[FDS_C_QY];
  const T *tin = a.in[0];
  const double dx = a.s[FDS_S_DX], hdx = a.s[FDS_S_HALF_DX];
  const bool lh = (t == 0), rh = (t == TX - 1);
  const bool owned = t >= G && t < TX - G && x0 + t < nx;
  T rad = T(1), radh = T(1);
  if (AXI) {
    rad = radius<T>(wrap_any(x0 + t, nx), dx, hdx);
    radh = radius<T>(wrap_any(x0 + TX, nx), dx, hdx);
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    for (int b = 0; b < 2; ++b) {
      sqt[s][b][t + 1] = T(0);
      sgq[s][b][t] = T(0);
      if (lh) {
        sqt[s][b][0] = T(0);
        sgq[s][b][TX] = T(0);
      }
    }
  }

  ThStage<T> st[S];
#pragma unroll
  for (int s = 0; s < S; ++s) th_stage_init(st[s], a, tile);
  ThHalo<T> h{}, nh{};
  ThIn<T> nin{};

  auto load_row = [&](int rho) {
    const int j = (y0 + rho) * nx + x0 + t;
    nin.t = ld(tin, j, n);
    nin.qx = ld(Qx, j, n);
    nin.qy = ld(Qy, j, n);
    nin.gx = ld(Gx, j, n);
    nin.gy = ld(Gy, j, n);
    if (KAHAN) nin.ct = ld(a.cin[0], j, n);
    if (lh) {
      nh.t = ld(tin, j - 1, n);
      nh.q = ld(Qx, j - 1, n);
      if (KAHAN) nh.ct = ld(a.cin[0], j - 1, n);
    }
    if (rh) {
      nh.t = ld(tin, j + 1, n);
      nh.q = ld(Qx, j + 1, n);
      nh.g = ld(Gx, j + 1, n);
      if (KAHAN) nh.ct = ld(a.cin[0], j + 1, n);
    }
  };

  load_row(-G - 1);
  const int iters = H + 2 * G + S + 1;
  for (int i = 0; i < iters; ++i) {
    const int buf = i & 1;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int rho = i - G - 1 - 2 * s;
      const StageRows R(G, H, s);
      if (rho < R.lo - 1 || rho > R.hi + 1) continue;
      if (s == 0) {
        st[0].in = nin;
        h = nh;
        if (rho < R.hi + 1) load_row(rho + 1);
        th_pre<T, KAHAN, TX, true>(st[0], h, a, G, rho, R, step, row, sqt[0][buf]);
      } else {
        th_pre<T, KAHAN, TX, false>(st[s], h, a, G, rho, R, step + s, row + s, sqt[s][buf]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int s = S - 1; s >= 0; --s) {
      const int rho = i - G - 1 - 2 * s;
      const StageRows R(G, H, s);
      if (rho < R.lo - 1 || rho > R.hi + 1) continue;
      T tn, ctn, fx, fy;
      if (s == 0)
        th_post<T, KAHAN, TX, AXI, true>(st[0], h, a, G, rho, R, step, row, sqt[0][buf],
                                         sgq[0][buf], sgq[0][buf ^ 1], rad, radh, tn, ctn, fx,
                                         fy);
      else
        th_post<T, KAHAN, TX, AXI, false>(st[s], h, a, G, rho, R, step + s, row + s,
                                          sqt[s][buf], sgq[s][buf], sgq[s][buf ^ 1], rad, radh,
                                          tn, ctn, fx, fy);
      if (s == S - 1) {
        if (owned && rho >= 1 && rho - 1 < H && y0 + rho - 1 < a.ny) {
          const int jp = (y0 + rho - 1) * nx + x0 + t;
          a.out[0][jp] = tn;
          if (KAHAN) a.cout[0][jp] = ctn;
        }
        if (store && owned && rho >= 0 && rho < H && y0 + rho < a.ny) {
          const int jv = (y0 + rho) * nx + x0 + t;
          a.out[1][jv] = fx;
          a.out[2][jv] = fy;
        }
      } else {
        ThIn<T> &nxt = st[s + 1 < S ? s + 1 : s].in;
        nxt = st[s].hold;
        nxt.t = tn;
        nxt.ct = ctn;
      }
      ThIn<T> &hd = st[s].hold;
      hd.qx = st[s].in.qx;
      hd.qy = st[s].in.qy;
      hd.gx = st[s].in.gx;
      hd.gy = st[s].in.gy;
    }
  }
}

// One tile of instance `a` in a batched streaming launch (lossless models only).
template <typename T, bool KAHAN, int TX, int S>
__device__ __forceinline__ void any_stream(const StepArgs<T> &a, int tile, int64_t step,
                                           int64_t row, StreamLds<T, TX, S> &L) {
  const int f = a.flags;
  if (f & kFlagThermal) {
    if (f & kFlagAxi) thermal_stream<T, KAHAN, TX, true, S>(a, tile, step, row, L);
    else thermal_stream<T, KAHAN, TX, false, S>(a, tile, step, row, L);
  } else if (f & kFlagAxi) {
    acoustic_stream<T, KAHAN, TX, true, S>(a, tile, step, row, L);
  } else {
    acoustic_stream<T, KAHAN, TX, false, S>(a, tile, step, row, L);
  }
}

}  // namespace fds

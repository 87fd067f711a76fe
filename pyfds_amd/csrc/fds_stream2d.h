// Row-streaming step kernels for the lossless 2D models (Acoustic2D / Acoustic3DAxi and
// Thermal2D / Thermal3DAxi).  A workgroup of TX threads owns TX columns [x0, x0+TX) of a strip
// of H rows [y0, y0+H) and marches down it: thread t keeps its column's values in registers,
// prefetches the next row while it computes the current one, and exchanges the x-neighbour
// products through two double-buffered LDS rows, so one row costs one barrier and the only
// re-read halo is one row above and one below the strip.
//
// Arithmetic is the tile kernel's, term by term.  Column-indexed differences use the product
// identity (-a)*b == -(a*b): the backward difference (-B[j-1] p[j-1]) + B[j] p[j] is formed from
// the products bp = B*p of both columns, and the forward difference from k*v products, so FP64
// results stay bit-identical to the reference.
//
// Events (boundaries, outputs) are the tile engine's lists for a TX x H tile: sorted by row,
// consumed with a uniform cursor, applied by the thread that owns the column (thread 0 for the
// left halo column -1, thread TX-1 for the right halo column TX).
//
// Row pipeline for row r = 0..H (row H only for vy / q_y): load row r+1; boundaries of p (T) in
// row r; record; backward differences -> v (q) of row r; their boundaries and outputs; write row
// r; forward products of row r; update p (T) of row r-1 from the products of rows r-1 and r.
#pragma once

#include "fds_common.h"
#include "fds_step2d.h"
#include "../../include/pyfds_amd.h"

namespace fds {

// Apply every event of row r of a region (row width w, origin (rmin, cmin)) whose column is
// `col` (or the halo column `hcol` when has_halo) to (x, c) / (hx, hc).
template <typename T, bool KAHAN>
__device__ __forceinline__ void stream_events(const EvView &ev, const ProgView &pg, int &cur,
                                              int end, int w, int rmin, int cmin, int r, int col,
                                              int64_t step, T &x, T &c, bool has_halo, int hcol,
                                              T &hx, T &hc) {
  int e = cur;
  while (e < end) {
    const int pos = ev.pos[e];
    const int er = pos / w + rmin;
    if (er != r) break;
    const int ec = pos - (er - rmin) * w + cmin;
    if (ec == col) apply_prog<T, KAHAN>(pg, ev.id[e], step, x, c);
    if (has_halo && ec == hcol) apply_prog<T, KAHAN>(pg, ev.id[e], step, hx, hc);
    ++e;
  }
  cur = e;
}

template <typename T, bool KAHAN>
__device__ __forceinline__ void stream_record(const OutView &o, int &cur, int end, int w,
                                              int rmin, int cmin, int r, int col, int64_t row,
                                              T x, T c) {
  int e = cur;
  while (e < end) {
    const int pos = o.ev.pos[e];
    const int er = pos / w + rmin;
    if (er != r) break;
    const int ec = pos - (er - rmin) * w + cmin;
    if (ec == col) o.buf[row * o.m + o.ev.id[e]] = out_value<T, KAHAN>(x, c);
    ++e;
  }
  cur = e;
}

template <typename T>
__device__ __forceinline__ T ld(const T *p, int j, int n) {
  return (j >= 0 && j < n) ? p[j] : T(0);
}

// ------------------------------------------------------------------------------------------
// Acoustic (lossless)
// ------------------------------------------------------------------------------------------
template <typename T, bool KAHAN, int TX, bool AXI>
__device__ void acoustic_stream(const StepArgs<T> &a, int tile, int64_t step, int64_t row) {
  const int H = a.strip_h;
  __shared__ T sbp[2][TX + 2];    // Bx*p of the row, columns -1..TX
  __shared__ T skv[2][TX + 1];    // forward x products of the row, columns 0..TX

  const int t = threadIdx.x;
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const int x0 = tx * TX, y0 = ty * H;
  const int nx = a.nx, n = (int)a.n;
  const bool iso_b = (a.flags & kFlagIsoB) != 0;
  const T *Kx = a.coef[FDS_C_KX], *Ky = a.coef[FDS_C_KY];
  const T *Bx = a.coef[FDS_C_BX], *By = iso_b ? Bx : a.coef[FDS_C_BY];
  const T *pin = a.in[0], *vxin = a.in[1], *vyin = a.in[2];
  const double dx = a.s[FDS_S_DX], hdx = a.s[FDS_S_HALF_DX];
  const bool lh = (t == 0), rh = (t == TX - 1);     // halo owners
  const bool colok = x0 + t < nx;
  // radius of this column and of the right halo column (Axi)
  T rad = T(1), radh = T(1);
  if (AXI) {
    rad = radius<T>(wrap_col(x0 + t, nx), dx, hdx);
    radh = radius<T>(wrap_col(x0 + TX, nx), dx, hdx);
  }
  // event cursors
  const int pe = a.bev[0].ptr[tile + 1], ve1 = a.bev[1].ptr[tile + 1], ve2 = a.bev[2].ptr[tile + 1];
  int pc = a.bev[0].ptr[tile], vc1 = a.bev[1].ptr[tile], vc2 = a.bev[2].ptr[tile];
  const int oe0 = a.outp[0].ev.ptr[tile + 1], oe1 = a.outp[1].ev.ptr[tile + 1];
  const int oe2 = a.outp[2].ev.ptr[tile + 1];
  int oc0 = a.outp[0].ev.ptr[tile], oc1 = a.outp[1].ev.ptr[tile], oc2 = a.outp[2].ev.ptr[tile];
  const int WP = TX + 2, WX = TX + 1, WY = TX;

  // ---- prologue: row -1 (p, By for vy of row 0)
  int j = (y0 - 1) * nx + x0 + t;
  T p = ld(pin, j, n), cpv = KAHAN ? ld(a.cin[0], j, n) : T(0);
  T dummy = T(0), dummyc = T(0);
  stream_events<T, KAHAN>(a.bev[0], a.prog[0], pc, pe, WP, -1, -1, -1, t, step, p, cpv, false,
                          0, dummy, dummyc);
  T bpy_prev = ld(By, j, n) * p;
  // current row 0 loads
  j += nx;
  T n_p = ld(pin, j, n), n_bx = ld(Bx, j, n), n_by = ld(By, j, n);
  T n_vx = ld(vxin, j, n), n_vy = ld(vyin, j, n), n_kx = ld(Kx, j, n), n_ky = ld(Ky, j, n);
  T n_cp = KAHAN ? ld(a.cin[0], j, n) : T(0), n_cvx = KAHAN ? ld(a.cin[1], j, n) : T(0);
  T n_cvy = KAHAN ? ld(a.cin[2], j, n) : T(0);
  T n_hp = T(0), n_hb = T(0), n_hvx = T(0), n_hcvx = T(0), n_hk = T(0);
  if (lh) { n_hp = ld(pin, j - 1, n); n_hb = ld(Bx, j - 1, n); }
  if (rh) {
    n_hp = ld(pin, j + 1, n); n_hb = ld(Bx, j + 1, n); n_hvx = ld(vxin, j + 1, n);
    n_hk = ld(Kx, j + 1, n);
    if (KAHAN) n_hcvx = ld(a.cin[1], j + 1, n);
  }
  T n_hcp = (KAHAN && (lh || rh)) ? ld(a.cin[0], lh ? j - 1 : j + 1, n) : T(0);

  T p_prev = T(0), cp_prev = T(0), kvy_prev = T(0);
  for (int r = 0; r <= H; ++r) {
    const int jr = j;   // flat index of (r, t)
    // this row's values
    T pr = n_p, bx = n_bx, by = n_by, vx = n_vx, vy = n_vy, kx = n_kx, ky = n_ky;
    T cp = n_cp, cvx = n_cvx, cvy = n_cvy;
    T hp = n_hp, hb = n_hb, hvx = n_hvx, hcvx = n_hcvx, hk = n_hk, hcp = n_hcp;
    // prefetch row r + 1
    if (r < H) {
      j += nx;
      n_p = ld(pin, j, n); n_by = ld(By, j, n); n_vy = ld(vyin, j, n); n_ky = ld(Ky, j, n);
      if (KAHAN) { n_cp = ld(a.cin[0], j, n); n_cvy = ld(a.cin[2], j, n); }
      if (r + 1 < H) {
        n_bx = ld(Bx, j, n); n_vx = ld(vxin, j, n); n_kx = ld(Kx, j, n);
        if (KAHAN) n_cvx = ld(a.cin[1], j, n);
        if (lh) { n_hp = ld(pin, j - 1, n); n_hb = ld(Bx, j - 1, n); }
        if (rh) {
          n_hp = ld(pin, j + 1, n); n_hb = ld(Bx, j + 1, n); n_hvx = ld(vxin, j + 1, n);
          n_hk = ld(Kx, j + 1, n);
          if (KAHAN) n_hcvx = ld(a.cin[1], j + 1, n);
        }
        if (KAHAN && (lh || rh)) n_hcp = ld(a.cin[0], lh ? j - 1 : j + 1, n);
      }
    }
    const int hcol = lh ? -1 : TX;
    // boundaries of p in row r (this column and the halo column), then record
    stream_events<T, KAHAN>(a.bev[0], a.prog[0], pc, pe, WP, -1, -1, r, t, step, pr, cp,
                            lh || rh, hcol, hp, hcp);
    if (r < H) stream_record<T, KAHAN>(a.outp[0], oc0, oe0, WP, -1, -1, r, t, row, pr, cp);
    const int buf = r & 1;
    const T bpx = bx * pr;
    if (r < H) {
      sbp[buf][t + 1] = bpx;
      if (lh) sbp[buf][0] = hb * hp;
    }
    const T bpy = by * pr;
    __syncthreads();
    // velocities of row r (vx only for r < H)
    T vxn = T(0), cvxn = T(0), hvxn = T(0), hcvxn = T(0);
    const bool in_r = jr < n;
    if (r < H && in_r) {
      const T d = (-sbp[buf][t]) + bpx;
      vxn = vx;
      cvxn = cvx;
      if (KAHAN) kahan_sub(vxn, cvxn, d); else vxn = vxn - d;
    }
    if (r < H && rh && jr + 1 < n) {
      const T d = (-bpx) + hb * hp;
      hvxn = hvx;
      hcvxn = hcvx;
      if (KAHAN) kahan_sub(hvxn, hcvxn, d); else hvxn = hvxn - d;
    }
    T vyn = T(0), cvyn = T(0);
    if (in_r) {
      const T d = (-bpy_prev) + bpy;
      vyn = vy;
      cvyn = cvy;
      if (KAHAN) kahan_sub(vyn, cvyn, d); else vyn = vyn - d;
    }
    if (r < H) {
      stream_events<T, KAHAN>(a.bev[1], a.prog[1], vc1, ve1, WX, 0, 0, r, t, step, vxn, cvxn,
                              rh, TX, hvxn, hcvxn);
      stream_record<T, KAHAN>(a.outp[1], oc1, oe1, WX, 0, 0, r, t, row, vxn, cvxn);
    }
    stream_events<T, KAHAN>(a.bev[2], a.prog[2], vc2, ve2, WY, 0, 0, r, t, step, vyn, cvyn,
                            false, 0, dummy, dummyc);
    if (r < H) {
      stream_record<T, KAHAN>(a.outp[2], oc2, oe2, WY, 0, 0, r, t, row, vyn, cvyn);
      if (colok && y0 + r < a.ny) {
        a.out[1][jr] = vxn;
        a.out[2][jr] = vyn;
        if (KAHAN) { a.cout[1][jr] = cvxn; a.cout[2][jr] = cvyn; }
      }
      // forward x products of row r
      T kvx;
      if (AXI) kvx = (kx / rad) * (vxn * rad);
      else kvx = kx * vxn;
      skv[buf][t] = kvx;
      if (rh) {
        if (AXI) skv[buf][TX] = (hk / radh) * (hvxn * radh);
        else skv[buf][TX] = hk * hvxn;
      }
    }
    const T kvy = ky * vyn;
    // pressure of row r - 1 (its x products were published before this row's barrier)
    if (r >= 1 && colok && y0 + r - 1 < a.ny) {
      const int pb = buf ^ 1;
      const T dxt = (-skv[pb][t]) + skv[pb][t + 1];
      const T dyt = (-kvy_prev) + kvy;
      const T d = dxt + dyt;
      T pn = p_prev, cpn = cp_prev;
      if (KAHAN) kahan_sub(pn, cpn, d); else pn = pn - d;
      a.out[0][jr - nx] = pn;
      if (KAHAN) a.cout[0][jr - nx] = cpn;
    }
    p_prev = pr;
    cp_prev = cp;
    kvy_prev = kvy;
    bpy_prev = bpy;
  }
}

// ------------------------------------------------------------------------------------------
// Thermal: q = -(Dbwd(Q) T) assigned, T -= Dfwd(G) q
// ------------------------------------------------------------------------------------------
template <typename T, bool KAHAN, int TX, bool AXI>
__device__ void thermal_stream(const StepArgs<T> &a, int tile, int64_t step, int64_t row) {
  const int H = a.strip_h;
  __shared__ T sqt[2][TX + 2];
  __shared__ T sgq[2][TX + 1];

  const int t = threadIdx.x;
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const int x0 = tx * TX, y0 = ty * H;
  const int nx = a.nx, n = (int)a.n;
  const bool iso_b = (a.flags & kFlagIsoB) != 0;
  const bool store = (a.flags & kFlagStore) != 0;
  const T *Gx = a.coef[FDS_C_GX], *Gy = a.coef[FDS_C_GY];
  const T *Qx = a.coef[FDS_C_QX], *Qy = iso_b ? Qx : a.coef[FDS_C_QY];
  const T *tin = a.in[0];
  const double dx = a.s[FDS_S_DX], hdx = a.s[FDS_S_HALF_DX];
  const bool lh = (t == 0), rh = (t == TX - 1);
  const bool colok = x0 + t < nx;
  T rad = T(1), radh = T(1);
  if (AXI) {
    rad = radius<T>(wrap_col(x0 + t, nx), dx, hdx);
    radh = radius<T>(wrap_col(x0 + TX, nx), dx, hdx);
  }
  const int pe = a.bev[0].ptr[tile + 1], ve1 = a.bev[1].ptr[tile + 1], ve2 = a.bev[2].ptr[tile + 1];
  int pc = a.bev[0].ptr[tile], vc1 = a.bev[1].ptr[tile], vc2 = a.bev[2].ptr[tile];
  const int oe0 = a.outp[0].ev.ptr[tile + 1], oe1 = a.outp[1].ev.ptr[tile + 1];
  const int oe2 = a.outp[2].ev.ptr[tile + 1];
  int oc0 = a.outp[0].ev.ptr[tile], oc1 = a.outp[1].ev.ptr[tile], oc2 = a.outp[2].ev.ptr[tile];
  const int WP = TX + 2, WX = TX + 1, WY = TX;

  int j = (y0 - 1) * nx + x0 + t;
  T tv = ld(tin, j, n), ctv = KAHAN ? ld(a.cin[0], j, n) : T(0);
  T dummy = T(0), dummyc = T(0);
  stream_events<T, KAHAN>(a.bev[0], a.prog[0], pc, pe, WP, -1, -1, -1, t, step, tv, ctv, false,
                          0, dummy, dummyc);
  T qty_prev = ld(Qy, j, n) * tv;
  j += nx;
  T n_t = ld(tin, j, n), n_qx = ld(Qx, j, n), n_qy = ld(Qy, j, n);
  T n_gx = ld(Gx, j, n), n_gy = ld(Gy, j, n), n_ct = KAHAN ? ld(a.cin[0], j, n) : T(0);
  T n_ht = T(0), n_hq = T(0), n_hg = T(0), n_hct = T(0);
  if (lh) { n_ht = ld(tin, j - 1, n); n_hq = ld(Qx, j - 1, n); }
  if (rh) { n_ht = ld(tin, j + 1, n); n_hq = ld(Qx, j + 1, n); n_hg = ld(Gx, j + 1, n); }
  if (KAHAN && (lh || rh)) n_hct = ld(a.cin[0], lh ? j - 1 : j + 1, n);

  T t_prev = T(0), ct_prev = T(0), gqy_prev = T(0);
  for (int r = 0; r <= H; ++r) {
    const int jr = j;
    T tr = n_t, qx = n_qx, qy = n_qy, gx = n_gx, gy = n_gy, ct = n_ct;
    T ht = n_ht, hq = n_hq, hg = n_hg, hct = n_hct;
    if (r < H) {
      j += nx;
      n_t = ld(tin, j, n); n_qy = ld(Qy, j, n); n_gy = ld(Gy, j, n);
      if (KAHAN) n_ct = ld(a.cin[0], j, n);
      if (r + 1 < H) {
        n_qx = ld(Qx, j, n); n_gx = ld(Gx, j, n);
        if (lh) { n_ht = ld(tin, j - 1, n); n_hq = ld(Qx, j - 1, n); }
        if (rh) { n_ht = ld(tin, j + 1, n); n_hq = ld(Qx, j + 1, n); n_hg = ld(Gx, j + 1, n); }
        if (KAHAN && (lh || rh)) n_hct = ld(a.cin[0], lh ? j - 1 : j + 1, n);
      }
    }
    const int hcol = lh ? -1 : TX;
    stream_events<T, KAHAN>(a.bev[0], a.prog[0], pc, pe, WP, -1, -1, r, t, step, tr, ct,
                            lh || rh, hcol, ht, hct);
    if (r < H) stream_record<T, KAHAN>(a.outp[0], oc0, oe0, WP, -1, -1, r, t, row, tr, ct);
    const int buf = r & 1;
    const T qtx = qx * tr;
    if (r < H) {
      sqt[buf][t + 1] = qtx;
      if (lh) sqt[buf][0] = hq * ht;
    }
    const T qty = qy * tr;
    __syncthreads();
    const bool in_r = jr < n;
    T fx = T(0), hfx = T(0), fy = T(0);
    T zero = T(0), zeroc = T(0);
    if (r < H && in_r) fx = -((-sqt[buf][t]) + qtx);
    if (r < H && rh && jr + 1 < n) hfx = -((-qtx) + hq * ht);
    if (in_r) fy = -((-qty_prev) + qty);
    if (r < H) {
      stream_events<T, false>(a.bev[1], a.prog[1], vc1, ve1, WX, 0, 0, r, t, step, fx, zeroc,
                              rh, TX, hfx, zero);
      stream_record<T, false>(a.outp[1], oc1, oe1, WX, 0, 0, r, t, row, fx, zeroc);
    }
    stream_events<T, false>(a.bev[2], a.prog[2], vc2, ve2, WY, 0, 0, r, t, step, fy, zeroc,
                            false, 0, dummy, dummyc);
    if (r < H) {
      stream_record<T, false>(a.outp[2], oc2, oe2, WY, 0, 0, r, t, row, fy, zeroc);
      if (store && colok && y0 + r < a.ny) {
        a.out[1][jr] = fx;
        a.out[2][jr] = fy;
      }
      T gqx;
      if (AXI) gqx = (gx / rad) * (fx * rad);
      else gqx = gx * fx;
      sgq[buf][t] = gqx;
      if (rh) {
        if (AXI) sgq[buf][TX] = (hg / radh) * (hfx * radh);
        else sgq[buf][TX] = hg * hfx;
      }
    }
    const T gqy = gy * fy;
    if (r >= 1 && colok && y0 + r - 1 < a.ny) {
      const int pb = buf ^ 1;
      const T dxt = (-sgq[pb][t]) + sgq[pb][t + 1];
      const T dyt = (-gqy_prev) + gqy;
      const T d = dxt + dyt;
      T tn = t_prev, ctn = ct_prev;
      if (KAHAN) kahan_sub(tn, ctn, d); else tn = tn - d;
      a.out[0][jr - nx] = tn;
      if (KAHAN) a.cout[0][jr - nx] = ctn;
    }
    t_prev = tr;
    ct_prev = ct;
    gqy_prev = gqy;
    qty_prev = qty;
  }
}

// One tile of instance `a` in a batched streaming launch (lossless models only).
template <typename T, bool KAHAN, int TX>
__device__ __forceinline__ void any_stream(const StepArgs<T> &a, int tile, int64_t step,
                                           int64_t row) {
  const int f = a.flags;
  if (f & kFlagThermal) {
    if (f & kFlagAxi) thermal_stream<T, KAHAN, TX, true>(a, tile, step, row);
    else thermal_stream<T, KAHAN, TX, false>(a, tile, step, row);
  } else if (f & kFlagAxi) {
    acoustic_stream<T, KAHAN, TX, true>(a, tile, step, row);
  } else {
    acoustic_stream<T, KAHAN, TX, false>(a, tile, step, row);
  }
}

}  // namespace fds

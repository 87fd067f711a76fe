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

#include <type_traits>

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

// Cell j of an array of n cells, zero outside [0, n) (the reference's zero ghosts at the array
// ends).  CHK = false for workgroups whose reads all fall inside the array.
template <bool CHK, typename T>
__device__ __forceinline__ T ld(const T *p, int j, int n) {
  if (!CHK) return p[j];
  return (j >= 0 && j < n) ? p[j] : T(0);
}

// Cell rb + off with rb uniform across the workgroup (a row base): unchecked loads address a
// uniform row pointer plus the per-thread offset.
template <bool CHK, typename T>
__device__ __forceinline__ T ldr(const T *p, int rb, int off, int n) {
  if (!CHK) return (p + rb)[off];
  return ld<true>(p, rb + off, n);
}

// Read of a state value that the launch reads once (NT: non-temporal, it does not displace
// the coefficients and ghost rows that later stages and neighbouring strips re-read).
template <bool NT, bool CHK, typename T>
__device__ __forceinline__ T ldn(const T *p, int rb, int off, int n) {
  if (!NT) return ldr<CHK>(p, rb, off, n);
  if (!CHK) return __builtin_nontemporal_load(p + rb + off);
  const int j = rb + off;
  return (j >= 0 && j < n) ? __builtin_nontemporal_load(p + j) : T(0);
}

// Store of a new state value (NT: non-temporal, nothing reads it again within the launch).
template <bool NT, typename T>
__device__ __forceinline__ void stn(T *p, int rb, int off, T v) {
  if (NT) __builtin_nontemporal_store(v, p + rb + off);
  else (p + rb)[off] = v;
}

// ------------------------------------------------------------------------------------------
// LDS rings (shapes with RING): every input row of a strip reaches the workgroup by direct
// global-to-LDS loads (global_load_lds_dwordx4, 16 bytes per lane), PF rows ahead of the row
// stage 0 consumes, so loads in flight hold no registers and the loop issues no load into a
// register (whose use the compiler would guard with a wait for all loads in flight).  The
// state arrays of a row go to slot r % RS of the state ring (RS = PF + 1, read by stage 0);
// the coefficient arrays to slot r % RC of the coefficient ring (RC = PF + 2S - 1), where the
// later stages read them 2s iterations after stage 0 (r = row ordinal in the strip).
//
// Window: CW elements (CH chunks of E) from the aligned flat index at or below column -1, so
// every chunk is 16-byte aligned: this needs nx to be a multiple of E and 16-byte-aligned
// arrays (the host checks both), and then every chunk lies wholly inside or wholly outside
// the array.  A chunk outside the array loads the array's first chunk instead (the window
// reaches up to E - 1 cells before column -1 and CW - TX - 2 - lead cells after column TX,
// beyond what decides whether a tile reads outside the array), and the reads of a tile with
// out-of-array cells turn those cells into zeros (as ld<true>).
//
// Completion: vector-memory loads leave the wave's vmcnt in issue order, so a row's loads are
// complete once at most NA * (rows issued after it) loads remain outstanding; stores and the
// event-list loads of tiles with events only make that wait stricter.  Empty asm statements
// with a memory clobber keep the compiler from moving memory operations across the wait.
// ------------------------------------------------------------------------------------------
// Ring load modes: kRingX4 moves 16-byte chunks of an aligned window (one load per array row
// for 64-wide tiles; rows must start on 16-byte boundaries), kRingDword one dword per lane of
// the exact window of columns -1 .. TX (any row start: batches of fields of any width).
constexpr int kRingX4 = 1, kRingDword = 2;

template <typename T, int TX, int MODE>
struct RingGeom {
  static constexpr int E = MODE == kRingX4 ? 16 / (int)sizeof(T) : 1;   // window alignment
  static constexpr int CW = MODE == kRingX4 ? (TX + 2 + E - 1 + E - 1) / E * E
                                            : (TX + 2 + 1) / 2 * 2;   // elements per array row
  static constexpr int CH = CW / E;                                   // kRingX4: lanes per row
  static constexpr int DW = (TX + 2) * (int)sizeof(T) / 4;            // kRingDword: dwords
  static constexpr int IPA = MODE == kRingX4 ? 1 : (DW + 63) / 64;     // loads per array row
  static_assert(MODE != kRingX4 || CH <= TX, "one ring load per array row needs CH <= TX");
};

// elements of the two rings of a strip kernel: NS state and NC coefficient arrays
template <typename T, int TX, int PF, int S, int NS, int NC, int MODE>
constexpr int ring_elems() {
  return ((PF + 1) * NS + (PF + 2 * S - 1) * NC) * RingGeom<T, TX, MODE>::CW;
}

template <int K>
__device__ __forceinline__ void vm_wait() {
  static_assert(K >= 0 && K < 64, "vmcnt range");
  // s_waitcnt fields: vmcnt[3:0] and [15:14]; expcnt[6:4] = 7 and lgkmcnt[11:8] = 15 (no wait)
  __builtin_amdgcn_s_waitcnt((K & 15) | (7 << 4) | (15 << 8) | ((K >> 4) << 14));
}

__device__ __forceinline__ void compiler_fence() { __asm__ __volatile__("" ::: "memory"); }

// Wait until at most rows * NA loads are outstanding (rows <= R).
template <int NA, int R>
__device__ __forceinline__ void ring_wait(int rows) {
  if constexpr (R == 0) {
    vm_wait<0>();
  } else {
    if (rows >= R) vm_wait<R * NA>();
    else ring_wait<NA, R - 1>(rows);
  }
}

// Ring loads of one row of NQ arrays (window start fs: aligned in kRingX4 mode, column -1 in
// kRingDword mode) to LDS at `dst` (array q at dst + q * CW); every tile clamps its chunks (or
// elements) into the array.
template <typename T>
using lds_ptr = __attribute__((address_space(3))) T *;

// The LDS address of a ring (converted once per strip: a generic-to-LDS conversion tests for
// null, so each conversion costs scalar instructions).
template <typename T>
__device__ __forceinline__ lds_ptr<T> ring_lds(T *ring) {
  return (lds_ptr<T>)ring;
}

template <typename T, int TX, int NQ, int MODE>
__device__ __forceinline__ void ring_issue(lds_ptr<T> dst, const T *const *src, int fs, int n) {
  using RG = RingGeom<T, TX, MODE>;
  const int t = threadIdx.x;
  if constexpr (MODE == kRingX4) {
    if (t < RG::CH) {
      int j = fs + t * RG::E;
      if (j < 0 || j >= n) j = 0;
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        __builtin_amdgcn_global_load_lds((const void *)(src[q] + j),
                                         (lds_ptr<void>)(dst + q * RG::CW), 16, 0, 0);
    }
  } else {
    constexpr int DPE = (int)sizeof(T) / 4;   // dwords per element
#pragma unroll
    for (int k = 0; k < RG::IPA; ++k) {
      const int d = k * 64 + t;               // dword of the window
      if (d < RG::DW) {
        int j = fs + d / DPE;
        if (j < 0 || j >= n) j = 0;
        const int half = d % DPE;
#pragma unroll
        for (int q = 0; q < NQ; ++q)
          __builtin_amdgcn_global_load_lds(
              (const void *)(reinterpret_cast<const uint32_t *>(src[q] + j) + half),
              (lds_ptr<void>)((lds_ptr<uint32_t>)(dst + q * RG::CW) + k * 64), 4, 0, 0);
      }
    }
  }
}

// Slot of row ordinal i + add in a ring of m slots (uniform integer arithmetic; unrolling the
// row loop far enough to make slots compile-time constants measured no faster: the larger
// loop body costs what the address arithmetic saves).
__device__ __forceinline__ int ring_slot(int i, int add, int m) {
  const int v = i + add;
  return (v % m + m) % m;
}

// Slot `add` rows after the one in slot `base` of a ring of m slots (-m < add < m): one
// wrap instead of a division.
__device__ __forceinline__ int wrap_slot(int base, int add, int m) {
  int v = base + add;
  v = v >= m ? v - m : v;
  return v < 0 ? v + m : v;
}

// Whether every read of a streaming tile (rows y0-G-1 .. y0+H+G, columns x0-1 .. x0+TX) falls
// inside the array, so its loads need no bounds checks.
__device__ __forceinline__ bool strip_inside(int x0, int y0, int H, int G, int TX, int nx,
                                             int64_t n) {
  const int64_t lo = (int64_t)(y0 - G - 1) * nx + x0 - 1;
  const int64_t hi = (int64_t)(y0 + H + G) * nx + x0 + TX;
  return lo >= 0 && hi < n;
}

// Whether tile `tile` has boundary or output events in any component.
template <typename T>
__device__ __forceinline__ bool tile_has_events(const StepArgs<T> &a, int tile) {
  bool any = false;
  for (int k = 0; k < 3; ++k) {
    any = any || a.bev[k].ptr[tile + 1] > a.bev[k].ptr[tile];
    any = any || a.outp[k].ev.ptr[tile + 1] > a.outp[k].ev.ptr[tile];
  }
  return any;
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

// Event cursors of one stage (one per component for boundaries and for outputs).
struct StageCursors {
  Cursor b[3], o[3];
};

template <typename T>
__device__ __forceinline__ StageCursors stage_cursors(const StepArgs<T> &a, int tile) {
  StageCursors c;
  for (int k = 0; k < 3; ++k) {
    c.b[k] = cursor(a.bev[k], tile);
    c.o[k] = cursor(a.outp[k].ev, tile);
  }
  return c;
}

// f(integral_constant<int, s>) for s = 0 .. N-1 (or N-1 .. 0), each s a compile-time constant.
template <int N, bool REV = false, typename F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, REV ? N - 1 - I : I>{}), ...);
}

template <int N, bool REV = false, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_impl<N, REV>(f, std::make_integer_sequence<int, N>{});
}

// The main loop runs iterations i = 0 .. iters-1 in blocks of U = lcm(2, RS) copies of the body,
// each copy with its iteration's parity Q = i % 2 and ring slot W = i % RS as compile-time
// constants, so values carried between iterations live in registers indexed by Q (or W) and
// never need copying.  Iterations past the end are skipped (a uniform branch).
template <int RS, typename F, int... J>
__device__ __forceinline__ void run_block(F &f, int i, int iters,
                                          std::integer_sequence<int, J...>) {
  ((i + J < iters
        ? (f(std::integral_constant<int, J % 2>{}, std::integral_constant<int, J % RS>{}, i + J),
           0)
        : 0),
   ...);
}

template <int RS, typename F>
__device__ __forceinline__ void run_loop(int iters, F &&f) {
  constexpr int U = RS % 2 == 0 ? RS : 2 * RS;
  for (int i = 0; i < iters; i += U) run_block<RS>(f, i, iters, std::make_integer_sequence<int, U>{});
}

// As run_block, with the body's steady flag ST (a compile-time argument after Q and W).
template <int RS, bool ST, typename F, int... J>
__device__ __forceinline__ void run_block_st(F &f, int i, int iters,
                                             std::integer_sequence<int, J...>) {
  ((i + J < iters
        ? (f(std::integral_constant<int, J % 2>{}, std::integral_constant<int, J % RS>{},
             std::integral_constant<bool, ST>{}, i + J),
           0)
        : 0),
   ...);
}

// As run_loop, for a body with a steady form (SF): the whole blocks of iterations inside
// [i0, i1] run with ST = true (the body's row-range tests known to hold), the others with
// ST = false; without SF only the general form exists.  The three ranges are three loops: a
// single loop that branches between the two forms per block is smaller code, but its
// register allocation covers both forms at once (the FP64 three-step acoustic kernel then
// needs 146 VGPRs, 3 waves per SIMD, against 115 and 4 with separate loops).
template <int RS, bool SF, typename F>
__device__ __forceinline__ void run_loop_steady(int iters, int i0, int i1, F &&f) {
  constexpr int U = RS % 2 == 0 ? RS : 2 * RS;
  constexpr auto seq = std::make_integer_sequence<int, U>{};
  if constexpr (!SF) {
    for (int i = 0; i < iters; i += U) run_block_st<RS, false>(f, i, iters, seq);
  } else {
    const int lo = i0 < 0 ? 0 : i0, hi = i1 + 1 < iters ? i1 + 1 : iters;
    const int s0 = (lo + U - 1) / U * U < iters ? (lo + U - 1) / U * U : iters;
    const int s1 = hi / U * U > s0 ? hi / U * U : s0;   // steady blocks: [s0, s1)
    int i = 0;
    for (; i < s0; i += U) run_block_st<RS, false>(f, i, iters, seq);
    for (; i < s1; i += U) run_block_st<RS, true>(f, i, iters, seq);
    for (; i < iters; i += U) run_block_st<RS, false>(f, i, iters, seq);
  }
}

// ------------------------------------------------------------------------------------------
// Acoustic (lossless): v -= Dbwd(B) p, then p -= Dfwd(K) v
//
// Stage s at an iteration of parity Q reads: stage 0 the row loaded into its ring slot; stage s > 0
// p from pn[s-1][Q^1] (finished by stage s-1 one iteration ago), v from vh[s-1][Q] (two
// iterations ago) and its previous row's p from pn[s-1][Q]; its coefficients it loads itself
// (one iteration ahead, from cache).  Consumers run their post-barrier part before producers,
// so a producer overwrites a slot only after its consumer is done with it.
// ------------------------------------------------------------------------------------------
template <typename T>
struct AcRow {         // one row of one column: state and coefficients
  T p, cp, vx, cvx, vy, cvy, bx, by, kx, ky;
};

template <typename T>
struct AcCoef {
  T bx, by, kx, ky;
};

template <typename T>
struct AcHalo {        // left thread: p, B of column -1; right thread: p, B, vx, K of column TX
  T p, cp, b, vx, cvx, k;
};

template <typename T>
struct AcVel {
  T vx, cvx, vy, cvy;
};

// S stages over one strip (S <= G + 1).  NT: state reads and writes bypass the caches'
// temporal allocation (pays off when the working set is far larger than the caches); EV: the
// tile has events; CHK: some reads fall outside the arrays; ISO: the y coefficients are the x
// ones (isotropic material arrays), so they are neither loaded nor held twice.
template <typename T, bool KAHAN, int TX, bool AXI, int S, bool NT, bool EV, bool CHK, bool ISO,
          int PF, int RING>
__device__ __forceinline__ void acoustic_strip(const StepArgs<T> &a, int tile, int64_t step, int64_t row,
                               StreamLds<T, TX, S> &L, T *ring) {
  auto &sbp = L.bwd;    // Bx*p of the row, columns -1..TX
  auto &skv = L.fwd;    // forward x products of the row, columns 0..TX
  constexpr bool WAVE = TX == 64;   // one wave: neighbours exchanged by DPP wave shifts
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
      if (WAVE) break;
      sbp[s][b][t + 1] = T(0);
      skv[s][b][t] = T(0);
      if (lh) {
        sbp[s][b][0] = T(0);
        skv[s][b][TX] = T(0);
      }
    }
  }

  StageCursors cur[S];
  if constexpr (EV) {
#pragma unroll
    for (int s = 0; s < S; ++s) cur[s] = stage_cursors(a, tile);
  }
  constexpr int RS = PF + 1;   // stage 0 loads its rows PF iterations ahead
  AcRow<T> nin[RS] = {};      // stage 0 input rows (register ring; RING: unused)
  AcHalo<T> nh[RS] = {};       // stage 0 halo
  T pp[2] = {}, cpp[2] = {};   // stage 0 input p after its events (previous-row p)
  AcCoef<T> cf[S][2] = {};     // stages > 0: coefficients
  T bpy[S][2] = {}, kvy[S][2] = {};
  T pn[S][2] = {}, cpn[S][2] = {};   // new p of the row above the stage's current row
  AcVel<T> vh[S][2] = {};            // new v of the stage's current row
  T kvd[S][2] = {};                  // one wave: forward x difference of the row

  auto load_row = [&](int rho, AcRow<T> &d, AcHalo<T> &h) __attribute__((always_inline)) {
    const int rb = (y0 + rho) * nx + x0;
    d.p = ldn<NT, CHK>(pin, rb, t, n);
    d.vx = ldn<NT, CHK>(vxin, rb, t, n);
    d.vy = ldn<NT, CHK>(vyin, rb, t, n);
    d.bx = ldr<CHK>(Bx, rb, t, n);
    d.by = ISO ? d.bx : ldr<CHK>(By, rb, t, n);
    d.kx = ldr<CHK>(Kx, rb, t, n);
    d.ky = ISO ? d.kx : ldr<CHK>(Ky, rb, t, n);
    if (KAHAN) {
      d.cp = ldn<NT, CHK>(a.cin[0], rb, t, n);
      d.cvx = ldn<NT, CHK>(a.cin[1], rb, t, n);
      d.cvy = ldn<NT, CHK>(a.cin[2], rb, t, n);
    }
    // ghost columns: every lane loads them (thread 0 column -1, the others column TX, one
    // cache line), so the row needs no branch; only the edge threads use the values
    const int off = lh ? -1 : TX;
    h.p = ldr<CHK>(pin, rb, off, n);
    h.b = ldr<CHK>(Bx, rb, off, n);
    if (KAHAN) h.cp = ldr<CHK>(a.cin[0], rb, off, n);
    h.vx = ldr<CHK>(vxin, rb, TX, n);
    h.k = ldr<CHK>(Kx, rb, TX, n);
    if (KAHAN) h.cvx = ldr<CHK>(a.cin[1], rb, TX, n);
  };
  auto load_coef = [&](int rho, AcCoef<T> &d) __attribute__((always_inline)) {
    const int rb = (y0 + rho) * nx + x0;
    d.bx = ldr<CHK>(Bx, rb, t, n);
    d.by = ISO ? d.bx : ldr<CHK>(By, rb, t, n);
    d.kx = ldr<CHK>(Kx, rb, t, n);
    d.ky = ISO ? d.kx : ldr<CHK>(Ky, rb, t, n);
  };

  // LDS rings (RING: kRingX4 or kRingDword): state p, vx, vy[, cp, cvx, cvy]; coefficients
  // Bx, Kx[, By, Ky]
  constexpr int RMODE = RING ? RING : kRingX4;
  using RG = RingGeom<T, TX, RMODE>;
  constexpr int NS = KAHAN ? 6 : 3, NC = ISO ? 2 : 4, NA = (NS + NC) * RG::IPA;
  constexpr int RC = PF + 2 * S - 1;
  constexpr int SP = 0, SVX = 1, SVY = 2, SCP = 3, CBX = 0, CKX = 1, CBY = ISO ? 0 : 2,
                CKY = ISO ? 1 : 3;
  const int lead = RMODE == kRingX4 ? ((x0 - 1) % RG::E + RG::E) % RG::E : 0;
  T *const cring = ring + RS * NS * RG::CW;
  const lds_ptr<T> lring = ring_lds(ring), lcring = lring + RS * NS * RG::CW;
  const T *ssrc[NS], *csrc[NC];
  if constexpr (RING) {
    ssrc[SP] = pin;
    ssrc[SVX] = vxin;
    ssrc[SVY] = vyin;
    if constexpr (KAHAN) {
      ssrc[SCP] = a.cin[0];
      ssrc[SCP + 1] = a.cin[1];
      ssrc[SCP + 2] = a.cin[2];
    }
    csrc[CBX] = Bx;
    csrc[CKX] = Kx;
    if constexpr (!ISO) {
      csrc[CBY] = By;
      csrc[CKY] = Ky;
    }
  }
  // row ordinal r = rho + G + 1 of the strip lives in state slot r % RS and coefficient slot
  // r % RC (at iteration i, stage s holds row ordinal i - 2s)
  auto sbase = [&](int ss) __attribute__((always_inline)) { return ring + ss * NS * RG::CW; };
  auto cbase = [&](int cs) __attribute__((always_inline)) { return cring + cs * NC * RG::CW; };
  // row rho into state slot ss and coefficient slot cs
  auto issue_row = [&](int rho, int ss, int cs) __attribute__((always_inline)) {
    const int fs = (y0 + rho) * nx + x0 - 1 - lead;
    ring_issue<T, TX, NS, RMODE>(lring + ss * NS * RG::CW, ssrc, fs, n);
    ring_issue<T, TX, NC, RMODE>(lcring + cs * NC * RG::CW, csrc, fs, n);
  };
  // value of ring array q (base: the row's slot) at column c of row rho
  auto rd = [&](const T *base, int q, int rho, int c) __attribute__((always_inline)) -> T {
    const T v = base[q * RG::CW + lead + 1 + c];
    if (!CHK) return v;
    const int j = (y0 + rho) * nx + x0 + c;
    return (j >= 0 && j < n) ? v : T(0);
  };
  auto read_row = [&](int rho, int ss, int cs, AcRow<T> &d, AcHalo<T> &h)
                      __attribute__((always_inline)) {
    const T *sb = sbase(ss), *cb = cbase(cs);
    d.p = rd(sb, SP, rho, t);
    d.vx = rd(sb, SVX, rho, t);
    d.vy = rd(sb, SVY, rho, t);
    d.bx = rd(cb, CBX, rho, t);
    d.by = ISO ? d.bx : rd(cb, CBY, rho, t);
    d.kx = rd(cb, CKX, rho, t);
    d.ky = ISO ? d.kx : rd(cb, CKY, rho, t);
    if (KAHAN) {
      d.cp = rd(sb, SCP, rho, t);
      d.cvx = rd(sb, SCP + 1, rho, t);
      d.cvy = rd(sb, SCP + 2, rho, t);
    }
    // ghost columns: every lane reads them (lane 0 column -1, the others column TX; LDS
    // broadcasts), so the row needs no branch; only the edge lanes use the values
    const int off = lh ? -1 : TX;
    h.p = rd(sb, SP, rho, off);
    h.b = rd(cb, CBX, rho, off);
    if (KAHAN) h.cp = rd(sb, SCP, rho, off);
    h.vx = rd(sb, SVX, rho, TX);
    h.k = rd(cb, CKX, rho, TX);
    if (KAHAN) h.cvx = rd(sb, SCP + 1, rho, TX);
  };

#pragma unroll
  for (int r = 0; r < PF; ++r) {
    if constexpr (RING) issue_row(-G - 1 + r, r % RS, r % RC);
    else load_row(-G - 1 + r, nin[r], nh[r]);
  }
  const int iters = H + 2 * G + S + 1;
  // RING: slots are computed, so the loop only unrolls by the parity
  // RING without events, G = S-1, a strip wholly inside the grid: iterations 3S-1 .. H+2G+1-PF
  // run the steady form of the body (ST): every stage's row is inside its range, stage 0
  // fetches the row PF ahead and the last stage's rows are 1 .. H-S, so the range tests and
  // the stores' row tests are compiled out
  run_loop_steady<RING ? 1 : RS, RING && !EV>(
      iters, G == S - 1 && y0 + H <= a.ny ? 3 * S - 1 : iters, H + 2 * G + 1 - PF,
      [&](auto Qc, auto Wc, auto STc, int i) __attribute__((always_inline)) {
    constexpr int Q = decltype(Qc)::value, W = decltype(Wc)::value;
    constexpr bool ST = decltype(STc)::value;
    // RING: slots of this iteration's row ordinal i in the state and coefficient rings (the
    // other rows' slots follow by one wrap)
    const int is0 = RING ? ring_slot(i, 0, RS) : 0, ic0 = RING ? ring_slot(i, 0, RC) : 0;
    if constexpr (RING) {
      // stage 0's row PF ahead, first thing in the iteration
      const StageRows R0(G, H, 0);
      if (ST || i - G - 1 + PF <= R0.hi + 1) {
        compiler_fence();
        issue_row(i - G - 1 + PF, wrap_slot(is0, PF, RS), wrap_slot(ic0, PF, RC));
      }
    }
    T bpx[S];
    AcRow<T> row0 = {};      // RING: stage 0 input row of this iteration, read from the rings
    AcHalo<T> h0 = {};
    AcRow<T> &in0 = RING ? row0 : nin[RING ? 0 : W];
    AcHalo<T> &hin = RING ? h0 : nh[RING ? 0 : W];
    // ---- before the barrier
    auto pre = [&](auto Sc) __attribute__((always_inline)) {
      constexpr int s = decltype(Sc)::value;
      const int rho = i - G - 1 - 2 * s;
      const StageRows R(G, H, s);
      if (!RING && s > 0 && (ST || (rho + 1 >= R.lo - 1 && rho + 1 <= R.hi + 1)))
        load_coef(rho + 1, cf[s][Q ^ 1]);
      if (!ST && (rho < R.lo - 1 || rho > R.hi + 1)) return;
      if constexpr (RING) {
        if (s == 0) {
          compiler_fence();
          // loads issued after this row's: the rows rho + 1 .. min(rho + PF, hi + 1)
          if constexpr (ST) vm_wait<PF * NA>(); else ring_wait<NA, PF>(R.hi + 1 - rho);
          compiler_fence();
          read_row(rho, is0, ic0, row0, h0);
        }
      } else {
        if (s == 0 && (ST || rho + PF <= R.hi + 1))
          load_row(rho + PF, nin[(W + PF) % RS], nh[(W + PF) % RS]);
      }
      T &p = s == 0 ? in0.p : pn[s > 0 ? s - 1 : 0][Q ^ 1];
      T &cp = s == 0 ? in0.cp : cpn[s > 0 ? s - 1 : 0][Q ^ 1];
      T cbx, cby;
      if constexpr (RING) {
        if (s == 0) {
          cbx = in0.bx;
          cby = in0.by;
        } else {
          const T *cb = cbase(wrap_slot(ic0, -2 * s, RC));
          cbx = rd(cb, CBX, rho, t);
          cby = ISO ? cbx : rd(cb, CBY, rho, t);
        }
      } else {
        cbx = s == 0 ? in0.bx : cf[s][Q].bx;
        cby = s == 0 ? in0.by : cf[s][Q].by;
      }
      AcHalo<T> &h = hin;
      if constexpr (EV) {
        stream_events<T, KAHAN>(a.bev[0], a.prog[0], cur[s].b[0], TX + 2, -G - 1, -1, rho, t,
                                step + s, p, cp, s == 0 && (lh || rh), lh ? -1 : TX, h.p, h.cp);
        stream_record<T, KAHAN>(a.outp[0], cur[s].o[0], TX + 2, -G - 1, -1, rho, t, row + s, p,
                                cp);
      }
      bpx[s] = cbx * p;
      bpy[s][Q] = cby * p;
      if (!WAVE && (ST || rho <= R.hi)) {
        sbp[s][Q][t + 1] = bpx[s];
        if (s == 0 && lh) sbp[s][Q][0] = h.b * h.p;
      }
      if (s == 0) {
        pp[Q] = p;
        cpp[Q] = cp;
      }
    };
    // ---- after the barrier (one wave: no barrier), consumers first
    auto post = [&](auto Sc) __attribute__((always_inline)) {
      constexpr int s = decltype(Sc)::value;
      const int rho = i - G - 1 - 2 * s;
      const StageRows R(G, H, s);
      if (!ST && (rho < R.lo - 1 || rho > R.hi + 1)) return;
      const int sp = s > 0 ? s - 1 : 0;
      const T ivx = s == 0 ? in0.vx : vh[sp][Q].vx;
      const T icvx = s == 0 ? in0.cvx : vh[sp][Q].cvx;
      const T ivy = s == 0 ? in0.vy : vh[sp][Q].vy;
      const T icvy = s == 0 ? in0.cvy : vh[sp][Q].cvy;
      T ckx, cky;
      if constexpr (RING) {
        if (s == 0) {
          ckx = in0.kx;
          cky = in0.ky;
        } else {
          const T *cb = cbase(wrap_slot(ic0, -2 * s, RC));
          ckx = rd(cb, CKX, rho, t);
          cky = ISO ? ckx : rd(cb, CKY, rho, t);
        }
      } else {
        ckx = s == 0 ? in0.kx : cf[s][Q].kx;
        cky = s == 0 ? in0.ky : cf[s][Q].ky;
      }
      const T pprev = s == 0 ? pp[Q ^ 1] : pn[sp][Q];
      const T cpprev = s == 0 ? cpp[Q ^ 1] : cpn[sp][Q];
      const AcHalo<T> &h = hin;
      T vxn = T(0), cvxn = T(0), vyn = T(0), cvyn = T(0);
      if (ST || (rho >= R.lo && rho <= R.hi)) {
        T left;
        if constexpr (WAVE) {
          left = wave_from_left(bpx[s]);
          if (s == 0 && lh) left = h.b * h.p;
        } else {
          left = sbp[s][Q][t];
        }
        const T d = (-left) + bpx[s];
        vxn = ivx;
        cvxn = icvx;
        if (KAHAN) kahan_sub(vxn, cvxn, d); else vxn = vxn - d;
        T hvxn = T(0), hcvxn = T(0);
        if (s == 0 && rh) {
          const T dh = (-bpx[s]) + h.b * h.p;
          hvxn = h.vx;
          hcvxn = h.cvx;
          if (KAHAN) kahan_sub(hvxn, hcvxn, dh); else hvxn = hvxn - dh;
        }
        if constexpr (EV) {
          stream_events<T, KAHAN>(a.bev[1], a.prog[1], cur[s].b[1], TX + 1, -G, 0, rho, t,
                                  step + s, vxn, cvxn, s == 0 && rh, TX, hvxn, hcvxn);
          stream_record<T, KAHAN>(a.outp[1], cur[s].o[1], TX + 1, -G, 0, rho, t, row + s, vxn,
                                  cvxn);
        }
        T kvx;
        if (AXI) kvx = (ckx / rad) * (vxn * rad);
        else kvx = ckx * vxn;
        T kvh = T(0);
        if (s == 0 && rh) {
          if (AXI) kvh = (h.k / radh) * (hvxn * radh);
          else kvh = h.k * hvxn;
        }
        if constexpr (WAVE) {
          T right = wave_from_right(kvx);
          if (s == 0 && rh) right = kvh;
          kvd[s][Q] = (-kvx) + right;   // x difference of row rho, used by next iteration
        } else {
          skv[s][Q][t] = kvx;
          if (s == 0 && rh) skv[s][Q][TX] = kvh;
        }
      }
      kvy[s][Q] = T(0);
      if (ST || rho >= R.lo) {
        const T d = (-bpy[s][Q ^ 1]) + bpy[s][Q];
        vyn = ivy;
        cvyn = icvy;
        if (KAHAN) kahan_sub(vyn, cvyn, d); else vyn = vyn - d;
        if constexpr (EV) {
          T dummy = T(0), dummyc = T(0);
          stream_events<T, KAHAN>(a.bev[2], a.prog[2], cur[s].b[2], TX, -G, 0, rho, t,
                                  step + s, vyn, cvyn, false, 0, dummy, dummyc);
          stream_record<T, KAHAN>(a.outp[2], cur[s].o[2], TX, -G, 0, rho, t, row + s, vyn,
                                  cvyn);
        }
        kvy[s][Q] = cky * vyn;
      }
      T pnew = T(0), cpnew = T(0);
      if (ST || rho > R.lo) {    // p of row rho-1 (its x products were published an iteration ago)
        const T dxt = WAVE ? kvd[s][Q ^ 1] : (-skv[s][Q ^ 1][t]) + skv[s][Q ^ 1][t + 1];
        const T dyt = (-kvy[s][Q ^ 1]) + kvy[s][Q];
        const T d = dxt + dyt;
        pnew = pprev;
        cpnew = cpprev;
        if (KAHAN) kahan_sub(pnew, cpnew, d); else pnew = pnew - d;
      }
      if (s == S - 1) {
        // last stage: write the owned cells
        // (steady: the last stage's row is 1 .. H-S of a strip inside the grid)
        if (owned && (ST || (rho >= 1 && rho - 1 < H && y0 + rho - 1 < a.ny))) {
          const int rb = (y0 + rho - 1) * nx + x0;
          stn<NT>(a.out[0], rb, t, pnew);
          if (KAHAN) stn<NT>(a.cout[0], rb, t, cpnew);
        }
        if (owned && (ST || (rho >= 0 && rho < H && y0 + rho < a.ny))) {
          const int rb = (y0 + rho) * nx + x0;
          stn<NT>(a.out[1], rb, t, vxn);
          stn<NT>(a.out[2], rb, t, vyn);
          if (KAHAN) {
            stn<NT>(a.cout[1], rb, t, cvxn);
            stn<NT>(a.cout[2], rb, t, cvyn);
          }
        }
      } else {
        pn[s][Q] = pnew;
        cpn[s][Q] = cpnew;
        vh[s][Q] = AcVel<T>{vxn, cvxn, vyn, cvyn};
      }
    };
    if constexpr (WAVE && RING && !EV) {
      // one wave and no barrier: stages S-1 .. 1 (both parts, consumers still first) run
      // while stage 0's row is in flight, and stage 0 waits for it last
      static_for<S - 1, true>([&](auto Sc) __attribute__((always_inline)) {
        constexpr int s = decltype(Sc)::value + 1;
        pre(std::integral_constant<int, s>{});
        post(std::integral_constant<int, s>{});
      });
      pre(std::integral_constant<int, 0>{});
      post(std::integral_constant<int, 0>{});
    } else {
      static_for<S>(pre);
      if (!WAVE) __syncthreads();
      static_for<S, true>(post);
    }
  });
}

// ------------------------------------------------------------------------------------------
// Thermal: q = -(Dbwd(Q) T) assigned, T -= Dfwd(G) q (same stage protocol; only T is handed
// on, the assigned flux needs no input)
// ------------------------------------------------------------------------------------------
template <typename T>
struct ThRow {
  T t, ct, qx, qy, gx, gy;
};

template <typename T>
struct ThCoef {
  T qx, qy, gx, gy;
};

template <typename T>
struct ThHalo {        // left thread: T, Q of column -1; right thread: T, Q, G of column TX
  T t, ct, q, g;
};

template <typename T, bool KAHAN, int TX, bool AXI, int S, bool NT, bool EV, bool CHK, bool ISO,
          int PF, int RING>
__device__ __forceinline__ void thermal_strip(const StepArgs<T> &a, int tile, int64_t step, int64_t row,
                              StreamLds<T, TX, S> &L, T *ring) {
  auto &sqt = L.bwd;    // Qx*T of the row
  auto &sgq = L.fwd;    // forward x products of the row
  constexpr bool WAVE = TX == 64;
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
      if (WAVE) break;
      sqt[s][b][t + 1] = T(0);
      sgq[s][b][t] = T(0);
      if (lh) {
        sqt[s][b][0] = T(0);
        sgq[s][b][TX] = T(0);
      }
    }
  }

  StageCursors cur[S];
  if constexpr (EV) {
#pragma unroll
    for (int s = 0; s < S; ++s) cur[s] = stage_cursors(a, tile);
  }
  constexpr int RS = PF + 1;
  ThRow<T> nin[RS] = {};
  ThHalo<T> nh[RS] = {};
  T tp[2] = {}, ctp[2] = {};
  ThCoef<T> cf[S][2] = {};
  T qty[S][2] = {}, gqy[S][2] = {};
  T tn[S][2] = {}, ctn[S][2] = {};
  T gqd[S][2] = {};                   // one wave: forward x difference of the row

  auto load_row = [&](int rho, ThRow<T> &d, ThHalo<T> &h) __attribute__((always_inline)) {
    const int rb = (y0 + rho) * nx + x0;
    d.t = ldn<NT, CHK>(tin, rb, t, n);
    d.qx = ldr<CHK>(Qx, rb, t, n);
    d.qy = ISO ? d.qx : ldr<CHK>(Qy, rb, t, n);
    d.gx = ldr<CHK>(Gx, rb, t, n);
    d.gy = ISO ? d.gx : ldr<CHK>(Gy, rb, t, n);
    if (KAHAN) d.ct = ldn<NT, CHK>(a.cin[0], rb, t, n);
    // ghost columns: every lane loads them (thread 0 column -1, the others column TX, one
    // cache line), so the row needs no branch; only the edge threads use the values
    const int off = lh ? -1 : TX;
    h.t = ldr<CHK>(tin, rb, off, n);
    h.q = ldr<CHK>(Qx, rb, off, n);
    if (KAHAN) h.ct = ldr<CHK>(a.cin[0], rb, off, n);
    h.g = ldr<CHK>(Gx, rb, TX, n);
  };
  auto load_coef = [&](int rho, ThCoef<T> &d) __attribute__((always_inline)) {
    const int rb = (y0 + rho) * nx + x0;
    d.qx = ldr<CHK>(Qx, rb, t, n);
    d.qy = ISO ? d.qx : ldr<CHK>(Qy, rb, t, n);
    d.gx = ldr<CHK>(Gx, rb, t, n);
    d.gy = ISO ? d.gx : ldr<CHK>(Gy, rb, t, n);
  };

  // LDS rings (RING: kRingX4 or kRingDword): state T[, cT]; coefficients Qx, Gx[, Qy, Gy]
  constexpr int RMODE = RING ? RING : kRingX4;
  using RG = RingGeom<T, TX, RMODE>;
  constexpr int NS = KAHAN ? 2 : 1, NC = ISO ? 2 : 4, NA = (NS + NC) * RG::IPA;
  constexpr int RC = PF + 2 * S - 1;
  constexpr int ST = 0, SCT = 1, CQX = 0, CGX = 1, CQY = ISO ? 0 : 2, CGY = ISO ? 1 : 3;
  const int lead = RMODE == kRingX4 ? ((x0 - 1) % RG::E + RG::E) % RG::E : 0;
  T *const cring = ring + RS * NS * RG::CW;
  const lds_ptr<T> lring = ring_lds(ring), lcring = lring + RS * NS * RG::CW;
  const T *ssrc[NS], *csrc[NC];
  if constexpr (RING) {
    ssrc[ST] = tin;
    if constexpr (KAHAN) ssrc[SCT] = a.cin[0];
    csrc[CQX] = Qx;
    csrc[CGX] = Gx;
    if constexpr (!ISO) {
      csrc[CQY] = Qy;
      csrc[CGY] = Gy;
    }
  }
  auto sbase = [&](int ss) __attribute__((always_inline)) { return ring + ss * NS * RG::CW; };
  auto cbase = [&](int cs) __attribute__((always_inline)) { return cring + cs * NC * RG::CW; };
  auto issue_row = [&](int rho, int ss, int cs) __attribute__((always_inline)) {
    const int fs = (y0 + rho) * nx + x0 - 1 - lead;
    ring_issue<T, TX, NS, RMODE>(lring + ss * NS * RG::CW, ssrc, fs, n);
    ring_issue<T, TX, NC, RMODE>(lcring + cs * NC * RG::CW, csrc, fs, n);
  };
  auto rd = [&](const T *base, int q, int rho, int c) __attribute__((always_inline)) -> T {
    const T v = base[q * RG::CW + lead + 1 + c];
    if (!CHK) return v;
    const int j = (y0 + rho) * nx + x0 + c;
    return (j >= 0 && j < n) ? v : T(0);
  };
  auto read_row = [&](int rho, int ss, int cs, ThRow<T> &d, ThHalo<T> &h)
                      __attribute__((always_inline)) {
    const T *sb = sbase(ss), *cb = cbase(cs);
    d.t = rd(sb, ST, rho, t);
    d.qx = rd(cb, CQX, rho, t);
    d.qy = ISO ? d.qx : rd(cb, CQY, rho, t);
    d.gx = rd(cb, CGX, rho, t);
    d.gy = ISO ? d.gx : rd(cb, CGY, rho, t);
    if (KAHAN) d.ct = rd(sb, SCT, rho, t);
    // ghost columns: every lane reads them (lane 0 column -1, the others column TX; LDS
    // broadcasts), so the row needs no branch; only the edge lanes use the values
    const int off = lh ? -1 : TX;
    h.t = rd(sb, ST, rho, off);
    h.q = rd(cb, CQX, rho, off);
    if (KAHAN) h.ct = rd(sb, SCT, rho, off);
    h.g = rd(cb, CGX, rho, TX);
  };

#pragma unroll
  for (int r = 0; r < PF; ++r) {
    if constexpr (RING) issue_row(-G - 1 + r, r % RS, r % RC);
    else load_row(-G - 1 + r, nin[r], nh[r]);
  }
  const int iters = H + 2 * G + S + 1;
  // RING: slots are computed, so the loop only unrolls by the parity
  // RING without events, G = S-1, a strip wholly inside the grid: iterations 3S-1 .. H+2G+1-PF
  // run the steady form of the body (ST): every stage's row is inside its range, stage 0
  // fetches the row PF ahead and the last stage's rows are 1 .. H-S, so the range tests and
  // the stores' row tests are compiled out
  run_loop_steady<RING ? 1 : RS, RING && !EV>(
      iters, G == S - 1 && y0 + H <= a.ny ? 3 * S - 1 : iters, H + 2 * G + 1 - PF,
      [&](auto Qc, auto Wc, auto STc, int i) __attribute__((always_inline)) {
    constexpr int Q = decltype(Qc)::value, W = decltype(Wc)::value;
    constexpr bool ST = decltype(STc)::value;
    // RING: slots of this iteration's row ordinal i in the state and coefficient rings (the
    // other rows' slots follow by one wrap)
    const int is0 = RING ? ring_slot(i, 0, RS) : 0, ic0 = RING ? ring_slot(i, 0, RC) : 0;
    if constexpr (RING) {
      // stage 0's row PF ahead, first thing in the iteration
      const StageRows R0(G, H, 0);
      if (ST || i - G - 1 + PF <= R0.hi + 1) {
        compiler_fence();
        issue_row(i - G - 1 + PF, wrap_slot(is0, PF, RS), wrap_slot(ic0, PF, RC));
      }
    }
    T qtx[S];
    ThRow<T> row0 = {};      // RING: stage 0 input row of this iteration, read from the rings
    ThHalo<T> h0 = {};
    ThRow<T> &in0 = RING ? row0 : nin[RING ? 0 : W];
    ThHalo<T> &hin = RING ? h0 : nh[RING ? 0 : W];
    auto pre = [&](auto Sc) __attribute__((always_inline)) {
      constexpr int s = decltype(Sc)::value;
      const int rho = i - G - 1 - 2 * s;
      const StageRows R(G, H, s);
      if (!RING && s > 0 && (ST || (rho + 1 >= R.lo - 1 && rho + 1 <= R.hi + 1)))
        load_coef(rho + 1, cf[s][Q ^ 1]);
      if (!ST && (rho < R.lo - 1 || rho > R.hi + 1)) return;
      if constexpr (RING) {
        if (s == 0) {
          compiler_fence();
          if constexpr (ST) vm_wait<PF * NA>(); else ring_wait<NA, PF>(R.hi + 1 - rho);
          compiler_fence();
          read_row(rho, is0, ic0, row0, h0);
        }
      } else {
        if (s == 0 && (ST || rho + PF <= R.hi + 1))
          load_row(rho + PF, nin[(W + PF) % RS], nh[(W + PF) % RS]);
      }
      T &tv = s == 0 ? in0.t : tn[s > 0 ? s - 1 : 0][Q ^ 1];
      T &ct = s == 0 ? in0.ct : ctn[s > 0 ? s - 1 : 0][Q ^ 1];
      T cqx, cqy;
      if constexpr (RING) {
        if (s == 0) {
          cqx = in0.qx;
          cqy = in0.qy;
        } else {
          const T *cb = cbase(wrap_slot(ic0, -2 * s, RC));
          cqx = rd(cb, CQX, rho, t);
          cqy = ISO ? cqx : rd(cb, CQY, rho, t);
        }
      } else {
        cqx = s == 0 ? in0.qx : cf[s][Q].qx;
        cqy = s == 0 ? in0.qy : cf[s][Q].qy;
      }
      ThHalo<T> &h = hin;
      if constexpr (EV) {
        stream_events<T, KAHAN>(a.bev[0], a.prog[0], cur[s].b[0], TX + 2, -G - 1, -1, rho, t,
                                step + s, tv, ct, s == 0 && (lh || rh), lh ? -1 : TX, h.t, h.ct);
        stream_record<T, KAHAN>(a.outp[0], cur[s].o[0], TX + 2, -G - 1, -1, rho, t, row + s, tv,
                                ct);
      }
      qtx[s] = cqx * tv;
      qty[s][Q] = cqy * tv;
      if (!WAVE && (ST || rho <= R.hi)) {
        sqt[s][Q][t + 1] = qtx[s];
        if (s == 0 && lh) sqt[s][Q][0] = h.q * h.t;
      }
      if (s == 0) {
        tp[Q] = tv;
        ctp[Q] = ct;
      }
    };
    auto post = [&](auto Sc) __attribute__((always_inline)) {
      constexpr int s = decltype(Sc)::value;
      const int rho = i - G - 1 - 2 * s;
      const StageRows R(G, H, s);
      if (!ST && (rho < R.lo - 1 || rho > R.hi + 1)) return;
      const int sp = s > 0 ? s - 1 : 0;
      T cgx, cgy;
      if constexpr (RING) {
        if (s == 0) {
          cgx = in0.gx;
          cgy = in0.gy;
        } else {
          const T *cb = cbase(wrap_slot(ic0, -2 * s, RC));
          cgx = rd(cb, CGX, rho, t);
          cgy = ISO ? cgx : rd(cb, CGY, rho, t);
        }
      } else {
        cgx = s == 0 ? in0.gx : cf[s][Q].gx;
        cgy = s == 0 ? in0.gy : cf[s][Q].gy;
      }
      const T tprev = s == 0 ? tp[Q ^ 1] : tn[sp][Q];
      const T ctprev = s == 0 ? ctp[Q ^ 1] : ctn[sp][Q];
      const ThHalo<T> &h = hin;
      T zero = T(0), zeroc = T(0);
      T fx = T(0), fy = T(0);
      if (ST || (rho >= R.lo && rho <= R.hi)) {
        T left;
        if constexpr (WAVE) {
          left = wave_from_left(qtx[s]);
          if (s == 0 && lh) left = h.q * h.t;
        } else {
          left = sqt[s][Q][t];
        }
        fx = -((-left) + qtx[s]);
        T hfx = T(0);
        if (s == 0 && rh) hfx = -((-qtx[s]) + h.q * h.t);
        if constexpr (EV) {
          stream_events<T, false>(a.bev[1], a.prog[1], cur[s].b[1], TX + 1, -G, 0, rho, t,
                                  step + s, fx, zeroc, s == 0 && rh, TX, hfx, zero);
          stream_record<T, false>(a.outp[1], cur[s].o[1], TX + 1, -G, 0, rho, t, row + s, fx,
                                  zeroc);
        }
        T gqx;
        if (AXI) gqx = (cgx / rad) * (fx * rad);
        else gqx = cgx * fx;
        T gqh = T(0);
        if (s == 0 && rh) {
          if (AXI) gqh = (h.g / radh) * (hfx * radh);
          else gqh = h.g * hfx;
        }
        if constexpr (WAVE) {
          T right = wave_from_right(gqx);
          if (s == 0 && rh) right = gqh;
          gqd[s][Q] = (-gqx) + right;
        } else {
          sgq[s][Q][t] = gqx;
          if (s == 0 && rh) sgq[s][Q][TX] = gqh;
        }
      }
      gqy[s][Q] = T(0);
      if (ST || rho >= R.lo) {
        fy = -((-qty[s][Q ^ 1]) + qty[s][Q]);
        if constexpr (EV) {
          T dummy = T(0), dummyc = T(0);
          stream_events<T, false>(a.bev[2], a.prog[2], cur[s].b[2], TX, -G, 0, rho, t,
                                  step + s, fy, zeroc, false, 0, dummy, dummyc);
          stream_record<T, false>(a.outp[2], cur[s].o[2], TX, -G, 0, rho, t, row + s, fy,
                                  zeroc);
        }
        gqy[s][Q] = cgy * fy;
      }
      T tnew = T(0), ctnew = T(0);
      if (ST || rho > R.lo) {
        const T dxt = WAVE ? gqd[s][Q ^ 1] : (-sgq[s][Q ^ 1][t]) + sgq[s][Q ^ 1][t + 1];
        const T dyt = (-gqy[s][Q ^ 1]) + gqy[s][Q];
        const T d = dxt + dyt;
        tnew = tprev;
        ctnew = ctprev;
        if (KAHAN) kahan_sub(tnew, ctnew, d); else tnew = tnew - d;
      }
      if (s == S - 1) {
        // (steady: the last stage's row is 1 .. H-S of a strip inside the grid)
        if (owned && (ST || (rho >= 1 && rho - 1 < H && y0 + rho - 1 < a.ny))) {
          const int rb = (y0 + rho - 1) * nx + x0;
          stn<NT>(a.out[0], rb, t, tnew);
          if (KAHAN) stn<NT>(a.cout[0], rb, t, ctnew);
        }
        if (store && owned && (ST || (rho >= 0 && rho < H && y0 + rho < a.ny))) {
          const int rb = (y0 + rho) * nx + x0;
          stn<NT>(a.out[1], rb, t, fx);
          stn<NT>(a.out[2], rb, t, fy);
        }
      } else {
        tn[s][Q] = tnew;
        ctn[s][Q] = ctnew;
      }
    };
    if constexpr (WAVE && RING && !EV) {
      // one wave and no barrier: stages S-1 .. 1 (both parts, consumers still first) run
      // while stage 0's row is in flight, and stage 0 waits for it last
      static_for<S - 1, true>([&](auto Sc) __attribute__((always_inline)) {
        constexpr int s = decltype(Sc)::value + 1;
        pre(std::integral_constant<int, s>{});
        post(std::integral_constant<int, s>{});
      });
      pre(std::integral_constant<int, 0>{});
      post(std::integral_constant<int, 0>{});
    } else {
      static_for<S>(pre);
      if (!WAVE) __syncthreads();
      static_for<S, true>(post);
    }
  });
}

// Path of a tile: bit 1 = it has events, bit 0 = all its reads fall inside the arrays.  The
// value is uniform; variants are selected with one switch (a flat switch keeps each variant's
// registers to itself, nested branches over the four inlined strips do not).
template <typename T, int TX>
__device__ __forceinline__ int strip_path(const StepArgs<T> &a, int tile) {
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const int G = a.ghost;
  const bool inside = strip_inside(tx * (TX - 2 * G) - G, ty * a.strip_h, a.strip_h, G, TX,
                                   a.nx, a.n);
  return __builtin_amdgcn_readfirstlane((tile_has_events(a, tile) ? 2 : 0) | (inside ? 1 : 0));
}

#define FDS_STRIP_CASES(FN, AXI, ISO, BASE)                                                   \
  case BASE + 0:                                                                                \
    FN<T, KAHAN, TX, AXI, S, NT, false, true, ISO, PF, RING>(a, tile, step, row, L, ring);      \
    break;                                                                                      \
  case BASE + 1:                                                                                \
    FN<T, KAHAN, TX, AXI, S, NT, false, false, ISO, PF, RING>(a, tile, step, row, L, ring);     \
    break;                                                                                      \
  case BASE + 2:                                                                                \
    FN<T, KAHAN, TX, AXI, S, NT, true, true, ISO, PF, RING>(a, tile, step, row, L, ring);       \
    break;                                                                                      \
  case BASE + 3:                                                                                \
    FN<T, KAHAN, TX, AXI, S, NT, true, false, ISO, PF, RING>(a, tile, step, row, L, ring);      \
    break;

// ISO kernels serve fields whose y coefficient arrays are the x ones (the host checks).  PF:
// rows of stage-0 input prefetched ahead (more loads in flight per wave, more registers).
template <typename T, bool KAHAN, int TX, bool AXI, int S, bool NT, bool ISO, int PF, int RING>
__device__ __forceinline__ void acoustic_stream(const StepArgs<T> &a, int tile, int64_t step,
                                                int64_t row, StreamLds<T, TX, S> &L, T *ring) {
  switch (strip_path<T, TX>(a, tile)) {
    FDS_STRIP_CASES(acoustic_strip, AXI, ISO, 0)
    default: break;
  }
}

template <typename T, bool KAHAN, int TX, bool AXI, int S, bool NT, bool ISO, int PF, int RING>
__device__ __forceinline__ void thermal_stream(const StepArgs<T> &a, int tile, int64_t step,
                                               int64_t row, StreamLds<T, TX, S> &L, T *ring) {
  switch (strip_path<T, TX>(a, tile)) {
    FDS_STRIP_CASES(thermal_strip, AXI, ISO, 0)
    default: break;
  }
}

// One tile of instance `a` in a batched streaming launch (lossless models only): one switch
// over model variant and path.
template <typename T, bool KAHAN, int TX, int S, bool NT, int PF, int RING>
__device__ __forceinline__ void any_stream(const StepArgs<T> &a, int tile, int64_t step,
                                           int64_t row, StreamLds<T, TX, S> &L, T *ring) {
  const int f = a.flags;
  const int model = ((f & kFlagThermal) ? 2 : 0) | ((f & kFlagAxi) ? 1 : 0);
  switch (__builtin_amdgcn_readfirstlane(model * 4 + strip_path<T, TX>(a, tile))) {
    FDS_STRIP_CASES(acoustic_strip, false, false, 0)
    FDS_STRIP_CASES(acoustic_strip, true, false, 4)
    FDS_STRIP_CASES(thermal_strip, false, false, 8)
    FDS_STRIP_CASES(thermal_strip, true, false, 12)
    default: break;
  }
}

#undef FDS_STRIP_CASES

}  // namespace fds

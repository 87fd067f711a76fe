// pyfds_amd C-ABI: contexts, uploads, per-tile event lists, the step loop and HIP graphs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pyfds_amd.h"
#include "fds_common.h"
#include "fds_kernels.h"
#include "fds_lbc.h"
#include "fds_line1d.h"
#include "fds_step2d.h"
#include "fds_step_nl.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t err_ = (expr);                                                                \
    if (err_ != hipSuccess)                                                                  \
      return fail(FDS_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(err_),       \
                  __FILE__, __LINE__);                                                       \
  } while (0)

// Owned device allocation.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t ensure(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    release();
    if (b == 0) b = 16;
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) bytes = b;
    return e;
  }
  template <typename U>
  U *as() const { return static_cast<U *>(p); }
};

// Pinned host memory (hipHostMalloc), grown on demand: the staging area that asynchronous
// runs copy their recorded outputs into.
struct HostBuf {
  void *p = nullptr;
  size_t bytes = 0;
  HostBuf() = default;
  HostBuf(const HostBuf &) = delete;
  HostBuf &operator=(const HostBuf &) = delete;
  ~HostBuf() { release(); }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t ensure(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    release();
    hipError_t e = hipHostMalloc(&p, std::max<size_t>(b, 16), hipHostMallocDefault);
    if (e == hipSuccess) bytes = std::max<size_t>(b, 16);
    return e;
  }
};

// A run enqueued by fds_run_async / fds_batch_run_async whose outputs are not taken yet:
// its step count and the staging slot (and event set) it uses.
struct Flight {
  int64_t rows = 0;
  int slot = 0;
};
constexpr int kMaxFlights = 2;

template <typename U>
hipError_t upload(DevBuf &d, const std::vector<U> &h, hipStream_t s) {
  hipError_t e = d.ensure(std::max<size_t>(h.size() * sizeof(U), 16));
  if (e != hipSuccess) return e;
  if (h.empty()) return hipSuccess;
  e = hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(U), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

int64_t floordiv(int64_t a, int64_t b) {
  int64_t q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
  return q;
}

int64_t ceildiv(int64_t a, int64_t b) { return -floordiv(-a, b); }

// Where a component's values live inside a tile's LDS arrays: rows [rmin, rmax] x cols
// [cmin, cmax] of the tile (relative to its origin), stored row-major with row width w.
struct Region {
  int rmin, rmax, cmin, cmax, w;
  int pos(int r, int c) const { return (r - rmin) * w + (c - cmin); }
};

struct Events {
  std::vector<int32_t> ptr, pos, id;
};

struct Comp {
  bool present = false;
  bool accumulated = false;   // updated with -= (Kahan compensation applies)
  DevBuf buf[2], cbuf[2];
  // boundary program (host copy kept for validation and event building)
  std::vector<int64_t> cells;
  std::vector<int32_t> ptr;
  std::vector<double> a, table;
  std::vector<int64_t> off, stride, len;
  DevBuf d_ptr, d_a, d_off, d_stride, d_table;
  DevBuf ev_ptr, ev_pos, ev_id;
  // outputs
  std::vector<int64_t> out_cells;
  DevBuf oev_ptr, oev_pos, oev_id;
  DevBuf lhead, ohead;   // resident 1D engine: per-thread first boundary cell / output point
  DevBuf obuf;
  int64_t obuf_rows = 0;
  HostBuf stage[kMaxFlights];   // outputs of asynchronous runs, one slot per run in flight
};

enum Family { kAcoustic, kThermal, kNonlinear };

}  // namespace

struct fds_ctx {
  int model = 0, form = 0, device = 0;
  int64_t nx = 0, ny = 0, n = 0;
  Family family = kAcoustic;
  bool one_d = false;     // 1D linear model run on the 2D kernels (nx = n, ny = 1)
  bool axi = false;
  hipStream_t stream = nullptr;
  Comp comp[3];
  int parity = 0;
  DevBuf coef[FDS_NUM_COEFF];
  bool coef_set[FDS_NUM_COEFF] = {};
  DevBuf zeros;           // n zeros of the state type (1D y coefficients, absent slots)
  double s[FDS_NUM_SCALARS] = {};
  bool lossy = false, iso_b = true, iso_k = true;
  bool committed = false;
  int tile = -1;          // requested shape index (-1: default)
  int shape = 0;          // shape index in use (tile table, then the streaming table)
  bool streaming = false; // row-streaming engine in use
  int strip_rows = 0;     // streaming strip height override (0: from the shape and the grid)
  int num_cu = 256;       // compute units of the device
  int fuse = 1;           // time steps per launch of the streaming shape
  int ghost = 0;          // ghost columns per side of a streaming tile (fuse - 1; lossy: 1)
  int ghost_rows = 0;     // ghost rows per side of a streaming strip (fuse - 1; lossy: 0)
  int sx = 0;             // x stride of the tiles (tx - 2*ghost)
  bool line = false;      // resident 1D engine in use (one workgroup, many steps per launch)
  int line_var = 0;       // its variant: index into fds::kLineCpt
  int line_threads = 0;   // its workgroup size
  int line_wgs = 1;       // workgroups (segments of the line)
  int line_seg = 0;       // cells owned per segment
  int line_ghost = 0;     // ghost cells per segment side = most steps per launch (0: one segment)
  int line_cfg_threads = -1, line_cfg_ghost = 0;  // requested segmentation (-1: automatic)
  int tx = 0, ty = 0, tiles_x = 0, tiles_y = 0, ntiles = 0;
  bool events_dirty = true;
  DevBuf dyn;
  DevBuf zero_ptr;        // ntiles+1 zeros: empty event lists
  const fds::KernelSet *ks = nullptr;
  size_t elem = 8;
  // graphs: key (chunk, parity) -> exec
  std::map<std::pair<int, int>, hipGraphExec_t> graphs;
  int graph_chunk = 0;              // steps of a full graph chunk (other graphs: run tails)
  uint64_t generation = 0, graph_generation = ~0ull;
  std::vector<hipGraph_t> graph_objs;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;   // bracket the launches of the last run
  double last_run_ms = 0.0;
  // asynchronous runs in flight (oldest first), the next staging slot, and per slot the
  // events after the run's first launch, its last launch and its output copies
  std::deque<Flight> flights;
  int next_slot = 0;
  int batch_flights = 0;          // asynchronous batch runs in flight that include this context
  hipEvent_t fev[kMaxFlights][3] = {};
  // flow along x (AcousticFlow2D): per-row shift period, applied after every step
  DevBuf flow_period;
  bool flow = false;
  // ThermoAcoustic1D loss coupling (acoustic context): density, loss factor, accumulator
  DevBuf couple_rho, couple_loss, couple_acc;
  // learned absorbing boundary (caller-owned bf16 weights), applied after every step
  const uint16_t *lbc_w1 = nullptr, *lbc_w2 = nullptr;
  int lbc_layer = 0;
  float lbc_cs[3] = {1.f, 1.f, 1.f};
  // non-finite detection: scans in every graph chunk and after every run; the result
  // (dyn[kDynNonfinite]) is copied to nf_stage[slot] (slot kMaxFlights: synchronous runs)
  bool nf_check = false;
  HostBuf nf_stage;
};

struct fds_batch {
  std::vector<fds_ctx *> ctxs;
  int device = 0, form = 0, tile = -1;
  size_t elem = 8;
  hipStream_t stream = nullptr;
  const fds::KernelSet *ks = nullptr;
  int shape = 0;
  bool streaming = false;
  int fuse = 1;
  int tx = 0, ty = 0, ntiles = 0;
  size_t lds = 0;
  DevBuf descs[2][2];          // [parity][store]
  DevBuf tile_inst, inst_base, dyn;
  std::map<std::pair<int, int>, hipGraphExec_t> graphs;
  std::vector<hipGraph_t> graph_objs;
  std::vector<uint64_t> build_gen;   // member generations the descriptors were built for
  std::vector<int> build_parity;     // member parities at build time
  int flip = 0;                      // members flipped this many times (mod 2) since build
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_run_ms = 0.0;
  std::deque<Flight> flights;     // asynchronous batch runs in flight (oldest first)
  int next_slot = 0;
  hipEvent_t fev[kMaxFlights][3] = {};
  // tile scheduling of streaming batches: 0 the hardware dispatcher (one workgroup per tile,
  // XCD-banded), 1 a persistent grid taking tiles from a work queue in `order`
  int schedule = 0;
  DevBuf order, queue;
  int nq = 0;             // work-queue counters the queue buffer holds (one per launch)
};

namespace {

bool is_kahan(const fds_ctx *c) { return c->form == FDS_FP32_KAHAN; }

void drop_graphs(fds_ctx *c) {
  for (auto &kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto g : c->graph_objs) (void)hipGraphDestroy(g);
  c->graphs.clear();
  c->graph_objs.clear();
}

void touch(fds_ctx *c) { c->generation++; }

Region region_of(const fds_ctx *c, int k) {
  const int tx = c->tx, ty = c->ty;
  if (c->family == kNonlinear) {
    if (k == 0) return Region{0, 0, -1, tx, tx + 2};
    if (k == 1) return Region{0, 0, 0, tx, tx + 1};
    return Region{0, 0, 0, tx - 1, tx};
  }
  // rows of all fused stages: G ghost rows above and below the strip
  const int G = c->ghost_rows;
  if (k == 0) return Region{-G - 1, ty + G, -1, tx, tx + 2};
  if (k == 1) return Region{-G, ty - 1 + G, 0, tx, tx + 1};
  return Region{-G, ty + G, 0, tx - 1, tx};
}

void build_bound_events(const fds_ctx *c, int k, Events &ev) {
  const Comp &cp = c->comp[k];
  const Region R = region_of(c, k);
  // tile (txi, tyi) has its column 0 at absolute column txi*sx - g and row 0 at tyi*ty
  const int64_t nx = c->nx, sx = c->sx, ty = c->ty, g = c->ghost;
  std::vector<std::pair<int32_t, std::pair<int32_t, int32_t>>> all;   // tile, (pos, u)
  const int64_t cmin_abs = R.cmin - g;
  const int64_t cmax_abs = (int64_t)(c->tiles_x - 1) * sx + R.cmax - g;
  const int64_t rmax_abs = (int64_t)(c->tiles_y - 1) * ty + R.rmax;
  for (size_t u = 0; u < cp.cells.size(); ++u) {
    const int64_t j = cp.cells[u];
    // absolute row Rw and column C with Rw*nx + C = j, C in [cmin_abs, cmax_abs]
    const int64_t rlo = std::max<int64_t>(R.rmin, ceildiv(j - cmax_abs, nx));
    const int64_t rhi = std::min<int64_t>(rmax_abs, floordiv(j - cmin_abs, nx));
    for (int64_t Rw = rlo; Rw <= rhi; ++Rw) {
      const int64_t C = j - Rw * nx;
      const int64_t ty_lo = std::max<int64_t>(0, ceildiv(Rw - R.rmax, ty));
      const int64_t ty_hi = std::min<int64_t>(c->tiles_y - 1, floordiv(Rw - R.rmin, ty));
      const int64_t tx_lo = std::max<int64_t>(0, ceildiv(C + g - R.cmax, sx));
      const int64_t tx_hi = std::min<int64_t>(c->tiles_x - 1, floordiv(C + g - R.cmin, sx));
      for (int64_t tyi = ty_lo; tyi <= ty_hi; ++tyi)
        for (int64_t txi = tx_lo; txi <= tx_hi; ++txi) {
          const int r = (int)(Rw - tyi * ty), col = (int)(C - txi * sx + g);
          all.push_back({(int32_t)(tyi * c->tiles_x + txi), {R.pos(r, col), (int32_t)u}});
        }
    }
  }
  std::sort(all.begin(), all.end());
  ev.ptr.assign(c->ntiles + 1, 0);
  ev.pos.clear();
  ev.id.clear();
  for (auto &t : all) {
    ev.ptr[t.first + 1]++;
    ev.pos.push_back(t.second.first);
    ev.id.push_back(t.second.second);
  }
  for (int t = 0; t < c->ntiles; ++t) ev.ptr[t + 1] += ev.ptr[t];
}

void build_output_events(const fds_ctx *c, int k, Events &ev) {
  const Comp &cp = c->comp[k];
  const Region R = region_of(c, k);
  std::vector<std::pair<int32_t, std::pair<int32_t, int32_t>>> all;
  for (size_t s = 0; s < cp.out_cells.size(); ++s) {
    const int64_t j = cp.out_cells[s];
    // the tile that owns the cell (owned columns of a streaming tile start at its ghost)
    const int64_t iy = j / c->nx, ix = j % c->nx;
    const int64_t tyi = iy / c->ty, txi = ix / c->sx;
    const int r = (int)(iy - tyi * c->ty), col = (int)(ix - txi * c->sx + c->ghost);
    all.push_back({(int32_t)(tyi * c->tiles_x + txi), {R.pos(r, col), (int32_t)s}});
  }
  std::sort(all.begin(), all.end());
  ev.ptr.assign(c->ntiles + 1, 0);
  ev.pos.clear();
  ev.id.clear();
  for (auto &t : all) {
    ev.ptr[t.first + 1]++;
    ev.pos.push_back(t.second.first);
    ev.id.push_back(t.second.second);
  }
  for (int t = 0; t < c->ntiles; ++t) ev.ptr[t + 1] += ev.ptr[t];
}

// Shape index -> (tx, ty); indices from kNumTiles on name the streaming shapes.
void shape_of(int t, int *tx, int *ty) {
  if (t < fds::kNumTiles) {
    *tx = fds::kTileShapes[t][0];
    *ty = fds::kTileShapes[t][1];
  } else {
    *tx = fds::kStreamShapes[t - fds::kNumTiles][0];
    *ty = fds::kStreamShapes[t - fds::kNumTiles][1];
  }
}

// The streaming engine covers the 2D models (lossy acoustics one step per launch); 1D models
// run on the resident engine or tiles.  Batches stream lossless members only.
bool can_stream(const fds_ctx *c) {
  return c->family != kNonlinear && c->ny > 1;
}

bool can_stream_batch(const fds_ctx *c) { return can_stream(c) && !c->lossy; }

// Resident 1D engine variant for n cells: `cpt` cells per thread if given, else the fewest
// cells per thread that fit (the most threads: a step is bound by the one CU's issue rate, so
// spreading the line over more waves wins; measured on 2000 cells: 0.59 us per step with 2
// cells per thread, 0.62 with 4, 0.70 with 8).  Returns the variant index or -1.
int line_variant(int64_t n, int cpt, int fam) {
  for (int i = 0; i < fds::kNumLineCpt; ++i) {
    const int64_t thr = (n + fds::kLineCpt[i] - 1) / fds::kLineCpt[i];
    if (thr > fds::line_max_threads(fam, i)) continue;
    if (cpt <= 0 || fds::kLineCpt[i] == cpt) return i;
  }
  return -1;
}

int choose_tiles(fds_ctx *c) {
  c->streaming = false;
  c->line = false;
  c->fuse = 1;
  c->ghost = 0;
  c->ghost_rows = 0;
  if ((c->one_d || c->family == kNonlinear) && c->tile < 0 && c->n <= fds::kLineMaxResident) {
    // segmented lines take up to kLineMaxResident cells; one segment must hold the whole
    // line.  Automatic
    // layout: one segment up to kLineSegmentAbove cells, else segments of the default shape
    int cfg_threads = c->line_cfg_threads, cfg_ghost = c->line_cfg_ghost, cpt_req = c->strip_rows;
    if (cfg_threads < 0) {
      cfg_threads = 0;
      if (c->strip_rows <= 0 && c->n > fds::kLineSegmentAbove) {
        const bool lng = c->n > fds::kLineSegLongAbove;
        cfg_threads = fds::kLineSegThreads;
        cfg_ghost = lng ? fds::kLineSegLongGhost : fds::kLineSegGhost;
        cpt_req = lng ? fds::kLineSegLongCpt : fds::kLineSegCpt;
      }
    }
    int v = -1;
    if (cfg_threads > 0) {
      for (int i = 0; i < fds::kNumLineCpt; ++i)
        if (fds::kLineCpt[i] == cpt_req) v = i;
    } else {
      v = line_variant(c->n, cpt_req,
                       c->family == kNonlinear ? fds::kLineNonlinear : fds::kLineAcoustic);
    }
    if (v < 0 && c->strip_rows > 0)
      return fail(FDS_E_ARG, "the resident 1D engine has no %d-cell-per-thread variant for %lld "
                  "cells", c->strip_rows, (long long)c->n);
    if (v >= 0) {
      // event lists are indexed by thread over all segments, like tiles
      c->line = true;
      c->line_var = v;
      const int cpt = fds::kLineCpt[v];
      if (cfg_threads > 0) {
        // W workgroups of cfg_threads threads: L owned + 2 G ghost cells each
        c->line_threads = cfg_threads;
        c->line_ghost = cfg_ghost;
        c->line_seg = c->line_threads * cpt - 2 * c->line_ghost;
        c->line_wgs = (int)((c->n + c->line_seg - 1) / c->line_seg);
      } else {
        c->line_threads = (int)(((c->n + cpt - 1) / cpt + 63) / 64 * 64);
        c->line_ghost = 0;
        c->line_seg = (int)c->n;
        c->line_wgs = 1;
      }
      c->shape = -1;
      c->tx = (int)c->n;
      c->ty = 1;
      c->sx = c->tx;
      c->tiles_x = c->ntiles = c->line_threads * c->line_wgs;
      c->tiles_y = 1;
      c->events_dirty = true;
      return 0;
    }
  }
  if (c->family == kNonlinear) {
    c->tx = fds::kNlTile;
    c->ty = 1;
    c->shape = 0;
  } else {
    int t = c->tile;
    if (t < 0) t = (c->ny == 1) ? fds::kTileRow : (c->lossy ? fds::kDefaultLossy2D
                                                             : fds::kDefaultLossless2D);
    if (t >= fds::kNumTiles && !can_stream(c))
      t = (c->ny == 1) ? fds::kTileRow : fds::kDefaultTile2D;
    // lossy acoustics streams one step per launch with its own row ring: shapes with deeper
    // prefetch have no lossy kernels
    if (t >= fds::kNumTiles && c->lossy && c->family == kAcoustic &&
        fds::stream_single_only(t - fds::kNumTiles))
      return fail(FDS_E_ARG, "shape %d (prefetch %d rows) serves lossless fields only", t,
                  fds::kStreamShapes[t - fds::kNumTiles][4]);
    c->shape = t;
    c->streaming = t >= fds::kNumTiles;
    shape_of(t, &c->tx, &c->ty);
    if (c->streaming) {
      // lossy acoustics: one step per launch, one ghost column per side, no ghost rows
      const bool lossy = c->lossy && c->family == kAcoustic;
      c->fuse = lossy ? 1 : fds::kStreamShapes[t - fds::kNumTiles][2];
      c->ghost = lossy ? 1 : c->fuse - 1;
      c->ghost_rows = lossy ? 0 : c->ghost;
      if (c->strip_rows > 0) {
        c->ty = c->strip_rows;
      } else {
        // a strip is walked row by row by its waves: small grids get shorter strips so that
        // there are enough waves to cover the memory latency (~4 per SIMD)
        const int64_t sx = c->tx - 2 * c->ghost;
        const int64_t waves = ((c->nx + sx - 1) / sx) * (c->tx / 64);
        const int64_t want = 4 * 4 * (int64_t)c->num_cu;
        const int64_t h = (c->ny * waves + want - 1) / want;
        const int hmin = 2 * c->fuse + 2;
        if (h < c->ty) c->ty = (int)std::max<int64_t>(hmin, h);
      }
    }
  }
  c->sx = c->tx - 2 * c->ghost;
  c->tiles_x = (int)((c->nx + c->sx - 1) / c->sx);
  c->tiles_y = (int)((c->ny + c->ty - 1) / c->ty);
  const int64_t nt = (int64_t)c->tiles_x * c->tiles_y;
  if (nt > (int64_t)std::numeric_limits<int32_t>::max() / 2)
    return fail(FDS_E_ARG, "grid too large for the tile index (%lld tiles)", (long long)nt);
  c->ntiles = (int)nt;
  c->events_dirty = true;
  return 0;
}

size_t lds_bytes(const fds_ctx *c) {
  if (c->family == kNonlinear || c->streaming || c->line) return 0;   // static LDS only
  const bool sep_k = c->axi || !c->iso_k;
  fds::Lds2D L = fds::lds2d_layout(c->tx, c->ty, !c->iso_b, sep_k, is_kahan(c));
  return (size_t)L.total * c->elem;
}

// Event lists of the resident 1D engine: CSR over the threads of all segments, pos = register
// slot of the cell in its thread's block of cpt cells; id = unique boundary cell or output
// slot.  Boundary cells go to every segment whose ghost-extended range holds them (ghost
// cells must evolve like the owner's), recorded points only to their owner.
void build_line_events(const fds_ctx *c, const std::vector<int64_t> &cells, bool outputs,
                       Events &ev) {
  const int64_t cpt = fds::kLineCpt[c->line_var];
  const int64_t L = c->line_seg, G = c->line_ghost, E = (int64_t)c->line_threads * cpt;
  const int64_t W = c->line_wgs;
  std::vector<std::pair<int32_t, std::pair<int32_t, int32_t>>> all;
  for (size_t u = 0; u < cells.size(); ++u) {
    const int64_t j = cells[u];
    const int64_t w_lo = outputs ? j / L : std::max<int64_t>(0, ceildiv(j + G - E + 1, L));
    const int64_t w_hi = outputs ? j / L : std::min<int64_t>(W - 1, floordiv(j + G, L));
    for (int64_t w = w_lo; w <= w_hi; ++w) {
      const int64_t local = j - (w * L - G);
      all.push_back({(int32_t)(w * c->line_threads + local / cpt),
                     {(int32_t)(local % cpt), (int32_t)u}});
    }
  }
  std::sort(all.begin(), all.end());
  ev.ptr.assign(c->ntiles + 1, 0);
  ev.pos.clear();
  ev.id.clear();
  for (auto &t : all) {
    ev.ptr[t.first + 1]++;
    ev.pos.push_back(t.second.first);
    ev.id.push_back(t.second.second);
  }
  for (int t = 0; t < c->ntiles; ++t) ev.ptr[t + 1] += ev.ptr[t];
}

// Per-thread heads of the resident 1D engine's event lists: the first boundary cell with its
// program entry (when the program has one entry), the first recorded point.
void build_line_heads(const Comp &cp, const Events &ev, const Events &oe, int nthreads,
                      std::vector<fds::LineEvHead> &bh, std::vector<fds::LineOutHead> &oh) {
  bh.assign(nthreads, fds::LineEvHead{0xff, 0, 0, 0.0});
  oh.assign(nthreads, fds::LineOutHead{0xff, 0});
  for (int t = 0; t < nthreads; ++t) {
    const int e0 = ev.ptr[t], e1 = ev.ptr[t + 1];
    if (e0 < e1) {
      const int u = ev.id[e0];
      const int q0 = cp.ptr[u];
      if (cp.ptr[u + 1] - q0 == 1) {
        bh[t].info = ev.pos[e0] | (e1 - e0 > 1 ? 0x100 : 0);
        bh[t].stride = (int32_t)cp.stride[q0];
        bh[t].off = cp.off[q0];
        bh[t].a = cp.a[q0];
      } else {
        bh[t].info = 0x1ff;
      }
    }
    const int o0 = oe.ptr[t], o1 = oe.ptr[t + 1];
    if (o0 < o1) oh[t] = fds::LineOutHead{oe.pos[o0] | (o1 - o0 > 1 ? 0x100 : 0), oe.id[o0]};
  }
}

int rebuild_events(fds_ctx *c) {
  if (!c->events_dirty) return 0;
  std::vector<int32_t> zp(c->ntiles + 1, 0);
  HIP_TRY(upload(c->zero_ptr, zp, c->stream));
  for (int k = 0; k < 3; ++k) {
    Comp &cp = c->comp[k];
    if (!cp.present) continue;
    Events ev;
    if (c->line) build_line_events(c, cp.cells, false, ev); else build_bound_events(c, k, ev);
    HIP_TRY(upload(cp.ev_ptr, ev.ptr, c->stream));
    HIP_TRY(upload(cp.ev_pos, ev.pos, c->stream));
    HIP_TRY(upload(cp.ev_id, ev.id, c->stream));
    Events oe;
    if (c->line) build_line_events(c, cp.out_cells, true, oe);
    else build_output_events(c, k, oe);
    HIP_TRY(upload(cp.oev_ptr, oe.ptr, c->stream));
    HIP_TRY(upload(cp.oev_pos, oe.pos, c->stream));
    HIP_TRY(upload(cp.oev_id, oe.id, c->stream));
    if (c->line) {
      std::vector<fds::LineEvHead> bh;
      std::vector<fds::LineOutHead> oh;
      build_line_heads(cp, ev, oe, c->ntiles, bh, oh);
      HIP_TRY(upload(cp.lhead, bh, c->stream));
      HIP_TRY(upload(cp.ohead, oh, c->stream));
    }
  }
  c->events_dirty = false;
  touch(c);
  return 0;
}

// Convert host float64 to the state type and copy to the device.
int to_device(fds_ctx *c, void *dst, const double *src, int64_t n) {
  if (c->elem == 8) {
    HIP_TRY(hipMemcpyAsync(dst, src, n * 8, hipMemcpyHostToDevice, c->stream));
  } else {
    std::vector<float> tmp(n);
    for (int64_t i = 0; i < n; ++i) tmp[i] = (float)src[i];
    HIP_TRY(hipMemcpyAsync(dst, tmp.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

template <typename T>
void fill_args(const fds_ctx *c, fds::StepArgs<T> &a, int parity) {
  std::memset(&a, 0, sizeof a);
  a.nx = (int32_t)c->nx;
  a.ny = (int32_t)c->ny;
  a.n = c->n;
  a.tiles_x = c->tiles_x;
  a.ntiles = c->ntiles;
  for (int k = 0; k < 3; ++k) {
    const Comp &cp = c->comp[k];
    const bool pp = cp.accumulated || c->family != kThermal || k == 0;
    const int rd = pp ? parity : 0, wr = pp ? parity ^ 1 : 0;
    a.in[k] = cp.buf[rd].as<const T>();
    a.out[k] = cp.buf[wr].as<T>();
    a.cin[k] = cp.cbuf[rd].as<const T>();
    a.cout[k] = cp.cbuf[wr].as<T>();
    const bool has_b = !cp.cells.empty();
    a.bev[k].ptr = has_b ? cp.ev_ptr.as<const int32_t>() : c->zero_ptr.as<const int32_t>();
    a.bev[k].pos = cp.ev_pos.as<const int32_t>();
    a.bev[k].id = cp.ev_id.as<const int32_t>();
    a.prog[k].ptr = cp.d_ptr.as<const int32_t>();
    a.prog[k].a = cp.d_a.as<const double>();
    a.prog[k].off = cp.d_off.as<const int64_t>();
    a.prog[k].stride = cp.d_stride.as<const int64_t>();
    a.prog[k].table = cp.d_table.as<const double>();
    const bool has_o = !cp.out_cells.empty();
    a.outp[k].ev.ptr = has_o ? cp.oev_ptr.as<const int32_t>() : c->zero_ptr.as<const int32_t>();
    a.outp[k].ev.pos = cp.oev_pos.as<const int32_t>();
    a.outp[k].ev.id = cp.oev_id.as<const int32_t>();
    a.outp[k].buf = cp.obuf.as<double>();
    a.outp[k].m = (int32_t)cp.out_cells.size();
    a.lhead[k] = cp.lhead.p;
    a.ohead[k] = cp.ohead.p;
  }
  for (int s = 0; s < FDS_NUM_COEFF; ++s)
    a.coef[s] = c->coef_set[s] ? c->coef[s].as<const T>() : c->zeros.as<const T>();
  if (c->family != kNonlinear) {
    // cleared y slots alias their x slot, except for 1D models whose y terms are zero
    const int pairs[3][2] = {{0, 1}, {2, 3}, {4, 5}};
    for (auto &pr : pairs)
      if (!c->coef_set[pr[1]]) a.coef[pr[1]] = c->one_d ? c->zeros.as<const T>() : a.coef[pr[0]];
  }
  for (int s = 0; s < FDS_NUM_SCALARS; ++s) a.s[s] = c->s[s];
  a.dyn = c->dyn.as<const int64_t>();
  a.flags = (c->iso_b ? fds::kFlagIsoB : 0) | (c->iso_k ? fds::kFlagIsoK : 0);
  a.strip_h = c->ty;
  a.ghost = c->ghost;
}

// ring_ok: the arrays suit an LDS-ring shape's 16-byte row loads (else its register-path twin
// runs, same bits).
fds::LaunchFn pick_launch(const fds_ctx *c, int stages, bool ring_ok = true) {
  const int t = c->shape;
  if (c->streaming && c->lossy) {
    // variant: axial-symmetric, isotropic (B, K and M y slots alias the x ones), plain
    const int v = c->axi ? 1 : (c->iso_b && c->iso_k && !c->coef_set[FDS_C_MY]) ? 2 : 0;
    const int s0 = t - fds::kNumTiles;
    return c->ks->stream_lossy[ring_ok ? s0 : fds::ring_fallback(s0)][v];
  }
  if (c->streaming) {
    const int s0 = t - fds::kNumTiles, f = stages - 1;
    const int s = ring_ok ? s0 : fds::ring_fallback(s0);
    // variant: axial-symmetric, isotropic (y coefficient slots alias the x ones), plain
    const int v = c->axi ? 1 : (c->iso_b && c->iso_k) ? 2 : 0;
    return c->family == kThermal ? c->ks->stream_thermal[s][v][f]
                                 : c->ks->stream_acoustic[s][v][f];
  }
  switch (c->family) {
    case kAcoustic: return c->ks->acoustic[t][c->axi ? 1 : 0][c->lossy ? 1 : 0];
    case kThermal: return c->ks->thermal[t][c->axi ? 1 : 0];
    default: return c->ks->nonlinear;
  }
}

// Flow along x (pyfds/acoustic_flow.py:49-57): after the step at `step`, every row r whose
// period divides the step (period 0 always; numpy's int % 0 is 0) moves one cell to +x and
// takes a zero at its start, in every state array (and compensation array) of the step's
// output.  One workgroup per row shifts it in place, chunk by chunk from the right end, so
// each chunk reads its left neighbour before that neighbour is overwritten.
template <typename T>
__global__ __launch_bounds__(256) void k_flow_shift(T *a0, T *a1, T *a2, T *c0, T *c1, T *c2,
                                                    int nx, const int64_t *__restrict__ period,
                                                    const int64_t *__restrict__ dyn, int k) {
  const int r = blockIdx.x;
  const int64_t step = dyn[0] + k;
  const int64_t p = period[r];
  const int64_t ap = p < 0 ? -p : p;
  if (!(ap == 0 || step % ap == 0)) return;
  T *rows[6] = {a0, a1, a2, c0, c1, c2};
  for (int q = 0; q < 6; ++q) {
    T *row = rows[q];
    if (!row) continue;
    row += (int64_t)r * nx;
    for (int s = ((nx - 1) / 256) * 256; s >= 0; s -= 256) {
      const int i = s + (int)threadIdx.x;
      const T v = (i >= 1 && i < nx) ? row[i - 1] : T(0);
      __syncthreads();
      if (i < nx) row[i] = v;
      __syncthreads();
    }
  }
}

// Step counters at the start of a run, written in stream order (no host round trip, so a run
// may be enqueued behind one still executing).
__global__ void k_set_dyn(int64_t *dyn, int64_t start) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    dyn[0] = start;
    dyn[1] = 0;
  }
}

__global__ void k_advance(int64_t *dyn, int64_t by) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    dyn[0] += by;
    dyn[1] += by;
  }
}

// Zero the first n work-queue counters (one per launch of a run).
__global__ __launch_bounds__(256) void k_queue_reset(unsigned *q, int n) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) q[i] = 0u;
}

// Non-finite detection.  dyn[kDynNonfinite] holds the smallest (step * 4 + component) over the
// scans that met a NaN or an infinity in a component's state (all bits set: none), where step
// counts the steps done when the scan ran.  A scan reads the state once (grid-stride loop);
// only lanes that meet a non-finite value touch the flag.
constexpr int kDynNonfinite = 4;
constexpr uint64_t kNoNonfinite = ~0ull;
constexpr int64_t kNonfiniteEvery = 50;   // resident 1D engine: steps between scans (at least)

__global__ void k_nonfinite_reset(int64_t *dyn) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    reinterpret_cast<unsigned long long *>(dyn)[kDynNonfinite] = kNoNonfinite;
}

template <typename T>
__global__ __launch_bounds__(256) void k_nonfinite_scan(const T *__restrict__ x, int64_t n,
                                                        int comp,
                                                        const int64_t *__restrict__ base,
                                                        int64_t add, int64_t *flag_dyn) {
  bool bad = false;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    bad |= !__builtin_isfinite(x[i]);
  if (bad)
    atomicMin(reinterpret_cast<unsigned long long *>(flag_dyn) + kDynNonfinite,
              (unsigned long long)((base[0] + add) * 4 + comp));
}

// Launch one step: reads state buffer `parity`, writes parity^1.
template <typename T>
void launch_lbc(fds_ctx *, const fds::StepArgs<T> &) {}

template <>
void launch_lbc<float>(fds_ctx *c, const fds::StepArgs<float> &a) {
  fds::LbcStep st;
  for (int k = 0; k < 3; ++k) {
    st.prev[k] = a.in[k];
    st.cur[k] = a.out[k];
    st.cur_comp[k] = is_kahan(c) ? a.cout[k] : nullptr;
    st.cs[k] = c->lbc_cs[k];
  }
  st.w1 = c->lbc_w1;
  st.w2 = c->lbc_w2;
  st.g = (int)c->nx;
  st.l = c->lbc_layer;
  fds::launch_lbc_step(st, c->stream);
}

// Time steps per launch: the streaming shapes' fused stage count (one step per launch when a
// per-step learned boundary follows the step kernel).
int step_group(const fds_ctx *c) {
  return (c->streaming && !c->lbc_w1 && !c->flow) ? c->fuse : 1;
}

// Launch `stages` fused steps (1 <= stages <= step_group(c)) starting at step k of the launch
// sequence: reads state buffer `parity`, writes parity^1.
template <typename T>
int launch_step(fds_ctx *c, int parity, int k, bool store, int stages) {
  fds::StepArgs<T> a;
  fill_args<T>(c, a, parity);
  a.k = k;
  if (store) a.flags |= fds::kFlagStore;
  bool ring_ok = true;
  if (c->streaming && fds::kStreamShapes[c->shape - fds::kNumTiles][5]) {
    // LDS-ring kernels load 16-byte chunks of rows: rows must start on 16-byte boundaries and
    // every array they read must be 16-byte aligned (whole device allocations are); otherwise
    // the register-path kernel of the same tiling runs
    ring_ok = c->nx % (16 / (int)c->elem) == 0;
    for (int k = 0; k < 3; ++k)
      ring_ok = ring_ok && ((uintptr_t)a.in[k] & 15) == 0 && ((uintptr_t)a.cin[k] & 15) == 0;
    for (int q = 0; q < FDS_NUM_COEFF; ++q) ring_ok = ring_ok && ((uintptr_t)a.coef[q] & 15) == 0;
  }
  const int nb = fds::xcd_grid(c->ntiles);
  pick_launch(c, stages, ring_ok)(&a, nb, lds_bytes(c), c->stream);
  HIP_TRY(hipGetLastError());
  if (c->flow) {
    hipLaunchKernelGGL(k_flow_shift<T>, dim3((unsigned)c->ny), dim3(256), 0, c->stream,
                       a.out[0], a.out[1], a.out[2], a.cout[0], a.cout[1], a.cout[2],
                       (int)c->nx, c->flow_period.as<const int64_t>(), a.dyn, k);
    HIP_TRY(hipGetLastError());
  }
  if (c->lbc_w1) {
    launch_lbc<T>(c, a);
    HIP_TRY(hipGetLastError());
  }
  return 0;
}

// Launch the resident 1D kernel for `steps` steps starting at step k of the run.
template <typename T>
int launch_line(fds_ctx *c, int parity, int k, int steps) {
  fds::StepArgs<T> a;
  fill_args<T>(c, a, parity);
  a.k = k;
  a.steps = steps;
  a.strip_h = c->line_seg;
  a.ghost = c->line_ghost;
  const int fam = c->family == kNonlinear ? fds::kLineNonlinear
                  : c->family == kThermal ? fds::kLineThermal
                  : (c->lossy ? fds::kLineAcousticLossy : fds::kLineAcoustic);
  c->ks->line[c->line_var][fam](&a, c->line_wgs, c->line_threads, c->stream);
  HIP_TRY(hipGetLastError());
  return 0;
}

int launch_step_any(fds_ctx *c, int parity, int k, bool store, int stages = 1) {
  return c->elem == 8 ? launch_step<double>(c, parity, k, store, stages)
                      : launch_step<float>(c, parity, k, store, stages);
}

// Scan component k's state (buffer `parity` where the component ping-pongs) for non-finite
// values on `stream`; a hit records step base[0] + add in flag_dyn.
int scan_nonfinite(const fds_ctx *c, int k, int parity, const int64_t *base, int64_t add,
                   int64_t *flag_dyn, hipStream_t stream) {
  const Comp &cp = c->comp[k];
  if (!cp.present) return 0;
  const bool pp = cp.accumulated || c->family != kThermal || k == 0;
  const void *x = cp.buf[pp ? parity : 0].p;
  const unsigned nb = (unsigned)std::min<int64_t>((c->n + 1023) / 1024, 2048);
  if (c->elem == 8)
    hipLaunchKernelGGL(k_nonfinite_scan<double>, dim3(nb), dim3(256), 0, stream,
                       static_cast<const double *>(x), c->n, k, base, add, flag_dyn);
  else
    hipLaunchKernelGGL(k_nonfinite_scan<float>, dim3(nb), dim3(256), 0, stream,
                       static_cast<const float *>(x), c->n, k, base, add, flag_dyn);
  HIP_TRY(hipGetLastError());
  return 0;
}

int scan_all_nonfinite(const fds_ctx *c, int parity, const int64_t *base, int64_t add,
                       hipStream_t stream) {
  for (int k = 0; k < 3; ++k)
    if (int rc = scan_nonfinite(c, k, parity, base, add, c->dyn.as<int64_t>(), stream))
      return rc;
  return 0;
}

// Copy the context's non-finite result to staging slot `slot` (stream order).
int stage_nonfinite(fds_ctx *c, int slot, hipStream_t stream) {
  if (!c->nf_check) return 0;
  HIP_TRY(c->nf_stage.ensure((kMaxFlights + 1) * sizeof(uint64_t)));
  HIP_TRY(hipMemcpyAsync(static_cast<uint64_t *>(c->nf_stage.p) + slot,
                         c->dyn.as<uint64_t>() + kDynNonfinite, sizeof(uint64_t),
                         hipMemcpyDeviceToHost, stream));
  return 0;
}

// After the copy of stage_nonfinite(c, slot) completed: FDS_E_NONFINITE if a scan met a NaN
// or an infinity.
int check_nonfinite(const fds_ctx *c, int slot) {
  if (!c->nf_check || !c->nf_stage.p) return 0;
  const uint64_t v = static_cast<const uint64_t *>(c->nf_stage.p)[slot];
  if (v == kNoNonfinite) return 0;
  return fail(FDS_E_NONFINITE, "non-finite value in component %d, found after step %lld",
              (int)(v & 3), (long long)(v >> 2));
}

int get_graph(fds_ctx *c, int chunk, int parity, hipGraphExec_t *out) {
  if (c->graph_generation != c->generation) {
    drop_graphs(c);
    c->graph_generation = c->generation;
  }
  auto key = std::make_pair(chunk, parity);
  auto it = c->graphs.find(key);
  if (it != c->graphs.end()) {
    *out = it->second;
    return 0;
  }
  // capture on an idle stream (an asynchronous run may still be executing)
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  const int grp = step_group(c);   // chunk is a multiple of grp
  for (int k = 0; k < chunk && rc == 0; k += grp)
    rc = launch_step_any(c, parity ^ ((k / grp) & 1), k, false, grp);
  if (rc == 0 && c->nf_check)       // the first component after the chunk
    rc = scan_nonfinite(c, 0, parity ^ ((chunk / grp) & 1), c->dyn.as<const int64_t>(), chunk,
                        c->dyn.as<int64_t>(), c->stream);
  if (rc == 0) {
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, c->stream, c->dyn.as<int64_t>(),
                       (int64_t)chunk);
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(FDS_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
  hipGraphExec_t ex = nullptr;
  e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(FDS_E_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
  }
  c->graph_objs.push_back(g);
  c->graphs[key] = ex;
  // the graph's device-side copy now, not on its first launch inside a run
  HIP_TRY(hipGraphUpload(ex, c->stream));
  *out = ex;
  return 0;
}

int check_ctx(fds_ctx *c) {
  if (!c) return fail(FDS_E_ARG, "null context");
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return fail(FDS_E_HIP, "hipSetDevice: %s", hipGetErrorString(e));
  return 0;
}

// Setters, synchronous runs and output reads wait until the asynchronous runs are taken.
int check_idle(fds_ctx *c) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->flights.empty() || c->batch_flights)
    return fail(FDS_E_STATE, "asynchronous runs in flight: take them first");
  return 0;
}

int check_comp(fds_ctx *c, int comp) {
  if (comp < 0 || comp > 2 || !c->comp[comp].present)
    return fail(FDS_E_ARG, "component %d does not exist for this model", comp);
  return 0;
}

}  // namespace

extern "C" {

const char *fds_last_error(void) { return g_last_error.c_str(); }

const char *fds_version(void) { return "pyfds_amd 0.1 (gfx950)"; }

int fds_tile_count(void) { return fds::kNumTiles + fds::kNumStream; }

int fds_engine(fds_ctx *c) {
  if (int rc = check_ctx(c)) return rc;
  if (c->line) return 2;
  if (c->family == kNonlinear) return 3;
  return c->streaming ? 1 : 0;
}

int fds_steps_per_launch(fds_ctx *c) {
  if (int rc = check_ctx(c)) return rc;
  return step_group(c);
}

int fds_tile_engine(int i) {
  if (i < 0 || i >= fds::kNumTiles + fds::kNumStream)
    return fail(FDS_E_ARG, "tile index %d out of range", i);
  return i >= fds::kNumTiles ? 1 : 0;
}

int fds_tile_traits(int i, int *steps, int *nontemporal, int *prefetch, int *lds_ring) {
  if (i < 0 || i >= fds::kNumTiles + fds::kNumStream)
    return fail(FDS_E_ARG, "tile index %d out of range", i);
  if (!steps || !nontemporal || !prefetch || !lds_ring) return fail(FDS_E_ARG, "null argument");
  const bool st = i >= fds::kNumTiles;
  *steps = st ? fds::kStreamShapes[i - fds::kNumTiles][2] : 1;
  *nontemporal = st ? fds::kStreamShapes[i - fds::kNumTiles][3] : 0;
  *prefetch = st ? fds::kStreamShapes[i - fds::kNumTiles][4] : 0;
  *lds_ring = st ? fds::kStreamShapes[i - fds::kNumTiles][5] : 0;
  return 0;
}

int fds_tile_shape(int i, int *tx, int *ty) {
  if (i < 0 || i >= fds::kNumTiles + fds::kNumStream)
    return fail(FDS_E_ARG, "tile index %d out of range", i);
  shape_of(i, tx, ty);
  return 0;
}

int fds_create(fds_ctx **out, int model, int form, int64_t nx, int64_t ny, int device) {
  g_last_error.clear();
  if (!out) return fail(FDS_E_ARG, "null output pointer");
  *out = nullptr;
  if (form < FDS_FP64 || form > FDS_FP32_KAHAN) return fail(FDS_E_ARG, "unknown form %d", form);
  if (nx < 1 || ny < 1) return fail(FDS_E_ARG, "grid must be at least 1x1");
  // kernels index cells with 32-bit ints, including look-ahead rows past the last strip
  if (nx * (ny + 40) + 1024 > (int64_t)std::numeric_limits<int32_t>::max())
    return fail(FDS_E_ARG, "grid of %lld cells exceeds the 32-bit cell index",
                (long long)(nx * ny));
  auto *c = new fds_ctx();
  c->model = model;
  c->form = form;
  c->device = device;
  c->n = nx * ny;
  c->nx = nx;
  c->ny = ny;
  switch (model) {
    case FDS_ACOUSTIC1D: c->family = kAcoustic; c->one_d = true; break;
    case FDS_ACOUSTIC2D: c->family = kAcoustic; break;
    case FDS_ACOUSTIC_AXI: c->family = kAcoustic; c->axi = true; break;
    case FDS_THERMAL1D: c->family = kThermal; c->one_d = true; break;
    case FDS_THERMAL2D: c->family = kThermal; break;
    case FDS_THERMAL_AXI: c->family = kThermal; c->axi = true; break;
    case FDS_IDEALGAS1D:
    case FDS_ACOUSTIC2ND1D: c->family = kNonlinear; break;
    default: delete c; return fail(FDS_E_ARG, "unknown model %d", model);
  }
  if ((c->one_d || c->family == kNonlinear) && ny != 1) {
    delete c;
    return fail(FDS_E_ARG, "1D models need ny = 1");
  }
  if (c->one_d || c->family == kNonlinear) {
    c->nx = c->n;
    c->ny = 1;
  }
  c->elem = (form == FDS_FP64) ? 8 : 4;
  c->ks = form == FDS_FP64 ? &fds::kernels_f64()
                           : (form == FDS_FP32 ? &fds::kernels_f32() : &fds::kernels_k32());
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    delete c;
    return fail(FDS_E_HIP, "hipSetDevice(%d): %s", device, hipGetErrorString(e));
  }
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(FDS_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
      cus > 0)
    c->num_cu = cus;
  // components: all three exist for every model (1D linear models carry an inert third one)
  for (int k = 0; k < 3; ++k) {
    Comp &cp = c->comp[k];
    cp.present = true;
    cp.accumulated = (c->family != kThermal) ? !(c->family == kNonlinear && k == 0) : (k == 0);
    const int nbuf = (c->family == kThermal && k > 0) ? 1 : 2;
    for (int b = 0; b < nbuf; ++b) {
      if (cp.buf[b].ensure(c->n * c->elem) != hipSuccess ||
          hipMemsetAsync(cp.buf[b].p, 0, c->n * c->elem, c->stream) != hipSuccess) {
        fds_destroy(c);
        return fail(FDS_E_HIP, "out of device memory for %lld cells", (long long)c->n);
      }
      if (form == FDS_FP32_KAHAN && cp.accumulated) {
        if (cp.cbuf[b].ensure(c->n * c->elem) != hipSuccess ||
            hipMemsetAsync(cp.cbuf[b].p, 0, c->n * c->elem, c->stream) != hipSuccess) {
          fds_destroy(c);
          return fail(FDS_E_HIP, "out of device memory (compensation)");
        }
      }
    }
  }
  if (c->zeros.ensure(c->n * c->elem) != hipSuccess ||
      hipMemsetAsync(c->zeros.p, 0, c->n * c->elem, c->stream) != hipSuccess ||
      c->dyn.ensure(64) != hipSuccess) {
    fds_destroy(c);
    return fail(FDS_E_HIP, "out of device memory");
  }
  c->s[FDS_S_DX] = 1.0;
  c->s[FDS_S_HALF_DX] = 0.5;
  int rc = choose_tiles(c);
  if (rc) {
    fds_destroy(c);
    return rc;
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    fds_destroy(c);
    return fail(FDS_E_HIP, "stream synchronize failed");
  }
  *out = c;
  return 0;
}

int fds_destroy(fds_ctx *c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graphs(c);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  for (auto &set : c->fev)
    for (auto e : set)
      if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int fds_set_scalars(fds_ctx *c, const double *s, int n) {
  if (int rc = check_idle(c)) return rc;
  if (n < 0 || n > FDS_NUM_SCALARS) return fail(FDS_E_ARG, "bad scalar count %d", n);
  for (int i = 0; i < n; ++i) c->s[i] = s[i];
  touch(c);
  return 0;
}

int fds_set_coeff(fds_ctx *c, int slot, const double *data, int64_t n) {
  if (int rc = check_idle(c)) return rc;
  if (slot < 0 || slot >= FDS_NUM_COEFF) return fail(FDS_E_ARG, "bad coefficient slot %d", slot);
  if (n == 0) {
    c->coef[slot].release();
    c->coef_set[slot] = false;
  } else {
    if (n != c->n) return fail(FDS_E_ARG, "coefficient needs %lld values, got %lld",
                               (long long)c->n, (long long)n);
    HIP_TRY(c->coef[slot].ensure(n * c->elem));
    if (int rc = to_device(c, c->coef[slot].p, data, n)) return rc;
    c->coef_set[slot] = true;
  }
  c->committed = false;
  touch(c);
  return 0;
}

int fds_commit(fds_ctx *c) {
  if (int rc = check_idle(c)) return rc;
  if (c->family == kAcoustic) {
    if (!c->coef_set[FDS_C_KX] || !c->coef_set[FDS_C_BX])
      return fail(FDS_E_STATE, "acoustic coefficients K and B must be set");
    c->lossy = c->coef_set[FDS_C_MX];
    if (c->axi && c->lossy && (!c->coef_set[FDS_C_MC] || !c->coef_set[FDS_C_MT]))
      return fail(FDS_E_STATE, "lossy Axi field needs the MC and MT coefficients");
  } else if (c->family == kThermal) {
    if (!c->coef_set[FDS_C_GX] || !c->coef_set[FDS_C_QX])
      return fail(FDS_E_STATE, "thermal coefficients G and Q must be set");
  } else {
    if (!c->coef_set[FDS_C_RHO0]) return fail(FDS_E_STATE, "static density must be set");
  }
  c->iso_b = !c->one_d && !c->coef_set[FDS_C_BY];   // (thermal: QY)
  c->iso_k = !c->one_d && !c->coef_set[FDS_C_KY];   // (thermal: GY)
  if (c->family != kNonlinear) {
    // the engine depends on the lossy flag just derived
    if (int rc = choose_tiles(c)) return rc;
    const size_t lds = lds_bytes(c);
    if (lds > 65536) {
      // fall back to the default shape, which always fits
      c->tile = -1;
      if (int rc = choose_tiles(c)) return rc;
    }
  }
  c->committed = true;
  touch(c);
  return 0;
}

int fds_set_state(fds_ctx *c, int comp, const double *data, int64_t n) {
  if (int rc = check_idle(c)) return rc;
  if (int rc = check_comp(c, comp)) return rc;
  if (n != c->n) return fail(FDS_E_ARG, "state needs %lld values", (long long)c->n);
  Comp &cp = c->comp[comp];
  const int b = (cp.buf[1].p && (cp.accumulated || c->family != kThermal || comp == 0)) ? c->parity : 0;
  if (c->form == FDS_FP32_KAHAN && cp.accumulated) {
    std::vector<float> x(n), cc(n);
    for (int64_t i = 0; i < n; ++i) {
      x[i] = (float)data[i];
      cc[i] = (float)((double)x[i] - data[i]);
    }
    HIP_TRY(hipMemcpyAsync(cp.buf[b].p, x.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(cp.cbuf[b].p, cc.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
  }
  return to_device(c, cp.buf[b].p, data, n);
}

int fds_get_state(fds_ctx *c, int comp, double *data, int64_t n) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_comp(c, comp)) return rc;
  // the context's own runs in flight are ordered before the copy on its stream; a batch run
  // in flight is on the batch's stream
  if (c->batch_flights) return fail(FDS_E_STATE, "a batch run of this context is in flight");
  if (n != c->n) return fail(FDS_E_ARG, "state holds %lld values", (long long)c->n);
  Comp &cp = c->comp[comp];
  const int b = (cp.buf[1].p && (cp.accumulated || c->family != kThermal || comp == 0)) ? c->parity : 0;
  if (c->elem == 8) {
    HIP_TRY(hipMemcpyAsync(data, cp.buf[b].p, n * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
  }
  std::vector<float> x(n), cc;
  HIP_TRY(hipMemcpyAsync(x.data(), cp.buf[b].p, n * 4, hipMemcpyDeviceToHost, c->stream));
  const bool kahan = c->form == FDS_FP32_KAHAN && cp.accumulated;
  if (kahan) {
    cc.resize(n);
    HIP_TRY(hipMemcpyAsync(cc.data(), cp.cbuf[b].p, n * 4, hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int64_t i = 0; i < n; ++i) data[i] = kahan ? (double)x[i] - (double)cc[i] : (double)x[i];
  return 0;
}

int fds_set_boundaries(fds_ctx *c, int comp, int64_t ncell, const int64_t *cells,
                       const int32_t *ptr, int64_t nent, const double *a, const int64_t *off,
                       const int64_t *stride, const int64_t *len, int64_t ntable,
                       const double *table) {
  if (int rc = check_idle(c)) return rc;
  if (int rc = check_comp(c, comp)) return rc;
  Comp &cp = c->comp[comp];
  if (ncell < 0 || nent < 0 || ntable < 0) return fail(FDS_E_ARG, "negative size");
  if (ncell > 0 && (int64_t)ptr[ncell] != nent) return fail(FDS_E_ARG, "ptr[ncell] != nent");
  for (int64_t u = 0; u < ncell; ++u) {
    if (cells[u] < 0 || cells[u] >= c->n)
      return fail(FDS_E_INDEX, "boundary cell %lld out of range", (long long)cells[u]);
    if (ptr[u] > ptr[u + 1]) return fail(FDS_E_ARG, "ptr not monotone");
  }
  for (int64_t e = 0; e < nent; ++e) {
    if (off[e] < 0 || stride[e] < 0 || len[e] < 1 ||
        (stride[e] > 0 && off[e] + (len[e] - 1) * stride[e] >= ntable) ||
        (stride[e] == 0 && off[e] >= ntable))
      return fail(FDS_E_ARG, "boundary entry %lld addresses outside the table", (long long)e);
  }
  cp.cells.assign(cells, cells + ncell);
  cp.ptr.assign(ptr, ptr + (ncell > 0 ? ncell + 1 : 0));
  cp.a.assign(a, a + nent);
  cp.off.assign(off, off + nent);
  cp.stride.assign(stride, stride + nent);
  cp.len.assign(len, len + nent);
  cp.table.assign(table, table + ntable);
  HIP_TRY(upload(cp.d_ptr, cp.ptr, c->stream));
  HIP_TRY(upload(cp.d_a, cp.a, c->stream));
  HIP_TRY(upload(cp.d_off, cp.off, c->stream));
  HIP_TRY(upload(cp.d_stride, cp.stride, c->stream));
  HIP_TRY(upload(cp.d_table, cp.table, c->stream));
  c->events_dirty = true;
  touch(c);
  return 0;
}

int fds_set_outputs(fds_ctx *c, int comp, int64_t m, const int64_t *cells) {
  if (int rc = check_idle(c)) return rc;
  if (int rc = check_comp(c, comp)) return rc;
  if (m < 0) return fail(FDS_E_ARG, "negative output count");
  for (int64_t k = 0; k < m; ++k)
    if (cells[k] < 0 || cells[k] >= c->n)
      return fail(FDS_E_INDEX, "output cell %lld out of range", (long long)cells[k]);
  Comp &cp = c->comp[comp];
  cp.out_cells.assign(cells, cells + m);
  c->events_dirty = true;
  touch(c);
  return 0;
}

int fds_set_tile(fds_ctx *c, int t) {
  if (int rc = check_idle(c)) return rc;
  if (t < -1 || t >= fds::kNumTiles + fds::kNumStream)
    return fail(FDS_E_ARG, "tile index %d out of range", t);
  const int old = c->tile;
  c->tile = t;
  if (int rc = choose_tiles(c)) return rc;
  if (c->committed && lds_bytes(c) > 65536) {
    c->tile = old;
    choose_tiles(c);
    return fail(FDS_E_ARG, "tile %d needs more than 64 KiB of LDS for this field", t);
  }
  touch(c);
  return 0;
}

int fds_set_line_shape(fds_ctx *c, int cells_per_thread, int threads, int ghost) {
  if (int rc = check_idle(c)) return rc;
  if (!c->one_d && c->family != kNonlinear)
    return fail(FDS_E_ARG, "the resident engine serves 1D models only");
  const int fam = c->family == kNonlinear ? fds::kLineNonlinear : fds::kLineAcoustic;
  int v = -1;
  for (int i = 0; i < fds::kNumLineCpt; ++i)
    if (fds::kLineCpt[i] == cells_per_thread) v = i;
  if (v < 0) return fail(FDS_E_ARG, "cells per thread must be 1, 2, 4 or 8");
  if (threads != 0) {
    if (threads < 64 || threads % 64 || threads > fds::line_max_threads(fam, v))
      return fail(FDS_E_ARG, "segment workgroup of %d threads: a multiple of 64 up to %d",
                  threads, fds::line_max_threads(fam, v));
    if (ghost < 2 || threads * cells_per_thread - 2 * ghost < 1)
      return fail(FDS_E_ARG, "ghost cells must be at least 2 and leave owned cells");
  }
  const int o_rows = c->strip_rows, o_t = c->line_cfg_threads, o_g = c->line_cfg_ghost;
  c->strip_rows = cells_per_thread;
  c->line_cfg_threads = threads;
  c->line_cfg_ghost = threads ? ghost : 0;
  if (int rc = choose_tiles(c)) {
    c->strip_rows = o_rows;
    c->line_cfg_threads = o_t;
    c->line_cfg_ghost = o_g;
    choose_tiles(c);
    return rc;
  }
  touch(c);
  return 0;
}

int fds_set_strip_height(fds_ctx *c, int rows) {
  if (int rc = check_idle(c)) return rc;
  if (rows < 0 || rows > 65536) return fail(FDS_E_ARG, "strip height %d out of range", rows);
  const int old = c->strip_rows;
  c->strip_rows = rows;
  if (int rc = choose_tiles(c)) {
    c->strip_rows = old;
    choose_tiles(c);
    return rc;
  }
  touch(c);
  return 0;
}

}  // extern "C"

namespace {

// Common start of a run of nsteps steps from `start`: signal lengths, event lists, output
// rows, the device step counter.
int prepare_run(fds_ctx *c, int64_t start, int64_t nsteps) {
  // every boundary entry must have a value for every step of the run
  for (int k = 0; k < 3; ++k) {
    const Comp &cp = c->comp[k];
    for (size_t e = 0; e < cp.len.size(); ++e)
      if (start + nsteps > cp.len[e])
        return fail(FDS_E_INDEX, "boundary signal of component %d has %lld samples, step %lld "
                    "requested", k, (long long)cp.len[e], (long long)(start + nsteps - 1));
  }
  if (int rc = rebuild_events(c)) return rc;
  // output rows for this run
  for (int k = 0; k < 3; ++k) {
    Comp &cp = c->comp[k];
    const size_t need = std::max<size_t>(cp.out_cells.size() * nsteps * 8, 16);
    if (cp.obuf.bytes < need) {
      // grow geometrically (and to at least 4096 rows) so repeated runs keep their graphs
      const size_t rows = std::max<int64_t>(nsteps, 4096);
      HIP_TRY(cp.obuf.ensure(std::max(std::max(need, cp.obuf.bytes * 2),
                                      cp.out_cells.size() * rows * 8)));
      touch(c);
    }
    cp.obuf_rows = cp.out_cells.empty() ? 0 : nsteps;
  }
  hipLaunchKernelGGL(k_set_dyn, dim3(1), dim3(64), 0, c->stream, c->dyn.as<int64_t>(), start);
  HIP_TRY(hipGetLastError());
  if (c->nf_check) {
    hipLaunchKernelGGL(k_nonfinite_reset, dim3(1), dim3(64), 0, c->stream, c->dyn.as<int64_t>());
    HIP_TRY(hipGetLastError());
  }
  if (!c->ev0) HIP_TRY(hipEventCreate(&c->ev0));
  if (!c->ev1) HIP_TRY(hipEventCreate(&c->ev1));
  if (c->line) {
    // resident 1D engine: the signal table is addressed with 32-bit indices
    for (int k = 0; k < 3; ++k)
      if (c->comp[k].table.size() > (size_t)std::numeric_limits<int32_t>::max())
        return fail(FDS_E_ARG, "boundary table of component %d too large for the resident 1D "
                    "engine (select tile %d)", k, fds::kTileRow);
  }
  return 0;
}

template <typename T>
int launch_couple(fds_ctx *ac, fds_ctx *th, int k, int steps, int threads,
                  const fds::LineCouple<T> &cp) {
  fds::StepArgs<T> a, h;
  fill_args<T>(ac, a, ac->parity);
  fill_args<T>(th, h, th->parity);
  a.k = h.k = k;
  a.steps = h.steps = steps;
  a.strip_h = h.strip_h = ac->line_seg;
  a.ghost = h.ghost = ac->line_ghost;
  ac->ks->couple[ac->line_var][ac->lossy ? 1 : 0](&a, &h, &cp, ac->line_wgs, threads,
                                                   ac->stream);
  HIP_TRY(hipGetLastError());
  return 0;
}

template <typename T>
int couple_run(fds_ctx *ac, fds_ctx *th, const double *rho, const double *loss, double dt,
               int64_t stepping, int accumulate, double *accum, int64_t start, int64_t nsteps,
               int threads) {
  const int64_t n = ac->n;
  std::vector<T> r(n), l(n);
  for (int64_t i = 0; i < n; ++i) {
    r[i] = (T)rho[i];
    l[i] = (T)loss[i];
  }
  std::vector<double> acc(accumulate ? accum : nullptr, accumulate ? accum + n : nullptr);
  if (acc.empty()) acc.assign(1, 0.0);
  HIP_TRY(upload(ac->couple_rho, r, ac->stream));
  HIP_TRY(upload(ac->couple_loss, l, ac->stream));
  HIP_TRY(upload(ac->couple_acc, acc, ac->stream));
  fds::LineCouple<T> cp;
  cp.rho = ac->couple_rho.as<const T>();
  cp.loss = ac->couple_loss.as<const T>();
  cp.acc = ac->couple_acc.as<double>();
  cp.dt = (T)dt;
  cp.stepping = stepping;
  cp.accumulate = accumulate ? 1 : 0;
  HIP_TRY(hipEventRecord(ac->ev0, ac->stream));
  // segments: the loss transfer reads the velocity's left neighbour, so wrong values reach
  // one cell further into the temperature than into the fields: ghost - 2 steps per launch
  const int per = ac->line_wgs > 1 ? std::min(ac->line_ghost - 2, fds::kLineSteps)
                                   : fds::kLineSteps;
  int64_t done = 0;
  while (done < nsteps) {
    const int st = (int)std::min<int64_t>(nsteps - done, per);
    if (int rc = launch_couple<T>(ac, th, (int)done, st, threads, cp)) return rc;
    ac->parity ^= 1;
    th->parity ^= 1;
    done += st;
  }
  HIP_TRY(hipEventRecord(ac->ev1, ac->stream));
  HIP_TRY(hipStreamSynchronize(ac->stream));
  if (accumulate)
    HIP_TRY(hipMemcpy(accum, ac->couple_acc.p, n * sizeof(double), hipMemcpyDeviceToHost));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, ac->ev0, ac->ev1));
  ac->last_run_ms = th->last_run_ms = ms;
  return 0;
}

// The steps of a prepared run (prepare_run) enqueued on the context's stream, between events
// e0 and e1.
// Capture the chunk graphs a run of nsteps steps replays (at most two: one per buffer parity
// it starts a chunk on).  Capturing is set-up: it waits for the stream to be idle, then builds
// the graph on the host.
// Graph of the whole launches before the last one of a run's tail of `rest` steps (at most one
// chunk): rest - 1 rounded down to whole launches, when that is at least two launches.  0: the
// tail runs as plain launches (also once kMaxShortGraphs tail lengths hold graphs).
constexpr int kMaxShortGraphs = 8;

int short_graph_steps(const fds_ctx *c, int64_t rest, int parity) {
  const int g = step_group(c);
  const int64_t len = (rest - 1) / g * g;
  if (len < 2 * g) return 0;
  if (c->graphs.count(std::make_pair((int)len, parity))) return (int)len;
  int shorts = 0;
  for (const auto &kv : c->graphs) shorts += kv.first.first != c->graph_chunk;
  return shorts < 2 * kMaxShortGraphs ? (int)len : 0;
}

// both_parities: also the graphs a run starting from the other state buffer replays (runs
// prepared ahead of other runs of odd launch counts).
int capture_run_graphs(fds_ctx *c, int64_t nsteps, int chunk, bool both_parities = false) {
  if (c->line || chunk <= 0) return 0;
  if (both_parities) {
    c->parity ^= 1;
    const int rc = capture_run_graphs(c, nsteps, chunk);
    c->parity ^= 1;
    if (rc) return rc;
  }
  const int g = step_group(c);
  const int ch = std::max(g, chunk - chunk % g);
  c->graph_chunk = ch;
  int parity = c->parity;
  for (int64_t left = nsteps, n = 0; left > ch && n < 2; left -= ch, ++n) {
    hipGraphExec_t ex;
    if (int rc = get_graph(c, ch, parity, &ex)) return rc;
    parity ^= ((ch / g) & 1);
  }
  // the tail after the full chunks (a run shorter than a chunk is all tail)
  const int64_t tail = (nsteps - 1) % ch + 1, full = (nsteps - tail) / ch;
  parity = c->parity ^ (int)((full * (ch / g)) & 1);
  if (const int len = short_graph_steps(c, tail, parity)) {
    hipGraphExec_t ex;
    if (int rc = get_graph(c, len, parity, &ex)) return rc;
  }
  return 0;
}

int enqueue_steps(fds_ctx *c, int64_t nsteps, int chunk, hipEvent_t e0, hipEvent_t e1) {
  // graphs first, so that the events bracket the steps only
  if (int rc = capture_run_graphs(c, nsteps, chunk)) return rc;
  HIP_TRY(hipEventRecord(e0, c->stream));
  int64_t done = 0;
  if (c->line) {
    // one segment: up to kLineSteps steps per launch.  Several: a wrong value enters each
    // segment from its open ends, one cell per step from the left and, where the update of
    // the second component reads its right neighbour's old value (viscous and convective
    // terms, the nonlinear mass flux), one more cell on the right: ghost - 1 steps per launch.
    const int per = c->line_wgs > 1 ? std::min(c->line_ghost - 1, fds::kLineSteps)
                                    : fds::kLineSteps;
    int64_t scanned = 0;
    while (done < nsteps) {
      const int st = (int)std::min<int64_t>(nsteps - done, per);
      const int rc = c->elem == 8 ? launch_line<double>(c, c->parity, (int)done, st)
                                  : launch_line<float>(c, c->parity, (int)done, st);
      if (rc) return rc;
      c->parity ^= 1;
      done += st;
      // the first component every kNonfiniteEvery steps or more (launch granularity)
      if (c->nf_check && done < nsteps && done - scanned >= kNonfiniteEvery) {
        if (int rc2 = scan_nonfinite(c, 0, c->parity, c->dyn.as<const int64_t>(), done,
                                     c->dyn.as<int64_t>(), c->stream))
          return rc2;
        scanned = done;
      }
    }
    chunk = 0;
  }
  const int g = step_group(c);
  int64_t graph_done = 0;    // steps the device step counter advanced by (graph chunks)
  if (chunk > 0) {
    // graph-replayed full chunks (whole launches of g steps), never including the run's last
    // step (which stores the assigned components)
    chunk = std::max(g, chunk - chunk % g);
    while (nsteps - done > chunk) {
      hipGraphExec_t ex;
      if (int rc = get_graph(c, chunk, c->parity, &ex)) return rc;
      HIP_TRY(hipGraphLaunch(ex, c->stream));
      c->parity ^= ((chunk / g) & 1);
      done += chunk;
      graph_done += chunk;
    }
    // the tail's whole launches but the last as one graph too (one launch call, not several)
    if (const int len = short_graph_steps(c, nsteps - done, c->parity)) {
      hipGraphExec_t ex;
      if (int rc = get_graph(c, len, c->parity, &ex)) return rc;
      HIP_TRY(hipGraphLaunch(ex, c->stream));
      c->parity ^= ((len / g) & 1);
      done += len;
      graph_done += len;
    }
  }
  // the rest as plain launches of g stages, then one launch of the remaining stages (k counts
  // from the device step counter's current base)
  for (int64_t k = 0; done + k < nsteps;) {
    const int st = (int)std::min<int64_t>(g, nsteps - done - k);
    if (int rc = launch_step_any(c, c->parity, (int)k, done + k + st == nsteps, st)) return rc;
    c->parity ^= 1;
    k += st;
  }
  if (c->nf_check)   // every component after the run
    if (int rc = scan_all_nonfinite(c, c->parity, c->dyn.as<const int64_t>(),
                                    nsteps - graph_done, c->stream))
      return rc;
  HIP_TRY(hipEventRecord(e1, c->stream));
  return 0;
}

// True when prepare_run(c, ., nsteps) neither rebuilds the event tables nor grows an output
// buffer (so it frees no device memory a run in flight may use).
bool run_steady(const fds_ctx *c, int64_t nsteps) {
  if (c->events_dirty || (!c->graphs.empty() && c->graph_generation != c->generation))
    return false;
  for (int k = 0; k < 3; ++k) {
    const Comp &cp = c->comp[k];
    if (cp.obuf.bytes < std::max<size_t>(cp.out_cells.size() * nsteps * 8, 16)) return false;
  }
  return true;
}

int ensure_events(hipEvent_t (&ev)[3]) {
  for (int i = 0; i < 3; ++i)
    if (!ev[i]) HIP_TRY(hipEventCreate(&ev[i]));
  return 0;
}

// Enqueue the copy of a run's recorded outputs (rows x outputs per component) into staging
// slot `slot` on stream s.
int stage_outputs(fds_ctx *c, int slot, int64_t rows, hipStream_t s) {
  for (int k = 0; k < 3; ++k) {
    Comp &cp = c->comp[k];
    const size_t bytes = cp.out_cells.size() * (size_t)rows * 8;
    if (!bytes) continue;
    HIP_TRY(cp.stage[slot].ensure(bytes));
    HIP_TRY(hipMemcpyAsync(cp.stage[slot].p, cp.obuf.p, bytes, hipMemcpyDeviceToHost, s));
  }
  return 0;
}

// Copy a finished run's staged outputs to the caller (null destinations are skipped).
void take_outputs(fds_ctx *c, const Flight &f, double *const *dst) {
  for (int k = 0; k < 3; ++k) {
    const Comp &cp = c->comp[k];
    const size_t bytes = cp.out_cells.size() * (size_t)f.rows * 8;
    if (bytes && dst[k]) std::memcpy(dst[k], cp.stage[f.slot].p, bytes);
  }
}

}  // namespace

extern "C" {

int fds_couple_loss_run(fds_ctx *ac, fds_ctx *th, const double *rho, const double *loss,
                        double dt, int64_t stepping, int accumulate, double *accum,
                        int64_t start, int64_t nsteps) {
  if (int rc = check_idle(ac)) return rc;
  if (int rc = check_idle(th)) return rc;
  if (ac->model != FDS_ACOUSTIC1D || th->model != FDS_THERMAL1D)
    return fail(FDS_E_ARG, "the loss coupling joins an Acoustic1D and a Thermal1D context");
  if (ac->device != th->device || ac->form != th->form || ac->n != th->n)
    return fail(FDS_E_ARG, "coupled contexts must share device, form and cell count");
  if (!ac->committed || !th->committed) return fail(FDS_E_STATE, "coefficients not committed");
  if (ac->tile >= 0 || th->tile >= 0)
    return fail(FDS_E_ARG, "coupled contexts must use the resident 1D engine (tile -1)");
  if (!rho || !loss || (accumulate && !accum)) return fail(FDS_E_ARG, "null coupling input");
  if (stepping < 1) return fail(FDS_E_ARG, "coupling stepping must be at least 1");
  if (nsteps < 0 || start < 0) return fail(FDS_E_ARG, "negative step count");
  if (nsteps == 0) return 0;
  // one layout for both contexts: one workgroup of one cell per thread for short lines (the
  // coupled kernel holds two fields, so it has half the single-field threads from 2 cells per
  // thread on), else the single-field default segments
  const bool seg = ac->n > fds::kLineSegmentAbove;
  const int cpt = seg ? fds::kLineSegCpt : 1;
  const int cfg_t = seg ? fds::kLineSegThreads : 0, cfg_g = seg ? fds::kLineSegGhost : 0;
  static_assert(fds::kLineSegThreads <= fds::kCoupleMaxThreads[1] && fds::kLineSegCpt == 2 &&
                    fds::kLineSegmentAbove <= fds::kCoupleMaxThreads[0] &&
                    fds::kLineSegGhost > 2,
                "coupled layouts must fit the coupled kernel's launch bounds");
  for (fds_ctx *c : {ac, th}) {
    if (!c->line || c->strip_rows != cpt || c->line_cfg_threads != cfg_t ||
        c->line_cfg_ghost != cfg_g) {
      c->strip_rows = cpt;
      c->line_cfg_threads = cfg_t;
      c->line_cfg_ghost = cfg_g;
      if (int rc = choose_tiles(c)) return rc;
      touch(c);
    }
    if (!c->line)
      return fail(FDS_E_ARG, "coupled runs hold at most %lld cells",
                  (long long)fds::kLineMaxResident);
    if (int rc = prepare_run(c, start, nsteps)) return rc;
  }
  // the coupled kernel runs on the acoustic context's stream and reads the thermal context's
  // step counters, which prepare_run wrote on the thermal stream
  HIP_TRY(hipStreamSynchronize(th->stream));
  const int threads = ac->line_threads;
  return ac->elem == 8 ? couple_run<double>(ac, th, rho, loss, dt, stepping, accumulate, accum,
                                            start, nsteps, threads)
                       : couple_run<float>(ac, th, rho, loss, dt, stepping, accumulate, accum,
                                           start, nsteps, threads);
}

int fds_run(fds_ctx *c, int64_t start, int64_t nsteps, int chunk) {
  if (int rc = check_idle(c)) return rc;
  if (!c->committed) return fail(FDS_E_STATE, "coefficients not committed");
  if (nsteps < 0 || start < 0) return fail(FDS_E_ARG, "negative step count");
  if (nsteps == 0) return 0;
  if (int rc = prepare_run(c, start, nsteps)) return rc;
  if (int rc = enqueue_steps(c, nsteps, chunk, c->ev0, c->ev1)) return rc;
  if (int rc = stage_nonfinite(c, kMaxFlights, c->stream)) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_run_ms = ms;
  return check_nonfinite(c, kMaxFlights);
}

int fds_prepare_run(fds_ctx *c, int64_t start, int64_t nsteps, int chunk) {
  if (int rc = check_idle(c)) return rc;
  if (!c->committed) return fail(FDS_E_STATE, "coefficients not committed");
  if (nsteps < 1 || start < 0) return fail(FDS_E_ARG, "a run needs steps");
  if (int rc = prepare_run(c, start, nsteps)) return rc;
  if (int rc = capture_run_graphs(c, nsteps, chunk, true)) return rc;
  // pinned staging and events of asynchronous runs of this length (both slots)
  for (int slot = 0; slot < kMaxFlights; ++slot) {
    if (int rc = ensure_events(c->fev[slot])) return rc;
    for (int k = 0; k < 3; ++k) {
      const size_t bytes = c->comp[k].out_cells.size() * (size_t)nsteps * 8;
      if (bytes) HIP_TRY(c->comp[k].stage[slot].ensure(bytes));
    }
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int fds_run_async(fds_ctx *c, int64_t start, int64_t nsteps, int chunk) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->committed) return fail(FDS_E_STATE, "coefficients not committed");
  if (nsteps < 1 || start < 0) return fail(FDS_E_ARG, "an asynchronous run needs steps");
  if (c->batch_flights) return fail(FDS_E_STATE, "a batch run of this context is in flight");
  if ((int)c->flights.size() >= kMaxFlights)
    return fail(FDS_E_STATE, "%d runs in flight: take the oldest first", kMaxFlights);
  // a run that rebuilds tables or grows the output rows frees device memory the runs in
  // flight may read: let them finish first
  if (!run_steady(c, nsteps)) HIP_TRY(hipStreamSynchronize(c->stream));
  if (int rc = prepare_run(c, start, nsteps)) return rc;
  const int slot = c->next_slot;
  if (int rc = ensure_events(c->fev[slot])) return rc;
  if (int rc = enqueue_steps(c, nsteps, chunk, c->fev[slot][0], c->fev[slot][1])) return rc;
  if (int rc = stage_outputs(c, slot, nsteps, c->stream)) return rc;
  if (int rc = stage_nonfinite(c, slot, c->stream)) return rc;
  HIP_TRY(hipEventRecord(c->fev[slot][2], c->stream));
  if (c->flights.empty()) c->last_run_ms = 0.0;   // a new sequence of asynchronous runs
  c->flights.push_back(Flight{nsteps, slot});
  c->next_slot = (slot + 1) % kMaxFlights;
  return 0;
}

int fds_run_take(fds_ctx *c, double *dst0, double *dst1, double *dst2) {
  if (int rc = check_ctx(c)) return rc;
  if (c->flights.empty()) return fail(FDS_E_STATE, "no asynchronous run in flight");
  const Flight f = c->flights.front();
  HIP_TRY(hipEventSynchronize(c->fev[f.slot][2]));
  c->flights.pop_front();
  double *dst[3] = {dst0, dst1, dst2};
  take_outputs(c, f, dst);
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, c->fev[f.slot][0], c->fev[f.slot][1]));
  c->last_run_ms += ms;
  return check_nonfinite(c, f.slot);
}

int fds_set_nonfinite_check(fds_ctx *c, int enable) {
  if (int rc = check_idle(c)) return rc;
  if (c->nf_check != (enable != 0)) {
    c->nf_check = enable != 0;
    touch(c);   // the chunk graphs hold the scans
  }
  return 0;
}

int fds_set_flow(fds_ctx *c, const int64_t *period, int64_t nrows) {
  if (int rc = check_idle(c)) return rc;
  if (!period || nrows == 0) {
    c->flow = false;
    touch(c);
    return 0;
  }
  if (c->family != kAcoustic || c->one_d)
    return fail(FDS_E_ARG, "flow is implemented for the 2D acoustic models");
  if (nrows != c->ny) return fail(FDS_E_ARG, "flow needs %lld row periods, got %lld",
                                  (long long)c->ny, (long long)nrows);
  std::vector<int64_t> p(period, period + nrows);
  HIP_TRY(upload(c->flow_period, p, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->flow = true;
  touch(c);
  return 0;
}

int fds_set_learned_boundary(fds_ctx *c, const uint16_t *w1, const uint16_t *w2, int layer,
                             const float *comp_scale) {
  if (int rc = check_idle(c)) return rc;
  if (!w1) {
    c->lbc_w1 = c->lbc_w2 = nullptr;
    touch(c);
    return 0;
  }
  if (!w2 || !comp_scale) return fail(FDS_E_ARG, "weights and component scales are required");
  if (c->family != kAcoustic || c->axi || c->one_d)
    return fail(FDS_E_ARG, "the learned boundary is implemented for Acoustic2D");
  if (c->elem != 4) return fail(FDS_E_ARG, "the learned boundary needs an FP32 form");
  if (c->nx != c->ny) return fail(FDS_E_ARG, "the learned boundary needs a square grid");
  if (layer != fds::kLbcDepth || c->nx < 2 * layer + fds::kLbcBand + 1)
    return fail(FDS_E_ARG, "layer depth must be %d on a grid wider than %d", fds::kLbcDepth,
                2 * fds::kLbcDepth + fds::kLbcBand);
  for (int k = 0; k < 3; ++k)
    if (!(comp_scale[k] > 0.f)) return fail(FDS_E_ARG, "component scales must be positive");
  c->lbc_w1 = w1;
  c->lbc_w2 = w2;
  c->lbc_layer = layer;
  for (int k = 0; k < 3; ++k) c->lbc_cs[k] = comp_scale[k];
  touch(c);
  return 0;
}

int fds_last_run_ms(fds_ctx *c, double *ms) {
  if (int rc = check_ctx(c)) return rc;
  if (!ms) return fail(FDS_E_ARG, "null output pointer");
  *ms = c->last_run_ms;
  return 0;
}

int fds_read_outputs(fds_ctx *c, int comp, int64_t row0, int64_t nrows, double *dst) {
  if (int rc = check_idle(c)) return rc;
  if (int rc = check_comp(c, comp)) return rc;
  Comp &cp = c->comp[comp];
  const int64_t m = (int64_t)cp.out_cells.size();
  if (m == 0 || nrows == 0) return 0;
  if (row0 < 0 || nrows < 0 || row0 + nrows > cp.obuf_rows)
    return fail(FDS_E_INDEX, "output rows [%lld, %lld) not recorded", (long long)row0,
                (long long)(row0 + nrows));
  HIP_TRY(hipMemcpyAsync(dst, cp.obuf.as<double>() + row0 * m, nrows * m * 8,
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

}  // extern "C"

namespace {

bool batch_stale(const fds_batch *b) {
  if (b->build_gen.size() != b->ctxs.size()) return true;
  for (size_t i = 0; i < b->ctxs.size(); ++i) {
    const fds_ctx *c = b->ctxs[i];
    if (c->generation != b->build_gen[i] || (c->parity ^ b->build_parity[i]) != b->flip)
      return true;
  }
  return false;
}

void batch_drop_graphs(fds_batch *b) {
  for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto g : b->graph_objs) (void)hipGraphDestroy(g);
  b->graphs.clear();
  b->graph_objs.clear();
}

template <typename T>
int batch_build(fds_batch *b, int64_t nsteps) {
  for (auto *c : b->ctxs) {
    if (int rc = rebuild_events(c)) return rc;
    for (int k = 0; k < 3; ++k) {
      Comp &cp = c->comp[k];
      const size_t need = std::max<size_t>(cp.out_cells.size() * nsteps * 8, 16);
      if (cp.obuf.bytes < need) {
        const size_t rows = std::max<int64_t>(nsteps, 4096);
        HIP_TRY(cp.obuf.ensure(std::max(std::max(need, cp.obuf.bytes * 2),
                                        cp.out_cells.size() * rows * 8)));
        touch(c);
      }
      cp.obuf_rows = cp.out_cells.empty() ? 0 : nsteps;
    }
  }
  if (!batch_stale(b)) return 0;
  const int ni = (int)b->ctxs.size();
  for (int parity = 0; parity < 2; ++parity)
    for (int store = 0; store < 2; ++store) {
      std::vector<fds::StepArgs<T>> d(ni);
      for (int i = 0; i < ni; ++i) {
        fds_ctx *c = b->ctxs[i];
        // members flip in lock step: descriptor set `parity` reads each member's buffer
        // (its parity at build time) ^ parity
        fill_args<T>(c, d[i], c->parity ^ parity);
        d[i].dyn = b->dyn.as<const int64_t>();
        if (c->axi) d[i].flags |= fds::kFlagAxi;
        if (c->lossy) d[i].flags |= fds::kFlagLossy;
        if (c->family == kThermal) d[i].flags |= fds::kFlagThermal;
        if (store) d[i].flags |= fds::kFlagStore;
      }
      HIP_TRY(upload(b->descs[parity][store], d, b->stream));
    }
  batch_drop_graphs(b);
  b->build_gen.clear();
  b->build_parity.clear();
  for (auto *c : b->ctxs) {
    b->build_gen.push_back(c->generation);
    b->build_parity.push_back(c->parity);
  }
  b->flip = 0;
  return 0;
}

int batch_launch(fds_batch *b, int rel_parity, int k, bool store, int stages = 1) {
  fds::BatchLaunch L;
  L.descs = b->descs[rel_parity][store ? 1 : 0].p;
  L.tile_inst = b->tile_inst.as<const int32_t>();
  L.inst_base = b->inst_base.as<const int32_t>();
  L.ntiles = b->ntiles;
  L.k = k;
  L.order = b->order.as<const int32_t>();
  L.queue = b->queue.as<unsigned>();
  L.nq = b->nq;
  L.spl = b->fuse;
  L.num_cu = b->ctxs[0]->num_cu;
  // the work queue serves full launches; remainder launches (fewer stages) and other shapes
  // run on the dispatcher kernels
  if (b->streaming && b->schedule == 1 && stages == b->fuse)
    b->ks->stream_batch_wq[b->shape - fds::kNumTiles][stages - 1](L, fds::xcd_grid(b->ntiles),
                                                                    0, b->stream);
  else if (b->streaming)
    b->ks->stream_batch[b->shape - fds::kNumTiles][stages - 1](L, fds::xcd_grid(b->ntiles), 0,
                                                                 b->stream);
  else
    b->ks->batch[b->shape](L, fds::xcd_grid(b->ntiles), b->lds, b->stream);
  HIP_TRY(hipGetLastError());
  return 0;
}

int batch_graph(fds_batch *b, int chunk, int rel_parity, hipGraphExec_t *out) {
  auto key = std::make_pair(chunk, rel_parity);
  auto it = b->graphs.find(key);
  if (it != b->graphs.end()) {
    *out = it->second;
    return 0;
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  const int grp = b->fuse;   // chunk is a multiple of grp
  for (int k = 0; k < chunk && rc == 0; k += grp)
    rc = batch_launch(b, rel_parity ^ ((k / grp) & 1), k, false, grp);
  if (rc == 0)
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, b->stream, b->dyn.as<int64_t>(),
                       (int64_t)chunk);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(FDS_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
  hipGraphExec_t ex = nullptr;
  e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(FDS_E_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
  }
  b->graph_objs.push_back(g);
  b->graphs[key] = ex;
  HIP_TRY(hipGraphUpload(ex, b->stream));
  *out = ex;
  return 0;
}

// Checks a batch run's members, (re)builds the descriptors and output rows for nsteps steps
// and enqueues the step counters' start.
int batch_prepare(fds_batch *b, int64_t start, int64_t nsteps) {
  for (auto *c : b->ctxs) {
    if (!c->committed) return fail(FDS_E_STATE, "coefficients not committed");
    if (!c->flights.empty()) return fail(FDS_E_STATE, "a member has asynchronous runs in flight");
    for (int k = 0; k < 3; ++k)
      for (size_t e = 0; e < c->comp[k].len.size(); ++e)
        if (start + nsteps > c->comp[k].len[e])
          return fail(FDS_E_INDEX, "boundary signal too short for step %lld",
                      (long long)(start + nsteps - 1));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  int rc = b->elem == 8 ? batch_build<double>(b, nsteps) : batch_build<float>(b, nsteps);
  if (rc) return rc;
  hipLaunchKernelGGL(k_set_dyn, dim3(1), dim3(64), 0, b->stream, b->dyn.as<int64_t>(), start);
  HIP_TRY(hipGetLastError());
  if (b->schedule == 1) {   // one empty counter per launch of the run
    const int64_t need = nsteps / std::max(1, b->fuse) + 2;
    if (need > std::numeric_limits<int32_t>::max())
      return fail(FDS_E_ARG, "run too long for the work queue");
    if (b->nq < need) {
      // a larger buffer (batch_steady made runs in flight finish first)
      HIP_TRY(b->queue.ensure((size_t)need * 2 * sizeof(unsigned)));
      b->nq = (int)(b->queue.bytes / sizeof(unsigned));
      batch_drop_graphs(b);   // captured launches hold the old buffer
    }
    const unsigned grid = (unsigned)std::min<int64_t>((need + 255) / 256, 1024);
    hipLaunchKernelGGL(k_queue_reset, dim3(grid), dim3(256), 0, b->stream,
                       b->queue.as<unsigned>(), (int)need);
    HIP_TRY(hipGetLastError());
  }
  for (auto *c : b->ctxs)
    if (c->nf_check) {
      hipLaunchKernelGGL(k_nonfinite_reset, dim3(1), dim3(64), 0, b->stream,
                         c->dyn.as<int64_t>());
      HIP_TRY(hipGetLastError());
    }
  return 0;
}

// True when batch_prepare rebuilds nothing (no device memory freed under a run in flight).
bool batch_steady(const fds_batch *b, int64_t nsteps) {
  for (const auto *c : b->ctxs)
    if (!run_steady(c, nsteps)) return false;
  if (b->schedule == 1 && b->nq < nsteps / std::max(1, b->fuse) + 2) return false;
  return !batch_stale(b);
}

// The steps of a prepared batch run between events e0 and e1; the members' buffer parities
// follow the launches.
int batch_enqueue(fds_batch *b, int64_t nsteps, int chunk, hipEvent_t e0, hipEvent_t e1) {
  HIP_TRY(hipEventRecord(e0, b->stream));
  int rel = b->flip;
  int64_t done = 0;
  const int g = b->fuse;
  if (chunk > 0) {
    chunk = std::max(g, chunk - chunk % g);
    while (nsteps - done > chunk) {
      hipGraphExec_t ex;
      if (int rc2 = batch_graph(b, chunk, rel, &ex)) return rc2;
      HIP_TRY(hipGraphLaunch(ex, b->stream));
      rel ^= ((chunk / g) & 1);
      done += chunk;
    }
  }
  const int64_t graph_done = done;   // steps the batch step counter advanced by
  // full launches, then the remainder one stage at a time (batch kernels exist for the shape's
  // stage count and for one stage)
  for (int64_t k = 0; done + k < nsteps;) {
    const int st = nsteps - done - k >= g ? g : 1;
    if (int rc2 = batch_launch(b, rel, (int)k, done + k + st == nsteps, st)) return rc2;
    rel ^= 1;
    k += st;
  }
  HIP_TRY(hipEventRecord(e1, b->stream));
  // members advanced nsteps steps in lock step, flipping buffers once per launch
  if (rel != b->flip) {
    for (auto *c : b->ctxs) c->parity ^= 1;
    b->flip ^= 1;
  }
  // members that check for non-finite values: every component after the run
  for (auto *c : b->ctxs)
    if (c->nf_check)
      if (int rc2 = scan_all_nonfinite(c, c->parity, b->dyn.as<const int64_t>(),
                                       nsteps - graph_done, b->stream))
        return rc2;
  return 0;
}

// The first member whose non-finite check of staging slot `slot` failed, or 0.
int batch_check_nonfinite(const fds_batch *b, int slot) {
  for (size_t i = 0; i < b->ctxs.size(); ++i)
    if (int rc = check_nonfinite(b->ctxs[i], slot)) {
      const std::string msg = g_last_error;
      return fail(rc, "batch member %d: %s", (int)i, msg.c_str());
    }
  return 0;
}

}  // namespace

extern "C" {

int fds_batch_create(fds_batch **out, fds_ctx *const *ctxs, int n, int tile) {
  if (!out || !ctxs || n < 1) return fail(FDS_E_ARG, "batch needs at least one context");
  *out = nullptr;
  auto *b = new fds_batch();
  b->ctxs.assign(ctxs, ctxs + n);
  fds_ctx *c0 = ctxs[0];
  b->device = c0->device;
  b->form = c0->form;
  b->elem = c0->elem;
  b->ks = c0->ks;
  b->tile = tile;
  for (auto *c : b->ctxs) {
    if (!c || !c->committed) { delete b; return fail(FDS_E_STATE, "batch member not committed"); }
    if (!c->flights.empty() || c->batch_flights) {
      delete b;
      return fail(FDS_E_STATE, "a batch member has asynchronous runs in flight");
    }
    if (c->device != b->device || c->form != b->form) {
      delete b;
      return fail(FDS_E_ARG, "batch members must share device and form");
    }
    if (c->family == kNonlinear) {
      delete b;
      return fail(FDS_E_ARG, "nonlinear 1D models cannot be batched");
    }
    // the batch launches only the step kernels: a member's per-step flow shift or learned
    // boundary would be skipped
    if (c->flow || c->lbc_w1) {
      delete b;
      return fail(FDS_E_ARG, "members with flow (AcousticFlow2D) or a learned boundary "
                             "cannot be batched");
    }
  }
  hipError_t e = hipSetDevice(b->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete b; return fail(FDS_E_HIP, "stream: %s", hipGetErrorString(e)); }
  // one tile shape for every member (re-derives each member's events); the streaming
  // engine only when it covers every member
  bool all_stream = true;
  for (auto *c : b->ctxs) all_stream = all_stream && can_stream_batch(c);
  int t = tile < 0 ? (all_stream ? fds::kDefaultLossless2D : fds::kDefaultTile2D) : tile;
  if (t >= fds::kNumTiles + fds::kNumStream || t < 0) {
    fds_batch_destroy(b);
    return fail(FDS_E_ARG, "tile index %d out of range", t);
  }
  if (t >= fds::kNumTiles && !all_stream) t = fds::kDefaultTile2D;
  if (t >= fds::kNumTiles && fds::stream_single_only(t - fds::kNumTiles)) {
    fds_batch_destroy(b);
    return fail(FDS_E_ARG, "shape %d (prefetch %d rows) serves single fields only", t,
                fds::kStreamShapes[t - fds::kNumTiles][4]);
  }
  std::vector<int32_t> tinst, base;
  size_t lds = 0;
  for (int i = 0; i < n; ++i) {
    fds_ctx *c = b->ctxs[i];
    c->tile = t;
    if (int rc = choose_tiles(c)) { fds_batch_destroy(b); return rc; }
    touch(c);
    lds = std::max(lds, lds_bytes(c));
    base.push_back((int32_t)tinst.size());
    for (int k = 0; k < c->ntiles; ++k) tinst.push_back(i);
  }
  if (lds > 65536) {
    fds_batch_destroy(b);
    return fail(FDS_E_ARG, "tile %d needs more than 64 KiB of LDS for some member", t);
  }
  b->shape = t;
  b->streaming = t >= fds::kNumTiles;
  b->fuse = b->streaming ? fds::kStreamShapes[t - fds::kNumTiles][2] : 1;
  shape_of(t, &b->tx, &b->ty);
  b->lds = lds;
  b->ntiles = (int)tinst.size();
  // work-queue order: largest tiles first (owned cells plus ghosts; edge tiles are smaller)
  std::vector<int64_t> cost(b->ntiles);
  for (int g = 0; g < b->ntiles; ++g) {
    const fds_ctx *c = b->ctxs[tinst[g]];
    const int t = g - base[tinst[g]];
    const int64_t ty = t / std::max(1, c->tiles_x), tx = t - ty * c->tiles_x;
    const int64_t sx = std::max(1, c->sx), h = std::max(1, c->ty);
    const int64_t cols = std::min<int64_t>(sx, c->nx - tx * sx);
    const int64_t rows = std::min<int64_t>(h, c->ny - ty * h);
    cost[g] = (cols + 2 * c->ghost) * (rows + 2 * c->ghost_rows);
  }
  std::vector<int32_t> order(b->ntiles);
  for (int g = 0; g < b->ntiles; ++g) order[g] = g;
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t x, int32_t y) { return cost[x] > cost[y]; });
  if (upload(b->tile_inst, tinst, b->stream) != hipSuccess ||
      upload(b->inst_base, base, b->stream) != hipSuccess || b->dyn.ensure(64) != hipSuccess ||
      upload(b->order, order, b->stream) != hipSuccess) {
    fds_batch_destroy(b);
    return fail(FDS_E_HIP, "out of device memory for the batch tables");
  }
  *out = b;
  return 0;
}

int fds_batch_set_schedule(fds_batch *b, int schedule) {
  if (!b) return fail(FDS_E_ARG, "null batch");
  if (schedule != 0 && schedule != 1) return fail(FDS_E_ARG, "schedule is 0 or 1");
  if (!b->flights.empty()) return fail(FDS_E_STATE, "asynchronous runs in flight");
  if (schedule == 1 && !b->streaming)
    return fail(FDS_E_ARG, "the work-queue schedule needs a streaming shape");
  if (b->schedule != schedule) {
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    batch_drop_graphs(b);
    b->schedule = schedule;
  }
  return 0;
}

int fds_batch_destroy(fds_batch *b) {
  if (!b) return 0;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_drop_graphs(b);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  for (auto &set : b->fev)
    for (auto e : set)
      if (e) (void)hipEventDestroy(e);
  // members no longer wait for this batch's runs
  for (size_t i = 0; i < b->flights.size(); ++i)
    for (auto *c : b->ctxs) c->batch_flights--;
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

int fds_batch_run(fds_batch *b, int64_t start, int64_t nsteps, int chunk) {
  if (!b) return fail(FDS_E_ARG, "null batch");
  HIP_TRY(hipSetDevice(b->device));
  if (nsteps < 0 || start < 0) return fail(FDS_E_ARG, "negative step count");
  if (nsteps == 0) return 0;
  if (!b->flights.empty()) return fail(FDS_E_STATE, "asynchronous runs in flight: take them first");
  if (int rc = batch_prepare(b, start, nsteps)) return rc;
  if (!b->ev0) HIP_TRY(hipEventCreate(&b->ev0));
  if (!b->ev1) HIP_TRY(hipEventCreate(&b->ev1));
  if (int rc = batch_enqueue(b, nsteps, chunk, b->ev0, b->ev1)) return rc;
  for (auto *c : b->ctxs)
    if (int rc = stage_nonfinite(c, kMaxFlights, b->stream)) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_run_ms = ms;
  return batch_check_nonfinite(b, kMaxFlights);
}

int fds_batch_run_async(fds_batch *b, int64_t start, int64_t nsteps, int chunk) {
  if (!b) return fail(FDS_E_ARG, "null batch");
  HIP_TRY(hipSetDevice(b->device));
  if (nsteps < 1 || start < 0) return fail(FDS_E_ARG, "an asynchronous run needs steps");
  if ((int)b->flights.size() >= kMaxFlights)
    return fail(FDS_E_STATE, "%d runs in flight: take the oldest first", kMaxFlights);
  if (!batch_steady(b, nsteps)) HIP_TRY(hipStreamSynchronize(b->stream));
  if (int rc = batch_prepare(b, start, nsteps)) return rc;
  const int slot = b->next_slot;
  if (int rc = ensure_events(b->fev[slot])) return rc;
  if (int rc = batch_enqueue(b, nsteps, chunk, b->fev[slot][0], b->fev[slot][1])) return rc;
  for (auto *c : b->ctxs) {
    if (int rc = stage_outputs(c, slot, nsteps, b->stream)) return rc;
    if (int rc = stage_nonfinite(c, slot, b->stream)) return rc;
  }
  HIP_TRY(hipEventRecord(b->fev[slot][2], b->stream));
  if (b->flights.empty()) b->last_run_ms = 0.0;
  b->flights.push_back(Flight{nsteps, slot});
  for (auto *c : b->ctxs) c->batch_flights++;
  b->next_slot = (slot + 1) % kMaxFlights;
  return 0;
}

int fds_batch_take(fds_batch *b, double *const *dst) {
  if (!b) return fail(FDS_E_ARG, "null batch");
  HIP_TRY(hipSetDevice(b->device));
  if (b->flights.empty()) return fail(FDS_E_STATE, "no asynchronous run in flight");
  const Flight f = b->flights.front();
  HIP_TRY(hipEventSynchronize(b->fev[f.slot][2]));
  b->flights.pop_front();
  for (size_t i = 0; i < b->ctxs.size(); ++i) {
    b->ctxs[i]->batch_flights--;
    if (dst) take_outputs(b->ctxs[i], f, dst + 3 * i);
  }
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, b->fev[f.slot][0], b->fev[f.slot][1]));
  b->last_run_ms += ms;
  return batch_check_nonfinite(b, f.slot);
}

int fds_batch_last_run_ms(fds_batch *b, double *ms) {
  if (!b || !ms) return fail(FDS_E_ARG, "null argument");
  *ms = b->last_run_ms;
  return 0;
}

int fds_batch_steps_per_launch(fds_batch *b) {
  if (!b) return fail(FDS_E_ARG, "null batch");
  return b->fuse;
}

}  // extern "C"

namespace {

template <typename TA, typename TB>
__global__ __launch_bounds__(256) void k_window_sq_diff(const TA *__restrict__ a, int anx,
                                                        const TB *__restrict__ b, int bnx,
                                                        int wx, int wy, int ax0, int ay0,
                                                        int bx0, int by0,
                                                        double *__restrict__ out) {
  __shared__ double r0[256], r1[256];
  double num = 0.0, den = 0.0;
  const int64_t n = (int64_t)wx * wy;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const int iy = (int)(i / wx), ix = (int)(i % wx);
    const double va = (double)a[(int64_t)(ay0 + iy) * anx + ax0 + ix];
    const double vb = (double)b[(int64_t)(by0 + iy) * bnx + bx0 + ix];
    const double d = va - vb;
    num += d * d;
    den += vb * vb;
  }
  r0[threadIdx.x] = num;
  r1[threadIdx.x] = den;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      r0[threadIdx.x] += r0[threadIdx.x + s];
      r1[threadIdx.x] += r1[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = r0[0];
    out[2 * blockIdx.x + 1] = r1[0];
  }
}

const void *current_state(const fds_ctx *c, int comp) {
  const Comp &cp = c->comp[comp];
  const int b = (cp.buf[1].p && (cp.accumulated || c->family != kThermal || comp == 0)) ? c->parity : 0;
  return cp.buf[b].p;
}

}  // namespace

namespace {

constexpr int kWindowBlocks = 512;   // workgroups of the window reduction (partial sums each)

// Partial sums of the no-allocation window reduction: one array per GPU, held by the process
// until it ends.  Host code uses its device address from hipGetSymbolAddress (window_scratch),
// never the array's name.
static __device__ __align__(256) double scratch[2 * kWindowBlocks];

// Device address of `scratch` on `device`, looked up once per device (with it current).
int window_scratch(int device, double **out) {
  static std::mutex mu;
  static std::map<int, double *> ptrs;
  std::lock_guard<std::mutex> lock(mu);
  auto it = ptrs.find(device);
  if (it == ptrs.end()) {
    double *p = nullptr;
    HIP_TRY(hipSetDevice(device));
    const hipError_t e = hipGetSymbolAddress(reinterpret_cast<void **>(&p), HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(FDS_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    it = ptrs.emplace(device, p).first;
  }
  *out = it->second;
  return 0;
}

// The window reduction of fds_window_sq_diff into `part` (2 * kWindowBlocks doubles, device),
// summed on the host in block order.
int window_sq_diff(fds_ctx *a, int comp_a, fds_ctx *b, int comp_b, int wx, int wy, int ax0,
                   int ay0, int bx0, int by0, double *part, double *num, double *den) {
  const int blocks = kWindowBlocks;
  const void *pa = current_state(a, comp_a), *pb = current_state(b, comp_b);
  const int anx = (int)a->nx, bnx = (int)b->nx;
  double *o = part;
  if (a->elem == 8 && b->elem == 8)
    hipLaunchKernelGGL((k_window_sq_diff<double, double>), dim3(blocks), dim3(256), 0, a->stream,
                       (const double *)pa, anx, (const double *)pb, bnx, wx, wy, ax0, ay0, bx0, by0, o);
  else if (a->elem == 4 && b->elem == 4)
    hipLaunchKernelGGL((k_window_sq_diff<float, float>), dim3(blocks), dim3(256), 0, a->stream,
                       (const float *)pa, anx, (const float *)pb, bnx, wx, wy, ax0, ay0, bx0, by0, o);
  else if (a->elem == 4)
    hipLaunchKernelGGL((k_window_sq_diff<float, double>), dim3(blocks), dim3(256), 0, a->stream,
                       (const float *)pa, anx, (const double *)pb, bnx, wx, wy, ax0, ay0, bx0, by0, o);
  else
    hipLaunchKernelGGL((k_window_sq_diff<double, float>), dim3(blocks), dim3(256), 0, a->stream,
                       (const double *)pa, anx, (const float *)pb, bnx, wx, wy, ax0, ay0, bx0, by0, o);
  HIP_TRY(hipGetLastError());
  double h[2 * kWindowBlocks];
  HIP_TRY(hipMemcpyAsync(h, o, sizeof h, hipMemcpyDeviceToHost, a->stream));
  HIP_TRY(hipStreamSynchronize(a->stream));
  double s0 = 0, s1 = 0;
  for (int i = 0; i < blocks; ++i) {
    s0 += h[2 * i];
    s1 += h[2 * i + 1];
  }
  *num = s0;
  *den = s1;
  return 0;
}

int window_checks(fds_ctx *a, int comp_a, fds_ctx *b, int comp_b, int wx, int wy, int ax0,
                  int ay0, int bx0, int by0, double *num, double *den) {
  if (int rc = check_idle(a)) return rc;
  if (!b || b->device != a->device) return fail(FDS_E_ARG, "contexts must share a device");
  if (int rc = check_idle(b)) return rc;
  if (int rc = check_comp(a, comp_a)) return rc;
  if (int rc = check_comp(b, comp_b)) return rc;
  if (!num || !den) return fail(FDS_E_ARG, "null output");
  if (wx < 1 || wy < 1 || ax0 < 0 || ay0 < 0 || bx0 < 0 || by0 < 0 || ax0 + wx > a->nx ||
      ay0 + wy > a->ny || bx0 + wx > b->nx || by0 + wy > b->ny)
    return fail(FDS_E_INDEX, "window outside a grid");
  HIP_TRY(hipStreamSynchronize(b->stream));
  return 0;
}

}  // namespace

extern "C" int fds_window_sq_diff(fds_ctx *a, int comp_a, fds_ctx *b, int comp_b, int wx, int wy,
                                  int ax0, int ay0, int bx0, int by0, double *num, double *den) {
  if (int rc = window_checks(a, comp_a, b, comp_b, wx, wy, ax0, ay0, bx0, by0, num, den))
    return rc;
  DevBuf part;
  HIP_TRY(part.ensure(2 * kWindowBlocks * sizeof(double)));
  return window_sq_diff(a, comp_a, b, comp_b, wx, wy, ax0, ay0, bx0, by0, part.as<double>(),
                        num, den);
}

extern "C" int fds_window_sq_diff_static(fds_ctx *a, int comp_a, fds_ctx *b, int comp_b, int wx,
                                         int wy, int ax0, int ay0, int bx0, int by0, double *num,
                                         double *den) {
  if (int rc = window_checks(a, comp_a, b, comp_b, wx, wy, ax0, ay0, bx0, by0, num, den))
    return rc;
  double *part = nullptr;
  if (int rc = window_scratch(a->device, &part)) return rc;
  return window_sq_diff(a, comp_a, b, comp_b, wx, wy, ax0, ay0, bx0, by0, part, num, den);
}

extern "C" int fds_output_device_ptr(fds_ctx *c, int comp, double **ptr, int64_t *rows, int64_t *m) {
  if (int rc = check_idle(c)) return rc;
  if (int rc = check_comp(c, comp)) return rc;
  *ptr = c->comp[comp].obuf.as<double>();
  *rows = c->comp[comp].obuf_rows;
  *m = (int64_t)c->comp[comp].out_cells.size();
  return 0;
}

#include "fds_couple.inc"

// Workgroup schedules of the operator kernels and the autotuner over them
// (include/sfk.h, "workgroup schedule and autotuner").  The reference picks
// its loop blocking by hand per degree (legacy/src/test_dg_integrate.cc:64-66);
// here the schedule of a launch is (cells per workgroup, block order,
// continuous scatter) and sfk_tune measures every candidate.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "sfk.h"
#include "sfk_internal.h"
#include "sfk_kernels.h"
#include "sfk_op.h"

using namespace sfk;

namespace {

constexpr int kMaxThreads = 256; // __launch_bounds__ of the operator kernels

#define TUNE_FAIL(code, msg) \
  do {                       \
    set_error(msg);          \
    return code;             \
  } while (0)

// the kernel's default cells per workgroup (mirrors the device constants)
int default_cb(const sfk_op *op)
{
  const int n = op->d.degree + 1;
  if (op->d.kind == SFK_DG_SIPG && op->d.geometry == SFK_GEOM_PER_Q) {
    static const int tab[10] = {0, 0, 64, 28, 12, 8, 3, 3, 2, 1};
    return n < 10 ? tab[n] : 1;
  }
  return kMaxThreads / (n * n) > 0 ? kMaxThreads / (n * n) : 1;
}

// LDS bytes per cell of the operator kernel behind sfk_apply
size_t lds_per_cell(const sfk_op *op)
{
  const size_t n = op->d.degree + 1, n2 = n * n, n3 = n2 * n;
  const size_t es = elem_size(op);
  switch (op->d.kind) {
  case SFK_DG_SIPG:
    return (op->d.geometry == SFK_GEOM_PER_Q ? 4 * n3 + 12 * n2 : 2 * n3 + 12 * n2) * es;
  case SFK_DG_ADVECT: return advect_lds_per_cell((int)n, es == 4);
  case SFK_DG_PRECOMP: return precomp_lds_per_cell((int)n, es == 4);
  default: return 2 * n3 * es;
  }
}

// LDS per workgroup of the current device (gfx950: 160 KiB when no device
// can be queried)
int lds_limit()
{
  int dev = 0, lds_max = 163840;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
      lds_max = v;
  }
  return lds_max;
}

bool sipg_per_q(const sfk_op *op)
{
  return op->d.kind == SFK_DG_SIPG && op->d.geometry == SFK_GEOM_PER_Q;
}

bool tp_ok(const sfk_op *op)
{
  return op->d.kind == SFK_CONT_LAPLACE && !(op->d.flags & SFK_FLAG_VARIABLE_COEF) && op->d.degree <= kMaxDegreeTp;
}

int max_cb(const sfk_op *op)
{
  const int n = op->d.degree + 1;
  const int lds_max = lds_limit();
  int m = kMaxThreads / (n * n);
  const int by_lds = (int)(size_t(lds_max) / lds_per_cell(op));
  if (by_lds < m) m = by_lds;
  return m > 0 ? m : 1;
}

// ---- tuning cache: the file named by SFK_TUNE_CACHE holds one line per
// tuned operator, "<key> <cb> <order> <scatter> <shape> <zc> <cheby shape>
// <cheby zc> <cheby order> <seconds>", the key naming the device and every
// descriptor field the kernels and their timing depend on.  Lines are
// appended; the last line of a key wins.
std::mutex g_cache_mutex;

std::string cache_key(const sfk_op *op)
{
  std::string dev = "nodevice";
  int id = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&id) == hipSuccess && hipGetDeviceProperties(&prop, id) == hipSuccess) {
    dev = prop.gcnArchName;
    dev += ":" + std::to_string(prop.multiProcessorCount);
  }
  for (char &c : dev)
    if (c == ' ' || c == '\t') c = '_';
  const sfk_op_desc &d = op->d;
  std::ostringstream k;
  k << dev << "/k" << d.kind << "/p" << d.degree << "/f" << d.precision << "/b" << d.basis << "/g" << d.geometry
    << "/fl" << d.flags << "/n" << d.n_cells[0] << "x" << d.n_cells[1] << "x" << d.n_cells[2] << "/r" << d.n_rhs;
  return k.str();
}

struct CacheEntry {
  sfk_schedule s;
  int cheby[3];
  double seconds;
};

bool cache_lookup(const char *path, const std::string &key, CacheEntry &out)
{
  std::lock_guard<std::mutex> lock(g_cache_mutex);
  FILE *f = std::fopen(path, "r");
  if (!f) return false;
  bool found = false;
  char line[512];
  while (std::fgets(line, sizeof line, f)) {
    std::istringstream in(line);
    std::string k;
    CacheEntry e{};
    if (!(in >> k >> e.s.cells_per_block >> e.s.block_order >> e.s.cont_scatter >> e.s.tile_shape >>
          e.s.z_cells >> e.cheby[0] >> e.cheby[1] >> e.cheby[2] >> e.seconds))
      continue; // not a cache line
    if (k == key) {
      out = e;
      found = true;
    }
  }
  std::fclose(f);
  return found;
}

void cache_store(const char *path, const std::string &key, const CacheEntry &e)
{
  std::lock_guard<std::mutex> lock(g_cache_mutex);
  FILE *f = std::fopen(path, "a");
  if (!f) return; // an unwritable cache only costs the next run its tuning
  std::fprintf(f, "%s %d %d %d %d %d %d %d %d %.9g\n", key.c_str(), e.s.cells_per_block, e.s.block_order,
               e.s.cont_scatter, e.s.tile_shape, e.s.z_cells, e.cheby[0], e.cheby[1], e.cheby[2], e.seconds);
  std::fclose(f);
}

} // namespace

extern "C" {

sfk_status sfk_op_get_schedule(const sfk_op *op, sfk_schedule *s)
{
  if (!op || !s) TUNE_FAIL(SFK_ERR_INVALID_ARG, "sfk_op_get_schedule: null argument");
  s->cells_per_block = op->sched_cb > 0 ? op->sched_cb : default_cb(op);
  s->block_order = op->sched_order;
  s->cont_scatter = op->cont_scatter;
  s->tile_shape = op->tp_shape;
  s->z_cells = op->tp_zc;
  return SFK_OK;
}

sfk_status sfk_op_get_smoother_schedule(const sfk_op *op, sfk_schedule *s)
{
  if (!op || !s) TUNE_FAIL(SFK_ERR_INVALID_ARG, "sfk_op_get_smoother_schedule: null argument");
  const sfk_status st = sfk_op_get_schedule(op, s);
  if (st != SFK_OK) return st;
  if (op->cheby_shape >= 0) {
    s->block_order = op->cheby_order;
    s->tile_shape = op->cheby_shape;
    s->z_cells = op->cheby_zc;
  }
  if (op->cheby_cb > 0) {
    s->cells_per_block = op->cheby_cb;
    s->block_order = op->cheby_order;
  }
  return SFK_OK;
}

sfk_status sfk_op_tile_shape(const sfk_op *op, int shape, int *count, int *bx, int *by)
{
  if (!op) TUNE_FAIL(SFK_ERR_INVALID_ARG, "sfk_op_tile_shape: null handle");
  const int n = op->d.degree + 1, c = tp_shape_count(n);
  if (count) *count = c;
  if (shape < 0 || shape >= c) TUNE_FAIL(SFK_ERR_INVALID_ARG, "tile shape must be in 0.." + std::to_string(c - 1));
  int x = 0, y = 0;
  tp_shape(n, shape, &x, &y);
  if (bx) *bx = x;
  if (by) *by = y;
  return SFK_OK;
}

sfk_status sfk_op_max_cells_per_block(const sfk_op *op, int *max_cells)
{
  if (!op || !max_cells) TUNE_FAIL(SFK_ERR_INVALID_ARG, "sfk_op_max_cells_per_block: null argument");
  *max_cells = max_cb(op);
  return SFK_OK;
}

sfk_status sfk_op_set_schedule(sfk_op *op, const sfk_schedule *s)
{
  if (!op || !s) TUNE_FAIL(SFK_ERR_INVALID_ARG, "sfk_op_set_schedule: null argument");
  const int m = max_cb(op);
  if (s->cells_per_block < 0 || s->cells_per_block > m)
    TUNE_FAIL(SFK_ERR_INVALID_ARG, "cells_per_block must be in 0.." + std::to_string(m));
  if (s->block_order < 0 || s->block_order > 2) TUNE_FAIL(SFK_ERR_INVALID_ARG, "block_order must be 0, 1 or 2");
  if (s->cont_scatter < -1 || s->cont_scatter > 2)
    TUNE_FAIL(SFK_ERR_INVALID_ARG, "cont_scatter must be -1, 0, 1 or 2");
  if (s->cont_scatter == 2 && !tp_ok(op))
    TUNE_FAIL(SFK_ERR_UNSUPPORTED, "cont_scatter 2: continuous Laplacian with a constant coefficient only");
  if (s->tile_shape < 0 || s->tile_shape >= tp_shape_count(op->d.degree + 1))
    TUNE_FAIL(SFK_ERR_INVALID_ARG, "tile_shape out of range");
  if (s->z_cells < 0) TUNE_FAIL(SFK_ERR_INVALID_ARG, "z_cells must be >= 0");
  if (s->cont_scatter == 2 || s->cont_scatter == -1) {
    const size_t need = tp_lds_bytes(op->d.degree + 1, s->tile_shape, op->d.precision == SFK_FP32);
    if (tp_ok(op) && need > (size_t)lds_limit())
      TUNE_FAIL(SFK_ERR_UNSUPPORTED, "tile shape needs more LDS than the device has per workgroup");
  }
  op->sched_cb = s->cells_per_block;
  op->sched_order = s->block_order;
  op->cont_scatter = s->cont_scatter;
  op->tp_shape = s->tile_shape;
  op->tp_zc = s->z_cells;
  op->cheby_shape = -1; // the Chebyshev launches follow the new schedule
  op->cheby_cb = 0;
  return SFK_OK;
}

sfk_status sfk_tune(sfk_op *op, int repeats, sfk_schedule *best, double *best_seconds, void *stream)
{
  if (!op) TUNE_FAIL(SFK_ERR_INVALID_ARG, "sfk_tune: null handle");
  if (repeats < 1 || repeats > 10000) TUNE_FAIL(SFK_ERR_INVALID_ARG, "sfk_tune: repeats must be in 1..10000");
  const char *cache = std::getenv("SFK_TUNE_CACHE");
  const std::string key = cache && *cache ? cache_key(op) : std::string();
  if (!key.empty()) {
    CacheEntry e;
    if (cache_lookup(cache, key, e)) {
      // validated like a hand-set schedule; a stale line falls through to tuning
      if (sfk_op_set_schedule(op, &e.s) == SFK_OK) {
        // per-q SIPG: the Chebyshev step's cells per block in the z-cells
        // field (shape -1), orders 0..2
        const bool dg_cheby = sipg_per_q(op) && e.cheby[0] < 0 && e.cheby[1] > 0;
        const bool cheby_ok =
            dg_cheby ? e.cheby[1] <= max_cb(op) && e.cheby[2] >= 0 && e.cheby[2] <= 2
                     : e.cheby[0] < 0 || (tp_ok(op) && e.cheby[0] < tp_shape_count(op->d.degree + 1) &&
                                          e.cheby[1] >= 0 && (e.cheby[2] == 0 || e.cheby[2] == 1));
        if (cheby_ok) {
          op->cheby_shape = dg_cheby ? -1 : e.cheby[0];
          op->cheby_zc = dg_cheby ? 0 : e.cheby[1];
          op->cheby_cb = dg_cheby ? e.cheby[1] : 0;
          op->cheby_order = e.cheby[2];
          if (best) *best = e.s;
          if (best_seconds) *best_seconds = e.seconds;
          return SFK_OK;
        }
      }
    }
  }
  hipStream_t s = (hipStream_t)stream;
  const size_t bytes = size_t(op->nvec) * op->d.n_rhs * elem_size(op);
  void *x = nullptr, *y = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  sfk_status status = SFK_OK;
  const int saved_cb = op->sched_cb, saved_order = op->sched_order, saved_scatter = op->cont_scatter;
  const int saved_shape = op->tp_shape, saved_zc = op->tp_zc;
  const int saved_cshape = op->cheby_shape, saved_czc = op->cheby_zc, saved_corder = op->cheby_order;
  const int saved_ccb = op->cheby_cb;
  sfk_schedule win{saved_cb > 0 ? saved_cb : default_cb(op), saved_order, saved_scatter, saved_shape, saved_zc};
  double win_t = -1.;
  if (hipMalloc(&x, bytes) != hipSuccess || hipMalloc(&y, bytes) != hipSuccess) {
    set_error("sfk_tune: scratch allocation failed");
    status = SFK_ERR_ALLOC;
  }
  if (status == SFK_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) {
    set_error("sfk_tune: event creation failed");
    status = SFK_ERR_HIP;
  }
  if (status == SFK_OK)
    status = sfk_vec_fill(op->d.precision, x, op->nvec * op->d.n_rhs, 1.0, stream);
  // time one candidate (already set on the handle)
  auto trial = [&](const sfk_schedule &cand) {
    status = sfk_apply(op, x, y, stream); // warm-up (and LDS attribute)
    if (status != SFK_OK) return;
    if (hipEventRecord(e0, s) != hipSuccess) status = SFK_ERR_HIP;
    for (int r = 0; r < repeats && status == SFK_OK; ++r) status = sfk_apply(op, x, y, stream);
    if (status != SFK_OK) return;
    float ms = 0.f;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      set_error("sfk_tune: event timing failed");
      status = SFK_ERR_HIP;
      return;
    }
    const double t = 1e-3 * ms / repeats;
    if (win_t < 0. || t < win_t) {
      win_t = t;
      win = cand;
    }
  };
  if (status == SFK_OK) {
    const int m = max_cb(op);
    std::vector<int> scatters = {op->cont_scatter};
    // order 2 (slab traversal) exists in the DG kernels only
    const int max_order = op->d.kind == SFK_DG_SIPG || op->d.kind == SFK_DG_ADVECT ? 2 : 1;
    if (op->d.kind == SFK_CONT_LAPLACE) scatters = {0, 1};
    for (int sc : scatters)
      for (int order = 0; order <= max_order && status == SFK_OK; ++order)
        for (int cb = 1; cb <= m && status == SFK_OK; ++cb) {
          op->sched_cb = cb;
          op->sched_order = order;
          op->cont_scatter = sc;
          trial(sfk_schedule{cb, order, sc, saved_shape, saved_zc});
        }
    if (tp_ok(op)) {
      // tensor-product kernel: tile shapes x cells in z per workgroup
      const int n = op->d.degree + 1, nz = op->d.n_cells[2];
      const int lds_max = lds_limit();
      std::vector<int> zcs;
      for (int z = 1; z < nz; z *= 2) zcs.push_back(z);
      zcs.push_back(nz);
      for (int sh = 0; sh < tp_shape_count(n) && status == SFK_OK; ++sh) {
        if (tp_lds_bytes(n, sh, op->d.precision == SFK_FP32) > (size_t)lds_max) continue;
        for (int order = 0; order <= 1 && status == SFK_OK; ++order)
          for (int zc : zcs) {
            if (status != SFK_OK) break;
            op->sched_order = order;
            op->cont_scatter = 2;
            op->tp_shape = sh;
            op->tp_zc = zc;
            trial(sfk_schedule{saved_cb, order, 2, sh, zc});
          }
      }
    }
  }
  // the fused Chebyshev launch of the tensor-product kernel in the solver's
  // three-term form (src x, rhs x, previous iterate pv, dst y: four
  // streams, more registers than the matvec, so its best tile differs)
  double cwin_t = -1.;
  int cwin[3] = {-1, 0, 1};
  if (status == SFK_OK && tp_ok(op)) {
    void *pv = nullptr;
    if (hipMalloc(&pv, bytes) != hipSuccess) {
      set_error("sfk_tune: scratch allocation failed");
      status = SFK_ERR_ALLOC;
    }
    if (status == SFK_OK) status = sfk_vec_fill(op->d.precision, pv, op->nvec * op->d.n_rhs, 0.0, stream);
    if (status == SFK_OK) status = ensure_device_tables(op);
    const int n = op->d.degree + 1, nz = op->d.n_cells[2];
    std::vector<int> zcs;
    for (int z = 1; z < nz; z *= 2) zcs.push_back(z);
    zcs.push_back(nz);
    for (int sh = 0; sh < tp_shape_count(n) && status == SFK_OK; ++sh) {
      if (tp_lds_bytes(n, sh, op->d.precision == SFK_FP32) > (size_t)lds_limit()) continue;
      for (int order = 0; order <= 1 && status == SFK_OK; ++order)
        for (int zc : zcs) {
          if (status != SFK_OK) break;
          op->cheby_shape = sh;
          op->cheby_order = order;
          op->cheby_zc = zc;
          TpLaunch T;
          fill_tp_launch(op, T, true);
          T.cheby = 2;
          T.src = x;
          T.rhs = x;
          T.prev = pv;
          T.dst = y;
          T.c_np = 0.5;
          T.c_tm = 0.25;
          float ms = 0.f;
          bool ok = launch_cont_tp(T, s) == hipSuccess && hipEventRecord(e0, s) == hipSuccess;
          for (int r = 0; r < repeats && ok; ++r) ok = launch_cont_tp(T, s) == hipSuccess;
          ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
               hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
          if (!ok) {
            set_error("sfk_tune: Chebyshev launch or timing failed");
            status = SFK_ERR_HIP;
            break;
          }
          const double t = 1e-3 * ms / repeats;
          if (cwin_t < 0. || t < cwin_t) {
            cwin_t = t;
            cwin[0] = sh;
            cwin[1] = zc;
            cwin[2] = order;
          }
        }
    }
    if (pv) (void)hipFree(pv);
  }
  // the per-q SIPG kernel's fused Chebyshev step (six streams and the
  // metric: its best cells per block and order differ from the matvec's)
  int dwin[2] = {0, 1};
  if (status == SFK_OK && sipg_per_q(op)) {
    void *tv = nullptr;
    if (hipMalloc(&tv, bytes) != hipSuccess) {
      set_error("sfk_tune: scratch allocation failed");
      status = SFK_ERR_ALLOC;
    }
    if (status == SFK_OK) status = sfk_vec_fill(op->d.precision, tv, op->nvec, 0.0, stream);
    double dwin_t = -1.;
    const int m = max_cb(op);
    for (int order = 0; order <= 2 && status == SFK_OK; ++order)
      for (int cb = 1; cb <= m && status == SFK_OK; ++cb) {
        op->cheby_cb = cb;
        op->cheby_order = order;
        // rhs = diag = input = x (ones), tmp tv (read: c_tm != 0)
        status = sfk_cheby_step(op, x, x, x, y, tv, 0.5, 0.25, stream); // warm-up
        if (status != SFK_OK) break;
        float ms = 0.f;
        bool ok = hipEventRecord(e0, s) == hipSuccess;
        for (int r = 0; r < repeats && ok && status == SFK_OK; ++r)
          status = sfk_cheby_step(op, x, x, x, y, tv, 0.5, 0.25, stream);
        ok = ok && status == SFK_OK && hipEventRecord(e1, s) == hipSuccess &&
             hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
        if (!ok) {
          if (status == SFK_OK) {
            set_error("sfk_tune: Chebyshev timing failed");
            status = SFK_ERR_HIP;
          }
          break;
        }
        const double t = 1e-3 * ms / repeats;
        if (dwin_t < 0. || t < dwin_t) {
          dwin_t = t;
          dwin[0] = cb;
          dwin[1] = order;
        }
      }
    if (tv) (void)hipFree(tv);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (x) (void)hipFree(x);
  if (y) (void)hipFree(y);
  if (status != SFK_OK || win_t < 0.) {
    op->sched_cb = saved_cb;
    op->sched_order = saved_order;
    op->cont_scatter = saved_scatter;
    op->tp_shape = saved_shape;
    op->tp_zc = saved_zc;
    op->cheby_shape = saved_cshape;
    op->cheby_zc = saved_czc;
    op->cheby_order = saved_corder;
    op->cheby_cb = saved_ccb;
    return status != SFK_OK ? status : SFK_ERR_HIP;
  }
  op->cheby_shape = cwin[0];
  op->cheby_zc = cwin[1];
  op->cheby_order = cwin[2];
  op->cheby_cb = 0;
  if (dwin[0] > 0) {
    op->cheby_cb = dwin[0];
    op->cheby_order = dwin[1];
  }
  if (!key.empty()) {
    const CacheEntry ce = dwin[0] > 0 ? CacheEntry{win, {-1, dwin[0], dwin[1]}, win_t}
                                      : CacheEntry{win, {cwin[0], cwin[1], cwin[2]}, win_t};
    cache_store(cache, key, ce);
  }
  op->sched_cb = win.cells_per_block;
  op->sched_order = win.block_order;
  op->cont_scatter = win.cont_scatter;
  op->tp_shape = win.tile_shape;
  op->tp_zc = win.z_cells;
  if (best) *best = win;
  if (best_seconds) *best_seconds = win_t;
  return SFK_OK;
}

} // extern "C"

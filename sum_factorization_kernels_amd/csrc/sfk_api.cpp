// extern "C" boundary (include/sfk.h): handles, validation, device data
// ownership and dispatch to the kernel launchers.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include "sfk.h"
#include "sfk_internal.h"
#include "sfk_kernels.h"
#include "sfk_op.h"

namespace sfk {
namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string &msg) { g_last_error = msg; }
} // namespace sfk

using namespace sfk;

#define SFK_CHECK_HIP(expr)                                                                 \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " in " + #expr);      \
      return SFK_ERR_HIP;                                                                   \
    }                                                                                       \
  } while (0)

#define SFK_ENSURE_TABLES(op)                             \
  do {                                                    \
    const sfk_status st_ = sfk::ensure_device_tables(op); \
    if (st_ != SFK_OK) return st_;                        \
  } while (0)

#define SFK_REQUIRE(cond, code, msg) \
  do {                               \
    if (!(cond)) {                   \
      set_error(msg);                \
      return code;                   \
    }                                \
  } while (0)

namespace sfk {

size_t elem_size(const sfk_op *op) { return op->d.precision == SFK_FP32 ? 4 : 8; }

sfk_status upload(sfk_op *op, void **dev, const double *host, size_t count, size_t expected, const char *what)
{
  SFK_REQUIRE(host != nullptr, SFK_ERR_INVALID_ARG, std::string(what) + ": null host pointer");
  SFK_REQUIRE(count == expected, SFK_ERR_INVALID_ARG,
              std::string(what) + ": expected " + std::to_string(expected) + " values, got " + std::to_string(count));
  const size_t es = elem_size(op);
  if (*dev == nullptr) {
    hipError_t e = hipMalloc(dev, count * es);
    if (e != hipSuccess) {
      *dev = nullptr;
      set_error(std::string(what) + ": hipMalloc failed");
      return SFK_ERR_ALLOC;
    }
  }
  if (es == 8) {
    SFK_CHECK_HIP(hipMemcpy(*dev, host, count * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<float> tmp(count);
    for (size_t i = 0; i < count; ++i) tmp[i] = (float)host[i];
    SFK_CHECK_HIP(hipMemcpy(*dev, tmp.data(), count * 4, hipMemcpyHostToDevice));
  }
  return SFK_OK;
}

// the operator's 1D tables in device memory, uploaded on the first launch
// the precomputed DG Laplacian's kernels read their own table type
static hipError_t upload_precomp(sfk_op *op, void **dev)
{
  Precomp1D pc;
  if (!build_precomp_1d(op->d.degree, (op->d.flags & SFK_FLAG_REF_PRECOMP_COUPLING) != 0, pc))
    return hipErrorInvalidValue;
  // the reference's constants vanish for p >= 2 (not for p = 1, where
  // phi_{n-2} = phi_0): only then can the kernel skip the neighbour layers
  op->pc_couple = pc.v1 != 0. || pc.v2 != 0.;
  const double *ij = op->d.inv_jac;
  const double det = ij[0] * ij[1] * ij[2];
  const double f[3] = {ij[0] * ij[0] / det, ij[1] * ij[1] / det, ij[2] * ij[2] / det};
  return upload_precomp_tables(op->d.degree + 1, pc, f, op->d.precision == SFK_FP32, dev);
}

// the continuous Laplacian with a constant coefficient has the
// tensor-product kernel as well
static bool has_tp(const sfk_op *op)
{
  return op->d.kind == SFK_CONT_LAPLACE && !(op->d.flags & SFK_FLAG_VARIABLE_COEF) &&
         op->d.degree <= kMaxDegreeTp;
}

sfk_status ensure_device_tables(sfk_op *op)
{
  if (op->dtab && (op->tptab || !has_tp(op))) return SFK_OK;
  hipError_t e = hipSuccess;
  if (!op->dtab)
    e = op->d.kind == SFK_DG_PRECOMP ? upload_precomp(op, &op->dtab)
                                     : upload_device_tables(op->tab, op->d.precision == SFK_FP32, &op->dtab);
  if (e == hipSuccess && has_tp(op) && !op->tptab)
    e = upload_tp_tables(op->tab, op->d.inv_jac, op->d.precision == SFK_FP32, &op->tptab);
  if (e != hipSuccess) {
    set_error(std::string("device tables: ") + hipGetErrorString(e));
    return SFK_ERR_HIP;
  }
  return SFK_OK;
}

sfk_status ensure_scratch(sfk_op *op, size_t bytes)
{
  if (op->scratch_bytes >= bytes) return SFK_OK;
  if (op->scratch) (void)hipFree(op->scratch);
  op->scratch = nullptr;
  op->scratch_bytes = 0;
  if (hipMalloc(&op->scratch, bytes) != hipSuccess) {
    op->scratch = nullptr;
    set_error("scratch allocation failed");
    return SFK_ERR_ALLOC;
  }
  op->scratch_bytes = bytes;
  return SFK_OK;
}

// the continuous kernel a launch of this operator uses: the tensor-product
// kernel unless the schedule asks for a scatter variant
bool uses_tp(const sfk_op *op)
{
  return has_tp(op) && (op->cont_scatter == 2 || op->cont_scatter == -1);
}

// default cells in z per workgroup of the tensor-product kernel: about 1024
// workgroups (4 per CU) unless the mesh has fewer tile columns x cells in z
int tp_default_zc(const sfk_op *op, int shape)
{
  int bx = 0, by = 0;
  tp_shape(op->d.degree + 1, shape, &bx, &by);
  if (bx <= 0 || by <= 0) return 0;
  const long long tiles = (long long)((op->d.n_cells[0] + bx - 1) / bx) * ((op->d.n_cells[1] + by - 1) / by);
  const int nz = op->d.n_cells[2];
  long long zc = tiles * nz / 1024;
  if (zc < 1) zc = 1;
  if (zc > nz) zc = nz;
  return (int)zc;
}

void fill_tp_launch(const sfk_op *op, TpLaunch &L, bool cheby)
{
  L.n = op->d.degree + 1;
  L.fp32 = op->d.precision == SFK_FP32;
  L.tptab = op->tptab;
  for (int d = 0; d < 3; ++d) L.nc[d] = op->d.n_cells[d];
  L.nrhs = op->d.n_rhs;
  L.vstride = op->nvec;
  const bool own = cheby && op->cheby_shape >= 0;
  L.shape = own ? op->cheby_shape : op->tp_shape;
  const int zc = own ? op->cheby_zc : op->tp_zc;
  L.zc = zc > 0 ? zc : tp_default_zc(op, L.shape);
  L.order = own ? op->cheby_order : op->sched_order;
}

void fill_cell_launch(const sfk_op *op, CellLaunch &L)
{
  L.tables = &op->tab;
  L.fp32 = op->d.precision == SFK_FP32;
  L.ncells = op->ncells;
  for (int d = 0; d < 3; ++d) {
    L.nc[d] = op->d.n_cells[d];
    L.inv_jac[d] = op->d.inv_jac[d];
  }
  L.nrhs = op->d.n_rhs;
  L.vstride = op->nvec;
  L.coef = op->coef;
  L.dtab = op->dtab;
  L.cont_scatter = op->cont_scatter;
  L.cb = op->sched_cb;
  L.order = op->sched_order;
}

} // namespace sfk

extern "C" {

const char *sfk_last_error(void) { return sfk::g_last_error.c_str(); }
const char *sfk_version(void) { return "sfk-mi355x 0.1 (gfx950)"; }

sfk_status sfk_op_create(const sfk_op_desc *desc, sfk_op **out)
{
  SFK_REQUIRE(desc && out, SFK_ERR_INVALID_ARG, "sfk_op_create: null argument");
  *out = nullptr;
  const sfk_op_desc &d = *desc;
  SFK_REQUIRE(d.kind >= SFK_CELL_LAPLACE && d.kind <= SFK_DG_PRECOMP, SFK_ERR_INVALID_ARG, "invalid operator kind");
  SFK_REQUIRE(d.degree >= 1 && d.degree <= kMaxDegree, SFK_ERR_INVALID_ARG,
              "degree must be in 1.." + std::to_string(kMaxDegree));
  SFK_REQUIRE(d.precision == SFK_FP64 || d.precision == SFK_FP32, SFK_ERR_INVALID_ARG, "invalid precision");
  SFK_REQUIRE(d.geometry == SFK_GEOM_CARTESIAN || d.geometry == SFK_GEOM_PER_Q, SFK_ERR_INVALID_ARG,
              "invalid geometry");
  for (int k = 0; k < 3; ++k) SFK_REQUIRE(d.n_cells[k] >= 1, SFK_ERR_INVALID_ARG, "n_cells must be >= 1");
  SFK_REQUIRE(d.n_rhs >= 1 && d.n_rhs <= 64, SFK_ERR_INVALID_ARG, "n_rhs must be in 1..64");
  if (d.geometry == SFK_GEOM_CARTESIAN)
    for (int k = 0; k < 3; ++k)
      SFK_REQUIRE(std::isfinite(d.inv_jac[k]) && d.inv_jac[k] > 0, SFK_ERR_INVALID_ARG,
                  "inv_jac must be finite and positive");
  if (d.kind == SFK_DG_SIPG) {
    SFK_REQUIRE(d.basis == SFK_BASIS_HERMITE, SFK_ERR_UNSUPPORTED, "DG-SIPG uses the Hermite-like basis");
    SFK_REQUIRE(d.n_rhs == 1, SFK_ERR_UNSUPPORTED, "DG-SIPG: n_rhs must be 1");
    SFK_REQUIRE(!(d.flags & SFK_FLAG_REF_X_SWEEP) || d.geometry == SFK_GEOM_CARTESIAN, SFK_ERR_UNSUPPORTED,
                "SFK_FLAG_REF_X_SWEEP requires Cartesian geometry");
  }
  if (d.kind == SFK_DG_ADVECT) {
    SFK_REQUIRE(d.basis == SFK_BASIS_GLL, SFK_ERR_UNSUPPORTED, "advection uses the GLL basis");
    SFK_REQUIRE(d.geometry == SFK_GEOM_CARTESIAN, SFK_ERR_UNSUPPORTED, "advection: Cartesian geometry only");
    SFK_REQUIRE(d.n_rhs == 1, SFK_ERR_UNSUPPORTED, "advection: n_rhs must be 1");
  }
  if (d.kind == SFK_CONT_LAPLACE) {
    SFK_REQUIRE(d.basis == SFK_BASIS_GLL, SFK_ERR_UNSUPPORTED, "continuous Q_p uses the GLL basis");
    SFK_REQUIRE(d.geometry == SFK_GEOM_CARTESIAN, SFK_ERR_UNSUPPORTED, "continuous Q_p: affine Cartesian mesh only");
  }
  if (d.kind == SFK_CELL_LAPLACE)
    SFK_REQUIRE(d.geometry == SFK_GEOM_CARTESIAN, SFK_ERR_UNSUPPORTED,
                "cell Laplacian: Cartesian only (as the reference, evaluation_cell_laplacian.h:66-82)");
  if (d.flags & SFK_FLAG_VARIABLE_COEF)
    SFK_REQUIRE(d.kind == SFK_CONT_LAPLACE || d.kind == SFK_CELL_LAPLACE, SFK_ERR_UNSUPPORTED,
                "variable coefficient: Laplacian operators only");
  if (d.flags & SFK_FLAG_VARIABLE_VELOCITY)
    SFK_REQUIRE(d.kind == SFK_DG_ADVECT, SFK_ERR_UNSUPPORTED, "variable velocity: advection only");
  if (d.kind == SFK_DG_PRECOMP) {
    SFK_REQUIRE(d.basis == SFK_BASIS_HERMITE, SFK_ERR_UNSUPPORTED,
                "precomputed DG Laplacian uses the Hermite-like basis");
    SFK_REQUIRE(d.geometry == SFK_GEOM_CARTESIAN, SFK_ERR_UNSUPPORTED, "precomputed DG Laplacian: Cartesian only");
    SFK_REQUIRE(d.degree <= kMaxDegreeTp, SFK_ERR_UNSUPPORTED,
                "precomputed DG Laplacian: degree 1.." + std::to_string(kMaxDegreeTp));
  }
  if (d.flags & SFK_FLAG_REF_PRECOMP_COUPLING)
    SFK_REQUIRE(d.kind == SFK_DG_PRECOMP, SFK_ERR_UNSUPPORTED,
                "SFK_FLAG_REF_PRECOMP_COUPLING: precomputed DG Laplacian only");

  const long long ncells = (long long)d.n_cells[0] * d.n_cells[1] * d.n_cells[2];
  SFK_REQUIRE(ncells <= (1LL << 31), SFK_ERR_INVALID_ARG, "too many cells");
  sfk_op *op = new sfk_op();
  op->d = d;
  // per-q SIPG: slab traversal by default (its re-read neighbour layers and
  // face metric stay in L2: HBM traffic 1.18x -> 1.07x the algorithmic bytes)
  if (d.kind == SFK_DG_SIPG && d.geometry == SFK_GEOM_PER_Q) op->sched_order = 2;
  if (!build_tables(d.degree, d.basis, op->tab)) {
    delete op;
    set_error("table construction failed");
    return SFK_ERR_INVALID_ARG;
  }
  const long long n = d.degree + 1;
  op->ncells = ncells;
  if (d.kind == SFK_CONT_LAPLACE) {
    for (int k = 0; k < 3; ++k) op->grid[k] = d.n_cells[k] * d.degree + 1;
    op->nvec = (long long)op->grid[0] * op->grid[1] * op->grid[2];
  } else {
    op->nvec = op->ncells * n * n * n;
  }
  op->ncq = op->ncells * n * n * n;
  op->nfq = op->ncells * n * n;
  // tables to the device now when a device is present (a first launch
  // inside a graph capture must not allocate); without one the upload
  // waits for the first launch
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0) (void)ensure_device_tables(op);
  (void)hipGetLastError();
  *out = op;
  return SFK_OK;
}

sfk_status sfk_op_destroy(sfk_op *op)
{
  if (!op) return SFK_OK;
  for (void *p : {op->cell_metric, op->face_metric, op->coef, op->cell_vel, op->face_vel, op->scratch, op->dtab,
                  op->tptab})
    if (p) (void)hipFree(p);
  delete op;
  return SFK_OK;
}

sfk_status sfk_op_vector_size(const sfk_op *op, long long *n)
{
  SFK_REQUIRE(op && n, SFK_ERR_INVALID_ARG, "null argument");
  *n = op->nvec;
  return SFK_OK;
}

sfk_status sfk_op_qdata_sizes(const sfk_op *op, long long *n_cell_q, long long *n_face_q)
{
  SFK_REQUIRE(op, SFK_ERR_INVALID_ARG, "null argument");
  if (n_cell_q) *n_cell_q = op->ncq;
  if (n_face_q) *n_face_q = op->nfq;
  return SFK_OK;
}

sfk_status sfk_op_set_cell_metric(sfk_op *op, const double *host, size_t count)
{
  SFK_REQUIRE(op, SFK_ERR_INVALID_ARG, "null handle");
  SFK_REQUIRE(op->d.geometry == SFK_GEOM_PER_Q, SFK_ERR_INVALID_ARG, "cell metric needs SFK_GEOM_PER_Q");
  return upload(op, &op->cell_metric, host, count, (size_t)op->ncq * 6, "cell metric");
}

sfk_status sfk_op_set_face_metric(sfk_op *op, const double *host, size_t count)
{
  SFK_REQUIRE(op, SFK_ERR_INVALID_ARG, "null handle");
  SFK_REQUIRE(op->d.geometry == SFK_GEOM_PER_Q && op->d.kind == SFK_DG_SIPG, SFK_ERR_INVALID_ARG,
              "face metric needs DG-SIPG with SFK_GEOM_PER_Q");
  return upload(op, &op->face_metric, host, count, (size_t)op->nfq * 3 * 7, "face metric");
}

sfk_status sfk_op_set_coefficient(sfk_op *op, const double *host, size_t count)
{
  SFK_REQUIRE(op, SFK_ERR_INVALID_ARG, "null handle");
  SFK_REQUIRE(op->d.flags & SFK_FLAG_VARIABLE_COEF, SFK_ERR_INVALID_ARG, "operator lacks SFK_FLAG_VARIABLE_COEF");
  return upload(op, &op->coef, host, count, (size_t)op->ncq, "coefficient");
}

sfk_status sfk_op_set_velocity(sfk_op *op, const double *cell_host, size_t cell_count, const double *face_host,
                               size_t face_count)
{
  SFK_REQUIRE(op, SFK_ERR_INVALID_ARG, "null handle");
  SFK_REQUIRE(op->d.flags & SFK_FLAG_VARIABLE_VELOCITY, SFK_ERR_INVALID_ARG,
              "operator lacks SFK_FLAG_VARIABLE_VELOCITY");
  sfk_status s = upload(op, &op->cell_vel, cell_host, cell_count, (size_t)op->ncq * 3, "cell velocity");
  if (s != SFK_OK) return s;
  return upload(op, &op->face_vel, face_host, face_count, (size_t)op->nfq * 3, "face velocity");
}

// device-to-device copy into handle-owned storage (allocated on first use)
static sfk_status copy_device(sfk_op *op, void **dev, const void *src, size_t count, size_t expected,
                              const char *what, hipStream_t s)
{
  SFK_REQUIRE(src != nullptr, SFK_ERR_INVALID_ARG, std::string(what) + ": null device pointer");
  SFK_REQUIRE(count == expected, SFK_ERR_INVALID_ARG,
              std::string(what) + ": expected " + std::to_string(expected) + " values, got " + std::to_string(count));
  const size_t bytes = count * elem_size(op);
  if (*dev == nullptr && hipMalloc(dev, bytes) != hipSuccess) {
    *dev = nullptr;
    set_error(std::string(what) + ": hipMalloc failed");
    return SFK_ERR_ALLOC;
  }
  SFK_CHECK_HIP(hipMemcpyAsync(*dev, src, bytes, hipMemcpyDeviceToDevice, s));
  return SFK_OK;
}

sfk_status sfk_op_set_coefficient_device(sfk_op *op, const void *dev, size_t count, void *stream)
{
  SFK_REQUIRE(op, SFK_ERR_INVALID_ARG, "null handle");
  SFK_REQUIRE(op->d.flags & SFK_FLAG_VARIABLE_COEF, SFK_ERR_INVALID_ARG, "operator lacks SFK_FLAG_VARIABLE_COEF");
  return copy_device(op, &op->coef, dev, count, (size_t)op->ncq, "coefficient", (hipStream_t)stream);
}

sfk_status sfk_op_set_velocity_device(sfk_op *op, const void *cell_dev, size_t cell_count, const void *face_dev,
                                      size_t face_count, void *stream)
{
  SFK_REQUIRE(op, SFK_ERR_INVALID_ARG, "null handle");
  SFK_REQUIRE(op->d.flags & SFK_FLAG_VARIABLE_VELOCITY, SFK_ERR_INVALID_ARG,
              "operator lacks SFK_FLAG_VARIABLE_VELOCITY");
  sfk_status st = copy_device(op, &op->cell_vel, cell_dev, cell_count, (size_t)op->ncq * 3, "cell velocity",
                              (hipStream_t)stream);
  if (st != SFK_OK) return st;
  return copy_device(op, &op->face_vel, face_dev, face_count, (size_t)op->nfq * 3, "face velocity",
                     (hipStream_t)stream);
}

sfk_status sfk_op_set_sine_deformation(sfk_op *op, double amp)
{
  SFK_REQUIRE(op, SFK_ERR_INVALID_ARG, "null handle");
  SFK_REQUIRE(op->d.geometry == SFK_GEOM_PER_Q, SFK_ERR_INVALID_ARG, "needs SFK_GEOM_PER_Q");
  SFK_REQUIRE(std::isfinite(amp) && std::fabs(amp) < 0.3, SFK_ERR_INVALID_ARG,
              "amplitude must keep the mapping invertible (|amp| < 0.3)");
  const size_t es = elem_size(op);
  if (!op->cell_metric) SFK_CHECK_HIP(hipMalloc(&op->cell_metric, (size_t)op->ncq * 6 * es));
  if (op->d.kind == SFK_DG_SIPG && !op->face_metric)
    SFK_CHECK_HIP(hipMalloc(&op->face_metric, (size_t)op->nfq * 3 * 7 * es));
  SFK_CHECK_HIP(launch_sine_metric(op->tab, op->d.precision == SFK_FP32, op->d.n_cells, amp, op->cell_metric,
                                   op->d.kind == SFK_DG_SIPG ? op->face_metric : nullptr, nullptr));
  SFK_CHECK_HIP(hipDeviceSynchronize());
  return SFK_OK;
}

sfk_status sfk_apply(sfk_op *op, const void *src, void *dst, void *stream)
{
  SFK_REQUIRE(op && src && dst, SFK_ERR_INVALID_ARG, "sfk_apply: null argument");
  SFK_REQUIRE(src != dst, SFK_ERR_INVALID_ARG, "sfk_apply: src and dst must differ");
  hipStream_t s = (hipStream_t)stream;
  const sfk_op_desc &d = op->d;
  const bool fp32 = d.precision == SFK_FP32;
  switch (d.kind) {
  case SFK_CELL_LAPLACE:
  case SFK_CONT_LAPLACE: {
    SFK_REQUIRE(!(d.flags & SFK_FLAG_VARIABLE_COEF) || op->coef, SFK_ERR_NOT_READY, "coefficient not set");
    SFK_ENSURE_TABLES(op);
    if (uses_tp(op)) {
      TpLaunch T;
      fill_tp_launch(op, T);
      T.src = src;
      T.dst = dst;
      SFK_CHECK_HIP(launch_cont_tp(T, s));
      return SFK_OK;
    }
    CellLaunch L;
    fill_cell_launch(op, L);
    L.mode = d.kind == SFK_CELL_LAPLACE ? 0 : 1;
    L.src = src;
    L.dst = dst;
    SFK_CHECK_HIP(launch_cell(L, s));
    return SFK_OK;
  }
  case SFK_DG_SIPG: {
    if (d.geometry == SFK_GEOM_CARTESIAN) {
      CellLaunch L;
      SFK_ENSURE_TABLES(op);
      fill_cell_launch(op, L);
      L.mode = 2;
      L.refx = (d.flags & SFK_FLAG_REF_X_SWEEP) != 0;
      L.src = src;
      L.dst = dst;
      SFK_CHECK_HIP(launch_cell(L, s));
    } else {
      SFK_REQUIRE(op->cell_metric && op->face_metric, SFK_ERR_NOT_READY, "metric not set");
      SipgLaunch L;
      SFK_ENSURE_TABLES(op);
      fill_sipg_launch(op, L);
      L.src = src;
      L.dst = dst;
      SFK_CHECK_HIP(launch_sipg_general(L, s));
    }
    return SFK_OK;
  }
  case SFK_DG_ADVECT: {
    SFK_REQUIRE(!(d.flags & SFK_FLAG_VARIABLE_VELOCITY) || op->cell_vel, SFK_ERR_NOT_READY, "velocity not set");
    AdvectLaunch L;
    SFK_ENSURE_TABLES(op);
    fill_advect_launch(op, L);
    L.kind = 0;
    L.src = src;
    L.dst = dst;
    SFK_CHECK_HIP(launch_advect(L, s));
    return SFK_OK;
  }
  case SFK_DG_PRECOMP: {
    PrecompLaunch L;
    SFK_ENSURE_TABLES(op);
    fill_precomp_launch(op, L);
    L.src = src;
    L.dst = dst;
    SFK_CHECK_HIP(launch_precomp(L, s));
    return SFK_OK;
  }
  }
  set_error("unknown operator kind");
  return SFK_ERR_INVALID_ARG;
}

sfk_status sfk_apply_transpose(sfk_op *op, const void *src, void *dst, void *stream)
{
  SFK_REQUIRE(op && src && dst, SFK_ERR_INVALID_ARG, "sfk_apply_transpose: null argument");
  SFK_REQUIRE(src != dst, SFK_ERR_INVALID_ARG, "sfk_apply_transpose: src and dst must differ");
  if (op->d.kind != SFK_DG_ADVECT) {
    SFK_REQUIRE(!(op->d.kind == SFK_DG_SIPG && (op->d.flags & SFK_FLAG_REF_X_SWEEP)), SFK_ERR_UNSUPPORTED,
                "the reference-compatible SIPG operator is not symmetric");
    return sfk_apply(op, src, dst, stream);
  }
  hipStream_t s = (hipStream_t)stream;
  SFK_REQUIRE(!(op->d.flags & SFK_FLAG_VARIABLE_VELOCITY) || op->cell_vel, SFK_ERR_NOT_READY, "velocity not set");
  sfk_status st = ensure_scratch(op, (size_t)op->nvec * elem_size(op));
  if (st != SFK_OK) return st;
  // A^T w = L^T M^{-1} w: cell-local mass inverse, then the transposed weak form
  AdvectLaunch L;
  SFK_ENSURE_TABLES(op);
  fill_advect_launch(op, L);
  L.kind = 2;
  L.src = src;
  L.dst = op->scratch;
  SFK_CHECK_HIP(launch_advect(L, s));
  L.kind = 1;
  L.src = op->scratch;
  L.dst = dst;
  SFK_CHECK_HIP(launch_advect(L, s));
  return SFK_OK;
}

sfk_status sfk_cheby_step(sfk_op *op, const void *rhs, const void *sol_old, const void *diag, void *sol_new, void *tmp,
                          double c_np, double c_tm, void *stream)
{
  SFK_REQUIRE(op && rhs && sol_old && diag && sol_new && tmp, SFK_ERR_INVALID_ARG, "sfk_cheby_step: null argument");
  SFK_REQUIRE(sol_old != sol_new, SFK_ERR_INVALID_ARG, "sfk_cheby_step: sol_old and sol_new must differ");
  hipStream_t s = (hipStream_t)stream;
  const sfk_op_desc &d = op->d;
  const bool fp32 = d.precision == SFK_FP32;
  if (d.kind == SFK_DG_SIPG) {
    if (d.geometry == SFK_GEOM_CARTESIAN) {
      CellLaunch L;
      SFK_ENSURE_TABLES(op);
      fill_cell_launch(op, L);
      L.mode = 2;
      L.cheby = true;
      L.refx = (d.flags & SFK_FLAG_REF_X_SWEEP) != 0;
      L.src = sol_old;
      L.dst = sol_new;
      L.rhs = rhs;
      L.diag = diag;
      L.tm = tmp;
      L.c_np = c_np;
      L.c_tm = c_tm;
      SFK_CHECK_HIP(launch_cell(L, s));
    } else {
      SFK_REQUIRE(op->cell_metric && op->face_metric, SFK_ERR_NOT_READY, "metric not set");
      SipgLaunch L;
      SFK_ENSURE_TABLES(op);
      fill_sipg_launch(op, L, true);
      L.cheby = true;
      L.src = sol_old;
      L.dst = sol_new;
      L.rhs = rhs;
      L.diag = diag;
      L.tm = tmp;
      L.c_np = c_np;
      L.c_tm = c_tm;
      SFK_CHECK_HIP(launch_sipg_general(L, s));
    }
    return SFK_OK;
  }
  if (d.kind == SFK_DG_PRECOMP) {
    SFK_REQUIRE(d.n_rhs == 1, SFK_ERR_UNSUPPORTED, "sfk_cheby_step: n_rhs must be 1");
    PrecompLaunch L;
    SFK_ENSURE_TABLES(op);
    fill_precomp_launch(op, L);
    L.cheby = true;
    L.src = sol_old;
    L.dst = sol_new;
    L.rhs = rhs;
    L.diag = diag;
    L.tm = tmp;
    L.c_np = c_np;
    L.c_tm = c_tm;
    SFK_CHECK_HIP(launch_precomp(L, s));
    return SFK_OK;
  }
  if (d.kind == SFK_CONT_LAPLACE || d.kind == SFK_CELL_LAPLACE) {
    SFK_REQUIRE(d.n_rhs == 1, SFK_ERR_UNSUPPORTED, "sfk_cheby_step: n_rhs must be 1");
    SFK_ENSURE_TABLES(op);
    if (uses_tp(op)) {
      // the update fused into the tensor-product kernel's stores
      TpLaunch T;
      fill_tp_launch(op, T, true);
      T.cheby = 1;
      T.src = sol_old;
      T.dst = sol_new;
      T.rhs = rhs;
      T.diag = diag;
      T.tm = tmp;
      T.c_np = c_np;
      T.c_tm = c_tm;
      SFK_CHECK_HIP(launch_cont_tp(T, s));
      return SFK_OK;
    }
    sfk_status st = ensure_scratch(op, (size_t)op->nvec * elem_size(op));
    if (st != SFK_OK) return st;
    st = sfk_apply(op, sol_old, op->scratch, stream);
    if (st != SFK_OK) return st;
    SFK_CHECK_HIP(vec_cheby_update(fp32, op->nvec, op->scratch, rhs, sol_old, diag, sol_new, tmp, c_np, c_tm, s));
    return SFK_OK;
  }
  set_error("sfk_cheby_step: not supported for advection");
  return SFK_ERR_UNSUPPORTED;
}

sfk_status sfk_rk_stage(sfk_op *op, const void *src, const void *upd, void *vec_np, void *vec_tm, double c_np,
                        double c_tm, void *stream)
{
  SFK_REQUIRE(op && src && upd && vec_np, SFK_ERR_INVALID_ARG, "sfk_rk_stage: null argument");
  SFK_REQUIRE(op->d.kind == SFK_DG_ADVECT, SFK_ERR_UNSUPPORTED, "sfk_rk_stage: advection only");
  SFK_REQUIRE(c_tm == 0. || vec_tm != nullptr, SFK_ERR_INVALID_ARG, "sfk_rk_stage: vec_tm required when c_tm != 0");
  SFK_REQUIRE(src != vec_np && (c_tm == 0. || src != vec_tm), SFK_ERR_INVALID_ARG,
              "sfk_rk_stage: the stage input must not be overwritten");
  SFK_REQUIRE(!(op->d.flags & SFK_FLAG_VARIABLE_VELOCITY) || op->cell_vel, SFK_ERR_NOT_READY, "velocity not set");
  AdvectLaunch L;
  SFK_ENSURE_TABLES(op);
  fill_advect_launch(op, L);
  L.kind = 0;
  L.stage = true;
  L.src = src;
  L.upd = upd;
  L.dst = vec_np;
  L.tm = vec_tm;
  L.c_np = c_np;
  L.c_tm = c_tm;
  SFK_CHECK_HIP(launch_advect(L, (hipStream_t)stream));
  return SFK_OK;
}

sfk_status sfk_compute_diagonal(sfk_op *op, void *diag, void *stream)
{
  SFK_REQUIRE(op && diag, SFK_ERR_INVALID_ARG, "sfk_compute_diagonal: null argument");
  hipStream_t s = (hipStream_t)stream;
  const sfk_op_desc &d = op->d;
  const bool fp32 = d.precision == SFK_FP32;
  if (d.kind == SFK_CELL_LAPLACE || d.kind == SFK_CONT_LAPLACE) {
    SFK_REQUIRE(!(d.flags & SFK_FLAG_VARIABLE_COEF) || op->coef, SFK_ERR_NOT_READY, "coefficient not set");
    CellLaunch L;
    SFK_ENSURE_TABLES(op);
    fill_cell_launch(op, L);
    L.mode = d.kind == SFK_CELL_LAPLACE ? 0 : 1;
    L.dst = diag;
    if (L.mode == 1) SFK_CHECK_HIP(vec_fill(fp32, diag, op->nvec, 0., s));
    SFK_CHECK_HIP(launch_cell_diagonal(L, s));
    return SFK_OK;
  }
  // periodic DG with one cell in a direction: the cell is its own neighbour
  // there, and the diagonal kernels take the neighbour data as zero (the
  // cell's face coupling to itself reaches the diagonal for the GLL basis and
  // for low degrees of the Hermite-like basis)
  const bool self_neighbour = d.n_cells[0] == 1 || d.n_cells[1] == 1 || d.n_cells[2] == 1;
  SFK_REQUIRE(!((d.kind == SFK_DG_SIPG || d.kind == SFK_DG_PRECOMP) && self_neighbour), SFK_ERR_UNSUPPORTED,
              "sfk_compute_diagonal: a DG mesh with one cell in a direction (the cell is its own periodic "
              "neighbour) is not supported");
  if (d.kind == SFK_DG_SIPG) {
    SFK_REQUIRE(d.geometry == SFK_GEOM_PER_Q, SFK_ERR_UNSUPPORTED,
                "DG-SIPG diagonal: per-q geometry (use the identity-mapped metric for Cartesian meshes)");
    SFK_REQUIRE(op->cell_metric && op->face_metric, SFK_ERR_NOT_READY, "metric not set");
    SipgLaunch L;
    SFK_ENSURE_TABLES(op);
    fill_sipg_launch(op, L);
    L.diagonal = true;
    L.dst = diag;
    SFK_CHECK_HIP(launch_sipg_general(L, s));
    return SFK_OK;
  }
  if (d.kind == SFK_DG_PRECOMP) {
    PrecompLaunch L;
    SFK_ENSURE_TABLES(op);
    fill_precomp_launch(op, L);
    L.dst = diag;
    SFK_CHECK_HIP(launch_precomp_diagonal(L, s));
    return SFK_OK;
  }
  set_error("sfk_compute_diagonal: not supported for advection");
  return SFK_ERR_UNSUPPORTED;
}

sfk_status sfk_op_rhs_constant(sfk_op *op, double f, void *b, void *stream)
{
  SFK_REQUIRE(op && b, SFK_ERR_INVALID_ARG, "sfk_op_rhs_constant: null argument");
  SFK_REQUIRE(op->d.kind == SFK_CONT_LAPLACE, SFK_ERR_UNSUPPORTED, "sfk_op_rhs_constant: continuous Laplacian only");
  hipStream_t s = (hipStream_t)stream;
  CellLaunch L;
  SFK_ENSURE_TABLES(op);
  fill_cell_launch(op, L);
  L.mode = 1;
  L.dst = b;
  L.load = true;
  L.load_value = f;
  SFK_CHECK_HIP(vec_fill(op->d.precision == SFK_FP32, b, op->nvec, 0., s));
  SFK_CHECK_HIP(launch_cell_diagonal(L, s));
  return SFK_OK;
}

sfk_status sfk_grad_velocity(sfk_op *op, const void *u, const void *lambda, void *grad_cell, void *grad_face,
                             void *stream)
{
  SFK_REQUIRE(op && u && lambda && grad_cell && grad_face, SFK_ERR_INVALID_ARG, "sfk_grad_velocity: null argument");
  SFK_REQUIRE(op->d.kind == SFK_DG_ADVECT && (op->d.flags & SFK_FLAG_VARIABLE_VELOCITY), SFK_ERR_UNSUPPORTED,
              "sfk_grad_velocity: advection with SFK_FLAG_VARIABLE_VELOCITY only");
  SFK_REQUIRE(op->cell_vel, SFK_ERR_NOT_READY, "velocity not set");
  hipStream_t s = (hipStream_t)stream;
  sfk_status st = ensure_scratch(op, (size_t)op->nvec * elem_size(op));
  if (st != SFK_OK) return st;
  // mu = M^{-1} lambda, then dJ/da = mu^T dL/da u
  AdvectLaunch L;
  SFK_ENSURE_TABLES(op);
  fill_advect_launch(op, L);
  L.kind = 2;
  L.src = lambda;
  L.dst = op->scratch;
  SFK_CHECK_HIP(launch_advect(L, s));
  L.kind = 3;
  L.src = u;
  L.aux = op->scratch;
  L.dst = grad_cell;
  L.dst2 = grad_face;
  SFK_CHECK_HIP(launch_advect(L, s));
  return SFK_OK;
}

sfk_status sfk_grad_coefficient(sfk_op *op, const void *u, const void *lambda, void *grad_k, void *stream)
{
  SFK_REQUIRE(op && u && lambda && grad_k, SFK_ERR_INVALID_ARG, "sfk_grad_coefficient: null argument");
  SFK_REQUIRE((op->d.kind == SFK_CONT_LAPLACE) && (op->d.flags & SFK_FLAG_VARIABLE_COEF), SFK_ERR_UNSUPPORTED,
              "sfk_grad_coefficient: continuous Laplacian with SFK_FLAG_VARIABLE_COEF only");
  SFK_ENSURE_TABLES(op);
  CoefGradLaunch L;
  L.tables = &op->tab;
  L.dtab = op->dtab;
  L.fp32 = op->d.precision == SFK_FP32;
  L.u = u;
  L.lambda = lambda;
  L.grad = grad_k;
  L.ncells = op->ncells;
  for (int d = 0; d < 3; ++d) {
    L.nc[d] = op->d.n_cells[d];
    L.inv_jac[d] = op->d.inv_jac[d];
  }
  L.nrhs = op->d.n_rhs;
  L.vstride = op->nvec;
  SFK_CHECK_HIP(launch_coef_gradient(L, (hipStream_t)stream));
  return SFK_OK;
}

sfk_status sfk_apply_1d(int nn, const double *matrix_host, int sym, int forward, long long n_lines,
                        long long line_stride, int stride, const double *in, double *out, void *stream)
{
  SFK_REQUIRE(nn >= 2 && nn <= kMaxDegree + 1, SFK_ERR_INVALID_ARG, "nn out of range");
  SFK_REQUIRE(matrix_host && in && out && n_lines >= 0 && (sym == 1 || sym == -1), SFK_ERR_INVALID_ARG,
              "sfk_apply_1d: invalid argument");
  std::vector<double> M(matrix_host, matrix_host + nn * nn);
  SFK_CHECK_HIP(apply_1d(nn, M, sym, forward != 0, n_lines, line_stride, stride, in, out, (hipStream_t)stream));
  return SFK_OK;
}

sfk_status sfk_tables(int degree, int basis, double *S, double *D, double *Sinv, double *w, double *face,
                      double *hermite_d)
{
  Tables1D t;
  SFK_REQUIRE(build_tables(degree, basis, t), SFK_ERR_INVALID_ARG, "invalid degree or basis");
  const int n = t.n;
  if (S) std::memcpy(S, t.S.data(), sizeof(double) * n * n);
  if (D) std::memcpy(D, t.D.data(), sizeof(double) * n * n);
  if (Sinv) std::memcpy(Sinv, t.Sinv.data(), sizeof(double) * n * n);
  if (w) std::memcpy(w, t.weights.data(), sizeof(double) * n);
  if (face)
    for (int i = 0; i < n; ++i) {
      face[i] = t.fv0[i];
      face[n + i] = t.fv1[i];
      face[2 * n + i] = t.fd0[i];
      face[3 * n + i] = t.fd1[i];
    }
  if (hermite_d) *hermite_d = t.hermite_d;
  return SFK_OK;
}

sfk_status sfk_quadrature(int n, double *gauss_points, double *gauss_weights, double *gll)
{
  SFK_REQUIRE(n >= 2 && n <= 32, SFK_ERR_INVALID_ARG, "n out of range");
  std::vector<double> p, w;
  gauss_rule(n, p, w);
  if (gauss_points) std::memcpy(gauss_points, p.data(), sizeof(double) * n);
  if (gauss_weights) std::memcpy(gauss_weights, w.data(), sizeof(double) * n);
  if (gll) {
    std::vector<double> g = gll_points(n);
    std::memcpy(gll, g.data(), sizeof(double) * n);
  }
  return SFK_OK;
}

sfk_status sfk_layout_to_canonical(int degree, const int n_cells[3], int lanes, const double *in, double *out,
                                   void *stream)
{
  SFK_REQUIRE(n_cells && in && out && lanes >= 1 && degree >= 1, SFK_ERR_INVALID_ARG, "invalid argument");
  SFK_REQUIRE(n_cells[0] % lanes == 0, SFK_ERR_INVALID_ARG, "n_cells[0] must be a multiple of lanes");
  SFK_CHECK_HIP(layout_convert(degree, n_cells, lanes, in, out, true, (hipStream_t)stream));
  return SFK_OK;
}

sfk_status sfk_layout_to_interleaved(int degree, const int n_cells[3], int lanes, const double *in, double *out,
                                     void *stream)
{
  SFK_REQUIRE(n_cells && in && out && lanes >= 1 && degree >= 1, SFK_ERR_INVALID_ARG, "invalid argument");
  SFK_REQUIRE(n_cells[0] % lanes == 0, SFK_ERR_INVALID_ARG, "n_cells[0] must be a multiple of lanes");
  SFK_CHECK_HIP(layout_convert(degree, n_cells, lanes, in, out, false, (hipStream_t)stream));
  return SFK_OK;
}

sfk_status sfk_vec_cheby_update(int precision, long long n, const void *au, const void *rhs, const void *sol_old,
                                const void *diag, void *sol_new, void *tmp, double c_np, double c_tm, void *stream)
{
  SFK_REQUIRE(au && rhs && sol_old && diag && sol_new && tmp && n >= 0, SFK_ERR_INVALID_ARG,
              "sfk_vec_cheby_update: invalid argument");
  SFK_CHECK_HIP(vec_cheby_update(precision == SFK_FP32, n, au, rhs, sol_old, diag, sol_new, tmp, c_np, c_tm,
                                 (hipStream_t)stream));
  return SFK_OK;
}

sfk_status sfk_vec_alloc(size_t bytes, void **ptr)
{
  SFK_REQUIRE(ptr, SFK_ERR_INVALID_ARG, "sfk_vec_alloc: null argument");
  *ptr = nullptr;
  if (bytes == 0) return SFK_OK;
  if (hipMalloc(ptr, bytes) != hipSuccess) {
    *ptr = nullptr;
    set_error("sfk_vec_alloc: hipMalloc of " + std::to_string(bytes) + " bytes failed");
    return SFK_ERR_ALLOC;
  }
  return SFK_OK;
}

sfk_status sfk_vec_free(void *ptr)
{
  if (ptr) SFK_CHECK_HIP(hipFree(ptr));
  return SFK_OK;
}

sfk_status sfk_memcpy(void *dst, const void *src, size_t bytes, int kind)
{
  SFK_REQUIRE((dst && src) || bytes == 0, SFK_ERR_INVALID_ARG, "sfk_memcpy: null argument");
  SFK_REQUIRE(kind >= 0 && kind <= 2, SFK_ERR_INVALID_ARG, "sfk_memcpy: kind must be 0, 1 or 2");
  if (bytes == 0) return SFK_OK;
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
  SFK_CHECK_HIP(hipMemcpy(dst, src, bytes, k));
  return SFK_OK;
}

sfk_status sfk_synchronize(void *stream)
{
  SFK_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
  return SFK_OK;
}

sfk_status sfk_vec_fill(int precision, void *x, long long n, double value, void *stream)
{
  SFK_REQUIRE(x && n >= 0, SFK_ERR_INVALID_ARG, "invalid argument");
  SFK_CHECK_HIP(vec_fill(precision == SFK_FP32, x, n, value, (hipStream_t)stream));
  return SFK_OK;
}

sfk_status sfk_vec_dot(int precision, const void *x, const void *y, long long n, double *result, void *stream)
{
  SFK_REQUIRE(x && y && result && n >= 0, SFK_ERR_INVALID_ARG, "invalid argument");
  void *buf = nullptr;
  constexpr int kPartials = 2048;
  SFK_CHECK_HIP(hipMalloc(&buf, sizeof(double) * (kPartials + 1)));
  hipError_t e = vec_dot(precision == SFK_FP32, x, y, n, (double *)buf + kPartials, (double *)buf, kPartials,
                         (hipStream_t)stream);
  if (e == hipSuccess) e = hipMemcpyAsync(result, (double *)buf + kPartials, sizeof(double), hipMemcpyDeviceToHost,
                                          (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
  (void)hipFree(buf);
  SFK_CHECK_HIP(e);
  return SFK_OK;
}

} // extern "C"

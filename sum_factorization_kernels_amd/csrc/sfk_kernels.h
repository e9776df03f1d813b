// Launch descriptors of the HIP kernels (host side).  Each launcher picks
// the compile-time degree instantiation and the precision at run time.
#pragma once

#include <hip/hip_runtime.h>

#include "sfk_internal.h"

namespace sfk {

// Workgroup schedule fields of the operator launches: cb = cells per
// workgroup (0: the degree's default), order = block order (0 linear, 1 one
// contiguous range of cells per XCD; logical_block in sfk_device.h).

// kernels whose dynamic LDS exceeds the default 64 KiB need the limit
// raised (gfx950 has 160 KiB of LDS per CU)
inline hipError_t allow_dynamic_lds(const void *kern, size_t bytes)
{
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

struct CellLaunch {
  const Tables1D *tables = nullptr;
  bool fp32 = false;
  int mode = 0; // 0 element-local, 1 continuous (Dirichlet), 2 DG-SIPG Cartesian
  bool cheby = false, refx = false;
  const void *src = nullptr;
  void *dst = nullptr;
  const void *rhs = nullptr, *diag = nullptr;
  void *tm = nullptr;
  double c_np = 0., c_tm = 0.;
  const void *coef = nullptr;
  long long ncells = 0;
  int nc[3] = {1, 1, 1};
  double inv_jac[3] = {1., 1., 1.};
  int nrhs = 1;
  long long vstride = 0;
  bool load = false;       // launch_cell_diagonal: load vector of a constant instead
  double load_value = 0.;
  int cont_scatter = -1;   // continuous: -1 automatic, 0 zero fill + atomics, 1 colored
  int cb = 0, order = 1;   // schedule
  const void *dtab = nullptr; // DevTab<N,T> of the operator in device memory
};

hipError_t launch_cell(const CellLaunch &L, hipStream_t stream);

// continuous Laplacian, constant coefficient: tensor-product (Kronecker
// sum) kernel on the node grid (sfk_cont_tp.hip)
struct TpLaunch {
  int n = 0;            // points per direction (degree + 1)
  bool fp32 = false;
  int cheby = 0;        // 0 matvec, 1 fused Chebyshev step (tm form), 2 three-term
                        // Chebyshev step with the operator's inverse diagonal
  const void *tptab = nullptr; // TpTab<n, T> in device memory
  const void *src = nullptr;
  void *dst = nullptr;
  const void *rhs = nullptr, *diag = nullptr;
  void *tm = nullptr;
  const void *prev = nullptr; // cheby 2: previous iterate (null: zero)
  double c_np = 0., c_tm = 0.;
  // cheby 2: the input (src_from_rhs) or the previous iterate
  // (prev_from_rhs) is s_rhs D^{-1} rhs, formed in the kernel (not read)
  bool src_from_rhs = false, prev_from_rhs = false;
  double s_rhs = 0.;
  // cheby 0, FP64, one right-hand side: also write the workgroups' partial
  // sums of src . dst here (tp_launch_blocks of them)
  double *dot_part = nullptr;
  int nc[3] = {1, 1, 1};
  int nrhs = 1;
  long long vstride = 0;
  int shape = 0;        // tile shape index (tp_shape)
  int zc = 0;           // cells in z per workgroup (0: all)
  int order = 1;        // block order
};
hipError_t launch_cont_tp(const TpLaunch &L, hipStream_t stream);
long long tp_launch_blocks(const TpLaunch &L); // workgroups of one right-hand side
hipError_t upload_tp_tables(const Tables1D &t, const double inv_jac[3], bool fp32, void **dev);
int tp_shape_count(int n);
void tp_shape(int n, int shape, int *bx, int *by);
size_t tp_lds_bytes(int n, int shape, bool fp32);
hipError_t launch_cell_diagonal(const CellLaunch &L, hipStream_t stream);

// DG-SIPG with per-quadrature-point metric (deformed mesh)
struct SipgLaunch {
  const Tables1D *tables = nullptr;
  bool fp32 = false;
  bool cheby = false;
  bool diagonal = false; // write the operator diagonal into dst
  const void *src = nullptr;
  void *dst = nullptr;
  const void *rhs = nullptr, *diag = nullptr;
  void *tm = nullptr;
  double c_np = 0., c_tm = 0.;
  const void *cell_metric = nullptr, *face_metric = nullptr;
  long long ncells = 0;
  int nc[3] = {1, 1, 1};
  int cb = 0, order = 1; // schedule
  const void *dtab = nullptr;
};
hipError_t launch_sipg_general(const SipgLaunch &L, hipStream_t stream);

// DG advection (forward, RK stage, adjoint, mass inverse, velocity gradient)
struct AdvectLaunch {
  const Tables1D *tables = nullptr;
  bool fp32 = false;
  int kind = 0; // 0 forward M^{-1}L u, 1 L^T y (adjoint after M^{-1}), 2 M^{-1} x, 3 velocity gradient
  bool stage = false;
  const void *src = nullptr, *upd = nullptr, *aux = nullptr;
  void *dst = nullptr, *tm = nullptr, *dst2 = nullptr;
  double c_np = 0., c_tm = 0.;
  const void *cell_vel = nullptr, *face_vel = nullptr;
  double vel[3] = {0., 0., 0.};
  long long ncells = 0;
  int nc[3] = {1, 1, 1};
  double inv_jac[3] = {1., 1., 1.};
  int cb = 0, order = 1; // schedule
  const void *dtab = nullptr;
};
hipError_t launch_advect(const AdvectLaunch &L, hipStream_t stream);
// LDS bytes per cell of launch_advect's operator kernels (n points per direction)
size_t advect_lds_per_cell(int n, bool fp32);

// DG Laplacian with precomputed 1D matrices (PrecomputedDGLaplacian)
struct PrecompLaunch {
  int n = 0;             // points per direction (degree + 1)
  bool fp32 = false;
  bool couple = true;    // neighbour coupling (false: both constants are zero)
  bool cheby = false;
  const void *src = nullptr;
  void *dst = nullptr;
  const void *rhs = nullptr, *diag = nullptr;
  void *tm = nullptr;
  double c_np = 0., c_tm = 0.;
  long long ncells = 0;
  int nc[3] = {1, 1, 1};
  int nrhs = 1;
  int cb = 0, order = 1;         // schedule
  const void *ptab = nullptr;    // the operator's tables in device memory
};
hipError_t launch_precomp(const PrecompLaunch &L, hipStream_t stream);
hipError_t launch_precomp_diagonal(const PrecompLaunch &L, hipStream_t stream);
// the tables of launch_precomp in a new device allocation (f: direction factors)
hipError_t upload_precomp_tables(int n, const Precomp1D &pc, const double f[3], bool fp32, void **dev);
// LDS bytes per cell of launch_precomp's kernel (n points per direction)
size_t precomp_lds_per_cell(int n, bool fp32);

// continuous Laplacian: coefficient gradient g_q = JxW grad u . grad lambda
struct CoefGradLaunch {
  const Tables1D *tables = nullptr;
  bool fp32 = false;
  const void *u = nullptr, *lambda = nullptr;
  void *grad = nullptr;
  long long ncells = 0;
  int nc[3] = {1, 1, 1};
  double inv_jac[3] = {1., 1., 1.};
  int nrhs = 1;
  long long vstride = 0;
  const void *dtab = nullptr;
};
hipError_t launch_coef_gradient(const CoefGradLaunch &L, hipStream_t stream);

// sine-deformed mesh metric (cell and face), computed on the device
hipError_t launch_sine_metric(const Tables1D &t, bool fp32, const int nc[3], double amp,
                              void *cell_metric, void *face_metric, hipStream_t stream);

// vector kernels
hipError_t vec_fill(bool fp32, void *x, long long n, double v, hipStream_t s);
hipError_t vec_copy_convert(bool to_fp32, const void *x, void *y, long long n, hipStream_t s);
// partial dot products: writes ceil(n / chunk) FP64 partials into `partials`
hipError_t vec_dot(bool fp32, const void *x, const void *y, long long n, double *result_dev,
                   double *partials, int n_partials, hipStream_t s);
// y = a*x + b*y (FP64 scalars from device memory if ptrs given)
// the operator's 1D tables (DevTab<n, T> of sfk_device.h) in a new device
// allocation, T = float or double
hipError_t upload_device_tables(const Tables1D &t, bool fp32, void **dev);

// result_dev[0] = x.y, result_dev[1] = x.w (FP64 vectors; partials: 2*n_partials)
// result_dev[0] = sum of count partial sums (of a fused kernel)
hipError_t vec_sum_partials(const double *partials, long long count, double *result_dev, hipStream_t s);
hipError_t vec_dot2(const void *x, const void *y, const void *w, long long n, double *result_dev, double *partials,
                    int n_partials, hipStream_t s);
// solver kernels (FP64 vectors): x += alpha p (x null: skipped), r -= alpha
// q, optionally rs = (float) r, result_dev[0] = r.r (an non-null: alpha =
// *an / *ad, device scalars); and z = (double) zs with result_dev[0] = z.r,
// result_dev[1] = z.q (partials: 2*n_partials)
hipError_t vec_cg_update(long long n, double alpha, const double *an, const double *ad, const void *p,
                         const void *q, void *x, void *r, void *rs, double *result_dev, double *partials,
                         int n_partials, hipStream_t s);
hipError_t vec_widen_dot2(long long n, const void *zs, void *z, const void *r, const void *q, double *result_dev,
                          double *partials, int n_partials, hipStream_t s);
hipError_t vec_axpby(bool fp32, long long n, double a, const void *x, double b, void *y, hipStream_t s);
// z = c a b (elementwise)
hipError_t vec_mul_scaled(bool fp32, long long n, double c, const void *a, const void *b, void *z, hipStream_t s);
// p = (double) zs + b p (zs FP32, p FP64)
hipError_t vec_axpby_widen(long long n, const void *zs, double b, void *p, hipStream_t s);
// x += a p, then p = z + b p with a = *an / *ad, b = -*bn / *ad (device
// scalars; z FP32 when z_fp32, else FP64; x, p FP64)
hipError_t vec_xp_update(bool z_fp32, long long n, const double *an, const double *ad, const double *bn,
                         const void *z, void *p, void *x, hipStream_t s);
hipError_t vec_cheby_update(bool fp32, long long n, const void *au, const void *rhs, const void *sol_old,
                            const void *diag, void *sol_new, void *tm, double c_np, double c_tm,
                            hipStream_t s);
hipError_t vec_zero_dirichlet(bool fp32, void *x, const int g[3], hipStream_t s);
hipError_t layout_convert(int degree, const int nc[3], int lanes, const double *in, double *out,
                          bool to_canonical, hipStream_t s);
hipError_t apply_1d(int nn, const std::vector<double> &M, int sym, bool forward, long long n_lines,
                    long long line_stride, int stride, const double *in, double *out, hipStream_t s);

} // namespace sfk

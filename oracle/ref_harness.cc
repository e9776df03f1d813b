// Thin C entry points around the REFERENCE implementation, compiled from the
// reference headers where they lie (legacy/include of the reference
// checkout; see oracle/Makefile).  Output goes to oracle/_ref/ only.
//
// TEST INFRASTRUCTURE ONLY: used in this container to write golden vectors
// (tests/golden/make_golden.py), to pin the C oracle, and as the
// "reference" CPU baseline in bench.py.  Never linked into libsfk.so.
//
// Vectors cross this boundary in the reference's own lane-interleaved
// layout (VectorizedArray<double> lanes = ref_n_lanes()).

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <limits>
#include <memory>
#include <vector>
#include <array>
#include <x86intrin.h>
#include <mpi.h>

// the evaluators keep tables and vectors as private members; the harness
// needs to read and overwrite them.  This translation unit is compiled
// WITHOUT OpenMP: the reference's `#pragma omp ... private(...)` clause
// would otherwise be macro-expanded (the timed, threaded build is
// ref_bench.cc, which uses the public interface only).
#define private public
#define READ_VECTOR
#include "gauss_formula.h"
#include "lagrange_polynomials.h"
#include "vectorization.h"
#include "aligned_vector.h"
#include "utilities.h"
#include "matrix_vector_kernel.h"
#include "evaluation_cell_laplacian.h"
#include "evaluation_dg_laplacian.h"
#include "evaluation_dg_advect.h"
#include "precomputed_dg_laplacian.h"
#undef private

typedef VectorizedArray<double> VA;
static const int L = VA::n_array_elements;

template <typename AV>
static void copy_va_lane0(const AV &v, double *out, std::size_t n)
{
  for (std::size_t i = 0; i < n; ++i) out[i] = v[i][0];
}

static void load_scalar(AlignedVector<double> &v, const double *in)
{
  std::memcpy(v.begin(), in, sizeof(double) * v.size());
}
static void store_scalar(const AlignedVector<double> &v, double *out)
{
  std::memcpy(out, v.begin(), sizeof(double) * v.size());
}

// ---------------------------------------------------------------- tables
template <int degree>
static int tables_dispatch(int kind, int p, double *values_eo, double *gradients_eo,
                           double *face_eo, double *inv_eo, double *hd)
{
  if (p != degree) {
    if constexpr (degree < 15)
      return tables_dispatch<degree + 1>(kind, p, values_eo, gradients_eo, face_eo, inv_eo, hd);
    else
      return -1;
  }
  const int n = degree + 1, h = (n + 1) / 2;
  if (kind == 0) {
    EvaluationCellLaplacian<3, degree, double> e;
    e.fill_shape_values();
    copy_va_lane0(e.shape_values, values_eo, n * h);
    copy_va_lane0(e.shape_gradients, gradients_eo, n * h);
    *hd = e.hermite_derivative_on_face[0];
  } else if (kind == 1) {
    EvaluationDGLaplacian<3, degree, double> e;
    e.fill_shape_values();
    copy_va_lane0(e.shape_values_eo, values_eo, n * h);
    copy_va_lane0(e.shape_gradients_eo, gradients_eo, n * h);
    copy_va_lane0(e.shape_values_on_face_eo, face_eo, 2 * n);
    *hd = e.hermite_derivative_on_face[0];
  } else {
    EvaluationDGAdvection<3, degree, double> e;
    e.fill_shape_values();
    copy_va_lane0(e.shape_values_eo, values_eo, n * h);
    copy_va_lane0(e.shape_gradients_eo, gradients_eo, n * h);
    copy_va_lane0(e.shape_values_on_face_eo, face_eo, 2 * n);
    copy_va_lane0(e.inv_shape_values_eo, inv_eo, n * h);
  }
  return 0;
}

// ------------------------------------------------------------ 1D kernel
template <int nn, int type, bool cor, bool add, int dg>
static void k1d(int stride, const double *c, const double *in, double *out, const double *a,
                const double *dgc, double *face)
{
  // the kernel only compiles for Number = VectorizedArray (its streaming
  // store branch); broadcast the scalar line into all lanes and read lane 0
  const int len = stride * (nn - 1) + 1;
  std::vector<VA> vc(nn * ((nn + 1) / 2)), vin(len), vout(len), vadd(len), vdg(2 * nn), vface(4);
  for (std::size_t i = 0; i < vc.size(); ++i) vc[i] = c[i];
  for (int i = 0; i < len; ++i) {
    vin[i] = in[i];
    vout[i] = out[i];
    vadd[i] = a ? a[i] : 0.;
  }
  for (int i = 0; i < 2 * nn; ++i) vdg[i] = dgc ? dgc[i] : 0.;
  for (int i = 0; i < 4; ++i) vface[i] = face ? face[i] : 0.;
  // the stride is a template argument in the reference; dispatch the few
  // strides the tests use (1, nn, nn*nn)
  if (stride == 1)
    apply_1d_matvec_kernel<nn, 1, type, cor, add, VA, VA, false, dg>(vc.data(), vin.data(), vout.data(),
                                                                      vadd.data(), vdg.data(), vface.data());
  else if (stride == nn)
    apply_1d_matvec_kernel<nn, nn, type, cor, add, VA, VA, false, dg>(vc.data(), vin.data(), vout.data(),
                                                                       vadd.data(), vdg.data(), vface.data());
  else
    apply_1d_matvec_kernel<nn, nn * nn, type, cor, add, VA, VA, false, dg>(
        vc.data(), vin.data(), vout.data(), vadd.data(), vdg.data(), vface.data());
  for (int i = 0; i < len; ++i) out[i] = vout[i][0];
  if (face)
    for (int i = 0; i < 4; ++i) face[i] = vface[i][0];
}

template <int nn>
static int k1d_dispatch(int n, int type, int cor, int add, int dg, int stride, const double *c,
                        const double *in, double *out, const double *a, const double *dgc,
                        double *face)
{
  if (n != nn) {
    if constexpr (nn < 12)
      return k1d_dispatch<nn + 1>(n, type, cor, add, dg, stride, c, in, out, a, dgc, face);
    else
      return -1;
  }
#define K(T, C, A, G)                                             \
  if (type == T && cor == C && add == A && dg == G) {             \
    k1d<nn, T, C, A, G>(stride, c, in, out, a, dgc, face);        \
    return 0;                                                     \
  }
  K(0, 1, 0, 0) K(0, 0, 0, 0) K(1, 1, 0, 0) K(1, 0, 0, 0) K(1, 0, 1, 0) K(0, 0, 1, 0)
  K(1, 1, 0, 1) K(1, 1, 0, 2) K(1, 0, 0, 1) K(1, 0, 0, 2) K(1, 0, 1, 1) K(1, 0, 1, 2)
#undef K
  return -2;
}

// -------------------------------------------------------- cell Laplacian
template <int degree>
static int cell_dispatch(int p, int n_batches, const double *inv_jac, const double *src, double *dst)
{
  if (p != degree) {
    if constexpr (degree < 15)
      return cell_dispatch<degree + 1>(p, n_batches, inv_jac, src, dst);
    else
      return -1;
  }
  auto e = std::make_unique<EvaluationCellLaplacian<3, degree, double>>();
  e->initialize(n_batches, true);
  double det = 1.;
  for (int d = 0; d < 3; ++d) {
    e->jacobian_data[d] = inv_jac[d];
    det *= inv_jac[d];
  }
  e->jxw_data[0] = 1. / det;
  const std::size_t nd = (std::size_t)n_batches * e->dofs_per_cell;
  for (std::size_t i = 0; i < nd; ++i)
    for (int v = 0; v < L; ++v) e->input_array[i][v] = src[i * L + v];
  e->matrix_vector_product();
  for (std::size_t i = 0; i < nd; ++i)
    for (int v = 0; v < L; ++v) dst[i * L + v] = e->output_array[i][v];
  return 0;
}

// --------------------------------------------------------- DG Laplacian
template <int degree>
static int dglap_dispatch(int p, const unsigned *nc, const double *inv_jac, int mode,
                          const double *sol_old, const double *rhs, const double *diag,
                          const double *tm_in, double c_np, double c_tm, double *out_new,
                          double *out_tm)
{
  if (p != degree) {
    if constexpr (degree < 15)
      return dglap_dispatch<degree + 1>(p, nc, inv_jac, mode, sol_old, rhs, diag, tm_in, c_np, c_tm,
                                        out_new, out_tm);
    else
      return -1;
  }
  auto e = std::make_unique<EvaluationDGLaplacian<3, degree, double>>();
  e->blx = std::max(20 / (degree + 1), 3);
  e->bly = std::max(2, 20 / (degree + 1));
  e->blz = std::max(2, 20 / (degree + 1));
  unsigned n_cells[3] = {nc[0], nc[1], nc[2]};
  e->initialize(n_cells);
  if (inv_jac) {
    double det = 1.;
    for (int d = 0; d < 3; ++d) {
      e->jacobian_data[d] = inv_jac[d];
      det *= inv_jac[d];
    }
    e->jxw_data[0] = 1. / det;
  }
  if (mode == 0) {
    load_scalar(e->sol_new, sol_old);
    e->do_matvec();
    store_scalar(e->sol_tmp, out_new);
  } else {
    load_scalar(e->sol_old, sol_old);
    load_scalar(e->sol_rhs, rhs);
    load_scalar(e->mat_diagonal, diag);
    load_scalar(e->sol_tmp, tm_in);
    e->template do_cheby_iter<true>(e->sol_rhs, e->sol_old, e->mat_diagonal, e->sol_new, e->sol_tmp,
                                    c_np, c_tm);
    store_scalar(e->sol_new, out_new);
    store_scalar(e->sol_tmp, out_tm);
  }
  return 0;
}

// ---------------------------------------------------------- DG advection
template <int degree>
static int advect_dispatch(int p, const unsigned *nc, const double *inv_jac, int mode,
                           const double *src, const double *upd, const double *tm_in, double c_np,
                           double c_tm, double *out_np, double *out_tm)
{
  if (p != degree) {
    if constexpr (degree < 15)
      return advect_dispatch<degree + 1>(p, nc, inv_jac, mode, src, upd, tm_in, c_np, c_tm, out_np,
                                         out_tm);
    else
      return -1;
  }
  auto e = std::make_unique<EvaluationDGAdvection<3, degree, double>>();
  e->blx = std::max(20 / (degree + 1), 3);
  e->bly = std::max(2, 20 / (degree + 1));
  e->blz = std::max(2, 20 / (degree + 1));
  unsigned n_cells[3] = {nc[0], nc[1], nc[2]};
  e->initialize(n_cells);
  if (inv_jac) {
    double det = 1.;
    for (int d = 0; d < 3; ++d) {
      e->jacobian_data[d] = inv_jac[d];
      det *= inv_jac[d];
    }
    e->jxw_data[0] = 1. / det;
  }
  load_scalar(e->sol_new, src);
  if (mode == 0) {
    e->do_matvec();
    store_scalar(e->sol_tmp, out_np);
  } else {
    load_scalar(e->sol_old, upd);
    load_scalar(e->sol_tmp, tm_in);
    AlignedVector<double> np;
    np.resize(e->sol_new.size());
    e->template do_rk_stage<true>(e->sol_new, e->sol_old, np, e->sol_tmp, 0., c_np, c_tm);
    store_scalar(np, out_np);
    store_scalar(e->sol_tmp, out_tm);
  }
  return 0;
}

// Known-answer tests of the advection class: max error of M^{-1}L(u_exact)
// against the analytic time derivative (the quantity evaluation_dg_advect.h
// :221-251 prints) and of the solution after n_steps LSRK steps (:253-283).
template <int degree>
static int advect_kat_dispatch(int p, const unsigned *nc, int n_steps, double *err_deriv,
                               double *err_sol, double *t_end, double *inv_jac_out)
{
  if (p != degree) {
    if constexpr (degree < 15)
      return advect_kat_dispatch<degree + 1>(p, nc, n_steps, err_deriv, err_sol, t_end, inv_jac_out);
    else
      return -1;
  }
  auto e = std::make_unique<EvaluationDGAdvection<3, degree, double>>();
  e->blx = std::max(20 / (degree + 1), 3);
  e->bly = std::max(2, 20 / (degree + 1));
  e->blz = std::max(2, 20 / (degree + 1));
  unsigned n_cells[3] = {nc[0], nc[1], nc[2]};
  e->initialize(n_cells);
  for (int d = 0; d < 3; ++d) inv_jac_out[d] = e->jacobian_data[d][0];
  e->do_matvec();
  std::vector<double> gl = get_gauss_lobatto_points(degree + 1);
  auto max_err = [&](const AlignedVector<double> &v, double time, bool deriv) {
    ExactSolution<3> exact(time);
    double err = 0;
    const unsigned ncx = e->n_cells[0], ncy = e->n_cells[1], ncz = e->n_cells[2];
    for (unsigned iz = 0; iz < ncz; ++iz)
      for (unsigned iy = 0; iy < ncy; ++iy)
        for (unsigned ixb = 0; ixb < ncx; ++ixb) {
          std::size_t c = 0;
          for (int lz = 0; lz <= degree; ++lz)
            for (int ly = 0; ly <= degree; ++ly)
              for (int lx = 0; lx <= degree; ++lx)
                for (int lane = 0; lane < L; ++lane, ++c) {
                  Point<3> pt;
                  pt.data[0] = (ixb * L + lane + gl[lx]) / e->jacobian_data[0][0];
                  pt.data[1] = (iy + gl[ly]) / e->jacobian_data[1][0];
                  pt.data[2] = (iz + gl[lz]) / e->jacobian_data[2][0];
                  const std::size_t idx = ((iz * ncy + iy) * ncx + ixb) * e->dofs_per_cell * L + c;
                  const double ex = deriv ? exact.time_derivative(pt) : exact.value(pt);
                  err = std::max(err, std::abs(ex - v[idx]));
                }
        }
    return err;
  };
  *err_deriv = max_err(e->sol_tmp, 0., true);
  // re-initialize so that the time loop starts from the exact solution
  e->initialize(n_cells);
  double time = 0.;
  for (int s = 0; s < n_steps; ++s) e->do_time_step(time);
  *err_sol = max_err(e->sol_new, time, false);
  *t_end = time;
  return 0;
}

// --------------------------------------------- precomputed DG Laplacian
template <int degree>
static int precomp_dispatch(int p, const unsigned *nc, const double *src, double *dst, double *mass_eo,
                            double *lap_eo, double *outer)
{
  if (p != degree) {
    if constexpr (degree < 15)
      return precomp_dispatch<degree + 1>(p, nc, src, dst, mass_eo, lap_eo, outer);
    else
      return -1;
  }
  auto e = std::make_unique<PrecomputedDGLaplacian<3, degree, double>>();
  // block sizes of legacy/src/test_dg_precomputed.cc
  e->blx = std::max(20 / (degree + 1), 3);
  e->bly = std::max(2, 20 / (degree + 1));
  e->blz = std::max(2, 20 / (degree + 1));
  unsigned n_cells[3] = {nc[0], nc[1], nc[2]};
  e->initialize(n_cells);
  if (src) {
    load_scalar(e->input_array, src);
    e->matrix_vector_product();
    store_scalar(e->output_array, dst);
  }
  const int n = degree + 1, h = (n + 1) / 2;
  copy_va_lane0(e->mass_1d_eo, mass_eo, n * h);
  copy_va_lane0(e->laplace_1d_eo, lap_eo, n * h);
  outer[0] = e->value_outer_1;
  outer[1] = e->value_outer_2;
  return 0;
}

extern "C" {

__attribute__((visibility("default"))) int ref_n_lanes() { return L; }

__attribute__((visibility("default"))) void ref_gauss_points(int n, double *out)
{
  std::vector<double> v = get_gauss_points(n);
  std::copy(v.begin(), v.end(), out);
}
__attribute__((visibility("default"))) void ref_gauss_weights(int n, double *out)
{
  std::vector<double> v = get_gauss_weights(n);
  std::copy(v.begin(), v.end(), out);
}
__attribute__((visibility("default"))) void ref_gll_points(int n, double *out)
{
  std::vector<double> v = get_gauss_lobatto_points(n);
  std::copy(v.begin(), v.end(), out);
}
__attribute__((visibility("default"))) double ref_hermite_value(int degree, int i, double x) { return HermiteLikePolynomialBasis(degree).value(i, x); }
__attribute__((visibility("default"))) double ref_hermite_derivative(int degree, int i, double x)
{
  return HermiteLikePolynomialBasis(degree).derivative(i, x);
}

__attribute__((visibility("default"))) int ref_tables(int kind, int degree, double *values_eo, double *gradients_eo, double *face_eo,
               double *inv_eo, double *hd)
{
  return tables_dispatch<1>(kind, degree, values_eo, gradients_eo, face_eo, inv_eo, hd);
}

__attribute__((visibility("default"))) int ref_apply_1d(int nn, int type, int contract_over_rows, int add_into, int do_dg, int stride,
                 const double *coeffs, const double *in, double *out, const double *add,
                 const double *dg, double *face)
{
  return k1d_dispatch<2>(nn, type, contract_over_rows, add_into, do_dg, stride, coeffs, in, out, add,
                         dg, face);
}

__attribute__((visibility("default"))) int ref_cell_laplacian(int degree, int n_batches, const double *inv_jac, const double *src, double *dst)
{
  return cell_dispatch<1>(degree, n_batches, inv_jac, src, dst);
}

__attribute__((visibility("default"))) int ref_dg_laplacian(int degree, const unsigned *n_cells, const double *inv_jac, int mode,
                     const double *sol_old, const double *rhs, const double *diag,
                     const double *tm_in, double c_np, double c_tm, double *out_new, double *out_tm)
{
  return dglap_dispatch<1>(degree, n_cells, inv_jac, mode, sol_old, rhs, diag, tm_in, c_np, c_tm,
                           out_new, out_tm);
}

__attribute__((visibility("default"))) int ref_dg_advect(int degree, const unsigned *n_cells, const double *inv_jac, int mode,
                  const double *src, const double *upd, const double *tm_in, double c_np,
                  double c_tm, double *out_np, double *out_tm)
{
  return advect_dispatch<1>(degree, n_cells, inv_jac, mode, src, upd, tm_in, c_np, c_tm, out_np,
                            out_tm);
}

// PrecomputedDGLaplacian (legacy/include/precomputed_dg_laplacian.h): dst =
// matrix_vector_product() of src (lane-interleaved, n_cells[0] counts
// cells); src = NULL only fills the tables.  mass_eo, lap_eo: n*ceil(n/2)
// even-odd tables; outer = {value_outer_1, value_outer_2}.
__attribute__((visibility("default"))) int ref_dg_precomputed(int degree, const unsigned *n_cells, const double *src,
                                                              double *dst, double *mass_eo, double *lap_eo,
                                                              double *outer)
{
  return precomp_dispatch<1>(degree, n_cells, src, dst, mass_eo, lap_eo, outer);
}

__attribute__((visibility("default"))) int ref_advect_kat(int degree, const unsigned *n_cells, int n_steps, double *err_deriv,
                   double *err_sol, double *t_end, double *inv_jac_out)
{
  return advect_kat_dispatch<1>(degree, n_cells, n_steps, err_deriv, err_sol, t_end, inv_jac_out);
}

} // extern "C"

// C++ evaluator classes with the reference's names and methods, running on
// the MI355X through the C ABI of include/sfk.h (no HIP types here: plain
// g++ can compile code that uses this header; link with libsfk.so).
//
//   EvaluationCellLaplacian<dim,degree,Number>  legacy/include/evaluation_cell_laplacian.h:26-908
//   EvaluationDGLaplacian<dim,degree,Number>    legacy/include/evaluation_dg_laplacian.h:33-716
//   EvaluationDGAdvection<dim,degree,Number>    legacy/include/evaluation_dg_advect.h:96-845
//   PrecomputedDGLaplacian<dim,degree,Number>   legacy/include/precomputed_dg_laplacian.h:33-389
//   EvaluationContinuousLaplacian<dim,degree,Number>  (no reference counterpart)
//
// As in the reference, an evaluator owns its vectors (here in device memory,
// canonical cell-major layout) and initialize() sets the initial data.  Only
// dim == 3 exists.  Differences in behaviour, all deliberate:
//  * errors throw sfk::Error (status code + message) instead of calling
//    std::terminate / abort;
//  * x-direction periodic neighbours are exact for every cell (the
//    reference's evaluation_dg_laplacian.h:270 reads the last cell's +x
//    neighbour from a wrong address);
//  * the SIPG operator applies S, not S^T, in its last x sweep
//    (evaluation_dg_laplacian.h:481-483); set ref_x_sweep = true before
//    initialize() to reproduce the reference's numbers instead;
//  * cells per x row need not be a multiple of a SIMD width (lanes = 1 on
//    the GPU), and block sizes blx/bly/blz are accepted but do not change
//    the GPU schedule.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "sfk.h"

namespace sfk {

struct Error : std::runtime_error {
  sfk_status status;
  Error(sfk_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

inline void check(sfk_status s)
{
  if (s != SFK_OK) throw Error(s, sfk_last_error());
}

template <typename Number>
constexpr int precision_of()
{
  return sizeof(Number) == 4 ? SFK_FP32 : SFK_FP64;
}

// device vector of Number (replaces AlignedVector, legacy/include/aligned_vector.h)
template <typename Number>
class DeviceVector {
public:
  DeviceVector() = default;
  explicit DeviceVector(std::size_t n) { resize(n); }
  DeviceVector(const DeviceVector &) = delete;
  DeviceVector &operator=(const DeviceVector &) = delete;
  DeviceVector(DeviceVector &&o) noexcept : ptr_(o.ptr_), n_(o.n_) { o.ptr_ = nullptr, o.n_ = 0; }
  ~DeviceVector() { sfk_vec_free(ptr_); }
  void resize(std::size_t n)
  {
    check(sfk_vec_free(ptr_));
    ptr_ = nullptr;
    check(sfk_vec_alloc(n * sizeof(Number), &ptr_));
    n_ = n;
  }
  std::size_t size() const { return n_; }
  void *data() { return ptr_; }
  const void *data() const { return ptr_; }
  void fill(double v) { check(sfk_vec_fill(precision_of<Number>(), ptr_, (long long)n_, v, nullptr)); }
  void upload(const std::vector<Number> &h)
  {
    if (h.size() != n_) throw Error(SFK_ERR_INVALID_ARG, "DeviceVector::upload: size mismatch");
    check(sfk_memcpy(ptr_, h.data(), n_ * sizeof(Number), 0));
  }
  std::vector<Number> download() const
  {
    std::vector<Number> h(n_);
    check(sfk_memcpy(h.data(), ptr_, n_ * sizeof(Number), 1));
    return h;
  }

private:
  void *ptr_ = nullptr;
  std::size_t n_ = 0;
};

class Operator {
public:
  explicit Operator(const sfk_op_desc &d) { check(sfk_op_create(&d, &op_)); }
  Operator(const Operator &) = delete;
  Operator &operator=(const Operator &) = delete;
  ~Operator() { sfk_op_destroy(op_); }
  sfk_op *get() const { return op_; }
  long long vector_size() const
  {
    long long n = 0;
    check(sfk_op_vector_size(op_, &n));
    return n;
  }
  // quadrature points of all cells, and face points per direction
  long long n_cell_q() const
  {
    long long c = 0, f = 0;
    check(sfk_op_qdata_sizes(op_, &c, &f));
    return c;
  }
  long long n_face_q() const
  {
    long long c = 0, f = 0;
    check(sfk_op_qdata_sizes(op_, &c, &f));
    return f;
  }

private:
  sfk_op *op_ = nullptr;
};

inline sfk_op_desc make_desc(int kind, int degree, int precision, int basis, const unsigned *n_cells,
                             const double *inv_jac)
{
  sfk_op_desc d{};
  d.kind = kind;
  d.degree = degree;
  d.precision = precision;
  d.basis = basis;
  d.geometry = SFK_GEOM_CARTESIAN;
  d.flags = 0;
  for (int k = 0; k < 3; ++k) {
    d.n_cells[k] = (int)n_cells[k];
    d.inv_jac[k] = inv_jac[k];
  }
  d.n_rhs = 1;
  return d;
}

// inverse Jacobian diagonal of the reference drivers (evaluation_cell_laplacian.h
// :803-813, evaluation_dg_laplacian.h:605-613)
inline std::vector<double> reference_diagonal_jacobian() { return {4., 2., 4. / 3.}; }

} // namespace sfk

// --------------------------------------------------------------------------
template <int dim, int degree, typename Number>
class EvaluationCellLaplacian {
  static_assert(dim == 3, "three-dimensional operators only");

public:
  static constexpr unsigned int dimension = dim;
  static constexpr unsigned int n_q_points = (degree + 1) * (degree + 1) * (degree + 1);
  static constexpr unsigned int dofs_per_cell = n_q_points;
  int basis = SFK_BASIS_HERMITE; // the reference class uses the Hermite-like basis (:822)

  // n_element_batches = number of cells (one cell per GPU "lane")
  void initialize(const unsigned int n_element_batches, const bool is_cartesian)
  {
    if (!is_cartesian)
      throw sfk::Error(SFK_ERR_UNSUPPORTED, "EvaluationCellLaplacian: Cartesian cells only (the reference aborts, :81)");
    const unsigned nc[3] = {n_element_batches, 1, 1};
    const std::vector<double> ij = sfk::reference_diagonal_jacobian();
    op_.reset(new sfk::Operator(
        sfk::make_desc(SFK_CELL_LAPLACE, degree, sfk::precision_of<Number>(), basis, nc, ij.data())));
    n_cells_ = n_element_batches;
    input_array.resize(std::size_t(n_cells_) * dofs_per_cell);
    output_array.resize(std::size_t(n_cells_) * dofs_per_cell);
    input_array.fill(0.);
    output_array.fill(0.);
  }
  std::size_t n_elements() const { return n_cells_; }

  // boundary-integral check of evaluation_cell_laplacian.h:108-214 for the
  // basis in use: A_K x_d must equal the boundary integral of the linear
  // function; returns the largest error over the three directions.
  double do_verification();

  void matrix_vector_product()
  {
    sfk::check(sfk_apply(op_->get(), input_array.data(), output_array.data(), nullptr));
    sfk::check(sfk_synchronize(nullptr));
  }

  sfk::DeviceVector<Number> input_array, output_array;

private:
  std::unique_ptr<sfk::Operator> op_;
  unsigned n_cells_ = 0;
};

// --------------------------------------------------------------------------
template <int dim, int degree, typename Number>
class EvaluationDGLaplacian {
  static_assert(dim == 3, "three-dimensional operators only");

public:
  static constexpr unsigned int dimension = dim;
  static constexpr unsigned int n_q_points = (degree + 1) * (degree + 1) * (degree + 1);
  static constexpr unsigned int dofs_per_cell = n_q_points;
  unsigned int blx = 1, bly = 1, blz = 1;
  bool ref_x_sweep = false;

  // n_cells_in[0] counts cells (the reference divides it by its SIMD width)
  void initialize(const unsigned int *n_cells_in)
  {
    for (int d = 0; d < 3; ++d) n_cells[d] = n_cells_in[d];
    const std::vector<double> ij = sfk::reference_diagonal_jacobian();
    sfk_op_desc d = sfk::make_desc(SFK_DG_SIPG, degree, sfk::precision_of<Number>(), SFK_BASIS_HERMITE, n_cells,
                                   ij.data());
    if (ref_x_sweep) d.flags |= SFK_FLAG_REF_X_SWEEP;
    op_.reset(new sfk::Operator(d));
    const std::size_t n = n_elements() * dofs_per_cell;
    for (auto *v : {&sol_old, &sol_new, &mat_diagonal, &sol_tmp, &sol_rhs}) v->resize(n);
    // initial data of evaluation_dg_laplacian.h:86-90
    sol_old.fill(1.);
    sol_new.fill(0.);
    mat_diagonal.fill(1.);
    sol_tmp.fill(0.);
    sol_rhs.fill(1.);
  }
  std::size_t n_elements() const { return std::size_t(n_cells[0]) * n_cells[1] * n_cells[2]; }
  void do_verification() {}
  void do_matvec() { do_cheby_iter<false>(sol_new, sol_new, sol_new, sol_tmp, sol_tmp, 0., 0.); }
  void do_chebyshev() { do_cheby_iter<true>(sol_rhs, sol_old, mat_diagonal, sol_new, sol_tmp, 0.5, 0.5); }

  // evaluate_chebyshev = false: vec_tm... = A sol_old written to sol_new
  // (do_matvec passes sol_tmp); true: the fused update of :485-502 with
  // `src` as the right-hand side
  template <bool evaluate_chebyshev = true>
  void do_cheby_iter(const sfk::DeviceVector<Number> &src, const sfk::DeviceVector<Number> &sol_old_in,
                     const sfk::DeviceVector<Number> &diag, sfk::DeviceVector<Number> &sol_new_out,
                     sfk::DeviceVector<Number> &vec_tm, const Number coefficient_np, const Number coefficient_tm)
  {
    if (evaluate_chebyshev)
      sfk::check(sfk_cheby_step(op_->get(), src.data(), sol_old_in.data(), diag.data(), sol_new_out.data(),
                                vec_tm.data(), coefficient_np, coefficient_tm, nullptr));
    else
      sfk::check(sfk_apply(op_->get(), sol_old_in.data(), sol_new_out.data() == sol_old_in.data()
                                                              ? vec_tm.data()
                                                              : sol_new_out.data(),
                           nullptr));
    sfk::check(sfk_synchronize(nullptr));
  }

  // the unfused vector update of evaluation_dg_laplacian.h:556-601
  void emulate_cheby_vector_updates();

  unsigned int n_cells[3] = {1, 1, 1};
  sfk::DeviceVector<Number> sol_old, sol_new, sol_rhs, sol_tmp, mat_diagonal;

private:
  std::unique_ptr<sfk::Operator> op_;
};

// --------------------------------------------------------------------------
// PrecomputedDGLaplacian (legacy/include/precomputed_dg_laplacian.h:33-389):
// DG Laplacian of the Hermite-like basis from precomputed 1D matrices on
// unit cells, periodic.  ref_coupling = true before initialize() reproduces
// the reference's neighbour coupling constants (which vanish); the default
// couples neighbouring cells through the SIPG face terms.
template <int dim, int degree, typename Number>
class PrecomputedDGLaplacian {
  static_assert(dim == 3, "three-dimensional operators only");

public:
  static constexpr unsigned int dimension = dim;
  static constexpr unsigned int n_q_points = (degree + 1) * (degree + 1) * (degree + 1);
  static constexpr unsigned int dofs_per_cell = n_q_points;
  unsigned int blx = 1, bly = 1, blz = 1;
  bool ref_coupling = false;

  // n_cells_in[0] counts cells (the reference divides it by its SIMD width)
  void initialize(const unsigned int *n_cells_in)
  {
    for (int d = 0; d < 3; ++d) n_cells[d] = n_cells_in[d];
    const double unit[3] = {1., 1., 1.};
    sfk_op_desc d = sfk::make_desc(SFK_DG_PRECOMP, degree, sfk::precision_of<Number>(), SFK_BASIS_HERMITE, n_cells,
                                   unit);
    if (ref_coupling) d.flags |= SFK_FLAG_REF_PRECOMP_COUPLING;
    op_.reset(new sfk::Operator(d));
    const std::size_t n = n_elements() * dofs_per_cell;
    input_array.resize(n);
    output_array.resize(n);
    // initial data of precomputed_dg_laplacian.h:57-80
    input_array.fill(1.);
    output_array.fill(0.);
  }
  std::size_t n_elements() const { return std::size_t(n_cells[0]) * n_cells[1] * n_cells[2]; }

  // The reference has no check (:90-93).  Here: max |A 1| / max |A x| for a
  // fixed non-constant x, which vanishes to rounding for the SIPG coupling
  // (the basis is a partition of unity); input_array and output_array are
  // overwritten.
  double do_verification()
  {
    std::vector<Number> x(input_array.size());
    for (std::size_t i = 0; i < x.size(); ++i) x[i] = Number(std::sin(0.37 * double(i % 1013)));
    input_array.upload(x);
    matrix_vector_product();
    double scale = 0.;
    for (Number v : output_array.download()) scale = std::max(scale, std::abs(double(v)));
    input_array.fill(1.);
    matrix_vector_product();
    double r = 0.;
    for (Number v : output_array.download()) r = std::max(r, std::abs(double(v)));
    return scale > 0. ? r / scale : r;
  }

  void matrix_vector_product()
  {
    sfk::check(sfk_apply(op_->get(), input_array.data(), output_array.data(), nullptr));
    sfk::check(sfk_synchronize(nullptr));
  }

  unsigned int n_cells[3] = {1, 1, 1};
  sfk::DeviceVector<Number> input_array, output_array;

private:
  std::unique_ptr<sfk::Operator> op_;
};

// --------------------------------------------------------------------------
template <int dim>
struct ExactSolution {
  // evaluation_dg_advect.h:37-92: sin(4 pi x') cos(4 pi y') cos(4 pi z'),
  // x' = x - t a, a = (1, 0.15, -0.05)
  double time, a[3] = {1., 0.15, -0.05};
  explicit ExactSolution(double t = 0.) : time(t) {}
  double value(const double *p) const
  {
    const double k = 4. * 3.14159265358979323846;
    return std::sin(k * (p[0] - time * a[0])) * std::cos(k * (p[1] - time * a[1])) *
           std::cos(k * (p[2] - time * a[2]));
  }
  double time_derivative(const double *p) const
  {
    const double k = 4. * 3.14159265358979323846;
    const double x = p[0] - time * a[0], y = p[1] - time * a[1], z = p[2] - time * a[2];
    const double r = -a[0] * k * std::cos(k * x) * std::cos(k * y) * std::cos(k * z);
    return r + k * std::sin(k * x) * (a[1] * std::sin(k * y) * std::cos(k * z) + a[2] * std::cos(k * y) * std::sin(k * z));
  }
};

template <int dim, int degree, typename Number>
class EvaluationDGAdvection {
  static_assert(dim == 3, "three-dimensional operators only");

public:
  static constexpr unsigned int dimension = dim;
  static constexpr unsigned int n_q_points = (degree + 1) * (degree + 1) * (degree + 1);
  static constexpr unsigned int dofs_per_cell = n_q_points;
  unsigned int blx = 1, bly = 1, blz = 1;

  // Per-point velocity (no reference counterpart): set before initialize()
  // to keep the velocity per cell and face quadrature point (initialized to
  // the constant field a; set_velocity replaces it), which the gradient
  // grad_velocity needs.
  bool variable_velocity = false;

  // domain [0,2]^3 (evaluation_dg_advect.h:719-727); sol_new = exact
  // solution at the GLL nodes, time step 0.2 h / |a| / p (:160-165)
  void initialize(const unsigned int *n_cells_in)
  {
    sfk_graph_destroy(graph_); // a captured step refers to the old vectors
    graph_ = nullptr;
    for (int d = 0; d < 3; ++d) {
      n_cells[d] = n_cells_in[d];
      inv_jac[d] = n_cells[d] / 2.;
    }
    sfk_op_desc d = sfk::make_desc(SFK_DG_ADVECT, degree, sfk::precision_of<Number>(), SFK_BASIS_GLL, n_cells, inv_jac);
    ExactSolution<dim> ex;
    for (int k = 0; k < 3; ++k) d.velocity[k] = ex.a[k];
    if (variable_velocity) d.flags |= SFK_FLAG_VARIABLE_VELOCITY;
    op_.reset(new sfk::Operator(d));
    if (variable_velocity) {
      // the constant field at every point: cell [cell][3][q] = a, face
      // [d][cell][q] = a . e_d (the upper face's unit normal)
      const std::size_t nq = dofs_per_cell, nf = (degree + 1) * (degree + 1), nc = n_elements();
      std::vector<double> cell(nc * 3 * nq), face(3 * nc * nf);
      for (std::size_t c = 0; c < nc; ++c)
        for (int k = 0; k < 3; ++k)
          for (std::size_t q = 0; q < nq; ++q) cell[(c * 3 + k) * nq + q] = ex.a[k];
      for (int k = 0; k < 3; ++k)
        for (std::size_t i = 0; i < nc * nf; ++i) face[k * nc * nf + i] = ex.a[k];
      set_velocity(cell, face);
    }
    const std::size_t n = n_elements() * dofs_per_cell;
    for (auto *v : {&sol_old, &sol_new, &sol_tmp}) v->resize(n);
    sol_new.upload(nodal(ex, false));
    sol_old.fill(0.);
    sol_tmp.fill(0.);
    const double tn = std::sqrt(ex.a[0] * ex.a[0] + ex.a[1] * ex.a[1] + ex.a[2] * ex.a[2]);
    time_step = 0.2 / inv_jac[0] / tn / degree;
  }
  std::size_t n_elements() const { return std::size_t(n_cells[0]) * n_cells[1] * n_cells[2]; }

  // five-stage, fourth-order low-storage RK (evaluation_dg_advect.h:190-214).
  // With use_graph, the five fused stages are captured once into a HIP
  // graph on the evaluator's own stream and every later step replays it
  // (one launch and one synchronisation per step instead of five).
  bool use_graph = false;
  void do_time_step(double &current_time)
  {
    if (use_graph) {
      if (!graph_ || graph_dt_ != time_step) capture_time_step();
      sfk::check(sfk_graph_launch(graph_, stream_));
      sfk::check(sfk_synchronize(stream_));
    } else {
      time_step_stages(nullptr, true);
    }
    current_time += time_step;
  }
  ~EvaluationDGAdvection()
  {
    sfk_graph_destroy(graph_);
    sfk_stream_destroy(stream_);
  }
  void do_matvec() { do_rk_stage<false>(sol_new, sol_new, sol_tmp, sol_tmp, 0., 0., 0.); }

  // max |du/dt_exact - M^{-1}L(u)| at the GLL nodes (:221-251)
  double verify_derivative()
  {
    do_matvec();
    const std::vector<Number> r = sol_tmp.download();
    const std::vector<Number> ex = nodal(ExactSolution<dim>(0.), true);
    double err = 0;
    for (std::size_t i = 0; i < r.size(); ++i) err = std::max(err, std::abs(double(ex[i]) - double(r[i])));
    return err;
  }
  // max error against the travelling wave at time t (:253-283)
  double compute_max_error(const double t)
  {
    const std::vector<Number> u = sol_new.download();
    const std::vector<Number> ex = nodal(ExactSolution<dim>(t), false);
    double err = 0;
    for (std::size_t i = 0; i < u.size(); ++i) err = std::max(err, std::abs(double(ex[i]) - double(u[i])));
    return err;
  }

  template <bool evaluate_stage = true>
  void do_rk_stage(const sfk::DeviceVector<Number> &src, const sfk::DeviceVector<Number> &update_vec,
                   sfk::DeviceVector<Number> &vec_np, sfk::DeviceVector<Number> &vec_tm, const double /*time*/,
                   const Number coefficient_np, const Number coefficient_tm)
  {
    if (evaluate_stage) {
      // the stage input must stay intact: stage through the spare vector
      // when the reference's call pattern writes the input (stage 5 writes
      // np over upd, which is allowed; src is never overwritten there)
      sfk::check(sfk_rk_stage(op_->get(), src.data(), update_vec.data(), vec_np.data(), vec_tm.data(),
                              coefficient_np, coefficient_tm, nullptr));
    } else {
      sfk::check(sfk_apply(op_->get(), src.data(), vec_np.data(), nullptr));
    }
    sfk::check(sfk_synchronize(nullptr));
  }

  // ---- adjoint / gradient API (no reference counterpart; the forward
  // operator is do_matvec's M^{-1} L, evaluation_dg_advect.h:285-676)
  // velocity per point, layouts of sfk_op_set_velocity: cell [cell][3][q],
  // face [d][cell][q] (a . n on the upper face in direction d)
  void set_velocity(const std::vector<double> &cell, const std::vector<double> &face)
  {
    sfk::check(sfk_op_set_velocity(op_->get(), cell.data(), cell.size(), face.data(), face.size()));
  }
  std::size_t n_cell_q() const { return (std::size_t)op_->n_cell_q(); }
  std::size_t n_face_q() const { return (std::size_t)op_->n_face_q(); }
  // dst = A^T src (the downwind transpose of M^{-1} L)
  void apply_transpose(const sfk::DeviceVector<Number> &src, sfk::DeviceVector<Number> &dst)
  {
    sfk::check(sfk_apply_transpose(op_->get(), src.data(), dst.data(), nullptr));
    sfk::check(sfk_synchronize(nullptr));
  }
  // gradient of lambda^T A(a) u with respect to the per-point velocity a
  // (with lambda = A u - f: the gradient of 0.5 |A u - f|^2); grad_cell has
  // 3 n_cell_q() entries, grad_face 3 n_face_q() (variable_velocity only)
  void grad_velocity(const sfk::DeviceVector<Number> &u, const sfk::DeviceVector<Number> &lambda,
                     sfk::DeviceVector<Number> &grad_cell, sfk::DeviceVector<Number> &grad_face)
  {
    sfk::check(sfk_grad_velocity(op_->get(), u.data(), lambda.data(), grad_cell.data(), grad_face.data(), nullptr));
    sfk::check(sfk_synchronize(nullptr));
  }

  unsigned int n_cells[3] = {1, 1, 1};
  double inv_jac[3] = {1., 1., 1.};
  double time_step = 0.;
  sfk::DeviceVector<Number> sol_old, sol_new, sol_tmp;

private:
  std::vector<Number> nodal(const ExactSolution<dim> &ex, bool derivative) const
  {
    std::vector<double> gp(degree + 1), gw(degree + 1), gl(degree + 1);
    sfk::check(sfk_quadrature(degree + 1, gp.data(), gw.data(), gl.data()));
    std::vector<Number> v(n_elements() * dofs_per_cell);
    std::size_t idx = 0;
    for (unsigned iz = 0; iz < n_cells[2]; ++iz)
      for (unsigned iy = 0; iy < n_cells[1]; ++iy)
        for (unsigned ix = 0; ix < n_cells[0]; ++ix)
          for (int lz = 0; lz <= degree; ++lz)
            for (int ly = 0; ly <= degree; ++ly)
              for (int lx = 0; lx <= degree; ++lx, ++idx) {
                const double p[3] = {(ix + gl[lx]) / inv_jac[0], (iy + gl[ly]) / inv_jac[1], (iz + gl[lz]) / inv_jac[2]};
                v[idx] = Number(derivative ? ex.time_derivative(p) : ex.value(p));
              }
    return v;
  }
  // the stage sequence of do_time_step (time only enters the reference's
  // exact solution, not the stages)
  sfk_status time_step_stages(void *stream, bool sync)
  {
    const double a21 = 970286171893. / 4311952581923., a32 = 6584761158862. / 12103376702013.,
                 a43 = 2251764453980. / 15575788980749., a54 = 26877169314380. / 34165994151039.;
    const double b1 = 1153189308089. / 22510343858157., b2 = 1772645290293. / 4653164025191.,
                 b3 = -1672844663538. / 4480602732383., b4 = 2114624349019. / 3568978502595.,
                 b5 = 5198255086312. / 14908931495163.;
    const double dt = time_step;
    sfk_op *op = op_->get();
    struct Stage {
      const void *src, *upd;
      void *np, *tm;
      double c_np, c_tm;
    } st[5] = {{sol_new.data(), sol_new.data(), sol_old.data(), sol_tmp.data(), a21 * dt, (b1 - a21) * dt},
               {sol_old.data(), sol_tmp.data(), sol_new.data(), sol_tmp.data(), a32 * dt, (b2 - a32) * dt},
               {sol_new.data(), sol_tmp.data(), sol_old.data(), sol_tmp.data(), a43 * dt, (b3 - a43) * dt},
               {sol_old.data(), sol_tmp.data(), sol_tmp.data(), sol_new.data(), a54 * dt, (b4 - a54) * dt},
               {sol_tmp.data(), sol_new.data(), sol_new.data(), sol_tmp.data(), b5 * dt, 0.}};
    for (const Stage &g : st) {
      const sfk_status s = sfk_rk_stage(op, g.src, g.upd, g.np, g.tm, g.c_np, g.c_tm, stream);
      if (s != SFK_OK) return s;
      if (sync) sfk::check(sfk_synchronize(stream));
    }
    return SFK_OK;
  }
  void capture_time_step()
  {
    sfk_graph_destroy(graph_);
    graph_ = nullptr;
    if (!stream_) sfk::check(sfk_stream_create(&stream_));
    sfk::check(sfk_graph_begin(stream_));
    const sfk_status s = time_step_stages(stream_, false);
    const sfk_status e = sfk_graph_end(stream_, &graph_); // always end the capture
    sfk::check(s);
    sfk::check(e);
    graph_dt_ = time_step;
  }
  std::unique_ptr<sfk::Operator> op_;
  void *stream_ = nullptr;
  sfk_graph *graph_ = nullptr;
  double graph_dt_ = 0.;
};

// --------------------------------------------------------------------------
// Continuous Q_p Laplacian -div(k grad u) with homogeneous Dirichlet rows on
// the uniform mesh of [0,1]^3 (no reference counterpart: the reference's
// cell Laplacian is element-local, evaluation_cell_laplacian.h:221-222).
// Vectors hold the (nx p+1)(ny p+1)(nz p+1) nodes, x fastest, n_rhs of them
// back to back; the coefficient k, when variable, is given per cell
// quadrature point.  Adjoint API: the operator is symmetric
// (apply_transpose == matrix_vector_product), and grad_coefficient returns
// d(lambda^T A(k) u)/dk per quadrature point.
template <int dim, int degree, typename Number>
class EvaluationContinuousLaplacian {
  static_assert(dim == 3, "three-dimensional operators only");

public:
  void initialize(const unsigned int *n_cells_in, unsigned n_rhs_in = 1, bool variable_coefficient = false)
  {
    double ij[3];
    for (int d = 0; d < 3; ++d) {
      n_cells[d] = n_cells_in[d];
      ij[d] = n_cells[d];
    }
    sfk_op_desc d = sfk::make_desc(SFK_CONT_LAPLACE, degree, sfk::precision_of<Number>(), SFK_BASIS_GLL, n_cells, ij);
    d.n_rhs = (int)n_rhs_in;
    if (variable_coefficient) d.flags |= SFK_FLAG_VARIABLE_COEF;
    op_.reset(new sfk::Operator(d));
    n_rhs = n_rhs_in;
    input_array.resize(vector_size() * n_rhs);
    output_array.resize(vector_size() * n_rhs);
    input_array.fill(0.);
    output_array.fill(0.);
  }
  std::size_t vector_size() const { return (std::size_t)op_->vector_size(); }
  std::size_t n_cell_q() const { return (std::size_t)op_->n_cell_q(); }
  void set_coefficient(const std::vector<double> &k)
  {
    sfk::check(sfk_op_set_coefficient(op_->get(), k.data(), k.size()));
  }
  // output_array = A input_array (all right-hand sides)
  void matrix_vector_product()
  {
    sfk::check(sfk_apply(op_->get(), input_array.data(), output_array.data(), nullptr));
    sfk::check(sfk_synchronize(nullptr));
  }
  void apply_transpose(const sfk::DeviceVector<Number> &src, sfk::DeviceVector<Number> &dst)
  {
    sfk::check(sfk_apply_transpose(op_->get(), src.data(), dst.data(), nullptr));
    sfk::check(sfk_synchronize(nullptr));
  }
  // grad[q] = sum over the right-hand sides of JxW_q grad u_q . grad lambda_q
  void grad_coefficient(const sfk::DeviceVector<Number> &u, const sfk::DeviceVector<Number> &lambda,
                        sfk::DeviceVector<Number> &grad)
  {
    sfk::check(sfk_grad_coefficient(op_->get(), u.data(), lambda.data(), grad.data(), nullptr));
    sfk::check(sfk_synchronize(nullptr));
  }
  // the autotuner's schedule for matrix_vector_product (seconds per apply)
  double tune(int repeats = 5)
  {
    double t = 0.;
    sfk::check(sfk_tune(op_->get(), repeats, nullptr, &t, nullptr));
    return t;
  }
  sfk_op *handle() const { return op_->get(); }

  unsigned int n_cells[3] = {1, 1, 1};
  unsigned n_rhs = 1;
  sfk::DeviceVector<Number> input_array, output_array;

private:
  std::unique_ptr<sfk::Operator> op_;
};

// ------------------------------------------------------------------ bodies
template <int dim, int degree, typename Number>
double EvaluationCellLaplacian<dim, degree, Number>::do_verification()
{
  // A_K applied to u = x_d (linear in each cell): the cell integral of
  // grad v . grad u equals the boundary integral of v du/dn, so
  // A_K u - (face term on x_d = 1) + (face term on x_d = 0) = 0.
  const int n = degree + 1;
  std::vector<double> gp(n), gw(n), gl(n), S(n * n);
  sfk::check(sfk_quadrature(n, gp.data(), gw.data(), gl.data()));
  sfk::check(sfk_tables(degree, basis, S.data(), nullptr, nullptr, nullptr, nullptr, nullptr));
  std::vector<double> b1(n, 0.), v0(n, 0.), v1(n, 0.);
  // 1D integrals and end values of the basis through its values at Gauss points
  for (int i = 0; i < n; ++i)
    for (int q = 0; q < n; ++q) b1[i] += S[i * n + q] * gw[q];
  {
    // end values: interpolate phi_i from the Gauss points (exact for degree p)
    std::vector<double> l0(n, 1.), l1(n, 1.);
    for (int q = 0; q < n; ++q)
      for (int m = 0; m < n; ++m)
        if (m != q) {
          l0[q] *= (0. - gp[m]) / (gp[q] - gp[m]);
          l1[q] *= (1. - gp[m]) / (gp[q] - gp[m]);
        }
    for (int i = 0; i < n; ++i)
      for (int q = 0; q < n; ++q) {
        v0[i] += l0[q] * S[i * n + q];
        v1[i] += l1[q] * S[i * n + q];
      }
  }
  const std::vector<double> ij = sfk::reference_diagonal_jacobian();
  // nodal values of x_d: for the GLL basis the node positions, for the
  // Hermite-like basis the coefficients of the linear function x
  std::vector<double> lin(n);
  if (basis == SFK_BASIS_GLL) {
    for (int i = 0; i < n; ++i) lin[i] = gl[i];
  } else {
    // x = sum_i c_i phi_i(x): phi_0 + phi_1 carry the value and slope at 0
    // (phi_1'(0) = -phi_0'(0)), solve on the Gauss points
    std::vector<double> A(n * n), rhs(n);
    for (int q = 0; q < n; ++q) {
      rhs[q] = gp[q];
      for (int i = 0; i < n; ++i) A[q * n + i] = S[i * n + q];
    }
    for (int c = 0; c < n; ++c) // Gaussian elimination with partial pivoting
    {
      int piv = c;
      for (int r = c + 1; r < n; ++r)
        if (std::abs(A[r * n + c]) > std::abs(A[piv * n + c])) piv = r;
      for (int k = 0; k < n; ++k) std::swap(A[c * n + k], A[piv * n + k]);
      std::swap(rhs[c], rhs[piv]);
      for (int r = c + 1; r < n; ++r) {
        const double f = A[r * n + c] / A[c * n + c];
        for (int k = c; k < n; ++k) A[r * n + k] -= f * A[c * n + k];
        rhs[r] -= f * rhs[c];
      }
    }
    for (int c = n - 1; c >= 0; --c) {
      double s = rhs[c];
      for (int k = c + 1; k < n; ++k) s -= A[c * n + k] * lin[k];
      lin[c] = s / A[c * n + c];
    }
  }
  double max_error = 0;
  for (int test = 0; test < 3; ++test) {
    std::vector<Number> in(n_cells_ * dofs_per_cell);
    for (unsigned c = 0; c < n_cells_; ++c)
      for (int k = 0, q = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < n; ++i, ++q) {
            const int idx[3] = {i, j, k};
            in[c * dofs_per_cell + q] = Number(lin[idx[test]] / ij[test]);
          }
    input_array.upload(in);
    matrix_vector_product();
    const std::vector<Number> out = output_array.download();
    int other[2], o = 0;
    for (int e = 0; e < 3; ++e)
      if (e != test) other[o++] = e;
    const double fac = 1. / (ij[other[0]] * ij[other[1]]);
    for (unsigned c = 0; c < n_cells_; ++c)
      for (int k = 0, q = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < n; ++i, ++q) {
            const int idx[3] = {i, j, k};
            const double ref = fac * (v1[idx[test]] - v0[idx[test]]) * b1[idx[other[0]]] * b1[idx[other[1]]];
            max_error = std::max(max_error, std::abs(double(out[c * dofs_per_cell + q]) - ref));
          }
  }
  return max_error;
}

template <int dim, int degree, typename Number>
void EvaluationDGLaplacian<dim, degree, Number>::emulate_cheby_vector_updates()
{
  // tmp = 0.6 tmp + 0.45 diag (new - rhs); new = old - tmp (:556-601)
  sfk::check(sfk_vec_cheby_update(sfk::precision_of<Number>(), (long long)sol_new.size(), sol_new.data(),
                                  sol_rhs.data(), sol_old.data(), mat_diagonal.data(), sol_new.data(),
                                  sol_tmp.data(), 0.45, 0.6, nullptr));
  sfk::check(sfk_synchronize(nullptr));
}

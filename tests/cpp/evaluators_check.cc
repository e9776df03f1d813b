// Runs the C++ evaluator classes of sfk_evaluators.h (the reference's API
// over the C ABI) on the GPU and writes what they compute into a directory,
// for tests/test_gpu_cpp_api.py to compare with the oracle and the
// reference's known answers:
//   EvaluationDGLaplacian<3,3,double>   do_matvec, do_chebyshev (4x3x2 cells)
//   EvaluationCellLaplacian<3,4,double> matrix_vector_product, do_verification
//   EvaluationDGAdvection<3,3,double>   verify_derivative, do_time_step with
//                                       and without use_graph, compute_max_error,
//                                       apply_transpose, grad_velocity
//   PrecomputedDGLaplacian<3,3,double>  do_verification
//   EvaluationContinuousLaplacian<3,4,double/float> matrix_vector_product
//                                       (variable coefficient, 2 right-hand
//                                       sides), grad_coefficient
//
//   evaluators_check <out_dir> <advect nx> <ny> <nz> <advect steps>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "sfk_evaluators.h"

static std::string g_out;

template <typename T>
static void dump(const char *name, const std::vector<T> &v)
{
  std::vector<double> d(v.begin(), v.end());
  FILE *f = std::fopen((g_out + "/" + name + ".bin").c_str(), "wb");
  if (!f || std::fwrite(d.data(), sizeof(double), d.size(), f) != d.size()) {
    std::fprintf(stderr, "cannot write %s\n", name);
    std::exit(2);
  }
  std::fclose(f);
}

// deterministic input pattern (the test regenerates it)
static std::vector<double> pattern(std::size_t n, double a)
{
  std::vector<double> v(n);
  for (std::size_t i = 0; i < n; ++i) v[i] = std::sin(a * double(i % 1013) + 0.1 * double(i / 1013));
  return v;
}

template <typename T>
static std::vector<T> as(const std::vector<double> &v)
{
  return std::vector<T>(v.begin(), v.end());
}

int main(int argc, char **argv)
{
  if (argc < 6) {
    std::fprintf(stderr, "usage: %s out_dir nx ny nz steps\n", argv[0]);
    return 2;
  }
  g_out = argv[1];
  FILE *txt = std::fopen((g_out + "/scalars.txt").c_str(), "w");
  try {
    {
      EvaluationDGLaplacian<3, 3, double> e;
      const unsigned nc[3] = {4, 3, 2};
      e.initialize(nc);
      const std::size_t n = e.n_elements() * e.dofs_per_cell;
      e.sol_new.upload(pattern(n, 0.37));
      e.do_matvec(); // A sol_new -> sol_tmp
      dump("dg_matvec", e.sol_tmp.download());
      e.sol_old.upload(pattern(n, 0.21));
      e.sol_rhs.upload(pattern(n, 0.53));
      std::vector<double> dg = pattern(n, 0.11);
      for (double &v : dg) v = 1.5 + 0.5 * v;
      e.mat_diagonal.upload(dg);
      e.sol_tmp.upload(pattern(n, 0.71));
      e.do_chebyshev();
      dump("dg_cheby_new", e.sol_new.download());
      dump("dg_cheby_tmp", e.sol_tmp.download());
    }
    {
      EvaluationCellLaplacian<3, 4, double> e;
      e.initialize(7, true);
      std::fprintf(txt, "cell_verification %.17g\n", e.do_verification());
      e.input_array.upload(pattern(7 * e.dofs_per_cell, 0.29));
      e.matrix_vector_product();
      dump("cell_matvec", e.output_array.download());
    }
    {
      const unsigned nc[3] = {(unsigned)std::atoi(argv[2]), (unsigned)std::atoi(argv[3]), (unsigned)std::atoi(argv[4])};
      const int steps = std::atoi(argv[5]);
      for (int graph = 0; graph < 2; ++graph) {
        EvaluationDGAdvection<3, 3, double> e;
        e.use_graph = graph != 0;
        e.initialize(nc);
        const double vd = e.verify_derivative();
        e.initialize(nc);
        double t = 0.;
        for (int s = 0; s < steps; ++s) e.do_time_step(t);
        std::fprintf(txt, "advect_graph%d verify %.17g time %.17g error %.17g\n", graph, vd, t, e.compute_max_error(t));
      }
    }
    {
      EvaluationDGAdvection<3, 3, double> e;
      e.variable_velocity = true;
      const unsigned nc[3] = {4, 4, 4};
      e.initialize(nc);
      const std::size_t n = e.n_elements() * e.dofs_per_cell;
      sfk::DeviceVector<double> u(n), lam(n), at(n), gc(3 * e.n_cell_q()), gf(3 * e.n_face_q());
      u.upload(pattern(n, 0.43));
      lam.upload(pattern(n, 0.17));
      e.apply_transpose(lam, at);
      e.grad_velocity(u, lam, gc, gf);
      dump("advect_transpose", at.download());
      dump("advect_grad_cell", gc.download());
      dump("advect_grad_face", gf.download());
    }
    {
      PrecomputedDGLaplacian<3, 3, double> e;
      const unsigned nc[3] = {5, 4, 3};
      e.initialize(nc);
      std::fprintf(txt, "precomputed_verification %.17g\n", e.do_verification());
    }
    {
      EvaluationContinuousLaplacian<3, 4, double> e;
      const unsigned nc[3] = {5, 4, 3};
      e.initialize(nc, 2, true);
      std::vector<double> k = pattern(e.n_cell_q(), 0.31);
      for (double &v : k) v = std::exp(0.5 * v);
      e.set_coefficient(k);
      e.input_array.upload(pattern(2 * e.vector_size(), 0.19));
      e.matrix_vector_product();
      dump("cont_matvec", e.output_array.download());
      sfk::DeviceVector<double> lam(2 * e.vector_size()), g(e.n_cell_q()), at(2 * e.vector_size());
      lam.upload(pattern(2 * e.vector_size(), 0.61));
      e.grad_coefficient(e.input_array, lam, g);
      e.apply_transpose(lam, at);
      dump("cont_grad", g.download());
      dump("cont_transpose", at.download());
    }
    {
      EvaluationContinuousLaplacian<3, 4, float> e;
      const unsigned nc[3] = {5, 4, 3};
      e.initialize(nc);
      e.tune(2);
      e.input_array.upload(as<float>(pattern(e.vector_size(), 0.23)));
      e.matrix_vector_product();
      dump("cont_fp32_matvec", e.output_array.download());
    }
  } catch (const sfk::Error &err) {
    std::fprintf(stderr, "sfk error %d: %s\n", (int)err.status, err.what());
    return 1;
  }
  std::fclose(txt);
  std::printf("evaluators_check done\n");
  return 0;
}

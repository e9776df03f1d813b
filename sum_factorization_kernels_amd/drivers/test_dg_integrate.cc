// DG-SIPG benchmark driver with the sizing, timing and report lines of the
// reference's legacy/src/test_dg_integrate.cc:29-253, running the
// EvaluationDGLaplacian mirror class on the GPU.
//
//   test_dg_integrate [degree=4] [vector_size=10000000] [n_tests=100] [variants=2]
//
// variants > 1 adds the fused Chebyshev step (CH), > 2 the unfused vector
// update (VU).  Timings: best of 5 repeats of n_tests calls, each call
// synchronised (the reference's calls are synchronous too).
#include <sys/time.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <type_traits>
#include <limits>

#include "sfk_evaluators.h"

static double wall()
{
  struct timeval t;
  gettimeofday(&t, nullptr);
  return t.tv_sec + 1e-6 * t.tv_usec;
}

template <int degree>
static void run_program(unsigned long vector_size_guess, unsigned n_tests, unsigned variants)
{
  const int dim = 3;
  EvaluationDGLaplacian<3, degree, double> evaluator;
  const unsigned n_dofs_cell = (degree + 1) * (degree + 1) * (degree + 1);
  const unsigned n_cells_tot = std::max<unsigned>(vector_size_guess / n_dofs_cell, 1U);
  // sizing of test_dg_integrate.cc:52-72 with one cell per lane
  unsigned n_cells[3];
  n_cells[0] = std::max(static_cast<unsigned>(1.00001 * std::pow((double)n_cells_tot, 1. / dim)), 1U);
  n_cells[1] = n_cells[0];
  n_cells[2] = std::max(n_cells_tot / (n_cells[0] * n_cells[1]), 1U);
  evaluator.blx = std::max(20 / (degree + 1), 3);
  evaluator.bly = std::max(2, 20 / (degree + 1));
  evaluator.blz = std::max(2, 20 / (degree + 1));
  n_cells[2] = std::max(1U, n_cells[2] / evaluator.blz) * evaluator.blz;
  n_cells[1] = std::max(1U, n_cells[1] / evaluator.bly) * evaluator.bly;
  n_cells[0] = std::max(n_cells_tot / (n_cells[1] * n_cells[2]), 1U);
  evaluator.initialize(n_cells);
  const std::size_t global_size = evaluator.n_elements() * evaluator.dofs_per_cell;
  std::cout << "\nPolynomial degree: " << degree << "\nVector size:       " << global_size << " [" << n_cells[0]
            << " x " << n_cells[1] << " x " << n_cells[2] << " times " << evaluator.dofs_per_cell << "]\n";

  auto time_best = [&](auto &&fn) {
    fn();
    double best = std::numeric_limits<double>::max();
    for (int i = 0; i < 5; ++i) {
      const double t0 = wall();
      for (unsigned t = 0; t < n_tests; ++t) fn();
      best = std::min(best, wall() - t0);
    }
    return best;
  };
  // operation count of test_dg_integrate.cc:129-138
  const std::size_t ops_interpolate = 2 * ((degree + 1) / 2) * 2 + degree + 1 + 2 * ((degree - 1) * (degree + 1) / 2);
  const std::size_t p1 = degree + 1;
  const std::size_t ops_cell = 4 * dim * ops_interpolate * p1 * p1 + 2 * dim * 2 * p1 * p1 +
                               2 * dim * 4 * ops_interpolate * p1 + 2 * dim * 2 * (p1 + 2 * p1 + 4) * p1 * p1 +
                               2 * dim * 12;
  auto report = [&](const char *tag, double best, std::size_t vectors, std::size_t ops_per_cell) {
    const double mem = double(global_size) * sizeof(double) * vectors * n_tests;
    const double ops = double(global_size / evaluator.dofs_per_cell) * ops_per_cell * n_tests;
    std::cout << tag << " degree " << std::setw(2) << degree << "  " << n_cells[0] << " x " << n_cells[1] << " x "
              << n_cells[2] << " elem " << evaluator.dofs_per_cell << ", block sizes: " << evaluator.blx << " x "
              << evaluator.bly << " x " << evaluator.blz << ", MDoFs/s: " << global_size * n_tests / best / 1e6
              << ", GB/s: " << mem / best * 1e-9 << " GFLOP/s: " << ops / best * 1e-9
              << " ops/dof: " << ops / global_size / n_tests << std::endl;
  };
  report("MV", time_best([&] { evaluator.do_matvec(); }), 2, ops_cell);
  if (variants > 1)
    report("CH", time_best([&] { evaluator.do_chebyshev(); }), 6, ops_cell + (2 * dim + 5) * p1 * p1 * p1);
  if (variants > 2)
    report("VU", time_best([&] { evaluator.emulate_cheby_vector_updates(); }), 7, 6 * evaluator.dofs_per_cell);
}

// runs f(std::integral_constant<int, degree>) for a run-time degree in 1..MAX
// (the reference's drivers recurse over their degree range the same way,
// RunTime<dim, degree, max_degree>)
template <int MAX, int D = 1, typename F>
static bool for_degree(int degree, F &&f)
{
  if (degree == D) {
    f(std::integral_constant<int, D>());
    return true;
  }
  if constexpr (D < MAX)
    return for_degree<MAX, D + 1>(degree, f);
  else
    return false;
}

int main(int argc, char **argv)
{
  int degree = argc > 1 ? std::atoi(argv[1]) : 4;
  const unsigned long vector_size = argc > 2 ? std::atol(argv[2]) : 10000000;
  const unsigned n_tests = argc > 3 ? std::atoi(argv[3]) : 100;
  const unsigned variants = argc > 4 ? std::atoi(argv[4]) : 2;
  std::cout << sfk_version() << std::endl;
  try {
    const bool known = for_degree<12>(degree, [&](auto d) { run_program<decltype(d)::value>(vector_size, n_tests, variants); });
    if (!known) {
      std::cerr << "degree must be in 1..12\n";
      return 1;
    }
  } catch (const sfk::Error &e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 2;
  }
  return 0;
}

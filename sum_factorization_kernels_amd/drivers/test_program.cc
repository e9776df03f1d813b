// Element-local cell Laplacian driver with the protocol and report line of
// the reference's legacy/src/test_program.cc:38-180 (READ_VECTOR form: every
// cell reads and writes its own data): boundary-integral self-check, warm-up,
// best of 3 runs of n_tests applies.
//
//   test_program [degree=4] [n_tests=100] [vector_size=10000000] [basis=0 Hermite | 1 GLL]
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <type_traits>
#include <limits>

#include "sfk_evaluators.h"

template <int degree>
static void run_program(unsigned n_tests, unsigned long vector_size, int basis)
{
  const int dim = 3;
  const unsigned n_cells = std::max<unsigned long>(vector_size / ((degree + 1) * (degree + 1) * (degree + 1)), 1UL);
  EvaluationCellLaplacian<3, degree, double> evaluator;
  evaluator.basis = basis;
  evaluator.initialize(n_cells, true);
  std::cout << "Boundary-integral self-check, max error: " << evaluator.do_verification() << std::endl;
  const std::size_t n_dofs = std::size_t(n_cells) * evaluator.dofs_per_cell;
  // operation count of test_program.cc:63-97 (no faces)
  const std::size_t ops_interpolate = 2 * ((degree + 1) / 2) * 2 + degree + 1 + 2 * ((degree - 1) * (degree + 1) / 2);
  const std::size_t p1 = degree + 1;
  const double ops = double(n_cells) * (4 * dim * ops_interpolate * p1 * p1 + dim * 2 * dim * p1 * p1 * p1);
  for (int i = 0; i < 5; ++i) evaluator.matrix_vector_product();
  double best = std::numeric_limits<double>::max();
  for (int r = 0; r < 3; ++r) {
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned t = 0; t < n_tests; ++t) evaluator.matrix_vector_product();
    const double avg = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / n_tests;
    best = std::min(best, avg);
  }
  std::cout << "Best result p" << degree << " --- DoFs/s " << std::setw(12) << n_dofs / best << "  GFlops/s "
            << std::setw(12) << 1e-9 * ops / best << "  GB/s " << std::setw(12) << 1e-9 * 2 * 8 * n_dofs / best
            << "  ops/dof  " << ops / n_dofs << std::endl;
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
  const int degree = argc > 1 ? std::atoi(argv[1]) : 4;
  const unsigned n_tests = argc > 2 ? std::atoi(argv[2]) : 100;
  const unsigned long vector_size = argc > 3 ? std::atol(argv[3]) : 10000000;
  const int basis = argc > 4 ? std::atoi(argv[4]) : 0;
  std::cout << sfk_version() << std::endl;
  try {
    const bool known = for_degree<15>(degree, [&](auto d) { run_program<decltype(d)::value>(n_tests, vector_size, basis); });
    if (!known) {
      std::cerr << "degree must be in 1..15\n";
      return 1;
    }
  } catch (const sfk::Error &e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 2;
  }
  return 0;
}

// DG advection benchmark driver with the protocol of the reference's
// legacy/src/test_dg_advect.cc:23-190 (degrees 1..8 on the GPU mirror
// class): derivative KAT, LSRK time steps, matvec throughput, and the error
// against the travelling wave.
//
//   test_dg_advect [vector_size=300000] [n_tests=10] [degree=-1 (all)] [graph=0]
// graph=1 replays each time step from a HIP graph of its five stages.
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
static void run_program(unsigned long vector_size_guess, unsigned n_tests, bool graph)
{
  EvaluationDGAdvection<3, degree, double> evaluator;
  evaluator.use_graph = graph;
  const unsigned n_dofs_cell = (degree + 1) * (degree + 1) * (degree + 1);
  const unsigned n_cells_tot = std::max<unsigned>(vector_size_guess / n_dofs_cell, 1U);
  unsigned n_cells[3];
  n_cells[0] = std::max(static_cast<unsigned>(1.00001 * std::cbrt((double)n_cells_tot)), 1U);
  n_cells[1] = n_cells[0];
  n_cells[2] = std::max(n_cells_tot / (n_cells[0] * n_cells[1]), 1U);
  evaluator.initialize(n_cells);
  const std::size_t global_size = evaluator.n_elements() * evaluator.dofs_per_cell;
  std::cout << "\nPolynomial degree: " << degree << "\nVector size:       " << global_size << " [" << n_cells[0]
            << " x " << n_cells[1] << " x " << n_cells[2] << "]\n";
  std::cout << "Error against analytic time derivative: " << evaluator.verify_derivative() << std::endl;
  evaluator.initialize(n_cells);
  double time = 0.;
  double best = std::numeric_limits<double>::max();
  for (int i = 0; i < 5; ++i) {
    const double t0 = wall();
    for (unsigned t = 0; t < n_tests; ++t) evaluator.do_time_step(time);
    best = std::min(best, wall() - t0);
  }
  const std::size_t p1 = degree + 1;
  // operation count of test_dg_advect.cc:168-173 (per stage: one operator)
  const std::size_t ops_interpolate = 2 * ((degree + 1) / 2) * 2 + degree + 1 + 2 * ((degree - 1) * (degree + 1) / 2);
  const std::size_t ops_cell = 6 * 3 * ops_interpolate * p1 * p1 + 3 * 4 * ops_interpolate * p1 +
                               3 * 2 * 7 * p1 * p1 + 5 * p1 * p1 * p1;
  std::cout << "RK degree " << std::setw(2) << degree << (graph ? " (graph)" : "") << "  " << n_cells[0] << " x " << n_cells[1] << " x "
            << n_cells[2] << ", MDoFs/s: " << 5. * global_size * n_tests / best / 1e6
            << ", GB/s: " << 5. * global_size * 8 * 3 * n_tests / best * 1e-9
            << " GFLOP/s: " << 5. * evaluator.n_elements() * ops_cell * n_tests / best * 1e-9 << std::endl;
  std::cout << "Error against analytic solution at t=" << time << ": " << evaluator.compute_max_error(time)
            << std::endl;
  best = std::numeric_limits<double>::max();
  for (int i = 0; i < 5; ++i) {
    const double t0 = wall();
    for (unsigned t = 0; t < n_tests; ++t) evaluator.do_matvec();
    best = std::min(best, wall() - t0);
  }
  std::cout << "MV degree " << std::setw(2) << degree << "  " << n_cells[0] << " x " << n_cells[1] << " x "
            << n_cells[2] << ", MDoFs/s: " << 1. * global_size * n_tests / best / 1e6
            << ", GB/s: " << 2. * global_size * 8 * n_tests / best * 1e-9
            << " GFLOP/s: " << 1. * evaluator.n_elements() * ops_cell * n_tests / best * 1e-9 << std::endl;
}

// runs f(std::integral_constant<int, degree>) for a run-time degree in 1..MAX
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
  const unsigned long vector_size = argc > 1 ? std::atol(argv[1]) : 300000;
  const unsigned n_tests = argc > 2 ? std::atoi(argv[2]) : 10;
  const int only = argc > 3 ? std::atoi(argv[3]) : -1;
  const bool graph = argc > 4 && std::atoi(argv[4]) != 0;
  std::cout << sfk_version() << std::endl;
  try {
    // degrees 1..14, as legacy/src/test_dg_advect.cc:18-19
    for (int degree = 1; degree <= 14; ++degree)
      if (only == -1 || only == degree)
        for_degree<14>(degree, [&](auto d) { run_program<decltype(d)::value>(vector_size, n_tests, graph); });
  } catch (const sfk::Error &e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 2;
  }
  return 0;
}

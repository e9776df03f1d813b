// Benchmark driver of the precomputed-matrix DG Laplacian with the sizing,
// timing and report line of the reference's legacy/src/test_dg_precomputed.cc
// (run_program :23-141), running the PrecomputedDGLaplacian mirror class on
// the GPU.
//
//   test_dg_precomputed [degree=3] [vector_size=10000000] [n_tests=100] [ref_coupling=0]
//
// Timings: best of 3 repeats of n_tests calls (as the reference), each call
// synchronised.  The verification line prints max |A 1| / max |A x|.
#include <sys/time.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <limits>

#include "sfk_evaluators.h"

static double wall()
{
  struct timeval t;
  gettimeofday(&t, nullptr);
  return t.tv_sec + 1e-6 * t.tv_usec;
}

template <int degree>
static void run_program(unsigned long vector_size_guess, unsigned n_tests, bool ref_coupling)
{
  const int dim = 3;
  PrecomputedDGLaplacian<3, degree, double> evaluator;
  evaluator.ref_coupling = ref_coupling;
  const unsigned n_dofs_cell = (degree + 1) * (degree + 1) * (degree + 1);
  const unsigned n_cells_tot = std::max<unsigned>(vector_size_guess / n_dofs_cell, 1U);
  // sizing of test_dg_precomputed.cc:38-61 with one cell per lane
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

  evaluator.matrix_vector_product();
  double best = std::numeric_limits<double>::max();
  for (int i = 0; i < 3; ++i) {
    const double t0 = wall();
    for (unsigned t = 0; t < n_tests; ++t) evaluator.matrix_vector_product();
    best = std::min(best, wall() - t0);
  }
  // byte and operation models of test_dg_precomputed.cc:103-109
  const double mem = double(global_size) * sizeof(double) * 2 * n_tests;
  const std::size_t p1 = degree + 1;
  const double ops = double(global_size / evaluator.dofs_per_cell) *
                     double((1 << dim) * (2 * ((degree + 1) / 2) * 2 + degree + 1 + 2 * ((degree - 1) * (degree + 1) / 2)) *
                                p1 * p1 +
                            2 * dim * 6 * p1 * p1) *
                     n_tests;
  std::cout << "Degree " << std::setw(2) << degree << "  " << n_cells[0] << " x " << n_cells[1] << " x " << n_cells[2]
            << " elem " << evaluator.dofs_per_cell << ", block sizes: " << evaluator.blx << " x " << evaluator.bly
            << " x " << evaluator.blz << ", MDoFs/s: " << global_size * n_tests / best / 1e6
            << ", GB/s: " << mem / best * 1e-9 << " GFLOP/s: " << ops / best * 1e-9 << std::endl;
  std::cout << "Verification (max |A 1| / max |A x|): " << evaluator.do_verification() << std::endl;
}

int main(int argc, char **argv)
{
  int degree = argc > 1 ? std::atoi(argv[1]) : 3;
  const unsigned long vector_size = argc > 2 ? std::atol(argv[2]) : 10000000;
  const unsigned n_tests = argc > 3 ? std::atoi(argv[3]) : 100;
  const bool ref_coupling = argc > 4 ? std::atoi(argv[4]) != 0 : false;
  std::cout << sfk_version() << std::endl;
  std::cout << "Number of threads: 1 GPU\nSIMD width:        1" << std::endl;
  try {
    switch (degree) {
    case 1: run_program<1>(vector_size, n_tests, ref_coupling); break;
    case 2: run_program<2>(vector_size, n_tests, ref_coupling); break;
    case 3: run_program<3>(vector_size, n_tests, ref_coupling); break;
    case 4: run_program<4>(vector_size, n_tests, ref_coupling); break;
    case 5: run_program<5>(vector_size, n_tests, ref_coupling); break;
    case 6: run_program<6>(vector_size, n_tests, ref_coupling); break;
    case 7: run_program<7>(vector_size, n_tests, ref_coupling); break;
    case 8: run_program<8>(vector_size, n_tests, ref_coupling); break;
    default: std::cerr << "degree must be in 1..8\n"; return 1;
    }
  } catch (const sfk::Error &e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 2;
  }
  return 0;
}

// Threaded timing entry point around the REFERENCE cell Laplacian
// (EvaluationCellLaplacian, legacy/include/evaluation_cell_laplacian.h),
// compiled with READ_VECTOR so that every matrix_vector_product streams its
// cells' input and output vectors (the memory-bound form of the reference's
// own driver, legacy/src/test_program.cc:9,47-52,101-150): one evaluator per
// OpenMP thread, each holding its share of the cells.  TEST INFRASTRUCTURE
// ONLY: a "reference" CPU baseline of bench.py.  Never linked into
// libsfk.so.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <vector>
#include <omp.h>
#include <mpi.h>

#include "evaluation_cell_laplacian.h"

// seconds of the slowest thread's best loop of n_tests products, each thread
// with n_batches cell batches (lanes cells each)
template <int p>
static double time_cell(unsigned n_batches, int n_tests, int repeats)
{
  const int nt = omp_get_max_threads();
  std::vector<double> best(nt, 1e300);
#pragma omp parallel
  {
    auto e = std::make_unique<EvaluationCellLaplacian<3, p, double>>();
    e->initialize(n_batches, true);
    e->matrix_vector_product();
    for (int r = 0; r < repeats; ++r) {
#pragma omp barrier
      const auto t0 = std::chrono::steady_clock::now();
      for (int t = 0; t < n_tests; ++t) e->matrix_vector_product();
      const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      best[omp_get_thread_num()] = std::min(best[omp_get_thread_num()], dt);
    }
  }
  return *std::max_element(best.begin(), best.end());
}

extern "C" {

__attribute__((visibility("default"))) int refcell_n_lanes() { return VectorizedArray<double>::n_array_elements; }

// cells per thread = n_batches * lanes; threads > 0 sets the thread count
__attribute__((visibility("default"))) int refbench_cell_laplacian(int degree, unsigned n_batches, int n_tests,
                                                                   int repeats, int threads, double *seconds)
{
  if (threads > 0) omp_set_num_threads(threads);
  switch (degree) {
  case 4: *seconds = time_cell<4>(n_batches, n_tests, repeats); return 0;
  case 6: *seconds = time_cell<6>(n_batches, n_tests, repeats); return 0;
  case 8: *seconds = time_cell<8>(n_batches, n_tests, repeats); return 0;
  default: return -1;
  }
}

} // extern "C"

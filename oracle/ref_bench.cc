// Threaded timing entry points around the REFERENCE DG-SIPG and DG
// advection evaluators, compiled with OpenMP from the reference headers where
// they lie (see oracle/Makefile).  Uses the evaluators' public interface only
// (blx/bly/blz, initialize, do_matvec, do_chebyshev), so the reference's
// OpenMP loops run unchanged.  TEST INFRASTRUCTURE ONLY: the "reference" CPU
// baselines of bench.py.  Never linked into libsfk.so.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <type_traits>
#include <vector>
#include <omp.h>
#include <mpi.h>

#include "evaluation_dg_laplacian.h"
#include "evaluation_dg_advect.h"

template <int p>
static double time_dg(const unsigned *n_cells, int n_tests, int repeats, int chebyshev)
{
  auto e = std::make_unique<EvaluationDGLaplacian<3, p, double>>();
  // block sizes of legacy/src/test_dg_integrate.cc:64-66
  e->blx = std::max(20 / (p + 1), 3);
  e->bly = std::max(2, 20 / (p + 1));
  e->blz = std::max(2, 20 / (p + 1));
  unsigned nc[3] = {n_cells[0], n_cells[1], n_cells[2]};
  e->initialize(nc);
  if (chebyshev) e->do_chebyshev(); else e->do_matvec();
  // best of `repeats` timed loops of n_tests calls, legacy/src/test_dg_integrate.cc:98-125
  double best = 1e300;
  for (int r = 0; r < repeats; ++r) {
    const auto t0 = std::chrono::steady_clock::now();
    for (int t = 0; t < n_tests; ++t) {
      if (chebyshev) e->do_chebyshev(); else e->do_matvec();
    }
    best = std::min(best, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  }
  return best;
}

// DG advection do_matvec (legacy/src/test_dg_advect.cc:51-53 block sizes)
template <int p>
static double time_advect(const unsigned *n_cells, int n_tests, int repeats)
{
  auto e = std::make_unique<EvaluationDGAdvection<3, p, double>>();
  e->blx = std::max(20 / (p + 1), 3);
  e->bly = std::max(2, 20 / (p + 1));
  e->blz = std::max(2, 20 / (p + 1));
  unsigned nc[3] = {n_cells[0], n_cells[1], n_cells[2]};
  e->initialize(nc);
  e->do_matvec();
  double best = 1e300;
  for (int r = 0; r < repeats; ++r) {
    const auto t0 = std::chrono::steady_clock::now();
    for (int t = 0; t < n_tests; ++t) e->do_matvec();
    best = std::min(best, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  }
  return best;
}

extern "C" {

__attribute__((visibility("default"))) int refbench_n_lanes()
{
  return VectorizedArray<double>::n_array_elements;
}

__attribute__((visibility("default"))) int refbench_max_threads() { return omp_get_max_threads(); }

// seconds of the best timed loop; n_cells[0] counts cells (multiple of the
// lane count).  threads > 0 sets the OpenMP thread count first.
__attribute__((visibility("default"))) int refbench_dg_laplacian(int degree, const unsigned *n_cells,
                                                                 int n_tests, int repeats, int threads,
                                                                 int chebyshev, double *seconds)
{
  if (threads > 0) omp_set_num_threads(threads);
  switch (degree) {
  case 3: *seconds = time_dg<3>(n_cells, n_tests, repeats, chebyshev); return 0;
  case 4: *seconds = time_dg<4>(n_cells, n_tests, repeats, chebyshev); return 0;
  case 5: *seconds = time_dg<5>(n_cells, n_tests, repeats, chebyshev); return 0;
  case 6: *seconds = time_dg<6>(n_cells, n_tests, repeats, chebyshev); return 0;
  case 8: *seconds = time_dg<8>(n_cells, n_tests, repeats, chebyshev); return 0;
  default: return -1;
  }
}

// DG advection matvec; n_cells[0] a multiple of the lane count
__attribute__((visibility("default"))) int refbench_dg_advect(int degree, const unsigned *n_cells, int n_tests,
                                                              int repeats, int threads, double *seconds)
{
  if (threads > 0) omp_set_num_threads(threads);
  switch (degree) {
  case 3: *seconds = time_advect<3>(n_cells, n_tests, repeats); return 0;
  case 5: *seconds = time_advect<5>(n_cells, n_tests, repeats); return 0;
  default: return -1;
  }
}

} // extern "C"

// Internal (host-side) declarations shared by the C ABI, the tables and the
// kernel launchers.  Not part of the public interface (include/sfk.h).
#pragma once

#include <string>
#include <vector>

#include "sfk.h"

namespace sfk {

// degrees 1..15: (p+1)^2 <= 256 threads per cell.  The tensor-product
// continuous kernel and the precomputed DG Laplacian stop at kMaxDegreeTp.
constexpr int kMaxDegree = 15;
constexpr int kMaxDegreeTp = 8;

struct Tables1D {
  int degree = 0, n = 0, basis = 0;
  std::vector<double> gauss, weights, gll;
  std::vector<double> S, D, Sinv;       // dense, [i*n+q]
  std::vector<double> fv0, fv1, fd0, fd1; // Gauss-Lagrange value/derivative at 0 and 1
  double hermite_d = 0.;
};

// 1D matrices of the precomputed DG Laplacian on the unit interval
// (legacy/include/precomputed_dg_laplacian.h:305-349): Hermite-like basis,
// Gauss quadrature, penalty p(p+1).  mass, lap: dense [i*n+j]; v1, v2:
// coupling of rows 0 and 1 to the left neighbour's dofs n-1 and n-2.
struct Precomp1D {
  std::vector<double> mass, lap;
  double v1 = 0., v2 = 0.;
};
// ref_coupling: the reference's coupling expressions (:331-340), which
// vanish; otherwise the SIPG face coupling.  False if the basis does not
// have the coupling pattern the kernels assume.
bool build_precomp_1d(int degree, bool ref_coupling, Precomp1D &pc);

void gauss_rule(int n, std::vector<double> &pts, std::vector<double> &wts);
std::vector<double> gll_points(int n);
bool build_tables(int degree, int basis, Tables1D &t);

void set_error(const std::string &msg);

} // namespace sfk

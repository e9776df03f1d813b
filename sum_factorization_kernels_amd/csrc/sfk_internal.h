// Internal (host-side) declarations shared by the C ABI, the tables and the
// kernel launchers.  Not part of the public interface (include/sfk.h).
#pragma once

#include <string>
#include <vector>

#include "sfk.h"

namespace sfk {

constexpr int kMaxDegree = 8;

struct Tables1D {
  int degree = 0, n = 0, basis = 0;
  std::vector<double> gauss, weights, gll;
  std::vector<double> S, D, Sinv;       // dense, [i*n+q]
  std::vector<double> fv0, fv1, fd0, fd1; // Gauss-Lagrange value/derivative at 0 and 1
  double hermite_d = 0.;
};

void gauss_rule(int n, std::vector<double> &pts, std::vector<double> &wts);
std::vector<double> gll_points(int n);
bool build_tables(int degree, int basis, Tables1D &t);

void set_error(const std::string &msg);

} // namespace sfk

// Host-side 1D tables: Gauss / Gauss-Lobatto quadrature, the Lagrange and
// Hermite-like bases, and the dense 1D matrices the kernels contract with.
// The algorithms (Newton on the Legendre polynomial, Newton with deflation
// on Jacobi polynomials, the Hermite-like basis with its secant-placed
// support point) are those of the reference's
// legacy/include/gauss_formula.h:35-175 and
// legacy/include/lagrange_polynomials.h:31-274, because the operators must
// produce the reference's numbers; the code is this library's own.

#include "sfk_internal.h"

#include <cfloat>
#include <cmath>
#include <vector>

namespace sfk {

namespace {
constexpr double kPi = 3.14159265358979323846;

// Newton iteration for the i-th (1-based) root of P_n, returning the root
// and weight mapped to [0,1] (gauss_formula.h:35-64)
void legendre_newton(int n, int i, double &half_root, double &weight)
{
  const double tol = 5. * DBL_EPSILON;
  double z = std::cos(kPi * (i - .25) / (n + .5));
  double pn, dp;
  do {
    double pkm1 = 0., pk = 1.;
    for (int k = 0; k < n; ++k) {
      const double pkm2 = pkm1;
      pkm1 = pk;
      pk = ((2. * k + 1.) * z * pkm1 - k * pkm2) / (k + 1);
    }
    pn = pk;
    dp = n * (z * pk - pkm1) / (z * z - 1);
    z -= pn / dp;
  } while (std::fabs(pn / dp) > tol);
  half_root = 0.5 * z;
  weight = 1. / ((1. - z * z) * dp * dp);
}

// Jacobi polynomial P_n^{(a,b)}(x) by the three-term recursion with integer
// coefficients (gauss_formula.h:97-123)
double jacobi(double x, int a, int b, int n)
{
  double pm1 = 1.0;
  if (n == 0) return pm1;
  double p = ((a + b + 2) * x + (a - b)) / 2;
  for (int k = 1; k < n; ++k) {
    const int v = 2 * k + a + b;
    const int c1 = 2 * (k + 1) * (k + a + b + 1) * v;
    const int c2 = (v + 1) * (a * a - b * b);
    const int c3 = v * (v + 1) * (v + 2);
    const int c4 = 2 * (k + a) * (k + b) * (v + 2);
    const double next = (double)((c2 + c3 * x) * p - c4 * pm1) / c1;
    pm1 = p;
    p = next;
  }
  return p;
}

// roots of P_{degree-a-b}^{(a,b)} on [-1,1], Newton with deflation and a
// long double iterate (gauss_formula.h:125-163)
std::vector<double> jacobi_roots(int degree, int a, int b)
{
  const double tol = 5. * DBL_EPSILON;
  const int m = degree - a - b;
  std::vector<double> x(m > 0 ? m : 0);
  for (int i = 0; i < m; ++i) x[i] = -std::cos((double)(2 * i + 1) / (2 * m) * kPi);
  for (int k = 0; k < m; ++k) {
    long double r = x[k];
    if (k > 0) r = (r + x[k - 1]) / 2;
    double delta;
    do {
      double s = 0.;
      for (int i = 0; i < k; ++i) s += 1. / (r - x[i]);
      const double dj = 0.5 * (a + b + m + 1) * jacobi((double)r, a + 1, b + 1, m - 1);
      const double f = jacobi((double)r, a, b, m);
      delta = f / (f * s - dj);
      r += delta;
    } while (std::fabs(delta) >= tol);
    x[k] = (double)r;
  }
  return x;
}

struct Lagrange {
  std::vector<double> pts, inv_den;
  explicit Lagrange(const std::vector<double> &p) : pts(p), inv_den(p.size())
  {
    for (size_t i = 0; i < p.size(); ++i) {
      double den = 1.;
      for (size_t j = 0; j < p.size(); ++j)
        if (j != i) den *= p[i] - p[j];
      inv_den[i] = 1. / den;
    }
  }
  double value(int i, double x) const
  {
    double v = inv_den[i];
    for (size_t k = 0; k < pts.size(); ++k)
      if ((int)k != i) v *= x - pts[k];
    return v;
  }
  double derivative(int i, double x) const
  {
    double v = inv_den[i], d = 0.;
    for (size_t k = 0; k < pts.size(); ++k)
      if ((int)k != i) {
        const double t = x - pts[k];
        d = d * t + v;
        v *= t;
      }
    return d;
  }
};

// Hermite-like basis (lagrange_polynomials.h:86-274): phi_0, phi_1 carry
// value and derivative at x=0 (phi_0'(0) = -phi_1'(0)), phi_{p-1}, phi_p at
// x=1, the others are nodal at the interior Jacobi(2,2) roots.
struct HermiteLike {
  int size = 0;
  std::vector<double> nodes, scale;
  double xstar = 0.;

  explicit HermiteLike(int degree) : size(degree + 1), nodes(degree + 1, 0.), scale(degree + 1, 1.)
  {
    if (degree < 3) return;
    std::vector<double> r = jacobi_roots(degree + 1, 2, 2);
    for (int i = 0; i < degree - 3; ++i) r[i] = 0.5 * r[i] + 0.5;
    nodes[0] = nodes[1] = 0.;
    for (int m = 0; m < degree - 3; ++m) nodes[m + 2] = r[m];
    nodes[degree - 1] = nodes[degree] = 1.;
    // one secant step on the integral of phi_0*phi_1, which is linear in
    // the support point x* (lagrange_polynomials.h:89-123)
    std::vector<double> gp, gw;
    gauss_rule(degree + 1, gp, gw);
    const double gl = 0., gr = 0.5;
    double il = 0., ir = 0.;
    for (int q = 0; q <= degree; ++q) {
      const double x = gp[q];
      double common = x;
      for (int j = 0; j < degree - 3; ++j) common *= (x - r[j]) * (x - r[j]);
      common *= (x - 1.) * (x - 1.) * (x - 1.) * (x - 1.);
      il += gw[q] * (common * (x - gl));
      ir += gw[q] * (common * (x - gr));
    }
    xstar = gr - (gr - gl) / (ir - il) * ir;
    for (int m = 2; m < degree - 1; ++m) scale[m] = 1. / value(m, nodes[m]);
    scale[0] = 1. / value(0, 0.);
    scale[degree] = 1. / value(degree, 1.);
    scale[1] = -derivative(0, 0.) / derivative(1, 0.);
    scale[degree - 1] = -(derivative(degree, 1.) / derivative(degree - 1, 1.));
  }

  double value(int i, double x) const
  {
    if (size == 1) return 1.;
    if (size == 2) return i == 0 ? 1. - x : x;
    if (size == 3) return i == 0 ? (1. - x) * (1. - x) : (i == 1 ? 2. * x * (1. - x) : x * x);
    double v = scale[i];
    if (i == 0) {
      v *= (x - xstar);
      for (int k = 2; k < size; ++k) v *= x - nodes[k];
    } else if (i == size - 1) {
      v *= (x - 1. + xstar);
      for (int k = 0; k < size - 2; ++k) v *= x - nodes[k];
    } else {
      for (int k = 0; k < size; ++k)
        if (k != i) v *= x - nodes[k];
    }
    return v;
  }

  double derivative(int i, double x) const
  {
    if (size == 1) return 0.;
    if (size == 2) return i == 0 ? -1. : 1.;
    if (size == 3) return i == 0 ? -2. * (1. - x) : (i == 1 ? 2. - 4. * x : 2. * x);
    double v = scale[i], d = 0.;
    auto step = [&](double t) {
      d = d * t + v;
      v *= t;
    };
    if (i == 0) {
      d = v;
      v *= (x - xstar);
      for (int k = 2; k < size; ++k) step(x - nodes[k]);
    } else if (i == size - 1) {
      d = v;
      v *= (x - 1. + xstar);
      for (int k = 0; k < size - 2; ++k) step(x - nodes[k]);
    } else {
      for (int k = 0; k < size; ++k)
        if (k != i) step(x - nodes[k]);
    }
    return d;
  }
};
} // namespace

void gauss_rule(int n, std::vector<double> &pts, std::vector<double> &wts)
{
  // points: every point is written twice in the reference loop
  // (gauss_formula.h:66-79) and the later assignment wins; keep that order
  pts.assign(n, 0.);
  wts.assign(n, 0.);
  for (int i = 1; i <= n; ++i) {
    double x, w;
    legendre_newton(n, i, x, w);
    pts[i - 1] = .5 - x;
    pts[n - i] = .5 + x;
  }
  for (int i = 1; i <= (n + 1) / 2; ++i) {
    double x, w;
    legendre_newton(n, i, x, w);
    wts[i - 1] = w;
    wts[n - i] = w;
  }
}

std::vector<double> gll_points(int n)
{
  std::vector<double> r = jacobi_roots(n, 1, 1);
  std::vector<double> p;
  p.push_back(-1.);
  p.insert(p.end(), r.begin(), r.end());
  p.push_back(1.);
  for (double &x : p) x = 0.5 + 0.5 * x;
  return p;
}

bool build_precomp_1d(int degree, bool ref_coupling, Precomp1D &pc)
{
  if (degree < 1 || degree > kMaxDegree) return false;
  const int n = degree + 1;
  std::vector<double> x, w;
  gauss_rule(n, x, w);
  HermiteLike hb(degree);
  std::vector<double> val(n * n), der(n * n), v0(n), v1(n), d0(n), d1(n);
  for (int i = 0; i < n; ++i) {
    for (int q = 0; q < n; ++q) {
      val[i * n + q] = hb.value(i, x[q]);
      der[i * n + q] = hb.derivative(i, x[q]);
    }
    v0[i] = hb.value(i, 0.);
    v1[i] = hb.value(i, 1.);
    d0[i] = hb.derivative(i, 0.);
    d1[i] = hb.derivative(i, 1.);
  }
  const double sigma = double(degree) * n;
  pc.mass.assign(n * n, 0.);
  pc.lap.assign(n * n, 0.);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double sm = 0., sl = 0.;
      for (int q = 0; q < n; ++q) {
        sm += val[i * n + q] * val[j * n + q] * w[q];
        sl += der[i * n + q] * der[j * n + q] * w[q];
      }
      // the cell's own traces on its faces x = 0 (outer normal -1) and x = 1
      sl += sigma * v0[i] * v0[j] + 0.5 * d0[i] * v0[j] + 0.5 * d0[j] * v0[i];
      sl += sigma * v1[i] * v1[j] - 0.5 * d1[i] * v1[j] - 0.5 * d1[j] * v1[i];
      pc.mass[i * n + j] = sm;
      pc.lap[i * n + j] = sl;
    }
  // C(i, j): test phi_i of a cell at its x = 0 against trial phi_j of its
  // left neighbour at x = 1: 1/2 phi_j'(1) phi_i(0) - 1/2 phi_j(1) phi_i'(0)
  // - sigma phi_i(0) phi_j(1).  Only rows 0, 1 and columns n-2, n-1 can be
  // non-zero, with C(0, n-2) = C(1, n-1) and C(1, n-2) = 0 (mirror
  // symmetry); the kernels rely on that pattern.
  auto C = [&](int i, int j) { return 0.5 * d1[j] * v0[i] - 0.5 * v1[j] * d0[i] - sigma * v0[i] * v1[j]; };
  const int a = n - 1, b = n - 2;
  const double s1 = std::fabs(C(0, a));
  if (std::fabs(C(0, b) - C(1, a)) > 1e-12 * s1 || (n > 2 && std::fabs(C(1, b)) > 1e-12 * s1)) return false;
  if (ref_coupling) {
    pc.v1 = -v1[0] * v0[a] * sigma + 0.5 * d1[0] * v0[a] - 0.5 * d0[a] * v1[0];
    pc.v2 = -v1[0] * v0[b] * sigma + 0.5 * d1[0] * v0[b] - 0.5 * d0[b] * v1[0];
  } else {
    pc.v1 = C(0, a);
    pc.v2 = C(0, b);
  }
  return true;
}

bool build_tables(int degree, int basis, Tables1D &t)
{
  if (degree < 1 || degree > kMaxDegree) return false;
  const int n = degree + 1;
  t.degree = degree;
  t.n = n;
  t.basis = basis;
  gauss_rule(n, t.gauss, t.weights);
  t.gll = gll_points(n);
  Lagrange lg(t.gauss), ll(t.gll);
  HermiteLike hb(degree);
  t.S.assign(n * n, 0.);
  t.D.assign(n * n, 0.);
  t.Sinv.assign(n * n, 0.);
  for (int i = 0; i < n; ++i)
    for (int q = 0; q < n; ++q) {
      t.S[i * n + q] = basis == 0 ? hb.value(i, t.gauss[q]) : ll.value(i, t.gauss[q]);
      t.D[i * n + q] = lg.derivative(i, t.gauss[q]);
      t.Sinv[i * n + q] = lg.value(i, t.gll[q]);
    }
  t.fv0.resize(n);
  t.fv1.resize(n);
  t.fd0.resize(n);
  t.fd1.resize(n);
  for (int i = 0; i < n; ++i) {
    t.fv0[i] = lg.value(i, 0.);
    t.fv1[i] = lg.value(i, 1.);
    t.fd0[i] = lg.derivative(i, 0.);
    t.fd1[i] = lg.derivative(i, 1.);
  }
  t.hermite_d = hb.derivative(0, 0.);
  return true;
}

} // namespace sfk

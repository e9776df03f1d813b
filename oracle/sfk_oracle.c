/*
 * CPU oracle (test infrastructure only -- never linked into libsfk.so).
 * See sfk_oracle.h for the layout and the list of entry points; every
 * function below cites the reference lines it restates.
 */
#include "sfk_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXN SFKO_MAXN
static const double SFKO_PI = 3.14159265358979323846;

/* ------------------------------------------------------------------ */
/* quadrature: gauss_formula.h:35-94                                    */
/* ------------------------------------------------------------------ */
static void gauss_inner(int n_points, int i, double *x_out, double *w_out)
{
  /* gauss_formula.h:35-64: Newton iteration on the Legendre polynomial */
  const double tolerance = DBL_EPSILON * 5.;
  double z = cos(SFKO_PI * (i - .25) / (n_points + .5));
  double pp, p1;
  do {
    p1 = 1.;
    double p2 = 0.;
    for (int j = 0; j < n_points; ++j) {
      const double p3 = p2;
      p2 = p1;
      p1 = ((2. * j + 1.) * z * p2 - j * p3) / (j + 1);
    }
    pp = n_points * (z * p1 - p2) / (z * z - 1);
    z = z - p1 / pp;
  } while (fabs(p1 / pp) > tolerance);
  *x_out = 0.5 * z;
  *w_out = 1. / ((1. - z * z) * pp * pp);
}

void sfko_gauss_points(int n, double *pts)
{
  /* gauss_formula.h:66-79 (every point is written twice; the later
   * assignment wins, exactly as in the reference loop) */
  for (int i = 1; i <= n; ++i) {
    double x, w;
    gauss_inner(n, i, &x, &w);
    pts[i - 1] = .5 - x;
    pts[n - i] = .5 + x;
  }
}

void sfko_gauss_weights(int n, double *wts)
{
  /* gauss_formula.h:81-94 */
  const int m = (n + 1) / 2;
  for (int i = 1; i <= m; ++i) {
    double x, w;
    gauss_inner(n, i, &x, &w);
    wts[i - 1] = w;
    wts[n - i] = w;
  }
}

double sfko_jacobi_polynomial(double x, int alpha, int beta, int n)
{
  /* gauss_formula.h:97-123: three-term recursion with integer coefficients */
  double p[64];
  p[0] = 1.0;
  if (n == 0) return p[0];
  p[1] = ((alpha + beta + 2) * x + (alpha - beta)) / 2;
  if (n == 1) return p[1];
  for (int i = 1; i <= n - 1; ++i) {
    const int v = 2 * i + alpha + beta;
    const int a1 = 2 * (i + 1) * (i + alpha + beta + 1) * v;
    const int a2 = (v + 1) * (alpha * alpha - beta * beta);
    const int a3 = v * (v + 1) * (v + 2);
    const int a4 = 2 * (i + alpha) * (i + beta) * (v + 2);
    p[i + 1] = (double)((a2 + a3 * x) * p[i] - a4 * p[i - 1]) / a1;
  }
  return p[n];
}

void sfko_zeros_of_jacobi_polynomial(int degree, int alpha, int beta, double *x)
{
  /* gauss_formula.h:125-163: Newton with deflation, iterate kept in long
   * double as in the reference */
  const double tolerance = DBL_EPSILON * 5.;
  const int m = degree - alpha - beta;
  for (int i = 0; i < m; ++i) x[i] = -cos((double)(2 * i + 1) / (2 * m) * SFKO_PI);
  double s, J_x, f, delta;
  for (int k = 0; k < m; ++k) {
    long double r = x[k];
    if (k > 0) r = (r + x[k - 1]) / 2;
    do {
      s = 0.;
      for (int i = 0; i < k; ++i) s += 1. / (r - x[i]);
      J_x = 0.5 * (alpha + beta + m + 1) * sfko_jacobi_polynomial((double)r, alpha + 1, beta + 1, m - 1);
      f = sfko_jacobi_polynomial((double)r, alpha, beta, m);
      delta = f / (f * s - J_x);
      r += delta;
    } while (fabs(delta) >= tolerance);
    x[k] = (double)r;
  }
}

void sfko_gll_points(int n, double *pts)
{
  /* gauss_formula.h:165-175: {-1} + Jacobi(1,1) roots + {+1}, mapped */
  double roots[MAXN];
  sfko_zeros_of_jacobi_polynomial(n, 1, 1, roots);
  pts[0] = -1.;
  for (int i = 0; i < n - 2; ++i) pts[i + 1] = roots[i];
  pts[n - 1] = 1.;
  for (int i = 0; i < n; ++i) pts[i] = 0.5 + 0.5 * pts[i];
}

/* ------------------------------------------------------------------ */
/* Lagrange basis: lagrange_polynomials.h:31-82                          */
/* ------------------------------------------------------------------ */
static double lagrange_denominator(int npts, const double *pts, int i)
{
  double denominator = 1.;
  for (int j = 0; j < npts; ++j)
    if (j != i) denominator *= pts[i] - pts[j];
  return 1. / denominator;
}

double sfko_lagrange_value(int npts, const double *pts, int i, double x)
{
  double value = lagrange_denominator(npts, pts, i);
  for (int k = 0; k < npts; ++k)
    if (k != i) value *= x - pts[k];
  return value;
}

double sfko_lagrange_derivative(int npts, const double *pts, int i, double x)
{
  double value = lagrange_denominator(npts, pts, i);
  double derivative = 0;
  for (int k = 0; k < npts; ++k)
    if (k != i) {
      const double v = x - pts[k];
      derivative = derivative * v + value;
      value *= v;
    }
  return derivative;
}

/* ------------------------------------------------------------------ */
/* Hermite-like basis: lagrange_polynomials.h:86-274                     */
/* ------------------------------------------------------------------ */
typedef struct {
  int size;
  double points[MAXN];
  double denom[MAXN];
  double zero_one;
} hermite_basis;

static double hermite_raw_value(const hermite_basis *b, int idx, double x)
{
  /* lagrange_polynomials.h:163-207 */
  if (b->size == 1) return 1.;
  if (b->size == 2) return idx == 0 ? 1. - x : x;
  if (b->size == 3) {
    if (idx == 0) return (1. - x) * (1. - x);
    if (idx == 1) return 2. * x * (1. - x);
    return x * x;
  }
  double value = b->denom[idx];
  if (idx == 0) {
    value *= (x - b->zero_one);
    for (int i = 2; i < b->size; ++i) value *= x - b->points[i];
  } else if (idx == b->size - 1) {
    value *= (x - 1. + b->zero_one);
    for (int i = 0; i < b->size - 2; ++i) value *= x - b->points[i];
  } else {
    for (int i = 0; i < b->size; ++i)
      if (i != idx) value *= x - b->points[i];
  }
  return value;
}

static double hermite_raw_derivative(const hermite_basis *b, int idx, double x)
{
  /* lagrange_polynomials.h:209-268 */
  if (b->size == 1) return 0;
  if (b->size == 2) return idx == 0 ? -1. : 1.;
  if (b->size == 3) {
    if (idx == 0) return -2. * (1. - x);
    if (idx == 1) return 2. - 4. * x;
    return 2. * x;
  }
  double value = b->denom[idx];
  double derivative = 0;
  if (idx == 0) {
    derivative = value;
    value *= (x - b->zero_one);
    for (int i = 2; i < b->size; ++i) {
      const double v = x - b->points[i];
      derivative = derivative * v + value;
      value *= v;
    }
  } else if (idx == b->size - 1) {
    derivative = value;
    value *= (x - 1. + b->zero_one);
    for (int i = 0; i < b->size - 2; ++i) {
      const double v = x - b->points[i];
      derivative = derivative * v + value;
      value *= v;
    }
  } else {
    for (int i = 0; i < b->size; ++i)
      if (i != idx) {
        const double v = x - b->points[i];
        derivative = derivative * v + value;
        value *= v;
      }
  }
  return derivative;
}

static double find_support_point_x_star(int degree, const double *jacobi_roots)
{
  /* lagrange_polynomials.h:89-123: one secant step on the (linear in x*)
   * integral of l_0*l_1 with a (degree+1)-point Gauss rule */
  double guess_left = 0, guess_right = 0.5;
  double points[MAXN + 1], weights[MAXN + 1];
  sfko_gauss_points(degree + 1, points);
  sfko_gauss_weights(degree + 1, weights);
  double integral_left = 0, integral_right = 0;
  for (int q = 0; q < degree + 1; ++q) {
    const double x = points[q];
    double poly_val_common = x;
    for (int j = 0; j < degree - 3; ++j)
      poly_val_common *= (x - jacobi_roots[j]) * (x - jacobi_roots[j]);
    poly_val_common *= (x - 1.) * (x - 1.) * (x - 1.) * (x - 1.);
    integral_left += weights[q] * (poly_val_common * (x - guess_left));
    integral_right += weights[q] * (poly_val_common * (x - guess_right));
  }
  return guess_right - (guess_right - guess_left) / (integral_right - integral_left) * integral_right;
}

static void hermite_init(hermite_basis *b, int degree)
{
  /* lagrange_polynomials.h:126-156 */
  memset(b, 0, sizeof(*b));
  b->size = degree + 1;
  if (degree >= 3) {
    double jacobi_roots[MAXN];
    sfko_zeros_of_jacobi_polynomial(degree + 1, 2, 2, jacobi_roots);
    for (int i = 0; i < degree - 3; ++i) jacobi_roots[i] = 0.5 * jacobi_roots[i] + 0.5;
    b->points[0] = 0.;
    b->points[1] = 0.;
    for (int m = 0; m < degree - 3; m++) b->points[m + 2] = jacobi_roots[m];
    b->points[degree - 1] = 1.;
    b->points[degree] = 1.;
    b->zero_one = find_support_point_x_star(degree, jacobi_roots);
    for (int m = 0; m <= degree; ++m) b->denom[m] = 1.;
    for (int m = 2; m < degree - 1; ++m) b->denom[m] = 1. / hermite_raw_value(b, m, b->points[m]);
    b->denom[0] = 1. / hermite_raw_value(b, 0, 0.);
    b->denom[degree] = 1. / hermite_raw_value(b, degree, 1.);
    b->denom[1] = -hermite_raw_derivative(b, 0, 0.) / hermite_raw_derivative(b, 1, 0.);
    b->denom[degree - 1] = -(hermite_raw_derivative(b, degree, 1.) / hermite_raw_derivative(b, degree - 1, 1.));
  }
}

double sfko_hermite_value(int degree, int i, double x)
{
  hermite_basis b;
  hermite_init(&b, degree);
  return hermite_raw_value(&b, i, x);
}

double sfko_hermite_derivative(int degree, int i, double x)
{
  hermite_basis b;
  hermite_init(&b, degree);
  return hermite_raw_derivative(&b, i, x);
}

/* ------------------------------------------------------------------ */
/* 1D tables: evaluation_dg_laplacian.h:615-695,                         */
/* evaluation_dg_advect.h:729-825, evaluation_cell_laplacian.h:815-886   */
/* ------------------------------------------------------------------ */
int sfko_fill_tables(int degree, int basis, sfko_tables *t)
{
  if (degree < 1 || degree + 1 > MAXN || (basis != 0 && basis != 1)) return -1;
  memset(t, 0, sizeof(*t));
  const int n = degree + 1, stride = (n + 1) / 2;
  t->degree = degree;
  t->n = n;
  t->basis = basis;
  sfko_gauss_points(n, t->pts);
  sfko_gauss_weights(n, t->w);
  sfko_gll_points(n, t->gll);
  hermite_basis hb;
  hermite_init(&hb, degree);
  /* basis values: Hermite-like or GLL Lagrange, evaluated at Gauss points */
#define PHI(i, x) (basis == 0 ? hermite_raw_value(&hb, (i), (x)) : sfko_lagrange_value(n, t->gll, (i), (x)))
  for (int i = 0; i < n; ++i)
    for (int q = 0; q < n; ++q) {
      t->S[i * n + q] = PHI(i, t->pts[q]);
      t->D[i * n + q] = sfko_lagrange_derivative(n, t->pts, i, t->pts[q]);
      t->Sinv[i * n + q] = sfko_lagrange_value(n, t->pts, i, t->gll[q]);
    }
  for (int i = 0; i < n; ++i) {
    t->fv0[i] = sfko_lagrange_value(n, t->pts, i, 0.);
    t->fv1[i] = sfko_lagrange_value(n, t->pts, i, 1.);
    t->fd0[i] = sfko_lagrange_derivative(n, t->pts, i, 0.);
    t->fd1[i] = sfko_lagrange_derivative(n, t->pts, i, 1.);
  }
  t->hd = hermite_raw_derivative(&hb, 0, 0.);
  /* even-odd forms, evaluation_dg_laplacian.h:623-659 */
  for (int i = 0; i < (degree + 1) / 2; ++i)
    for (int q = 0; q < stride; ++q) {
      const double p1 = PHI(i, t->pts[q]);
      const double p2 = PHI(i, t->pts[n - 1 - q]);
      t->values_eo[i * stride + q] = 0.5 * (p1 + p2);
      t->values_eo[(degree - i) * stride + q] = 0.5 * (p1 - p2);
      const double g1 = sfko_lagrange_derivative(n, t->pts, i, t->pts[q]);
      const double g2 = sfko_lagrange_derivative(n, t->pts, i, t->pts[n - 1 - q]);
      t->gradients_eo[i * stride + q] = 0.5 * (g1 + g2);
      t->gradients_eo[(degree - i) * stride + q] = 0.5 * (g1 - g2);
      const double s1 = sfko_lagrange_value(n, t->pts, i, t->gll[q]);
      const double s2 = sfko_lagrange_value(n, t->pts, i, t->gll[degree - q]);
      t->inv_values_eo[i * stride + q] = 0.5 * (s1 + s2);
      t->inv_values_eo[(degree - i) * stride + q] = 0.5 * (s1 - s2);
    }
  if (degree % 2 == 0)
    for (int q = 0; q < stride; ++q) {
      t->values_eo[degree / 2 * stride + q] = PHI(degree / 2, t->pts[q]);
      t->gradients_eo[degree / 2 * stride + q] = sfko_lagrange_derivative(n, t->pts, degree / 2, t->pts[q]);
      t->inv_values_eo[degree / 2 * stride + q] = sfko_lagrange_value(n, t->pts, degree / 2, t->gll[q]);
    }
  for (int i = 0; i < degree / 2 + 1; ++i) {
    const double v0 = t->fv0[i], v1 = t->fv1[i];
    t->face_eo[degree - i] = 0.5 * (v0 - v1);
    t->face_eo[i] = 0.5 * (v0 + v1);
    const double d0 = t->fd0[i], d1 = t->fd1[i];
    t->face_eo[degree + 1 + i] = 0.5 * (d0 + d1);
    t->face_eo[degree + 1 + degree - i] = 0.5 * (d0 - d1);
  }
#undef PHI
  return 0;
}

int sfko_export_eo_tables(int degree, int basis, double *values_eo, double *gradients_eo,
                          double *inv_values_eo, double *face_eo, double *hermite_d)
{
  sfko_tables t;
  if (sfko_fill_tables(degree, basis, &t)) return -1;
  const int n = degree + 1, h = (n + 1) / 2;
  memcpy(values_eo, t.values_eo, sizeof(double) * n * h);
  memcpy(gradients_eo, t.gradients_eo, sizeof(double) * n * h);
  memcpy(inv_values_eo, t.inv_values_eo, sizeof(double) * n * h);
  memcpy(face_eo, t.face_eo, sizeof(double) * 2 * n);
  *hermite_d = t.hd;
  return 0;
}

/* ------------------------------------------------------------------ */
/* 1D even-odd kernel: matrix_vector_kernel.h:30-185                     */
/* ------------------------------------------------------------------ */
void sfko_apply_1d_eo(int nn, int stride, int type, int contract_over_rows, int add_into,
                      int do_dg, const double *coefficients_eo, const double *in, double *out,
                      const double *array_for_add, const double *dg_coefficients,
                      double *array_face)
{
  const int mid = nn / 2;
  const int offset = (nn + 1) / 2;
  double xp[MAXN], xm[MAXN];
  for (int i = 0; i < mid; ++i) {
    if (contract_over_rows && type == 1) {
      xp[i] = in[stride * i] - in[stride * (nn - 1 - i)];
      xm[i] = in[stride * i] + in[stride * (nn - 1 - i)];
    } else {
      xp[i] = in[i * stride] + in[(nn - 1 - i) * stride];
      xm[i] = in[i * stride] - in[(nn - 1 - i) * stride];
    }
  }
  const double xmid = in[stride * mid];
  if (contract_over_rows) {
    for (int col = 0; col < mid; ++col) {
      double r0 = coefficients_eo[col] * xp[0];
      double r1 = coefficients_eo[(nn - 1) * offset + col] * xm[0];
      for (int ind = 1; ind < mid; ++ind) {
        r0 += coefficients_eo[ind * offset + col] * xp[ind];
        r1 += coefficients_eo[(nn - 1 - ind) * offset + col] * xm[ind];
      }
      if (nn % 2 == 1) {
        if (type == 1)
          r1 += coefficients_eo[mid * offset + col] * xmid;
        else
          r0 += coefficients_eo[mid * offset + col] * xmid;
      }
      const double t = r0;
      r0 = (add_into ? array_for_add[col * stride] + t : t) + r1;
      r1 = (add_into ? array_for_add[(nn - 1 - col) * stride] + t : t) - r1;
      out[col * stride] = r0;
      out[(nn - 1 - col) * stride] = r1;
    }
    if (nn % 2 == 1) {
      /* the reference reads array_for_add[mid] (not [mid*stride]) here,
       * matrix_vector_kernel.h:86; no in-scope caller uses add_into with
       * contract_over_rows, so the restatement keeps the same index */
      double r0 = add_into ? array_for_add[mid] + coefficients_eo[mid * offset + mid] * xmid
                           : coefficients_eo[mid * offset + mid] * xmid;
      for (int ind = 0; ind < mid; ++ind) r0 += coefficients_eo[ind * offset + mid] * xp[ind];
      out[mid * stride] = r0;
    }
    if (do_dg > 0) {
      double r0 = dg_coefficients[0] * xm[0];
      double r1 = dg_coefficients[nn - 1] * xp[0];
      for (int ind = 1; ind < mid; ++ind) {
        r0 += dg_coefficients[ind] * xm[ind];
        r1 += dg_coefficients[nn - 1 - ind] * xp[ind];
      }
      if (nn % 2 == 1) r0 += dg_coefficients[mid] * xmid;
      array_face[0] = r0 + r1;
      array_face[1] = r0 - r1;
    }
    if (do_dg > 1) {
      double r0 = dg_coefficients[nn] * xp[0];
      double r1 = dg_coefficients[nn + nn - 1] * xm[0];
      for (int ind = 1; ind < mid; ++ind) {
        r0 += dg_coefficients[nn + ind] * xp[ind];
        r1 += dg_coefficients[nn + nn - 1 - ind] * xm[ind];
      }
      if (nn % 2 == 1) r1 += dg_coefficients[nn + mid] * xmid;
      array_face[2] = r0 + r1;
      array_face[3] = r0 - r1;
    }
  } else {
    for (int col = 0; col < mid; ++col) {
      double r0 = coefficients_eo[col * offset] * xp[0];
      double r1 = coefficients_eo[(nn - 1 - col) * offset] * xm[0];
      for (int ind = 1; ind < mid; ++ind) {
        r0 += coefficients_eo[col * offset + ind] * xp[ind];
        r1 += coefficients_eo[(nn - 1 - col) * offset + ind] * xm[ind];
      }
      if (nn % 2 == 1 && type > 0) r0 += coefficients_eo[col * offset + mid] * xmid;
      if (do_dg > 0) {
        r1 += dg_coefficients[col] * array_face[0];
        r0 += dg_coefficients[nn - 1 - col] * array_face[1];
      }
      if (do_dg > 1) {
        r0 += dg_coefficients[nn + col] * array_face[2];
        r1 += dg_coefficients[nn + nn - 1 - col] * array_face[3];
      }
      const double t = r0;
      r0 = (add_into ? array_for_add[col * stride] + t : t) + r1;
      if (type != 1)
        r1 = (add_into ? array_for_add[(nn - 1 - col) * stride] + t : t) - r1;
      else
        r1 = (add_into ? array_for_add[(nn - 1 - col) * stride] + r1 : r1) - t;
      out[col * stride] = r0;
      out[(nn - 1 - col) * stride] = r1;
    }
    if (nn % 2 == 1) {
      double r0 = add_into ? array_for_add[mid * stride] + coefficients_eo[mid * offset + mid] * xmid
                           : coefficients_eo[mid * offset + mid] * xmid;
      if (type == 1 && add_into)
        r0 = array_for_add[mid * stride];
      else if (type == 1)
        r0 = 0.;
      if (type == 1)
        for (int ind = 0; ind < mid; ++ind) r0 += coefficients_eo[mid * offset + ind] * xm[ind];
      else
        for (int ind = 0; ind < mid; ++ind) r0 += coefficients_eo[mid * offset + ind] * xp[ind];
      if (do_dg > 0) r0 += dg_coefficients[mid] * array_face[0];
      if (do_dg > 1) r0 += dg_coefficients[nn + mid] * array_face[3];
      out[mid * stride] = r0;
    }
  }
}

/* ------------------------------------------------------------------ */
/* dense tensor-product helpers (sum factorization, one direction at a   */
/* time, as in every operator of the reference)                          */
/* ------------------------------------------------------------------ */
static inline int dstride(int n, int dir) { return dir == 0 ? 1 : (dir == 1 ? n : n * n); }

/* forward: out[..q..] = sum_i M[i*n+q] in[..i..]; backward (transpose):
 * out[..i..] = sum_q M[i*n+q] in[..q..].  dims = 2 or 3. */
static void apply_dir(int n, int dims, int dir, const double *M, int transpose, const double *in,
                      double *out)
{
  const int s = dstride(n, dir);
  const int total = dims == 3 ? n * n * n : n * n;
  for (int base = 0; base < total; ++base) {
    if ((base / s) % n != 0) continue; /* base has coordinate 0 along dir */
    for (int o = 0; o < n; ++o) {
      double acc = 0;
      for (int k = 0; k < n; ++k) acc += (transpose ? M[o * n + k] : M[k * n + o]) * in[base + k * s];
      out[base + o * s] = acc;
    }
  }
}

static void apply3(int n, const double *M, int transpose, const double *in, double *out,
                   double *tmp)
{
  if (!transpose) {
    apply_dir(n, 3, 0, M, 0, in, out);
    apply_dir(n, 3, 1, M, 0, out, tmp);
    apply_dir(n, 3, 2, M, 0, tmp, out);
  } else {
    apply_dir(n, 3, 2, M, 1, in, out);
    apply_dir(n, 3, 1, M, 1, out, tmp);
    apply_dir(n, 3, 0, M, 1, tmp, out);
  }
}

static void apply2(int n, const double *M, const double *in, double *out)
{
  double tmp[MAXN * MAXN];
  apply_dir(n, 2, 0, M, 0, in, tmp);
  apply_dir(n, 2, 1, M, 0, tmp, out);
}

/* face plane index: (higher remaining dir)*n + (lower remaining dir) */
static inline int vol_index(int n, int dir, int layer, int fa)
{
  const int lo = fa % n, hi = fa / n;
  if (dir == 0) return (hi * n + lo) * n + layer;
  if (dir == 1) return (hi * n + layer) * n + lo;
  return (layer * n + hi) * n + lo;
}

/* extract sum_q c[q] * vol[q along dir] into a face array */
static void extract_face(int n, int dir, const double *c, const double *vol, double *face)
{
  for (int fa = 0; fa < n * n; ++fa) {
    double acc = 0;
    for (int q = 0; q < n; ++q) acc += c[q] * vol[vol_index(n, dir, q, fa)];
    face[fa] = acc;
  }
}

static void inject_face(int n, int dir, const double *c, const double *face, double *vol)
{
  for (int fa = 0; fa < n * n; ++fa)
    for (int q = 0; q < n; ++q) vol[vol_index(n, dir, q, fa)] += c[q] * face[fa];
}

static inline long cell_id(const int nc[3], int ix, int iy, int iz)
{
  return ((long)iz * nc[1] + iy) * nc[0] + ix;
}

static long neighbor(const int nc[3], int ix, int iy, int iz, int dir, int side)
{
  int c[3] = {ix, iy, iz};
  c[dir] += side ? 1 : -1;
  if (c[dir] < 0) c[dir] += nc[dir];
  if (c[dir] >= nc[dir]) c[dir] -= nc[dir];
  return cell_id(nc, c[0], c[1], c[2]);
}

/* ------------------------------------------------------------------ */
/* element-local cell Laplacian: evaluation_cell_laplacian.h:264-503     */
/* ------------------------------------------------------------------ */
void sfko_cell_laplacian(int degree, int basis, long n_cells, const double inv_jac[3],
                         const double *src, double *dst)
{
  sfko_tables t;
  if (sfko_fill_tables(degree, basis, &t)) return;
  const int n = t.n, n3 = n * n * n;
  const double jxw = 1. / (inv_jac[0] * inv_jac[1] * inv_jac[2]);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < n_cells; ++c) {
    double uq[MAXN * MAXN * MAXN], g[MAXN * MAXN * MAXN], rq[MAXN * MAXN * MAXN],
        tmp[MAXN * MAXN * MAXN], acc[MAXN * MAXN * MAXN];
    apply3(n, t.S, 0, src + c * n3, uq, tmp);
    memset(rq, 0, sizeof(double) * n3);
    for (int d = 0; d < 3; ++d) {
      apply_dir(n, 3, d, t.D, 0, uq, g);
      const double m = jxw * inv_jac[d] * inv_jac[d]; /* merged_array, :368-371 */
      for (int q = 0; q < n3; ++q) g[q] *= m * t.w[q % n] * t.w[(q / n) % n] * t.w[q / (n * n)];
      apply_dir(n, 3, d, t.D, 1, g, acc);
      for (int q = 0; q < n3; ++q) rq[q] += acc[q];
    }
    apply3(n, t.S, 1, rq, dst + c * n3, tmp);
  }
}

/* ------------------------------------------------------------------ */
/* DG-SIPG, Cartesian periodic: evaluation_dg_laplacian.h:129-510         */
/* ------------------------------------------------------------------ */
static void sipg_cell(const sfko_tables *t, const int nc[3], const double inv_jac[3],
                      const double *src, int ix, int iy, int iz, double *out, int ref_x_sweep)
{
  const int n = t->n, n3 = n * n * n, p = t->degree;
  const double jxw = 1. / (inv_jac[0] * inv_jac[1] * inv_jac[2]);
  double uq[MAXN * MAXN * MAXN], g[MAXN * MAXN * MAXN], rq[MAXN * MAXN * MAXN],
      tmp[MAXN * MAXN * MAXN], acc[MAXN * MAXN * MAXN];
  const double *u = src + cell_id(nc, ix, iy, iz) * n3;
  apply3(n, t->S, 0, u, uq, tmp);
  memset(rq, 0, sizeof(double) * n3);
  /* cell term, :410-416 */
  for (int d = 0; d < 3; ++d) {
    apply_dir(n, 3, d, t->D, 0, uq, g);
    const double m = inv_jac[d] * inv_jac[d] * jxw;
    for (int q = 0; q < n3; ++q) g[q] *= m * t->w[q % n] * t->w[(q / n) % n] * t->w[q / (n * n)];
    apply_dir(n, 3, d, t->D, 1, g, acc);
    for (int q = 0; q < n3; ++q) rq[q] += acc[q];
  }
  /* face terms, :185-254 (z), :256-301 (x), :303-407 (y) */
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < 2; ++s) {
      double vm[MAXN * MAXN], dm[MAXN * MAXN], upn[MAXN * MAXN], dpn[MAXN * MAXN], up[MAXN * MAXN],
          dp[MAXN * MAXN], tv[MAXN * MAXN], td[MAXN * MAXN];
      extract_face(n, d, s ? t->fv1 : t->fv0, uq, vm);
      extract_face(n, d, s ? t->fd1 : t->fd0, uq, dm);
      const double *nb = src + neighbor(nc, ix, iy, iz, d, s) * n3;
      const int lf = s ? 0 : p, li = s ? 1 : p - 1;
      const double w0 = (s ? -1. : 1.) * t->hd;
      for (int fa = 0; fa < n * n; ++fa) {
        upn[fa] = nb[vol_index(n, d, lf, fa)];
        dpn[fa] = w0 * (nb[vol_index(n, d, li, fa)] - upn[fa]);
      }
      apply2(n, t->S, upn, up);
      apply2(n, t->S, dpn, dp);
      const double sigma = s ? -1. : 1.;
      const double tau = (p + 1) * (p + 1) * inv_jac[d];
      const double jxw_face = jxw * inv_jac[d];
      for (int fa = 0; fa < n * n; ++fa) {
        const double wf = t->w[fa % n] * t->w[fa / n];
        const double avg_grad = sigma * 0.5 * inv_jac[d] * (dm[fa] + dp[fa]);
        const double jump = vm[fa] - up[fa];
        tv[fa] = (avg_grad + jump * tau) * jxw_face * wf;
        td[fa] = (sigma * 0.5 * inv_jac[d] * jxw_face) * jump * wf;
      }
      inject_face(n, d, s ? t->fv1 : t->fv0, tv, rq);
      inject_face(n, d, s ? t->fd1 : t->fd0, td, rq);
    }
  if (!ref_x_sweep) {
    apply3(n, t->S, 1, rq, out, tmp);
  } else {
    /* the reference's last x sweep contracts over rows (S^T instead of S),
     * evaluation_dg_laplacian.h:481-483 */
    apply_dir(n, 3, 2, t->S, 1, rq, out);
    apply_dir(n, 3, 1, t->S, 1, out, tmp);
    apply_dir(n, 3, 0, t->S, 0, tmp, out);
  }
}

void sfko_dg_sipg(int degree, const int nc[3], const double inv_jac[3], const double *src,
                  double *dst)
{
  sfko_dg_sipg_ex(degree, nc, inv_jac, src, dst, 0);
}

void sfko_dg_sipg_ex(int degree, const int nc[3], const double inv_jac[3], const double *src,
                     double *dst, int ref_x_sweep)
{
  sfko_tables t;
  if (sfko_fill_tables(degree, 0, &t)) return;
  const int n3 = t.n * t.n * t.n;
  const long ncell = (long)nc[0] * nc[1] * nc[2];
#pragma omp parallel for schedule(static)
  for (long c = 0; c < ncell; ++c) {
    const int ix = c % nc[0], iy = (c / nc[0]) % nc[1], iz = c / ((long)nc[0] * nc[1]);
    sipg_cell(&t, nc, inv_jac, src, ix, iy, iz, dst + c * n3, ref_x_sweep);
  }
}

void sfko_dg_sipg_cheby(int degree, const int nc[3], const double inv_jac[3], const double *rhs,
                        const double *sol_old, const double *diag, double *sol_new, double *tm,
                        double c_np, double c_tm)
{
  sfko_dg_sipg_cheby_ex(degree, nc, inv_jac, rhs, sol_old, diag, sol_new, tm, c_np, c_tm, 0);
}

void sfko_dg_sipg_cheby_ex(int degree, const int nc[3], const double inv_jac[3], const double *rhs,
                           const double *sol_old, const double *diag, double *sol_new, double *tm,
                           double c_np, double c_tm, int ref_x_sweep)
{
  sfko_tables t;
  if (sfko_fill_tables(degree, 0, &t)) return;
  const int n3 = t.n * t.n * t.n;
  const long ncell = (long)nc[0] * nc[1] * nc[2];
#pragma omp parallel for schedule(static)
  for (long c = 0; c < ncell; ++c) {
    double au[MAXN * MAXN * MAXN];
    const int ix = c % nc[0], iy = (c / nc[0]) % nc[1], iz = c / ((long)nc[0] * nc[1]);
    sipg_cell(&t, nc, inv_jac, sol_old, ix, iy, iz, au, ref_x_sweep);
    for (int q = 0; q < n3; ++q) {
      const long i = c * n3 + q;
      const double res = au[q] - rhs[i];
      const double tmp = c_tm * tm[i] + c_np * diag[i] * res;
      tm[i] = tmp;
      sol_new[i] = sol_old[i] - tmp;
    }
  }
}

/* ------------------------------------------------------------------ */
/* DG upwind advection: evaluation_dg_advect.h:285-676                    */
/* ------------------------------------------------------------------ */
static void advect_cell(const sfko_tables *t, const int nc[3], const double inv_jac[3],
                        const double vel[3], const double *src, int ix, int iy, int iz,
                        double *out)
{
  const int n = t->n, n3 = n * n * n, p = t->degree;
  const double jxw = 1. / (inv_jac[0] * inv_jac[1] * inv_jac[2]);
  double uq[MAXN * MAXN * MAXN], g[MAXN * MAXN * MAXN], rq[MAXN * MAXN * MAXN],
      tmp[MAXN * MAXN * MAXN], acc[MAXN * MAXN * MAXN];
  const double *u = src + cell_id(nc, ix, iy, iz) * n3;
  apply3(n, t->S, 0, u, uq, tmp);
  memset(rq, 0, sizeof(double) * n3);
  /* cell term (grad v, a u), :567-574 */
  for (int d = 0; d < 3; ++d) {
    for (int q = 0; q < n3; ++q)
      g[q] = vel[d] * inv_jac[d] * (jxw * (t->w[q % n] * t->w[(q / n) % n] * t->w[q / (n * n)]) * uq[q]);
    apply_dir(n, 3, d, t->D, 1, g, acc);
    for (int q = 0; q < n3; ++q) rq[q] += acc[q];
  }
  /* upwind fluxes, :349-565 */
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < 2; ++s) {
      double vm[MAXN * MAXN], upn[MAXN * MAXN], up[MAXN * MAXN], fl[MAXN * MAXN];
      extract_face(n, d, s ? t->fv1 : t->fv0, uq, vm);
      const double *nb = src + neighbor(nc, ix, iy, iz, d, s) * n3;
      const int lf = s ? 0 : p;
      for (int fa = 0; fa < n * n; ++fa) upn[fa] = nb[vol_index(n, d, lf, fa)];
      apply2(n, t->S, upn, up);
      const double an = s ? vel[d] : -vel[d];
      const double jxw_face = jxw * inv_jac[d];
      for (int fa = 0; fa < n * n; ++fa) {
        const double wf = t->w[fa % n] * t->w[fa / n];
        fl[fa] = -0.5 * ((vm[fa] + up[fa]) * an + fabs(an) * (vm[fa] - up[fa])) * jxw_face * wf;
      }
      inject_face(n, d, s ? t->fv1 : t->fv0, fl, rq);
    }
  /* inverse mass via collocation, :619-649 */
  for (int q = 0; q < n3; ++q) rq[q] /= jxw * (t->w[q % n] * t->w[(q / n) % n] * t->w[q / (n * n)]);
  apply3(n, t->Sinv, 0, rq, out, tmp);
}

void sfko_dg_advect(int degree, const int nc[3], const double inv_jac[3], const double vel[3],
                    const double *src, double *dst)
{
  sfko_tables t;
  if (sfko_fill_tables(degree, 1, &t)) return;
  const int n3 = t.n * t.n * t.n;
  const long ncell = (long)nc[0] * nc[1] * nc[2];
#pragma omp parallel for schedule(static)
  for (long c = 0; c < ncell; ++c) {
    const int ix = c % nc[0], iy = (c / nc[0]) % nc[1], iz = c / ((long)nc[0] * nc[1]);
    advect_cell(&t, nc, inv_jac, vel, src, ix, iy, iz, dst + c * n3);
  }
}

void sfko_dg_advect_stage(int degree, const int nc[3], const double inv_jac[3],
                          const double vel[3], const double *src, const double *upd,
                          double *vec_np, double *vec_tm, double c_np, double c_tm)
{
  sfko_tables t;
  if (sfko_fill_tables(degree, 1, &t)) return;
  const int n3 = t.n * t.n * t.n;
  const long ncell = (long)nc[0] * nc[1] * nc[2];
  double *r = (double *)malloc(sizeof(double) * n3 * ncell);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < ncell; ++c) {
    const int ix = c % nc[0], iy = (c / nc[0]) % nc[1], iz = c / ((long)nc[0] * nc[1]);
    advect_cell(&t, nc, inv_jac, vel, src, ix, iy, iz, r + c * n3);
  }
  /* stage update after all cells (src may alias vec_tm when c_tm == 0) */
  for (long i = 0; i < ncell * n3; ++i) {
    const double vnp = upd[i] + c_np * r[i];
    vec_np[i] = vnp;
    if (c_tm != 0.) vec_tm[i] = vnp + c_tm * r[i];
  }
  free(r);
}

/* ------------------------------------------------------------------ */
/* layout converters                                                    */
/* ------------------------------------------------------------------ */
static long interleaved_index(int n3, const int nc[3], int lanes, int ix, int iy, int iz, int dof)
{
  const int nxb = nc[0] / lanes;
  return (((long)iz * nc[1] + iy) * nxb + ix / lanes) * (long)lanes * n3 + (long)dof * lanes + ix % lanes;
}

void sfko_interleaved_to_canonical(int degree, const int nc[3], int lanes, const double *in,
                                   double *out)
{
  const int n = degree + 1, n3 = n * n * n;
  for (int iz = 0; iz < nc[2]; ++iz)
    for (int iy = 0; iy < nc[1]; ++iy)
      for (int ix = 0; ix < nc[0]; ++ix)
        for (int q = 0; q < n3; ++q)
          out[cell_id(nc, ix, iy, iz) * n3 + q] = in[interleaved_index(n3, nc, lanes, ix, iy, iz, q)];
}

void sfko_canonical_to_interleaved(int degree, const int nc[3], int lanes, const double *in,
                                   double *out)
{
  const int n = degree + 1, n3 = n * n * n;
  for (int iz = 0; iz < nc[2]; ++iz)
    for (int iy = 0; iy < nc[1]; ++iy)
      for (int ix = 0; ix < nc[0]; ++ix)
        for (int q = 0; q < n3; ++q)
          out[interleaved_index(n3, nc, lanes, ix, iy, iz, q)] = in[cell_id(nc, ix, iy, iz) * n3 + q];
}

/*
 * Extended operators of the CPU oracle (test infrastructure only).  The
 * reference has none of these; each is the textbook form stated below, in
 * the same sum-factorization steps and tables as the reference operators in
 * sfk_oracle.c, and is pinned by the checks named at each function
 * (tests/test_oracle_extended.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sfk_oracle.h"

#define MAXN SFKO_MAXN

static inline int dstride(int n, int dir) { return dir == 0 ? 1 : (dir == 1 ? n : n * n); }

static void apply_dir(int n, int dims, int dir, const double *M, int transpose, const double *in, double *out)
{
  const int s = dstride(n, dir);
  const int total = dims == 3 ? n * n * n : n * n;
  for (int base = 0; base < total; ++base) {
    if ((base / s) % n != 0) continue;
    for (int o = 0; o < n; ++o) {
      double acc = 0;
      for (int k = 0; k < n; ++k) acc += (transpose ? M[o * n + k] : M[k * n + o]) * in[base + k * s];
      out[base + o * s] = acc;
    }
  }
}

static void apply3(int n, const double *M, int transpose, const double *in, double *out)
{
  double tmp[MAXN * MAXN * MAXN];
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

static inline int vol_index(int n, int dir, int layer, int fa)
{
  const int lo = fa % n, hi = fa / n;
  if (dir == 0) return (hi * n + lo) * n + layer;
  if (dir == 1) return (hi * n + layer) * n + lo;
  return (layer * n + hi) * n + lo;
}

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

static long neighbor(const int nc[3], long c, int dir, int side)
{
  int x[3] = {(int)(c % nc[0]), (int)((c / nc[0]) % nc[1]), (int)(c / ((long)nc[0] * nc[1]))};
  x[dir] += side ? 1 : -1;
  if (x[dir] < 0) x[dir] += nc[dir];
  if (x[dir] >= nc[dir]) x[dir] -= nc[dir];
  return cell_id(nc, x[0], x[1], x[2]);
}

static inline double w3(const sfko_tables *t, int q)
{
  const int n = t->n;
  return t->w[q % n] * t->w[(q / n) % n] * t->w[q / (n * n)];
}

/* in-plane directions of a face normal to dir: lo, hi */
static inline void plane_dirs(int dir, int *lo, int *hi)
{
  *lo = dir == 0 ? 1 : 0;
  *hi = dir == 2 ? 1 : 2;
}

/* ------------------------------------------------------------------ */
/* continuous Q_p Laplacian, homogeneous Dirichlet rows                  */
/* -div(k grad u) with k per cell q-point (coef may be NULL: k = 1);      */
/* affine Cartesian cells; node grid (nx p+1)(ny p+1)(nz p+1).  dst = 0   */
/* on boundary nodes and boundary entries of src are ignored.  Pinned by  */
/* a global assembly with scipy/numpy (symmetry, A*x for x = linear      */
/* functions, brute-force Kronecker element matrices).                    */
/* ------------------------------------------------------------------ */
void sfko_cont_laplacian(int degree, const int nc[3], const double inv_jac[3], const double *coef,
                         const double *src, double *dst)
{
  sfko_tables t;
  if (sfko_fill_tables(degree, 1, &t)) return;
  const int n = t.n, n3 = n * n * n, p = degree;
  const int gx = nc[0] * p + 1, gy = nc[1] * p + 1, gz = nc[2] * p + 1;
  const long ng = (long)gx * gy * gz, ncell = (long)nc[0] * nc[1] * nc[2];
  const double jxw = 1. / (inv_jac[0] * inv_jac[1] * inv_jac[2]);
  memset(dst, 0, sizeof(double) * ng);
  for (long c = 0; c < ncell; ++c) {
    const int cx = c % nc[0], cy = (c / nc[0]) % nc[1], cz = c / ((long)nc[0] * nc[1]);
    double u[MAXN * MAXN * MAXN], uq[MAXN * MAXN * MAXN], g[MAXN * MAXN * MAXN], rq[MAXN * MAXN * MAXN],
        acc[MAXN * MAXN * MAXN], out[MAXN * MAXN * MAXN];
    long gidx[MAXN * MAXN * MAXN];
    int bd[MAXN * MAXN * MAXN];
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const int X = cx * p + i, Y = cy * p + j, Z = cz * p + k, q = (k * n + j) * n + i;
          gidx[q] = ((long)Z * gy + Y) * gx + X;
          bd[q] = X == 0 || Y == 0 || Z == 0 || X == gx - 1 || Y == gy - 1 || Z == gz - 1;
          u[q] = bd[q] ? 0. : src[gidx[q]];
        }
    apply3(n, t.S, 0, u, uq);
    memset(rq, 0, sizeof(double) * n3);
    for (int d = 0; d < 3; ++d) {
      apply_dir(n, 3, d, t.D, 0, uq, g);
      const double m = jxw * inv_jac[d] * inv_jac[d];
      for (int q = 0; q < n3; ++q) g[q] *= m * w3(&t, q) * (coef ? coef[c * n3 + q] : 1.);
      apply_dir(n, 3, d, t.D, 1, g, acc);
      for (int q = 0; q < n3; ++q) rq[q] += acc[q];
    }
    apply3(n, t.S, 1, rq, out);
    for (int q = 0; q < n3; ++q)
      if (!bd[q]) dst[gidx[q]] += out[q];
  }
}

/* d/dk_q of lambda^T A(k) u = JxW_q grad u(x_q) . grad lambda(x_q), summed
 * over n_rhs (u, lambda) pairs stored contiguously.  Pinned by central
 * finite differences of lambda^T A(k) u. */
void sfko_cont_laplacian_coef_grad(int degree, const int nc[3], const double inv_jac[3], int n_rhs,
                                   const double *u_all, const double *lam_all, double *grad)
{
  sfko_tables t;
  if (sfko_fill_tables(degree, 1, &t)) return;
  const int n = t.n, n3 = n * n * n, p = degree;
  const int gx = nc[0] * p + 1, gy = nc[1] * p + 1, gz = nc[2] * p + 1;
  const long ng = (long)gx * gy * gz, ncell = (long)nc[0] * nc[1] * nc[2];
  const double jxw = 1. / (inv_jac[0] * inv_jac[1] * inv_jac[2]);
  memset(grad, 0, sizeof(double) * ncell * n3);
  for (int r = 0; r < n_rhs; ++r) {
    const double *src = u_all + r * ng, *lam = lam_all + r * ng;
    for (long c = 0; c < ncell; ++c) {
      const int cx = c % nc[0], cy = (c / nc[0]) % nc[1], cz = c / ((long)nc[0] * nc[1]);
      double u[MAXN * MAXN * MAXN], l[MAXN * MAXN * MAXN], uq[MAXN * MAXN * MAXN], lq[MAXN * MAXN * MAXN],
          gu[MAXN * MAXN * MAXN], gl[MAXN * MAXN * MAXN];
      for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < n; ++i) {
            const int X = cx * p + i, Y = cy * p + j, Z = cz * p + k, q = (k * n + j) * n + i;
            const long gi = ((long)Z * gy + Y) * gx + X;
            const int b = X == 0 || Y == 0 || Z == 0 || X == gx - 1 || Y == gy - 1 || Z == gz - 1;
            u[q] = b ? 0. : src[gi];
            l[q] = b ? 0. : lam[gi];
          }
      apply3(n, t.S, 0, u, uq);
      apply3(n, t.S, 0, l, lq);
      for (int d = 0; d < 3; ++d) {
        apply_dir(n, 3, d, t.D, 0, uq, gu);
        apply_dir(n, 3, d, t.D, 0, lq, gl);
        for (int q = 0; q < n3; ++q)
          grad[c * n3 + q] += jxw * w3(&t, q) * inv_jac[d] * inv_jac[d] * gu[q] * gl[q];
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* DG-SIPG with per-quadrature-point metric (deformed periodic mesh).     */
/*   a(u,v) = sum_K (G grad^ u, grad^ v)_K                                */
/*          + sum_F [ -{du/dn}[v] - {dv/dn}[u] + sigma [u][v] ]_F          */
/* cell metric [cell][6][q] (G00,G11,G22,G01,G02,G12) with G = JxW        */
/* J^{-1}J^{-T}; face metric [d][cell][7][fa] for the face between cell   */
/* and its upper neighbour in d: aL(3), aU(3), sigma, where               */
/* aX = JxW_f J_X^{-1} n (n from lower to upper cell) and JxW_f includes   */
/* the face quadrature weight.  The face value of the neighbour and its   */
/* normal derivative use the Hermite trick of evaluation_dg_laplacian.h   */
/* :196-203; tangential derivatives are in-plane collocation derivatives. */
/* Pinned by: Cartesian metric == sfko_dg_sipg, symmetry, A*1 = 0,         */
/* positive semi-definiteness.                                            */
/* ------------------------------------------------------------------ */
static void sipg_general_cell(const sfko_tables *t, const int nc[3], const double *cm, const double *fm,
                              const double *src, long c, double *out)
{
  const int n = t->n, n2 = n * n, n3 = n * n * n, p = t->degree;
  const long ncell = (long)nc[0] * nc[1] * nc[2];
  double uq[MAXN * MAXN * MAXN], G[3][MAXN * MAXN * MAXN], T[3][MAXN * MAXN * MAXN], rq[MAXN * MAXN * MAXN],
      acc[MAXN * MAXN * MAXN];
  apply3(n, t->S, 0, src + c * n3, uq);
  for (int d = 0; d < 3; ++d) apply_dir(n, 3, d, t->D, 0, uq, G[d]);
  const double *m = cm + c * 6 * n3;
  for (int q = 0; q < n3; ++q) {
    const double g00 = m[q], g11 = m[n3 + q], g22 = m[2 * n3 + q], g01 = m[3 * n3 + q], g02 = m[4 * n3 + q],
                 g12 = m[5 * n3 + q];
    T[0][q] = g00 * G[0][q] + g01 * G[1][q] + g02 * G[2][q];
    T[1][q] = g01 * G[0][q] + g11 * G[1][q] + g12 * G[2][q];
    T[2][q] = g02 * G[0][q] + g12 * G[1][q] + g22 * G[2][q];
  }
  memset(rq, 0, sizeof(double) * n3);
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < 2; ++s) {
      const double *fv = s ? t->fv1 : t->fv0, *fd = s ? t->fd1 : t->fd0;
      int lo, hi;
      plane_dirs(d, &lo, &hi);
      double um[MAXN * MAXN], gm[3][MAXN * MAXN];
      extract_face(n, d, fv, uq, um);
      extract_face(n, d, fd, uq, gm[d]);
      extract_face(n, d, fv, G[lo], gm[lo]);
      extract_face(n, d, fv, G[hi], gm[hi]);
      const long nb = neighbor(nc, c, d, s);
      const double *u_nb = src + nb * n3;
      const int lf = s ? 0 : p, li = s ? 1 : p - 1;
      const double w0 = (s ? -1. : 1.) * t->hd;
      double upn[MAXN * MAXN], dpn[MAXN * MAXN], up[MAXN * MAXN], gp[3][MAXN * MAXN];
      for (int fa = 0; fa < n2; ++fa) {
        upn[fa] = u_nb[vol_index(n, d, lf, fa)];
        dpn[fa] = w0 * (u_nb[vol_index(n, d, li, fa)] - upn[fa]);
      }
      apply2(n, t->S, upn, up);
      apply2(n, t->S, dpn, gp[d]);
      apply_dir(n, 2, 0, t->D, 0, up, gp[lo]);
      apply_dir(n, 2, 1, t->D, 0, up, gp[hi]);
      /* face data: own side is L for s = 1 (face of this cell), U for s = 0 */
      const long owner = s ? c : nb;
      const double *f = fm + ((long)d * ncell + owner) * 7 * n2;
      const double sn = s ? 1. : -1.;
      double tv[MAXN * MAXN], td[3][MAXN * MAXN];
      for (int fa = 0; fa < n2; ++fa) {
        const double *aL = f, *aU = f + 3 * n2;
        double ao[3], an[3];
        for (int e = 0; e < 3; ++e) {
          ao[e] = s ? aL[e * n2 + fa] : aU[e * n2 + fa];
          an[e] = s ? aU[e * n2 + fa] : aL[e * n2 + fa];
        }
        const double sigma = f[6 * n2 + fa];
        double dn = 0;
        for (int e = 0; e < 3; ++e) dn += ao[e] * gm[e][fa] + an[e] * gp[e][fa];
        const double avg = sn * 0.5 * dn; /* {du/dn_out} JxW_f */
        const double jump = um[fa] - up[fa];
        tv[fa] = -avg + sigma * jump;
        for (int e = 0; e < 3; ++e) td[e][fa] = -0.5 * jump * sn * ao[e];
      }
      inject_face(n, d, fv, tv, rq);
      for (int e = 0; e < 3; ++e) inject_face(n, d, fv, td[e], T[e]);
    }
  for (int d = 0; d < 3; ++d) {
    apply_dir(n, 3, d, t->D, 1, T[d], acc);
    for (int q = 0; q < n3; ++q) rq[q] += acc[q];
  }
  apply3(n, t->S, 1, rq, out);
}

void sfko_dg_sipg_general(int degree, const int nc[3], const double *cell_metric, const double *face_metric,
                          const double *src, double *dst)
{
  sfko_tables t;
  if (sfko_fill_tables(degree, 0, &t)) return;
  const int n3 = t.n * t.n * t.n;
  const long ncell = (long)nc[0] * nc[1] * nc[2];
#pragma omp parallel for schedule(static)
  for (long c = 0; c < ncell; ++c) sipg_general_cell(&t, nc, cell_metric, face_metric, src, c, dst + c * n3);
}

/* Cartesian metric in the general layout (pins the general operator to the
 * reference-derived Cartesian one): G = JxW diag(ij^2) w_q, aL = aU =
 * JxW_f ij_d e_d w_f, sigma = (p+1)^2 ij_d JxW_f w_f. */
void sfko_cartesian_metric(int degree, const int nc[3], const double inv_jac[3], double *cell_metric,
                           double *face_metric)
{
  sfko_tables t;
  if (sfko_fill_tables(degree, 0, &t)) return;
  const int n = t.n, n2 = n * n, n3 = n2 * n;
  const long ncell = (long)nc[0] * nc[1] * nc[2];
  const double jxw = 1. / (inv_jac[0] * inv_jac[1] * inv_jac[2]);
  for (long c = 0; c < ncell; ++c)
    for (int q = 0; q < n3; ++q) {
      double *m = cell_metric + c * 6 * n3;
      for (int e = 0; e < 3; ++e) m[e * n3 + q] = jxw * inv_jac[e] * inv_jac[e] * w3(&t, q);
      for (int e = 3; e < 6; ++e) m[e * n3 + q] = 0.;
    }
  for (int d = 0; d < 3; ++d)
    for (long c = 0; c < ncell; ++c)
      for (int fa = 0; fa < n2; ++fa) {
        double *f = face_metric + ((long)d * ncell + c) * 7 * n2;
        const double jf = jxw * inv_jac[d] * t.w[fa % n] * t.w[fa / n];
        for (int e = 0; e < 3; ++e) {
          f[e * n2 + fa] = e == d ? jf * inv_jac[d] : 0.;
          f[(3 + e) * n2 + fa] = e == d ? jf * inv_jac[d] : 0.;
        }
        f[6 * n2 + fa] = (degree + 1) * (degree + 1) * inv_jac[d] * jf;
      }
}

/* ------------------------------------------------------------------ */
/* DG advection with per-point velocity, its transpose and the velocity   */
/* gradient.  cell_vel [cell][3][q]; face_vel [d][cell][fa] = a.n at the  */
/* upper face of the cell in direction d (n pointing to the upper cell).  */
/* Pinned by: constant velocity == sfko_dg_advect (reference-derived),    */
/* <A u, v> = <u, A^T v>, central finite differences of 0.5|A u - f|^2.   */
/* ------------------------------------------------------------------ */
static double face_an(const int nc[3], const double *face_vel, long c, int d, int s, int fa, int n2)
{
  const long ncell = (long)nc[0] * nc[1] * nc[2];
  if (s) return face_vel[((long)d * ncell + c) * n2 + fa];
  return -face_vel[((long)d * ncell + neighbor(nc, c, d, 0)) * n2 + fa];
}

/* r = M^{-1} L(u) (transpose == 0) or L^T y (transpose == 1) */
static void advect_var_cell(const sfko_tables *t, const int nc[3], const double inv_jac[3], const double *cv,
                            const double *fvel, const double *src, long c, int transpose, double *out)
{
  const int n = t->n, n2 = n * n, n3 = n2 * n, p = t->degree;
  const double jxw = 1. / (inv_jac[0] * inv_jac[1] * inv_jac[2]);
  double uq[MAXN * MAXN * MAXN], g[MAXN * MAXN * MAXN], rq[MAXN * MAXN * MAXN], acc[MAXN * MAXN * MAXN];
  apply3(n, t->S, 0, src + c * n3, uq);
  memset(rq, 0, sizeof(double) * n3);
  for (int d = 0; d < 3; ++d) {
    const double *a = cv + (c * 3 + d) * n3;
    if (!transpose) {
      for (int q = 0; q < n3; ++q) g[q] = a[q] * inv_jac[d] * jxw * w3(t, q) * uq[q];
      apply_dir(n, 3, d, t->D, 1, g, acc);
    } else {
      apply_dir(n, 3, d, t->D, 0, uq, g);
      for (int q = 0; q < n3; ++q) acc[q] = a[q] * inv_jac[d] * jxw * w3(t, q) * g[q];
    }
    for (int q = 0; q < n3; ++q) rq[q] += acc[q];
  }
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < 2; ++s) {
      double vm[MAXN * MAXN], upn[MAXN * MAXN], up[MAXN * MAXN], fl[MAXN * MAXN];
      extract_face(n, d, s ? t->fv1 : t->fv0, uq, vm);
      const double *nb = src + neighbor(nc, c, d, s) * n3;
      for (int fa = 0; fa < n2; ++fa) upn[fa] = nb[vol_index(n, d, s ? 0 : p, fa)];
      apply2(n, t->S, upn, up);
      for (int fa = 0; fa < n2; ++fa) {
        const double an = face_an(nc, fvel, c, d, s, fa, n2);
        const double wf = jxw * inv_jac[d] * t->w[fa % n] * t->w[fa / n];
        if (!transpose)
          fl[fa] = -0.5 * ((vm[fa] + up[fa]) * an + fabs(an) * (vm[fa] - up[fa])) * wf;
        else
          fl[fa] = -(an > 0 ? an : 0.) * (vm[fa] - up[fa]) * wf;
      }
      inject_face(n, d, s ? t->fv1 : t->fv0, fl, rq);
    }
  if (!transpose) {
    for (int q = 0; q < n3; ++q) rq[q] /= jxw * w3(t, q);
    apply3(n, t->Sinv, 0, rq, out);
  } else {
    apply3(n, t->S, 1, rq, out);
  }
}

void sfko_dg_advect_var(int degree, const int nc[3], const double inv_jac[3], const double *cell_vel,
                        const double *face_vel, const double *src, double *dst)
{
  sfko_tables t;
  if (sfko_fill_tables(degree, 1, &t)) return;
  const int n3 = t.n * t.n * t.n;
  const long ncell = (long)nc[0] * nc[1] * nc[2];
#pragma omp parallel for schedule(static)
  for (long c = 0; c < ncell; ++c) advect_var_cell(&t, nc, inv_jac, cell_vel, face_vel, src, c, 0, dst + c * n3);
}

/* M^{-1} x per cell: Sinv^T W^{-1} Sinv x */
void sfko_dg_mass_inverse(int degree, const int nc[3], const double inv_jac[3], const double *src, double *dst)
{
  sfko_tables t;
  if (sfko_fill_tables(degree, 1, &t)) return;
  const int n = t.n, n3 = n * n * n;
  const long ncell = (long)nc[0] * nc[1] * nc[2];
  const double jxw = 1. / (inv_jac[0] * inv_jac[1] * inv_jac[2]);
  for (long c = 0; c < ncell; ++c) {
    double a[MAXN * MAXN * MAXN];
    apply3(n, t.Sinv, 1, src + c * n3, a);
    for (int q = 0; q < n3; ++q) a[q] /= jxw * w3(&t, q);
    apply3(n, t.Sinv, 0, a, dst + c * n3);
  }
}

/* A^T w = L^T M^{-1} w */
void sfko_dg_advect_var_transpose(int degree, const int nc[3], const double inv_jac[3], const double *cell_vel,
                                  const double *face_vel, const double *src, double *dst)
{
  sfko_tables t;
  if (sfko_fill_tables(degree, 1, &t)) return;
  const int n3 = t.n * t.n * t.n;
  const long ncell = (long)nc[0] * nc[1] * nc[2];
  double *y = (double *)malloc(sizeof(double) * ncell * n3);
  sfko_dg_mass_inverse(degree, nc, inv_jac, src, y);
  for (long c = 0; c < ncell; ++c) advect_var_cell(&t, nc, inv_jac, cell_vel, face_vel, y, c, 1, dst + c * n3);
  free(y);
}

/* gradient of J = 0.5 |A(a) u - f|^2 given lambda = A u - f:
 * grad_cell[cell][3][q], grad_face[d][cell][fa] (normal velocity b of the
 * upper face of the cell). */
void sfko_dg_advect_velocity_grad(int degree, const int nc[3], const double inv_jac[3], const double *face_vel,
                                  const double *u, const double *lambda, double *grad_cell, double *grad_face)
{
  sfko_tables t;
  if (sfko_fill_tables(degree, 1, &t)) return;
  const int n = t.n, n2 = n * n, n3 = n2 * n;
  const long ncell = (long)nc[0] * nc[1] * nc[2];
  const double jxw = 1. / (inv_jac[0] * inv_jac[1] * inv_jac[2]);
  double *mu = (double *)malloc(sizeof(double) * ncell * n3);
  sfko_dg_mass_inverse(degree, nc, inv_jac, lambda, mu);
  for (long c = 0; c < ncell; ++c) {
    double uq[MAXN * MAXN * MAXN], mq[MAXN * MAXN * MAXN], g[MAXN * MAXN * MAXN];
    apply3(n, t.S, 0, u + c * n3, uq);
    apply3(n, t.S, 0, mu + c * n3, mq);
    for (int d = 0; d < 3; ++d) {
      apply_dir(n, 3, d, t.D, 0, mq, g);
      for (int q = 0; q < n3; ++q) grad_cell[(c * 3 + d) * n3 + q] = g[q] * inv_jac[d] * uq[q] * jxw * w3(&t, q);
      double um[MAXN * MAXN], mm[MAXN * MAXN], upn[MAXN * MAXN], mpn[MAXN * MAXN], up[MAXN * MAXN],
          mp[MAXN * MAXN];
      extract_face(n, d, t.fv1, uq, um);
      extract_face(n, d, t.fv1, mq, mm);
      const long nb = neighbor(nc, c, d, 1);
      for (int fa = 0; fa < n2; ++fa) {
        upn[fa] = u[nb * n3 + vol_index(n, d, 0, fa)];
        mpn[fa] = mu[nb * n3 + vol_index(n, d, 0, fa)];
      }
      apply2(n, t.S, upn, up);
      apply2(n, t.S, mpn, mp);
      for (int fa = 0; fa < n2; ++fa) {
        const double b = face_vel[((long)d * ncell + c) * n2 + fa];
        const double sg = b > 0 ? 1. : (b < 0 ? -1. : 0.);
        const double uh = 0.5 * (um[fa] + up[fa]) + 0.5 * sg * (um[fa] - up[fa]);
        grad_face[((long)d * ncell + c) * n2 + fa] =
            -(mm[fa] - mp[fa]) * uh * (jxw * inv_jac[d] * t.w[fa % n] * t.w[fa / n]);
      }
    }
  }
  free(mu);
}

/* ------------------------------------------------------------------ */
/* PrecomputedDGLaplacian, legacy/include/precomputed_dg_laplacian.h    */
/* ------------------------------------------------------------------ */

/* 1D matrices of fill_shape_values (:305-349) on the unit interval,
 * Hermite-like basis, Gauss quadrature with n = p+1 points, penalty
 * p(p+1) (:310):
 *   mass[i][j] = sum_q phi_i phi_j w_q,
 *   lap[i][j]  = sum_q phi_i' phi_j' w_q + the cell's own interior-penalty
 *                face terms at x = 0 and x = 1 (:322-330).
 * Neighbour coupling v[0] = v1, v[1] = v2 (rows 0 and 1 of a cell against
 * the left neighbour's dofs n-1 and n-2, :139-148):
 *   ref_coupling != 0: the reference's expressions (:331-340), which
 *     evaluate phi_0 at x = 1 and phi_{n-1}, phi_{n-2} at x = 0, where
 *     these functions vanish (both constants are zero);
 *   ref_coupling == 0: the SIPG face term between the test function phi_0
 *     of a cell (its x = 0) and the trial function phi_j of its left
 *     neighbour (its x = 1):
 *       C_j = 1/2 phi_j'(1) phi_0(0) - 1/2 phi_j(1) phi_0'(0) - sigma phi_0(0) phi_j(1),
 *     v1 = C_{n-1}, v2 = C_{n-2}.
 * Pinned by a brute-force assembly of the 1D SIPG form with an
 * independent quadrature and by A 1 = 0 (tests/test_oracle_precomputed.py). */
void sfko_precomp_1d(int degree, int ref_coupling, double *mass, double *lap, double *v)
{
  const int n = degree + 1;
  double x[MAXN], w[MAXN], val[MAXN][MAXN], der[MAXN][MAXN];
  double v0[MAXN], v1[MAXN], d0[MAXN], d1[MAXN];
  sfko_gauss_points(n, x);
  sfko_gauss_weights(n, w);
  for (int i = 0; i < n; ++i) {
    for (int q = 0; q < n; ++q) {
      val[i][q] = sfko_hermite_value(degree, i, x[q]);
      der[i][q] = sfko_hermite_derivative(degree, i, x[q]);
    }
    v0[i] = sfko_hermite_value(degree, i, 0.);
    v1[i] = sfko_hermite_value(degree, i, 1.);
    d0[i] = sfko_hermite_derivative(degree, i, 0.);
    d1[i] = sfko_hermite_derivative(degree, i, 1.);
  }
  const double sigma = (double)(n - 1) * n;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double sm = 0, sl = 0;
      for (int q = 0; q < n; ++q) {
        sm += val[i][q] * val[j][q] * w[q];
        sl += der[i][q] * der[j][q] * w[q];
      }
      mass[i * n + j] = sm;
      sl += v0[i] * v0[j] * sigma + 0.5 * d0[i] * v0[j] + 0.5 * d0[j] * v0[i];
      sl += v1[i] * v1[j] * sigma - 0.5 * d1[i] * v1[j] - 0.5 * d1[j] * v1[i];
      lap[i * n + j] = sl;
    }
  const int a = n - 1, b = n >= 2 ? n - 2 : 0;
  if (ref_coupling) {
    v[0] = -v1[0] * v0[a] * sigma + 0.5 * d1[0] * v0[a] - 0.5 * d0[a] * v1[0];
    v[1] = -v1[0] * v0[b] * sigma + 0.5 * d1[0] * v0[b] - 0.5 * d0[b] * v1[0];
  } else {
    v[0] = 0.5 * d1[a] * v0[0] - 0.5 * v1[a] * d0[0] - sigma * v0[0] * v1[a];
    v[1] = 0.5 * d1[b] * v0[0] - 0.5 * v1[b] * d0[0] - sigma * v0[0] * v1[b];
  }
}

/* dst = A src of PrecomputedDGLaplacian::do_inner_loop (:97-278), periodic,
 * canonical layout, on Cartesian cells of size 1/inv_jac (the reference:
 * unit cells, inv_jac = 1):
 *   A = f_x M (x) M (x) L~_x + f_y M (x) L~_y (x) M + f_z L~_z (x) M (x) M,
 * L~_d = lap plus the neighbour coupling along d, f_d = inv_jac_d^2 /
 * (inv_jac_x inv_jac_y inv_jac_z) (f_d = 1 on unit cells). */
void sfko_dg_precomputed(int degree, const int nc[3], const double inv_jac[3], int ref_coupling,
                         const double *src, double *dst)
{
  const int n = degree + 1, n2 = n * n, n3 = n2 * n;
  double M[MAXN * MAXN], Lm[MAXN * MAXN], v[2];
  sfko_precomp_1d(degree, ref_coupling, M, Lm, v);
  const double det = inv_jac[0] * inv_jac[1] * inv_jac[2];
  double f[3];
  for (int d = 0; d < 3; ++d) f[d] = inv_jac[d] * inv_jac[d] / det;
  const long ncell = (long)nc[0] * nc[1] * nc[2];
  double a[3][MAXN * MAXN * MAXN], t1[MAXN * MAXN * MAXN], t2[MAXN * MAXN * MAXN];
  for (long c = 0; c < ncell; ++c) {
    const double *u = src + c * n3;
    for (int d = 0; d < 3; ++d) {
      apply_dir(n, 3, d, Lm, 0, u, a[d]);
      const double *ul = src + neighbor(nc, c, d, 0) * n3, *uh = src + neighbor(nc, c, d, 1) * n3;
      for (int fa = 0; fa < n2; ++fa) {
#define VI(layer) vol_index(n, d, (layer), fa)
        a[d][VI(0)] += v[0] * ul[VI(n - 1)] + v[1] * ul[VI(n >= 2 ? n - 2 : 0)];
        a[d][VI(1)] += v[1] * ul[VI(n - 1)];
        a[d][VI(n - 1)] += v[0] * uh[VI(0)] + v[1] * uh[VI(1)];
        a[d][VI(n - 2)] += v[1] * uh[VI(0)];
#undef VI
      }
      for (int i = 0; i < n3; ++i) a[d][i] *= f[d];
    }
    double *out = dst + c * n3;
    /* M_z M_y a_x */
    apply_dir(n, 3, 1, M, 0, a[0], t1);
    apply_dir(n, 3, 2, M, 0, t1, t2);
    for (int i = 0; i < n3; ++i) out[i] = t2[i];
    /* M_z M_x a_y */
    apply_dir(n, 3, 0, M, 0, a[1], t1);
    apply_dir(n, 3, 2, M, 0, t1, t2);
    for (int i = 0; i < n3; ++i) out[i] += t2[i];
    /* M_y M_x a_z */
    apply_dir(n, 3, 0, M, 0, a[2], t1);
    apply_dir(n, 3, 1, M, 0, t1, t2);
    for (int i = 0; i < n3; ++i) out[i] += t2[i];
  }
}

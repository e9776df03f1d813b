/*
 * CPU oracle for the matrix-free sum-factorization operators.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path
 * is compared against; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The shipped library
 * (libsfk.so) never links, loads or calls it.
 *
 * It restates, in plain C with runtime polynomial degree, the algorithms of
 * the reference headers (paths relative to the reference checkout):
 *   legacy/include/gauss_formula.h         quadrature points / weights
 *   legacy/include/lagrange_polynomials.h  Lagrange and Hermite-like bases
 *   legacy/include/matrix_vector_kernel.h  1D even-odd kernel
 *   legacy/include/evaluation_cell_laplacian.h   element-local Laplacian
 *   legacy/include/evaluation_dg_laplacian.h     SIPG + Chebyshev update
 *   legacy/include/evaluation_dg_advect.h        upwind advection + LSRK
 * plus the operators the reference does not contain (continuous Q_p with
 * homogeneous Dirichlet rows, per-quadrature-point metrics, variable
 * coefficient, variable velocity with adjoint and gradients).  Every entry
 * point states which checks pin it (tests/test_oracle_*.py).
 *
 * Layout ("canonical"): vector index = cell * n^3 + dof,
 * cell = (iz*ny + iy)*nx + ix, dof = (k*n + j)*n + i (x fastest).
 */
#ifndef SFK_ORACLE_H
#define SFK_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define SFKO_MAXN 16

/* ---- quadrature (gauss_formula.h:35-175) ---- */
void sfko_gauss_points(int n, double *pts);
void sfko_gauss_weights(int n, double *w);
void sfko_gll_points(int n, double *pts);
double sfko_jacobi_polynomial(double x, int alpha, int beta, int n);
/* writes degree-alpha-beta roots on [-1,1] */
void sfko_zeros_of_jacobi_polynomial(int degree, int alpha, int beta, double *roots);

/* ---- bases (lagrange_polynomials.h:31-274) ---- */
double sfko_lagrange_value(int npts, const double *pts, int i, double x);
double sfko_lagrange_derivative(int npts, const double *pts, int i, double x);
/* Hermite-like basis of the given degree: value/derivative of function i */
double sfko_hermite_value(int degree, int i, double x);
double sfko_hermite_derivative(int degree, int i, double x);

/* ---- 1D tables ----
 * basis: 0 = Hermite-like (evaluation_cell_laplacian.h:822,
 *            evaluation_dg_laplacian.h:621), 1 = Lagrange on Gauss-Lobatto
 *            points (evaluation_dg_advect.h:737-740).
 * Dense forms (row = basis function / Gauss-Lagrange polynomial, column =
 * point):  S[i*n+q] = phi_i(x_q), D[i*n+q] = l_i'(x_q) (collocation
 * derivative of the Gauss-point Lagrange basis), Sinv[i*n+g] = l_i(y_g)
 * (Gauss Lagrange at GLL points), fv0/fv1/fd0/fd1[i] = l_i(0), l_i(1),
 * l_i'(0), l_i'(1), w[q] Gauss weights, hd = phi_0'(0) of the Hermite basis.
 * Even-odd forms exactly as the reference stores them:
 *   values_eo, gradients_eo, inv_values_eo: n*ceil(n/2),
 *   face_eo: 2n (evaluation_dg_laplacian.h:615-659). */
typedef struct {
  int degree, n, basis;
  double pts[SFKO_MAXN], w[SFKO_MAXN], gll[SFKO_MAXN];
  double S[SFKO_MAXN * SFKO_MAXN], D[SFKO_MAXN * SFKO_MAXN], Sinv[SFKO_MAXN * SFKO_MAXN];
  double fv0[SFKO_MAXN], fv1[SFKO_MAXN], fd0[SFKO_MAXN], fd1[SFKO_MAXN];
  double hd;
  double values_eo[SFKO_MAXN * SFKO_MAXN], gradients_eo[SFKO_MAXN * SFKO_MAXN];
  double inv_values_eo[SFKO_MAXN * SFKO_MAXN], face_eo[2 * SFKO_MAXN];
} sfko_tables;

int sfko_fill_tables(int degree, int basis, sfko_tables *t);
/* flat export for ctypes: sizes n*h, n*h, n*h, 2n, 1 with h=(n+1)/2 */
int sfko_export_eo_tables(int degree, int basis, double *values_eo, double *gradients_eo,
                          double *inv_values_eo, double *face_eo, double *hermite_d);

/* ---- 1D even-odd kernel (matrix_vector_kernel.h:30-185) ----
 * Runtime restatement of apply_1d_matvec_kernel<nn,stride,type,
 * contract_over_rows,add_into,double,double,false,do_dg>.  add may be NULL
 * when add_into == 0; dg/face may be NULL when do_dg == 0. */
void sfko_apply_1d_eo(int nn, int stride, int type, int contract_over_rows, int add_into,
                      int do_dg, const double *coeffs_eo, const double *in, double *out,
                      const double *add, const double *dg, double *face);

/* ---- element-local cell Laplacian (evaluation_cell_laplacian.h:216-503)
 * Cartesian: inv_jac[d] = 1/h_d, JxW = w_q / (inv_jac0*inv_jac1*inv_jac2). */
void sfko_cell_laplacian(int degree, int basis, long n_cells, const double inv_jac[3],
                         const double *src, double *dst);

/* ---- DG-SIPG, Cartesian periodic (evaluation_dg_laplacian.h:129-510) ----
 * Correct periodic wrap in x (the reference's last-cell x-neighbour index
 * at evaluation_dg_laplacian.h:270 is not reproduced). */
void sfko_dg_sipg(int degree, const int nc[3], const double inv_jac[3], const double *src,
                  double *dst);
/* ref_x_sweep = 1 reproduces the reference's last x-direction sweep, which
 * contracts over rows (S^T u instead of S u, evaluation_dg_laplacian.h
 * :481-483) and makes the reference operator non-symmetric; 0 is the SIPG
 * operator (symmetric, A*1 = 0).  Only this flag differs between the two. */
void sfko_dg_sipg_ex(int degree, const int nc[3], const double inv_jac[3], const double *src,
                     double *dst, int ref_x_sweep);
/* fused Chebyshev update (evaluation_dg_laplacian.h:485-502):
 *   r = A sol_old - rhs; tm = c_tm*tm + c_np*diag*r; sol_new = sol_old - tm */
void sfko_dg_sipg_cheby(int degree, const int nc[3], const double inv_jac[3], const double *rhs,
                        const double *sol_old, const double *diag, double *sol_new, double *tm,
                        double c_np, double c_tm);
void sfko_dg_sipg_cheby_ex(int degree, const int nc[3], const double inv_jac[3], const double *rhs,
                           const double *sol_old, const double *diag, double *sol_new, double *tm,
                           double c_np, double c_tm, int ref_x_sweep);

/* ---- DG upwind advection (evaluation_dg_advect.h:285-676) ----
 * dst = M^{-1} [ (grad v, a u) - <v, (a.n) u_upwind> ], GLL basis, constant
 * velocity a, Cartesian periodic. */
void sfko_dg_advect(int degree, const int nc[3], const double inv_jac[3], const double vel[3],
                    const double *src, double *dst);
/* LSRK stage (evaluation_dg_advect.h:651-667):
 *   r = M^{-1}L(src); np = upd + c_np*r; if (c_tm != 0) tm = np + c_tm*r */
void sfko_dg_advect_stage(int degree, const int nc[3], const double inv_jac[3],
                          const double vel[3], const double *src, const double *upd,
                          double *vec_np, double *vec_tm, double c_np, double c_tm);

/* ---- extended operators (sfk_oracle_ext.c) ---- */
void sfko_cont_laplacian(int degree, const int nc[3], const double inv_jac[3], const double *coef,
                         const double *src, double *dst);
void sfko_cont_laplacian_coef_grad(int degree, const int nc[3], const double inv_jac[3], int n_rhs,
                                   const double *u_all, const double *lam_all, double *grad);
void sfko_dg_sipg_general(int degree, const int nc[3], const double *cell_metric, const double *face_metric,
                          const double *src, double *dst);
void sfko_cartesian_metric(int degree, const int nc[3], const double inv_jac[3], double *cell_metric,
                           double *face_metric);
void sfko_dg_advect_var(int degree, const int nc[3], const double inv_jac[3], const double *cell_vel,
                        const double *face_vel, const double *src, double *dst);
void sfko_dg_mass_inverse(int degree, const int nc[3], const double inv_jac[3], const double *src, double *dst);
void sfko_dg_advect_var_transpose(int degree, const int nc[3], const double inv_jac[3], const double *cell_vel,
                                  const double *face_vel, const double *src, double *dst);
void sfko_dg_advect_velocity_grad(int degree, const int nc[3], const double inv_jac[3], const double *face_vel,
                                  const double *u, const double *lambda, double *grad_cell, double *grad_face);

/* ---- PrecomputedDGLaplacian (precomputed_dg_laplacian.h:96-377) ----
 * 1D mass / Laplace(+penalty) matrices [i*n+j] and the neighbour coupling
 * v = {v1, v2}; ref_coupling selects the reference's (vanishing) coupling
 * constants instead of the SIPG ones. */
void sfko_precomp_1d(int degree, int ref_coupling, double *mass, double *lap, double *v);
void sfko_dg_precomputed(int degree, const int nc[3], const double inv_jac[3], int ref_coupling,
                         const double *src, double *dst);

/* ---- layout converters (vectorization.h lane interleaving,
 * evaluation_dg_laplacian.h:162-166): interleaved index
 *   ((iz*ny+iy)*(nx/L) + ixb)*L*n^3 + dof*L + lane,  ix = ixb*L + lane. */
void sfko_interleaved_to_canonical(int degree, const int nc[3], int lanes, const double *in,
                                   double *out);
void sfko_canonical_to_interleaved(int degree, const int nc[3], int lanes, const double *in,
                                   double *out);

#ifdef __cplusplus
}
#endif
#endif

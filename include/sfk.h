/*
 * sfk.h -- C ABI of the MI355X (gfx950) matrix-free sum-factorization
 * operators.  Plain C: handles, plain pointers, sizes and status codes; no
 * HIP, torch or C++ types cross this boundary (a HIP stream is passed as
 * `void *`, NULL = the default stream).
 *
 * Reference interfaces replaced (paths in the reference checkout):
 *   EvaluationCellLaplacian<dim,degree,Number>   legacy/include/evaluation_cell_laplacian.h:26-908
 *   EvaluationDGLaplacian<dim,degree,Number>     legacy/include/evaluation_dg_laplacian.h:33-716
 *   EvaluationDGAdvection<dim,degree,Number>     legacy/include/evaluation_dg_advect.h:96-845
 *   apply_1d_matvec_kernel<...>                  legacy/include/matrix_vector_kernel.h:30-185
 * The reference has no FFI: its "plugin API" is those class templates.
 * The C++ mirror classes in sum_factorization_kernels_amd/include/ keep
 * their names and methods and call only this ABI.
 *
 * Vector layout ("canonical"): a vector of an operator on n_cells[0] x
 * n_cells[1] x n_cells[2] cells holds cell-major blocks of (p+1)^3 values,
 * cell = (iz*ny + iy)*nx + ix, dof = (k*(p+1) + j)*(p+1) + i (x fastest).
 * Continuous Q_p vectors hold the (nx*p+1)(ny*p+1)(nz*p+1) nodes of the
 * structured grid, x fastest.  Converters to and from the reference's
 * lane-interleaved layout: sfk_layout_*.
 *
 * All device pointers are caller-owned device memory (hipMalloc, torch).
 * A handle is not thread-safe; use one handle per host thread.  Every
 * entry point returns an sfk_status; sfk_last_error() returns a
 * thread-local description of the last failure.
 */
#ifndef SFK_H
#define SFK_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  SFK_OK = 0,
  SFK_ERR_INVALID_ARG = 1,  /* bad degree, size, pointer or descriptor field */
  SFK_ERR_UNSUPPORTED = 2,  /* combination not built into this library */
  SFK_ERR_HIP = 3,          /* a HIP runtime call failed */
  SFK_ERR_ALLOC = 4,        /* device allocation failed */
  SFK_ERR_NONFINITE = 5,    /* non-finite value detected (solver) */
  SFK_ERR_NOT_READY = 6,    /* required data (metric, coefficient) not set */
  SFK_ERR_NO_CONVERGENCE = 7
} sfk_status;

typedef enum {
  SFK_CELL_LAPLACE = 0, /* element-local Laplacian, EvaluationCellLaplacian */
  SFK_CONT_LAPLACE = 1, /* continuous Q_p Laplacian, homogeneous Dirichlet  */
  SFK_DG_SIPG = 2,      /* DG symmetric interior penalty, periodic          */
  SFK_DG_ADVECT = 3,    /* DG upwind advection M^{-1}L(u), periodic         */
  SFK_DG_PRECOMP = 4    /* DG Laplacian from precomputed 1D matrices,
                           PrecomputedDGLaplacian (precomputed_dg_laplacian.h
                           :96-377): Hermite-like basis, Cartesian, periodic  */
} sfk_op_kind;

typedef enum { SFK_FP64 = 0, SFK_FP32 = 1 } sfk_precision;
typedef enum { SFK_BASIS_HERMITE = 0, SFK_BASIS_GLL = 1 } sfk_basis;
typedef enum { SFK_GEOM_CARTESIAN = 0, SFK_GEOM_PER_Q = 1 } sfk_geometry;

/* descriptor flags */
#define SFK_FLAG_REF_X_SWEEP 1u      /* DG-SIPG: reproduce the reference's last x sweep
                                        (evaluation_dg_laplacian.h:481-483, S^T) */
#define SFK_FLAG_VARIABLE_COEF 2u    /* continuous Laplacian: -div(k grad u), k per q-point */
#define SFK_FLAG_VARIABLE_VELOCITY 4u /* advection: velocity per cell and face q-point */
#define SFK_FLAG_REF_PRECOMP_COUPLING 8u /* SFK_DG_PRECOMP: the reference's neighbour
                                        coupling constants (precomputed_dg_laplacian.h
                                        :331-340), which evaluate the basis at the
                                        opposite ends and vanish (the operator is then
                                        block-diagonal per cell); default: the SIPG
                                        coupling */

typedef struct sfk_op_desc {
  int kind;          /* sfk_op_kind */
  int degree;        /* polynomial degree p, 1..15 (precomputed DG: 1..8) */
  int precision;     /* sfk_precision: arithmetic and vector element type */
  int basis;         /* sfk_basis (SFK_CELL_LAPLACE); DG-SIPG is Hermite-like,
                        advection and continuous Q_p are GLL */
  int geometry;      /* sfk_geometry */
  unsigned flags;    /* SFK_FLAG_* */
  int n_cells[3];    /* cells per direction (SFK_CELL_LAPLACE: {n, 1, 1}) */
  double inv_jac[3]; /* Cartesian: inverse Jacobian diagonal 1/h_d */
  double velocity[3];/* advection: constant velocity (without the flag) */
  int n_rhs;         /* right-hand sides per apply (contiguous vectors), >= 1 */
} sfk_op_desc;

typedef struct sfk_op sfk_op;

/* Replaces the evaluators' initialize() (evaluation_dg_laplacian.h:43,
 * evaluation_cell_laplacian.h:34, evaluation_dg_advect.h:107): builds the
 * 1D tables (fill_shape_values) and the launch plan. */
sfk_status sfk_op_create(const sfk_op_desc *desc, sfk_op **op);
sfk_status sfk_op_destroy(sfk_op *op);
/* length (elements) of one vector of this operator */
sfk_status sfk_op_vector_size(const sfk_op *op, long long *n);
/* number of quadrature points of one cell / all cells, faces per direction */
sfk_status sfk_op_qdata_sizes(const sfk_op *op, long long *n_cell_q, long long *n_face_q);

/* Geometry / coefficient data (host arrays, copied to device and kept by the
 * handle; stored at the handle's precision):
 *  cell metric, SFK_GEOM_PER_Q: 6 values per cell q-point, layout
 *    [cell][component][q], components (G00,G11,G22,G01,G02,G12) of the
 *    symmetric G = JxW * J^{-1} J^{-T} (JxW including the quadrature weight);
 *  face metric, SFK_GEOM_PER_Q (DG-SIPG): for every face direction d and
 *    every cell c, the face between c and its upper neighbour in d:
 *    7 values per face q-point, layout [d][cell][component][q], components
 *    (aL0,aL1,aL2, aU0,aU1,aU2, sigma) with aX = JxW_f J_X^{-1} n,
 *    n the unit normal pointing from c to its neighbour, JxW_f the face
 *    Jacobian times the quadrature weight and sigma = penalty * JxW_f;
 *  coefficient (SFK_FLAG_VARIABLE_COEF): k per cell q-point [cell][q];
 *  velocity (SFK_FLAG_VARIABLE_VELOCITY): cell part [cell][3][q]
 *    (a_x,a_y,a_z at cell q-points), face part [d][cell][q] = a.n at the
 *    upper face of cell c in direction d. */
sfk_status sfk_op_set_cell_metric(sfk_op *op, const double *host, size_t count);
sfk_status sfk_op_set_face_metric(sfk_op *op, const double *host, size_t count);
sfk_status sfk_op_set_coefficient(sfk_op *op, const double *host, size_t count);
sfk_status sfk_op_set_velocity(sfk_op *op, const double *cell_host, size_t cell_count,
                               const double *face_host, size_t face_count);
/* Fills the cell (and for DG-SIPG face) metric of the mapping
 *   x = xhat + amp * sin(pi xhat) sin(pi yhat) sin(pi zhat) * (1,1,1)
 * of the uniform mesh of [0,1]^3 on the device (setup, not timed).
 * Penalty: (p+1)^2 * |F| / |K| averaged over both sides. */
sfk_status sfk_op_set_sine_deformation(sfk_op *op, double amp);
/* Device-side forms of sfk_op_set_coefficient / sfk_op_set_velocity: the
 * arrays are device memory in the handle's precision (same layouts and
 * counts), copied stream-ordered on `stream` (for callers whose
 * coefficients live on the GPU, e.g. inside an autodiff graph). */
sfk_status sfk_op_set_coefficient_device(sfk_op *op, const void *dev, size_t count, void *stream);
sfk_status sfk_op_set_velocity_device(sfk_op *op, const void *cell_dev, size_t cell_count,
                                      const void *face_dev, size_t face_count, void *stream);

/* dst = A src (n_rhs vectors).  Replaces matrix_vector_product()
 * (evaluation_cell_laplacian.h:216), do_matvec() (evaluation_dg_laplacian.h
 * :121, evaluation_dg_advect.h:216). */
sfk_status sfk_apply(sfk_op *op, const void *src, void *dst, void *stream);
/* dst = A^T src: advection adjoint (downwind); symmetric operators apply A. */
sfk_status sfk_apply_transpose(sfk_op *op, const void *src, void *dst, void *stream);
/* Fused Chebyshev update of do_cheby_iter<true> (evaluation_dg_laplacian.h
 * :485-502, :512-554):  r = A sol_old - rhs;  tmp = c_tm*tmp + c_np*diag*r;
 * sol_new = sol_old - tmp.  Supported for DG-SIPG, the precomputed DG
 * Laplacian and the continuous Laplacian. */
sfk_status sfk_cheby_step(sfk_op *op, const void *rhs, const void *sol_old, const void *diag,
                          void *sol_new, void *tmp, double c_np, double c_tm, void *stream);
/* Fused low-storage RK stage of do_rk_stage<true> (evaluation_dg_advect.h
 * :651-667, :678-715):  r = M^{-1}L(src);  vec_np = upd + c_np*r;
 * if (c_tm != 0) vec_tm = vec_np + c_tm*r  (vec_tm untouched otherwise, so
 * it may alias src). */
sfk_status sfk_rk_stage(sfk_op *op, const void *src, const void *upd, void *vec_np, void *vec_tm,
                        double c_np, double c_tm, void *stream);
/* diag = diagonal of A (point-Jacobi); continuous Laplacian: 0 on Dirichlet
 * rows.  SFK_ERR_UNSUPPORTED for advection and for a DG-SIPG or precomputed
 * operator with one cell in some direction (the cell is then its own periodic
 * neighbour). */
sfk_status sfk_compute_diagonal(sfk_op *op, void *diag, void *stream);

/* Gradients (no reference counterpart):
 * advection: with r = A u - f, J = 0.5 |r|^2, writes dJ/d(velocity) into
 *   grad_cell [cell][3][q] and grad_face [d][cell][q] (layouts of
 *   sfk_op_set_velocity); lambda = A u - f is passed in by the caller;
 * variable-coefficient Laplacian: g_q = JxW_q grad u_q . grad lambda_q
 *   (JxW_q including the quadrature weight) summed over the n_rhs pairs,
 *   i.e. d(lambda^T A(k) u)/dk_q; for an objective J(u) with A(k) u = f and
 *   the adjoint solution lambda of A lambda = dJ/du, dJ/dk_q = -g_q. */
sfk_status sfk_grad_velocity(sfk_op *op, const void *u, const void *lambda, void *grad_cell,
                             void *grad_face, void *stream);
sfk_status sfk_grad_coefficient(sfk_op *op, const void *u, const void *lambda, void *grad_k,
                                void *stream);

/* b_i = integral of f phi_i (continuous Laplacian; Dirichlet rows 0) */
sfk_status sfk_op_rhs_constant(sfk_op *op, double f, void *b, void *stream);

/* ---- workgroup schedule and autotuner ----------------------------------
 * Replaces the hand-chosen loop blocking of the reference drivers
 * (evaluator.blx/bly/blz, legacy/src/test_dg_integrate.cc:64-66,
 * legacy/src/test_dg_advect.cc): on the GPU the schedule of an operator
 * kernel is the number of cells per workgroup, the order in which
 * workgroups walk the mesh and, for the continuous Laplacian, the scatter
 * strategy.  Results do not depend on the schedule (bit-identical for DG
 * and the colored continuous scatter; the atomic scatter sums in a
 * run-dependent order). */
typedef struct sfk_schedule {
  int cells_per_block; /* cells per workgroup; 0 = the degree's default */
  int block_order;     /* 0 linear, 1 one contiguous range of cells per XCD,
                          2 as 1 with the DG kernels walking slabs of 4
                          z-layers x, z, y (neighbour data re-read from L2) */
  int cont_scatter;    /* continuous Laplacian: -1 automatic, 0 zero fill +
                          atomic adds, 1 colored stores (8 launches), 2 the
                          tensor-product kernel (constant coefficient only:
                          assembled 1D mass/stiffness sweeps on the node
                          grid, no scatter; the automatic choice then) */
  int tile_shape;      /* cont_scatter 2: tile of cells in x/y per workgroup,
                          an index < the count sfk_op_tile_shape returns
                          (0 = default) */
  int z_cells;         /* cont_scatter 2: cells in z per workgroup (0 =
                          automatic) */
} sfk_schedule;

/* Current schedule, with defaults resolved (cells_per_block > 0). */
sfk_status sfk_op_get_schedule(const sfk_op *op, sfk_schedule *s);
/* Schedule of the fused Chebyshev launches (sfk_cheby_step and the smoother
 * of sfk_solve_cg_cheby on this operator) of the tensor-product kernel (tile
 * shape, z_cells, order) and of the per-q-metric SIPG kernel (cells per
 * block, order): the one sfk_tune picked for them, else the matvec
 * schedule. */
sfk_status sfk_op_get_smoother_schedule(const sfk_op *op, sfk_schedule *s);
/* Sets the schedule (of the matvec and of the Chebyshev launches);
 * SFK_ERR_INVALID_ARG if cells_per_block exceeds
 * sfk_op_max_cells_per_block; SFK_ERR_UNSUPPORTED if the tile shape needs
 * more LDS than a workgroup has in the operator's precision (the widest Q6
 * tile fits FP32 only; sfk_tune skips such shapes). */
sfk_status sfk_op_set_schedule(sfk_op *op, const sfk_schedule *s);
/* Largest cells_per_block of the operator (workgroup of at most 256 threads,
 * LDS within the device's per-workgroup limit). */
sfk_status sfk_op_max_cells_per_block(const sfk_op *op, int *max_cells);
/* Number of tile shapes of the continuous tensor-product kernel and the
 * shape's cells in x and y (bx, by may be NULL). */
sfk_status sfk_op_tile_shape(const sfk_op *op, int shape, int *count, int *bx, int *by);
/* Times sfk_apply for every candidate schedule (`repeats` applies each,
 * after a warm-up apply, on internal scratch vectors; stream-ordered on
 * `stream`), sets the fastest on the handle and returns it in *best with its
 * time per apply in *best_seconds (either may be NULL).  Candidates: every
 * cells_per_block in [1, max] x both block orders x, for the continuous
 * Laplacian, both scatters, and with a constant coefficient every tile shape
 * x z_cells (powers of two up to the cells in z) of the tensor-product
 * kernel; for that kernel the fused Chebyshev launch (three-term form, four
 * vector streams) is timed over the same tile shapes, z_cells and orders and
 * its fastest kept separately (sfk_op_get_smoother_schedule), and for the
 * per-q-metric SIPG kernel the fused Chebyshev step over cells_per_block and
 * the three orders, kept the same way.  Requires the
 * operator's data (metric, coefficient, velocity) to be set.
 * Tuning cache: when the environment variable SFK_TUNE_CACHE names a file,
 * an operator whose key (device architecture and CU count, kind, degree,
 * precision, basis, geometry, flags, cells, right-hand sides) has a line
 * there gets that line's schedules and time without any timing run (a line
 * that no longer validates is ignored); otherwise the result of the timing
 * runs is appended to the file. */
sfk_status sfk_tune(sfk_op *op, int repeats, sfk_schedule *best, double *best_seconds, void *stream);

/* ---- solver (continuous Laplacian) ------------------------------------
 * Preconditioned CG in FP64 with a Chebyshev(degree) point-Jacobi
 * preconditioner whose bounds come from a Lanczos estimate of
 * lambda_max(D^{-1}A) (lanczos_steps FP64 steps, upper bound *= 1.1 (safety), lower =
 * upper/1000: a preconditioner acting on most of the spectrum).  The preconditioner runs on `smoother_op` (FP32 or FP64 operator
 * on the same mesh); the residual, inner products, Lanczos and the final
 * residual stay FP64 on `op` (which must be FP64).  x: initial guess in,
 * solution out; b: right-hand side (FP64, Dirichlet rows 0). */
typedef struct sfk_solve_stats {
  int iterations;
  double initial_residual, final_residual; /* ||b - A x||_2 (FP64) */
  double lambda_max, lambda_min;           /* Chebyshev bounds used */
  double seconds;                          /* wall time of the solve */
  int matvecs_outer, matvecs_inner;
} sfk_solve_stats;

sfk_status sfk_solve_cg_cheby(sfk_op *op, sfk_op *smoother_op, const double *b, double *x,
                              double rel_tol, int max_iter, int cheby_degree, int lanczos_steps,
                              sfk_solve_stats *stats, void *stream);

/* ---- Chebyshev smoother sweep -----------------------------------------
 * `degree` fused Chebyshev steps (do_cheby_iter<true>,
 * evaluation_dg_laplacian.h:485-502, :512-554) with the coefficients of the
 * Chebyshev iteration for D^{-1}A on [lambda_min, lambda_max] (Saad,
 * Iterative Methods for Sparse Linear Systems, Alg. 12.1):
 *   step 0: tmp = (1/theta) D^{-1}(A x - rhs), x -= tmp
 *   step j: tmp = rho_j rho_{j-1} tmp + (2 rho_j/delta) D^{-1}(A x - rhs), x -= tmp
 * with theta, delta the centre and half-width of the interval.  x is updated
 * in place (spare: a second vector of the same size; tmp: work vector);
 * diag_inv = D^{-1}.  DG-SIPG and the Laplacians. */
sfk_status sfk_cheby_sweep(sfk_op *op, const void *rhs, void *x, void *spare, const void *diag_inv,
                           void *tmp, int degree, double lambda_min, double lambda_max, void *stream);
/* No counterpart in the reference: the Chebyshev preconditioner of
 * sfk_solve_cg_cheby, z = P(rhs) = the result of sfk_cheby_sweep from x = 0
 * with the same degree and bounds, in the three-term form of the
 * tensor-product kernel (four vector streams per step, D^{-1} formed in the
 * kernel, the first iterate never stored; no diagonal vector needed).
 * Continuous Laplacian with a constant coefficient (degree <= 8), n_rhs 1,
 * degree >= 2; spare: a second work vector (distinct from z). */
sfk_status sfk_cheby_precondition(sfk_op *op, const void *rhs, void *z, void *spare, int degree,
                                  double lambda_min, double lambda_max, void *stream);

/* ---- HIP graphs --------------------------------------------------------
 * Captures the sfk calls issued on a (non-default) stream between
 * sfk_graph_begin and sfk_graph_end into an executable graph that
 * sfk_graph_launch replays with the same pointers and scalars (e.g. a
 * Chebyshev sweep or the five stages of an LSRK time step).  Issue every
 * call once before capturing: an operator allocates its scratch space on
 * first use.  Graph nodes are kernels and device copies only (no memset
 * nodes). */
typedef struct sfk_graph sfk_graph;
sfk_status sfk_stream_create(void **stream); /* default priority */
sfk_status sfk_stream_destroy(void *stream);
sfk_status sfk_graph_begin(void *stream);
sfk_status sfk_graph_end(void *stream, sfk_graph **graph);
sfk_status sfk_graph_launch(const sfk_graph *graph, void *stream);
sfk_status sfk_graph_destroy(sfk_graph *graph);

/* ---- 1D kernel (a standalone batched form of apply_1d_matvec_kernel) ---
 * Applies the 1D even-odd kernel to n_lines lines of nn values, line l at
 * in + l*line_stride with element stride `stride` (device FP64 arrays);
 * dense matrix M[i*nn+q] (row = basis function i, column = point q),
 * symmetry: M[nn-1-i][nn-1-q] = sym * M[i][q], sym = +1 or -1.
 * forward: out[q] = sum_i M[i][q] in[i]; backward: out[i] = sum_q M[i][q] in[q]. */
sfk_status sfk_apply_1d(int nn, const double *matrix_host, int sym, int forward, long long n_lines,
                        long long line_stride, int stride, const double *in, double *out,
                        void *stream);

/* ---- 1D tables (fill_shape_values, evaluation_dg_laplacian.h:615-695) ----
 * Host copies of the tables this library uses, in dense form: S[i*n+q]
 * basis at Gauss points, D[i*n+q] collocation derivative, Sinv[i*n+g] Gauss
 * Lagrange at GLL points, w[q] Gauss weights, face[4n] = l_i(0), l_i(1),
 * l_i'(0), l_i'(1), hermite_d = phi_0'(0). */
sfk_status sfk_tables(int degree, int basis, double *S, double *D, double *Sinv, double *w,
                      double *face, double *hermite_d);
sfk_status sfk_quadrature(int n, double *gauss_points, double *gauss_weights, double *gll_points);

/* ---- layout converters (lane-interleaved VectorizedArray layout of the
 * reference, evaluation_dg_laplacian.h:162-166) -- device FP64 arrays ---- */
sfk_status sfk_layout_to_canonical(int degree, const int n_cells[3], int lanes, const double *in,
                                   double *out, void *stream);
sfk_status sfk_layout_to_interleaved(int degree, const int n_cells[3], int lanes, const double *in,
                                     double *out, void *stream);

/* ---- device memory (for C/C++ callers without a HIP toolchain) ---- */
sfk_status sfk_vec_alloc(size_t bytes, void **ptr);
sfk_status sfk_vec_free(void *ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device (synchronous) */
sfk_status sfk_memcpy(void *dst, const void *src, size_t bytes, int kind);
sfk_status sfk_synchronize(void *stream);

/* ---- vector helpers on device arrays (FP64 / FP32 per `precision`) ---- */
sfk_status sfk_vec_fill(int precision, void *x, long long n, double value, void *stream);
sfk_status sfk_vec_dot(int precision, const void *x, const void *y, long long n, double *result,
                       void *stream);
/* Allocation-free form of sfk_vec_dot.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  *result_dev is
 * device memory and receives x.y (FP64-accumulated), stream-ordered, with no
 * synchronisation.  The partial sums live in a static device array of 4096
 * doubles per device (one per chunk of 65536 elements), so n is limited to
 * 4096 * 65536 = 268435456 elements (larger n: SFK_ERR_INVALID_ARG, nothing
 * launched).  Calls that overlap on one device race on that array; the
 * result depends on n only (fixed chunks), not on the launch. */
sfk_status sfk_vec_dot_noalloc(int precision, const void *x, const void *y, long long n, double *result_dev,
                               void *stream);
/* unfused Chebyshev vector update (EvaluationDGLaplacian::
 * emulate_cheby_vector_updates, evaluation_dg_laplacian.h:556-601, with
 * au = A sol_old supplied by the caller):
 *   tmp = c_tm*tmp + c_np*diag*(au - rhs);  sol_new = sol_old - tmp
 * au may alias sol_new (elementwise). */
sfk_status sfk_vec_cheby_update(int precision, long long n, const void *au, const void *rhs,
                                const void *sol_old, const void *diag, void *sol_new, void *tmp,
                                double c_np, double c_tm, void *stream);

const char *sfk_last_error(void);
const char *sfk_version(void);

#ifdef __cplusplus
}
#endif
#endif

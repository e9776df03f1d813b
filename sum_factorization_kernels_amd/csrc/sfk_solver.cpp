// Chebyshev-preconditioned CG for the continuous Q_p Laplacian, all on the
// device (the reference has no solver).
//
// PCG in FP64 on `op` (iterate, residual, search direction, inner products),
// preconditioned by a Chebyshev(k) point-Jacobi polynomial applied on the
// smoother operator: FP64, or FP32 with FP32 vectors (the residual is
// rounded to FP32 on entry and the correction widened to FP64 on exit), so
// k-1 of the k matvecs per iteration run in the smoother's precision.
// Chebyshev bounds: lambda_max of D^{-1}A from `lanczos_steps` FP64 Lanczos
// steps on D^{-1/2} A D^{-1/2} (largest Ritz value by bisection on the
// tridiagonal matrix), times 1.1; lambda_min = lambda_max / 1000.
// The Chebyshev recurrence (Saad, Iterative Methods for Sparse Linear
// Systems, Alg. 12.1) is written in the fused update form of
// EvaluationDGLaplacian::do_cheby_iter (legacy/include/evaluation_dg_laplacian.h
// :485-502): tmp = c_tm*tmp + c_np*D^{-1}(A z - r), z = z - tmp.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cmath>
#include <string>
#include <vector>

#include "sfk.h"
#include "sfk_internal.h"
#include "sfk_kernels.h"
#include "sfk_op.h"

namespace sfk {
namespace {

constexpr int kPartials = 2048;
constexpr int kSlots = 8; // device result slots after the partial sums
constexpr double kChebyRange = 1000.; // lambda_max / lambda_min of the Chebyshev interval

struct DevBuf {
  void *p = nullptr;
  ~DevBuf()
  {
    if (p) (void)hipFree(p);
  }
};
struct PinnedBuf {
  void *p = nullptr;
  ~PinnedBuf()
  {
    if (p) (void)hipHostFree(p);
  }
};

__global__ void inv_diag_kernel(const double *d, double *inv, float *inv32, double *isq, long long n)
{
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double v = d[i];
    const double iv = v != 0. ? 1. / v : 0.;
    inv[i] = iv;
    if (inv32) inv32[i] = (float)iv;
    isq[i] = v != 0. ? 1. / sqrt(v) : 0.;
  }
}

__global__ void scale_kernel(const double *a, const double *b, double *c, long long n)
{
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    c[i] = a[i] * b[i];
}

// deterministic start vector for Lanczos (interior only: the inverse
// diagonal square root is 0 on Dirichlet rows)
__global__ void start_vector_kernel(const double *isq, double *v, long long n)
{
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    unsigned long long h = (unsigned long long)i * 0x9E3779B97F4A7C15ULL;
    h ^= h >> 29;
    h *= 0xBF58476D1CE4E5B9ULL;
    h ^= h >> 32;
    v[i] = isq[i] != 0. ? (double)(h & 0xFFFFFF) / double(0xFFFFFF) - 0.5 : 0.;
  }
}

unsigned grid_of(long long n) { return (unsigned)std::min<long long>((n + 255) / 256, 65535LL * 8); }

// largest eigenvalue of the symmetric tridiagonal (alpha, beta) by bisection
double tridiag_max_eig(const std::vector<double> &al, const std::vector<double> &be)
{
  const int m = (int)al.size();
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < m; ++i) {
    const double r = (i > 0 ? std::fabs(be[i - 1]) : 0.) + (i + 1 < m ? std::fabs(be[i]) : 0.);
    lo = std::min(lo, al[i] - r);
    hi = std::max(hi, al[i] + r);
  }
  auto count_below = [&](double x) { // Sturm count of eigenvalues < x
    int cnt = 0;
    double q = 1.;
    for (int i = 0; i < m; ++i) {
      q = al[i] - x - (i > 0 ? be[i - 1] * be[i - 1] / q : 0.);
      if (q == 0.) q = 1e-300;
      if (q < 0.) ++cnt;
    }
    return cnt;
  };
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (count_below(mid) < m) lo = mid;
    else hi = mid;
  }
  return hi;
}

struct Solver {
  sfk_op *op, *sm;
  hipStream_t s;
  long long n;
  bool fp32;
  double *partials = nullptr, *dres = nullptr;
  double *hres = nullptr; // pinned host copy of the result slots
  sfk_solve_stats st{};

  sfk_status dot(bool f32, const void *x, const void *y, double &out)
  {
    hipError_t e = vec_dot(f32, x, y, n, dres, partials, kPartials, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&out, dres, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_error(std::string("solver: ") + hipGetErrorString(e));
      return SFK_ERR_HIP;
    }
    if (!std::isfinite(out)) {
      set_error("solver: non-finite inner product");
      return SFK_ERR_NONFINITE;
    }
    return SFK_OK;
  }

  // read back `count` sums that a fused kernel left in dres
  sfk_status results(double *out, int count)
  {
    hipError_t e = hipMemcpyAsync(out, dres, count * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_error(std::string("solver: ") + hipGetErrorString(e));
      return SFK_ERR_HIP;
    }
    for (int i = 0; i < count; ++i)
      if (!std::isfinite(out[i])) {
        set_error("solver: non-finite inner product");
        return SFK_ERR_NONFINITE;
      }
    return SFK_OK;
  }

  // read back `count` result slots (no finiteness check: the caller checks
  // the slots it uses)
  sfk_status read_slots(double *out, int count)
  {
    hipError_t e = hipMemcpyAsync(hres, dres, count * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_error(std::string("solver: ") + hipGetErrorString(e));
      return SFK_ERR_HIP;
    }
    for (int i = 0; i < count; ++i) out[i] = hres[i];
    return SFK_OK;
  }

  // (x . y, x . w) in one pass, FP64 vectors
  sfk_status dot2(const void *x, const void *y, const void *w, double &xy, double &xw)
  {
    hipError_t e = vec_dot2(x, y, w, n, dres, partials, kPartials, s);
    double h[2] = {0., 0.};
    if (e == hipSuccess) e = hipMemcpyAsync(h, dres, 2 * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_error(std::string("solver: ") + hipGetErrorString(e));
      return SFK_ERR_HIP;
    }
    xy = h[0];
    xw = h[1];
    if (!std::isfinite(xy) || !std::isfinite(xw)) {
      set_error("solver: non-finite inner product");
      return SFK_ERR_NONFINITE;
    }
    return SFK_OK;
  }
};

#define SOLVER_HIP(expr)                                              \
  do {                                                                \
    hipError_t e_ = (expr);                                           \
    if (e_ != hipSuccess) {                                           \
      set_error(std::string("solver: ") + hipGetErrorString(e_));     \
      return SFK_ERR_HIP;                                             \
    }                                                                 \
  } while (0)
#define SOLVER_ST(expr)          \
  do {                           \
    sfk_status s_ = (expr);      \
    if (s_ != SFK_OK) return s_; \
  } while (0)

// z = P(rr): the Chebyshev(degree >= 2) polynomial in D^{-1}A on [lmin,
// lmax] applied to rr, in the three-term form on the tensor-product kernel:
// z_1 = (1/theta) D^{-1} rr (z_0 = 0), z_{j+1} = z_j + c_tm (z_j - z_{j-1})
// + c_np D^{-1}(rr - A z_j).  z_1 is never stored: the first step forms it
// from rr as its input and the second as its previous iterate.  Step j
// writes z_{j+1} into bufs[(j + 1) % 2] (z_2 into *z_io, z_3 into *spare,
// then each over z_{j-1}: the kernel allows dst == prev); the result ends
// in *z_io (the two pointers are swapped when it lands in the other buffer).
hipError_t cheby_tp3(sfk_op *sm, const void *rr, void **z_io, void **spare, int degree, double lmin, double lmax,
                     hipStream_t s)
{
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma1 = theta / delta;
  void *bufs[2] = {*z_io, *spare};
  double rho_old = 1. / sigma1;
  for (int j = 1; j < degree; ++j) {
    const double rho = 1. / (2. * sigma1 - rho_old);
    TpLaunch T;
    fill_tp_launch(sm, T, true);
    T.cheby = 2;
    T.rhs = rr;
    T.s_rhs = 1. / theta;
    T.src = j == 1 ? rr : bufs[j % 2];
    T.src_from_rhs = j == 1;
    T.prev = j == 1 ? nullptr : (j == 2 ? rr : bufs[(j + 1) % 2]); // j == 1: the zero z_0
    T.prev_from_rhs = j == 2;
    T.dst = bufs[(j + 1) % 2];
    T.c_np = 2. * rho / delta;
    T.c_tm = rho * rho_old;
    const hipError_t e = launch_cont_tp(T, s);
    if (e != hipSuccess) return e;
    rho_old = rho;
  }
  if (degree % 2 == 1) std::swap(*z_io, *spare); // z_k is in bufs[k % 2]
  return hipSuccess;
}

} // namespace
} // namespace sfk

using namespace sfk;

extern "C" sfk_status sfk_op_rhs_constant(sfk_op *op, double f, void *b, void *stream);

extern "C" sfk_status sfk_solve_cg_cheby(sfk_op *op, sfk_op *smoother_op, const double *b, double *x, double rel_tol,
                                         int max_iter, int cheby_degree, int lanczos_steps, sfk_solve_stats *stats,
                                         void *stream)
{
  if (!op || !smoother_op || !b || !x || !stats) {
    set_error("sfk_solve_cg_cheby: null argument");
    return SFK_ERR_INVALID_ARG;
  }
  if (op->d.kind != SFK_CONT_LAPLACE || smoother_op->d.kind != SFK_CONT_LAPLACE || op->d.precision != SFK_FP64 ||
      op->d.n_rhs != 1 || smoother_op->d.n_rhs != 1 || op->nvec != smoother_op->nvec ||
      op->d.degree != smoother_op->d.degree) {
    set_error("sfk_solve_cg_cheby: op must be an FP64 continuous Laplacian and smoother_op the same mesh");
    return SFK_ERR_INVALID_ARG;
  }
  if (!(rel_tol > 0) || max_iter < 1 || cheby_degree < 1 || cheby_degree > 32 || lanczos_steps < 2 ||
      lanczos_steps > 200) {
    set_error("sfk_solve_cg_cheby: invalid tolerance, iteration count, degree or Lanczos steps");
    return SFK_ERR_INVALID_ARG;
  }
  const auto t0 = std::chrono::steady_clock::now();
  Solver S{op, smoother_op, (hipStream_t)stream, op->nvec, smoother_op->d.precision == SFK_FP32};
  const long long n = S.n;
  const bool f32 = S.fp32;
  const size_t n8 = (size_t)n * 8, ns = (size_t)n * (f32 ? 4 : 8);
  // device work space: FP64 vectors, and smoother-precision vectors
  DevBuf red, d64, inv64, isq, r, w, lv0, lv1, lw, lt, e64;
  DevBuf rs, zs, qs, tms, inv_s;
  // the outer matvec on the tensor-product kernel also sums p . q (its
  // workgroups' partial sums, one per workgroup).  It runs on the schedule
  // tuned for the operator's fused Chebyshev launches: like them it holds
  // more per-node state than the plain matvec, which the widest tiles
  // cannot hold without spills (Q6 FP64 22 x 1 cells: 128 VGPRs + spills)
  const bool fused_pq = uses_tp(op) && op->d.n_rhs == 1;
  long long n_red = 2 * kPartials;
  if (fused_pq) {
    TpLaunch T;
    fill_tp_launch(op, T, true);
    if (tp_launch_blocks(T) > n_red) n_red = tp_launch_blocks(T);
  }
  SOLVER_HIP(hipMalloc(&red.p, sizeof(double) * (n_red + kSlots)));
  S.partials = (double *)red.p;
  S.dres = S.partials + n_red; // after the partial sums: the result slots
  SOLVER_HIP(vec_fill(false, S.dres, kSlots, 0., (hipStream_t)stream));
  PinnedBuf hres;
  SOLVER_HIP(hipHostMalloc(&hres.p, sizeof(double) * kSlots));
  S.hres = (double *)hres.p;
  for (DevBuf *bb : {&d64, &inv64, &isq, &r, &w, &lv0, &lv1, &lw, &lt, &e64})
    SOLVER_HIP(hipMalloc(&bb->p, n8));
  for (DevBuf *bb : {&rs, &zs, &qs, &tms, &inv_s})
    SOLVER_HIP(hipMalloc(&bb->p, ns));
  const hipStream_t s = S.s;
  // every work vector starts at 0 (a fresh allocation may hold NaN bit
  // patterns, and 0 * NaN would poison the first Chebyshev step)
  for (DevBuf *bb : {&d64, &inv64, &isq, &r, &w, &lv0, &lv1, &lw, &lt, &e64})
    SOLVER_HIP(vec_fill(false, bb->p, n, 0., s));
  for (DevBuf *bb : {&rs, &zs, &qs, &tms, &inv_s})
    SOLVER_HIP(vec_fill(f32, bb->p, n, 0., s));
  // point-Jacobi diagonal (FP64), inverse in both precisions
  SOLVER_ST(sfk_compute_diagonal(op, d64.p, stream));
  hipLaunchKernelGGL(inv_diag_kernel, dim3(grid_of(n)), dim3(256), 0, s, (const double *)d64.p, (double *)inv64.p,
                     f32 ? (float *)inv_s.p : nullptr, (double *)isq.p, n);
  SOLVER_HIP(hipGetLastError());
  const void *inv_sm = f32 ? inv_s.p : inv64.p;

  // ---- Lanczos on D^{-1/2} A D^{-1/2} (FP64)
  hipLaunchKernelGGL(start_vector_kernel, dim3(grid_of(n)), dim3(256), 0, s, (const double *)isq.p,
                     (double *)lv1.p, n);
  double nrm2;
  SOLVER_ST(S.dot(false, lv1.p, lv1.p, nrm2));
  SOLVER_HIP(vec_axpby(false, n, 0., lv1.p, 1. / std::sqrt(nrm2), lv1.p, s));
  SOLVER_HIP(vec_fill(false, lv0.p, n, 0., s));
  std::vector<double> al, be;
  double beta = 0.;
  for (int k = 0; k < lanczos_steps; ++k) {
    hipLaunchKernelGGL(scale_kernel, dim3(grid_of(n)), dim3(256), 0, s, (const double *)lv1.p, (const double *)isq.p,
                       (double *)lt.p, n);
    SOLVER_ST(sfk_apply(op, lt.p, lw.p, stream));
    hipLaunchKernelGGL(scale_kernel, dim3(grid_of(n)), dim3(256), 0, s, (const double *)lw.p, (const double *)isq.p,
                       (double *)lw.p, n);
    ++S.st.matvecs_outer;
    double alpha;
    SOLVER_ST(S.dot(false, lw.p, lv1.p, alpha));
    SOLVER_HIP(vec_axpby(false, n, -alpha, lv1.p, 1., lw.p, s));
    SOLVER_HIP(vec_axpby(false, n, -beta, lv0.p, 1., lw.p, s));
    al.push_back(alpha);
    SOLVER_ST(S.dot(false, lw.p, lw.p, nrm2));
    beta = std::sqrt(nrm2);
    if (k + 1 == lanczos_steps || beta < 1e-14 * std::fabs(alpha)) break;
    be.push_back(beta);
    std::swap(lv0.p, lv1.p);
    SOLVER_HIP(vec_axpby(false, n, 1. / beta, lw.p, 0., lv1.p, s));
  }
  const double lmax = 1.1 * tridiag_max_eig(al, be);
  // the polynomial is a preconditioner inside CG, not a multigrid smoother:
  // it should act on most of the spectrum.  On the Q6 64^3 problem the FP64
  // iteration count at degree 6 falls from 268 (lambda_max / 30) to 243
  // (/100), 235 (/300) and 232 (/1000).
  const double lmin = lmax / kChebyRange;
  S.st.lambda_max = lmax;
  S.st.lambda_min = lmin;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma1 = theta / delta;

  // the smoother's Chebyshev steps in three-term form on the tensor-product
  // kernel (four vector streams per step instead of six, the inverse
  // diagonal formed in the kernel) when the smoother operator has it
  const bool tp3 = uses_tp(S.sm);
  if (tp3) SOLVER_ST(ensure_device_tables(S.sm));
  // z = P(rr) on the smoother operator; *z_io and *spare are two distinct
  // buffers, the result ends in *z_io (pointers swapped as needed)
  auto cheby = [&](const void *rr, void **z_io, void **spare) -> sfk_status {
    if (tp3) {
      if (cheby_degree == 1) {
        SOLVER_HIP(vec_mul_scaled(f32, n, 1. / theta, inv_sm, rr, *z_io, s));
        return SFK_OK;
      }
      SOLVER_HIP(cheby_tp3(S.sm, rr, z_io, spare, cheby_degree, lmin, lmax, s));
      S.st.matvecs_inner += cheby_degree - 1;
      return SFK_OK;
    }
    void *z = *z_io;
    // step 0: tm = (1/theta) D^{-1}(0 - rr), z = 0 - tm (null: the zero
    // vector, not read)
    SOLVER_HIP(vec_cheby_update(f32, n, nullptr, rr, nullptr, inv_sm, z, tms.p, 1. / theta, 0., s));
    double rho_old = 1. / sigma1;
    for (int j = 1; j < cheby_degree; ++j) {
      const double rho = 1. / (2. * sigma1 - rho_old);
      SOLVER_ST(sfk_cheby_step(S.sm, rr, z, inv_sm, *spare, tms.p, 2. * rho / delta, rho * rho_old, stream));
      std::swap(z, *spare);
      rho_old = rho;
      ++S.st.matvecs_inner;
    }
    *z_io = z;
    return SFK_OK;
  };

  // ---- residual of the initial guess (FP64)
  double bb;
  SOLVER_ST(S.dot(false, b, b, bb));
  const double bnorm = std::sqrt(bb);
  if (bnorm == 0.) {
    SOLVER_HIP(vec_fill(false, x, n, 0., s));
    *stats = S.st;
    return SFK_OK;
  }
  SOLVER_ST(sfk_apply(op, x, r.p, stream));
  ++S.st.matvecs_outer;
  SOLVER_HIP(vec_axpby(false, n, 1., b, -1., r.p, s));
  double rr;
  SOLVER_ST(S.dot(false, r.p, r.p, rr));
  S.st.initial_residual = std::sqrt(rr);
  double res = S.st.initial_residual;
  int iters = 0;

  // preconditioner z = P(r): the Chebyshev polynomial on the smoother
  // operator, in the smoother's precision.  FP32: rs already holds (float) r.
  // With zrq_dev (the inner products z.r and z.q wanted, left in
  // zrq_dev[0..1] on the device), an FP32 z stays FP32 (*z32; widened where
  // it is used), otherwise it is widened into e64.  *z64 receives the FP64
  // buffer that holds z (null when only *z32 does).
  auto precond = [&](const void *r64, const void *qv, void **z64, double *zrq_dev,
                     const void **z32 = nullptr) -> sfk_status {
    if (f32) {
      void *z = zs.p, *spare = qs.p;
      SOLVER_ST(cheby(rs.p, &z, &spare));
      if (zrq_dev) {
        SOLVER_HIP(vec_widen_dot2(n, z, nullptr, r64, qv, zrq_dev, S.partials, kPartials, s));
        *z64 = nullptr;
        *z32 = z;
        return SFK_OK;
      }
      SOLVER_HIP(vec_copy_convert(false, z, e64.p, n, s));
      *z64 = e64.p;
    } else {
      void *z = e64.p, *spare = zs.p;
      SOLVER_ST(cheby(r64, &z, &spare));
      *z64 = z;
      if (zrq_dev) SOLVER_HIP(vec_dot2(z, r64, qv, n, zrq_dev, S.partials, kPartials, s));
    }
    return SFK_OK;
  };

  // PCG in FP64 (x, r, p, q, z and all inner products FP64).  The FP32
  // polynomial preconditioner is linear only up to FP32 rounding, so the
  // search direction uses the flexible (Polak-Ribiere) coefficient
  // beta = z_new . (r_new - r_old) / (z_old . r_old) = -(z_new . q) / (p . q).
  // (Orthogonalising against the last two directions as well, flexible CG
  // with a truncation of two, did not reduce the iteration count on the
  // Q6 64^3 problem and costs two more vector passes per iteration.)
  // The step coefficients stay on the device: the kernels that use alpha =
  // (z.r) / (p.q) and beta read the inner products from the result slots,
  // and the host reads p.q, r.r and z.r back once per iteration (for the
  // convergence test), instead of waiting for every inner product.
  // Slots: 0 p.q, 1 r.r, 2 + 2 (k & 1) and 3 + 2 (k & 1) z.r and z.q of
  // the search direction of iteration k (the next iteration's pair is
  // written while this one's z.r is still needed for x += alpha p).
  const double target = rel_tol * bnorm;
  double *sl = S.dres;
  auto zr_slot = [&](int k) { return sl + 2 + 2 * (k & 1); };
  if (res > target) {
    void *z64 = nullptr, *pv = w.p, *qv = lt.p;
    void *rs32 = f32 ? rs.p : nullptr; // the CG update also writes (float) r
    // a (re)start from the residual in r.p: p = z = P(r), z.r into the slot
    // of iteration k
    auto restart = [&](int k) -> sfk_status {
      if (f32) SOLVER_HIP(vec_copy_convert(true, r.p, rs.p, n, s));
      SOLVER_ST(precond(r.p, nullptr, &z64, nullptr));
      SOLVER_HIP(hipMemcpyAsync(pv, z64, n8, hipMemcpyDeviceToDevice, s));
      SOLVER_HIP(vec_dot(false, r.p, z64, n, zr_slot(k), S.partials, kPartials, s));
      return SFK_OK;
    };
    SOLVER_ST(restart(0));
    while (iters < max_iter) {
      // q = A p and p.q -> slot 0
      if (fused_pq) {
        SOLVER_ST(ensure_device_tables(op));
        TpLaunch T;
        fill_tp_launch(op, T, true);
        T.src = pv;
        T.dst = qv;
        T.dot_part = S.partials;
        SOLVER_HIP(launch_cont_tp(T, s));
        SOLVER_HIP(vec_sum_partials(S.partials, tp_launch_blocks(T), sl, s));
      } else {
        SOLVER_ST(sfk_apply(op, pv, qv, stream));
        SOLVER_HIP(vec_dot(false, pv, qv, n, sl, S.partials, kPartials, s));
      }
      ++S.st.matvecs_outer;
      // r -= alpha q with alpha = (z.r) / (p.q) from the slots, (rs =
      // (float) r), r.r -> slot 1 in one pass; x += alpha p is folded into
      // the pass that forms the next search direction (which reads p
      // anyway), or done below when the loop may end
      double *zr = zr_slot(iters);
      SOLVER_HIP(vec_cg_update(n, 0., zr, sl, nullptr, qv, nullptr, r.p, rs32, sl + 1, S.partials, kPartials, s));
      double h[2 + 4];
      SOLVER_ST(S.read_slots(h, 6));
      const double pq = h[0], rz = h[2 + 2 * (iters & 1)];
      rr = h[1];
      if (!std::isfinite(pq) || !std::isfinite(rr) || !std::isfinite(rz)) {
        set_error("solver: non-finite inner product");
        return SFK_ERR_NONFINITE;
      }
      if (!(pq > 0.)) {
        set_error("sfk_solve_cg_cheby: operator not positive definite on the search direction");
        return SFK_ERR_NONFINITE;
      }
      const double alpha = rz / pq;
      ++iters;
      res = std::sqrt(rr);
      if (res <= target || iters == max_iter) SOLVER_HIP(vec_axpby(false, n, alpha, pv, 1., x, s));
      if (res <= target) {
        // the recursively updated residual drifts from b - A x: confirm
        // with the true residual and restart from it if it is not there yet
        SOLVER_ST(sfk_apply(op, x, r.p, stream));
        ++S.st.matvecs_outer;
        SOLVER_HIP(vec_axpby(false, n, 1., b, -1., r.p, s));
        SOLVER_ST(S.dot(false, r.p, r.p, rr));
        res = std::sqrt(rr);
        if (res <= target) break;
        SOLVER_ST(restart(iters));
        continue;
      }
      // z = P(r); z.r and z.q -> the slots of the next iteration
      const void *z32 = nullptr;
      double *zrq = zr_slot(iters);
      SOLVER_ST(precond(r.p, qv, &z64, zrq, &z32));
      const void *zv = z64 ? (const void *)z64 : z32;
      if (iters == max_iter) {
        // the loop ends: x is up to date; p = z + beta p for completeness
        double hz[2];
        SOLVER_HIP(hipMemcpyAsync(hz, zrq, sizeof hz, hipMemcpyDeviceToHost, s));
        SOLVER_HIP(hipStreamSynchronize(s));
        const double beta_cg = -hz[1] / pq;
        if (z64) SOLVER_HIP(vec_axpby(false, n, 1., z64, beta_cg, pv, s));
        else SOLVER_HIP(vec_axpby_widen(n, z32, beta_cg, pv, s));
      } else {
        // x += alpha p, p = z + beta p: alpha = (z.r) / (p.q) of this
        // iteration's direction, beta = -(z_new . q) / (p.q)
        SOLVER_HIP(vec_xp_update(z64 == nullptr, n, zr, sl, zrq + 1, zv, pv, x, s));
      }
    }
  }
  // final FP64 residual
  SOLVER_ST(sfk_apply(op, x, r.p, stream));
  ++S.st.matvecs_outer;
  SOLVER_HIP(vec_axpby(false, n, 1., b, -1., r.p, s));
  SOLVER_ST(S.dot(false, r.p, r.p, rr));
  SOLVER_HIP(hipStreamSynchronize(s));
  S.st.final_residual = std::sqrt(rr);
  S.st.iterations = iters;
  S.st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  *stats = S.st;
  if (!(S.st.final_residual <= target)) {
    char msg[160];
    snprintf(msg, sizeof msg, "sfk_solve_cg_cheby: no convergence after %d iterations (relative residual %.3e)",
             S.st.iterations, S.st.final_residual / bnorm);
    set_error(msg);
    return SFK_ERR_NO_CONVERGENCE;
  }
  return SFK_OK;
}

extern "C" sfk_status sfk_cheby_precondition(sfk_op *op, const void *rhs, void *z, void *spare, int degree,
                                             double lambda_min, double lambda_max, void *stream)
{
  if (!op || !rhs || !z || !spare || z == spare || degree < 2 || degree > 32 || !(lambda_min > 0) ||
      !(lambda_max > lambda_min)) {
    set_error("sfk_cheby_precondition: invalid argument");
    return SFK_ERR_INVALID_ARG;
  }
  if (!uses_tp(op)) {
    set_error("sfk_cheby_precondition: continuous Laplacian on the tensor-product kernel only (constant "
              "coefficient, degree <= 8)");
    return SFK_ERR_UNSUPPORTED;
  }
  if (op->d.n_rhs != 1) {
    set_error("sfk_cheby_precondition: n_rhs must be 1");
    return SFK_ERR_UNSUPPORTED;
  }
  sfk_status st = ensure_device_tables(op);
  if (st != SFK_OK) return st;
  void *zz = z, *sp = spare;
  hipError_t e = cheby_tp3(op, rhs, &zz, &sp, degree, lambda_min, lambda_max, (hipStream_t)stream);
  // odd degree: the result is in the spare vector; copy it to z
  if (e == hipSuccess && zz != z)
    e = hipMemcpyAsync(z, zz, size_t(op->nvec) * elem_size(op), hipMemcpyDeviceToDevice, (hipStream_t)stream);
  if (e != hipSuccess) {
    set_error(std::string("sfk_cheby_precondition: ") + hipGetErrorString(e));
    return SFK_ERR_HIP;
  }
  return SFK_OK;
}

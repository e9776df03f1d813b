// Vector kernels: fill, convert, FP64-accumulated dot products, axpby,
// the Chebyshev vector update for operators whose matvec cannot fuse it
// (continuous Q_p: the assembled value of a node is complete only after all
// cells have scattered), Dirichlet masking, layout converters between the
// canonical and the reference's lane-interleaved layout, and the batched 1D
// even-odd kernel.  All grid-stride loops; every block exits.
#include "sfk_device.h"
#include "sfk_kernels.h"

namespace sfk {

namespace {
constexpr int kThreads = 256;

inline unsigned grid_for(long long n, int per_thread = 1)
{
  long long b = (n + (long long)kThreads * per_thread - 1) / ((long long)kThreads * per_thread);
  if (b < 1) b = 1;
  if (b > 65535LL * 8) b = 65535LL * 8;
  return (unsigned)b;
}

template <typename T>
__global__ void fill_kernel(T *x, long long n, T v)
{
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] = v;
}

template <typename S, typename D>
__global__ void convert_kernel(const S *x, D *y, long long n)
{
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = D(x[i]);
}

template <typename T>
__global__ void __launch_bounds__(kThreads) dot_partial_kernel(const T *x, const T *y, long long n, double *partials)
{
  __shared__ double red[kThreads];
  double s = 0.;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    s += double(x[i]) * double(y[i]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(kThreads) dot_final_kernel(const double *partials, int n, double *result)
{
  __shared__ double red[kThreads];
  double s = 0.;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += partials[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *result = red[0];
}

__global__ void __launch_bounds__(kThreads) dot2_partial_kernel(const double *x, const double *y, const double *w,
                                                                long long n, double *partials)
{
  __shared__ double red[2][kThreads];
  double s = 0., u = 0.;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double xv = x[i];
    s += xv * y[i];
    u += xv * w[i];
  }
  red[0][threadIdx.x] = s;
  red[1][threadIdx.x] = u;
  __syncthreads();
  for (int k = kThreads / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      red[0][threadIdx.x] += red[0][threadIdx.x + k];
      red[1][threadIdx.x] += red[1][threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = red[0][0];
    partials[gridDim.x + blockIdx.x] = red[1][0];
  }
}

// CG update x += alpha p (skipped when x is null), r -= alpha q (FP64),
// optionally rs = (float) r, with per-block partial sums of r.r; with an
// non-null, alpha = *an / *ad (device scalars: no host round trip)
__global__ void __launch_bounds__(kThreads) cg_update_kernel(long long n, double alpha, const double *an,
                                                             const double *ad, const double *p, const double *q,
                                                             double *x, double *r, float *rs, double *partials)
{
  __shared__ double red[kThreads];
  if (an) alpha = *an / *ad;
  double acc = 0.;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    if (x) x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    if (rs) rs[i] = float(ri);
    acc += ri * ri;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// z = (double) zs (not written when z is null) with per-block partial sums
// of z.r and z.q
__global__ void __launch_bounds__(kThreads) widen_dot2_kernel(long long n, const float *zs, double *z,
                                                              const double *r, const double *q, double *partials)
{
  __shared__ double red[2][kThreads];
  double a = 0., b = 0.;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double zi = double(zs[i]);
    if (z) z[i] = zi;
    a += zi * r[i];
    b += zi * q[i];
  }
  red[0][threadIdx.x] = a;
  red[1][threadIdx.x] = b;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = red[0][0];
    partials[gridDim.x + blockIdx.x] = red[1][0];
  }
}

template <typename T>
__global__ void axpby_kernel(long long n, T a, const T *x, T b, T *y)
{
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = b != T(0) ? a * x[i] + b * y[i] : a * x[i]; // b == 0: y is write-only
}

// p = (double) zs + b p
__global__ void axpby_widen_kernel(long long n, const float *zs, double b, double *p)
{
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = double(zs[i]) + b * p[i];
}

// x += a p, then p = z + b p (z FP32 or FP64; x, p FP64): the iterate's
// update of one CG iteration folded into the next search direction's pass.
// a = *an / *ad, b = -*bn / *ad (device scalars)
template <typename Z>
__global__ void xp_update_kernel(long long n, const double *an, const double *ad, const double *bn, const Z *z,
                                 double *p, double *x)
{
  const double a = *an / *ad, b = -*bn / *ad;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double pi = p[i];
    x[i] += a * pi;
    p[i] = double(z[i]) + b * pi;
  }
}

template <typename T>
__global__ void cheby_kernel(long long n, const T *au, const T *rhs, const T *sol_old, const T *diag,
                             T *sol_new, T *tm, T c_np, T c_tm)
{
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    // a null au or sol_old stands for the zero vector (the first step of a
    // recurrence started from zero): not read
    const T res = (au ? au[i] : T(0)) - rhs[i];
    // c_tm == 0 starts a recurrence: the previous tm is not read
    const T t = (c_tm != T(0) ? c_tm * tm[i] : T(0)) + c_np * diag[i] * res;
    tm[i] = t;
    sol_new[i] = (sol_old ? sol_old[i] : T(0)) - t;
  }
}

// z = c a b (elementwise)
template <typename T>
__global__ void mul_scaled_kernel(long long n, T c, const T *a, const T *b, T *z)
{
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    z[i] = c * a[i] * b[i];
}

template <typename T>
__global__ void dirichlet_kernel(T *x, int gx, int gy, int gz)
{
  const long long n = (long long)gx * gy * gz;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int X = (int)(i % gx);
    const long long r = i / gx;
    const int Y = (int)(r % gy), Z = (int)(r / gy);
    if (X == 0 || Y == 0 || Z == 0 || X == gx - 1 || Y == gy - 1 || Z == gz - 1) x[i] = T(0);
  }
}

__global__ void layout_kernel(int n3, int nx, int ny, int nz, int lanes, const double *in, double *out,
                              int to_canonical)
{
  const long long total = (long long)nx * ny * nz * n3;
  const int nxb = nx / lanes;
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < total; g += (long long)gridDim.x * blockDim.x) {
    const int dof = (int)(g % n3);
    const long long cell = g / n3;
    const int ix = (int)(cell % nx);
    const long long r = cell / nx;
    const int iy = (int)(r % ny), iz = (int)(r / ny);
    const long long il =
        (((long long)iz * ny + iy) * nxb + ix / lanes) * (long long)lanes * n3 + (long long)dof * lanes + ix % lanes;
    if (to_canonical) out[g] = in[il];
    else out[il] = in[g];
  }
}

template <int N>
__global__ void apply1d_kernel(Eo<N, double> t, int sym, long long n_lines, long long line_stride, int stride,
                               const double *in, double *out)
{
  for (long long li = blockIdx.x * (long long)blockDim.x + threadIdx.x; li < n_lines; li += (long long)gridDim.x * blockDim.x) {
    double u[N], v[N];
#pragma unroll
    for (int i = 0; i < N; ++i) u[i] = in[li * line_stride + (long long)i * stride];
    if (sym > 0) eo_apply<N, 1>(t, u, v);
    else eo_apply<N, -1>(t, u, v);
#pragma unroll
    for (int i = 0; i < N; ++i) out[li * line_stride + (long long)i * stride] = v[i];
  }
}
} // namespace

hipError_t vec_fill(bool fp32, void *x, long long n, double v, hipStream_t s)
{
  if (fp32) hipLaunchKernelGGL(fill_kernel<float>, dim3(grid_for(n, 4)), dim3(kThreads), 0, s, (float *)x, n, (float)v);
  else hipLaunchKernelGGL(fill_kernel<double>, dim3(grid_for(n, 4)), dim3(kThreads), 0, s, (double *)x, n, v);
  return hipGetLastError();
}

hipError_t vec_copy_convert(bool to_fp32, const void *x, void *y, long long n, hipStream_t s)
{
  if (to_fp32)
    hipLaunchKernelGGL((convert_kernel<double, float>), dim3(grid_for(n, 4)), dim3(kThreads), 0, s, (const double *)x, (float *)y, n);
  else
    hipLaunchKernelGGL((convert_kernel<float, double>), dim3(grid_for(n, 4)), dim3(kThreads), 0, s, (const float *)x, (double *)y, n);
  return hipGetLastError();
}

hipError_t vec_dot(bool fp32, const void *x, const void *y, long long n, double *result_dev, double *partials,
                   int n_partials, hipStream_t s)
{
  unsigned g = grid_for(n, 8);
  if ((int)g > n_partials) g = n_partials;
  if (fp32)
    hipLaunchKernelGGL(dot_partial_kernel<float>, dim3(g), dim3(kThreads), 0, s, (const float *)x, (const float *)y, n, partials);
  else
    hipLaunchKernelGGL(dot_partial_kernel<double>, dim3(g), dim3(kThreads), 0, s, (const double *)x, (const double *)y, n, partials);
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kThreads), 0, s, (const double *)partials, (int)g, result_dev);
  return hipGetLastError();
}

hipError_t vec_sum_partials(const double *partials, long long count, double *result_dev, hipStream_t s)
{
  if (count < 0 || count > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kThreads), 0, s, partials, (int)count, result_dev);
  return hipGetLastError();
}

hipError_t vec_dot2(const void *x, const void *y, const void *w, long long n, double *result_dev, double *partials,
                    int n_partials, hipStream_t s)
{
  unsigned g = grid_for(n, 8);
  if ((int)g > n_partials) g = n_partials;
  hipLaunchKernelGGL(dot2_partial_kernel, dim3(g), dim3(kThreads), 0, s, (const double *)x, (const double *)y,
                     (const double *)w, n, partials);
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kThreads), 0, s, (const double *)partials, (int)g, result_dev);
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kThreads), 0, s, (const double *)partials + g, (int)g,
                     result_dev + 1);
  return hipGetLastError();
}

hipError_t vec_cg_update(long long n, double alpha, const double *an, const double *ad, const void *p,
                         const void *q, void *x, void *r, void *rs, double *result_dev, double *partials,
                         int n_partials, hipStream_t s)
{
  unsigned g = grid_for(n, 8);
  if ((int)g > n_partials) g = n_partials;
  hipLaunchKernelGGL(cg_update_kernel, dim3(g), dim3(kThreads), 0, s, n, alpha, an, ad, (const double *)p,
                     (const double *)q, (double *)x, (double *)r, (float *)rs, partials);
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kThreads), 0, s, (const double *)partials, (int)g, result_dev);
  return hipGetLastError();
}

hipError_t vec_widen_dot2(long long n, const void *zs, void *z, const void *r, const void *q, double *result_dev,
                          double *partials, int n_partials, hipStream_t s)
{
  unsigned g = grid_for(n, 8);
  if ((int)g > n_partials) g = n_partials;
  hipLaunchKernelGGL(widen_dot2_kernel, dim3(g), dim3(kThreads), 0, s, n, (const float *)zs, (double *)z,
                     (const double *)r, (const double *)q, partials);
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kThreads), 0, s, (const double *)partials, (int)g, result_dev);
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kThreads), 0, s, (const double *)partials + g, (int)g,
                     result_dev + 1);
  return hipGetLastError();
}

hipError_t vec_xp_update(bool z_fp32, long long n, const double *an, const double *ad, const double *bn,
                         const void *z, void *p, void *x, hipStream_t s)
{
  if (z_fp32)
    hipLaunchKernelGGL(xp_update_kernel<float>, dim3(grid_for(n, 4)), dim3(kThreads), 0, s, n, an, ad, bn,
                       (const float *)z, (double *)p, (double *)x);
  else
    hipLaunchKernelGGL(xp_update_kernel<double>, dim3(grid_for(n, 4)), dim3(kThreads), 0, s, n, an, ad, bn,
                       (const double *)z, (double *)p, (double *)x);
  return hipGetLastError();
}

hipError_t vec_axpby_widen(long long n, const void *zs, double b, void *p, hipStream_t s)
{
  hipLaunchKernelGGL(axpby_widen_kernel, dim3(grid_for(n, 4)), dim3(kThreads), 0, s, n, (const float *)zs, b, (double *)p);
  return hipGetLastError();
}

hipError_t vec_axpby(bool fp32, long long n, double a, const void *x, double b, void *y, hipStream_t s)
{
  if (fp32)
    hipLaunchKernelGGL(axpby_kernel<float>, dim3(grid_for(n, 4)), dim3(kThreads), 0, s, n, (float)a, (const float *)x, (float)b, (float *)y);
  else
    hipLaunchKernelGGL(axpby_kernel<double>, dim3(grid_for(n, 4)), dim3(kThreads), 0, s, n, a, (const double *)x, b, (double *)y);
  return hipGetLastError();
}

hipError_t vec_mul_scaled(bool fp32, long long n, double c, const void *a, const void *b, void *z, hipStream_t s)
{
  if (fp32)
    hipLaunchKernelGGL(mul_scaled_kernel<float>, dim3(grid_for(n, 4)), dim3(kThreads), 0, s, n, (float)c,
                       (const float *)a, (const float *)b, (float *)z);
  else
    hipLaunchKernelGGL(mul_scaled_kernel<double>, dim3(grid_for(n, 4)), dim3(kThreads), 0, s, n, c,
                       (const double *)a, (const double *)b, (double *)z);
  return hipGetLastError();
}

hipError_t vec_cheby_update(bool fp32, long long n, const void *au, const void *rhs, const void *sol_old,
                            const void *diag, void *sol_new, void *tm, double c_np, double c_tm, hipStream_t s)
{
  if (fp32)
    hipLaunchKernelGGL(cheby_kernel<float>, dim3(grid_for(n, 4)), dim3(kThreads), 0, s, n, (const float *)au,
                       (const float *)rhs, (const float *)sol_old, (const float *)diag, (float *)sol_new, (float *)tm,
                       (float)c_np, (float)c_tm);
  else
    hipLaunchKernelGGL(cheby_kernel<double>, dim3(grid_for(n, 4)), dim3(kThreads), 0, s, n, (const double *)au,
                       (const double *)rhs, (const double *)sol_old, (const double *)diag, (double *)sol_new,
                       (double *)tm, c_np, c_tm);
  return hipGetLastError();
}

hipError_t vec_zero_dirichlet(bool fp32, void *x, const int g[3], hipStream_t s)
{
  const long long n = (long long)g[0] * g[1] * g[2];
  if (fp32) hipLaunchKernelGGL(dirichlet_kernel<float>, dim3(grid_for(n, 4)), dim3(kThreads), 0, s, (float *)x, g[0], g[1], g[2]);
  else hipLaunchKernelGGL(dirichlet_kernel<double>, dim3(grid_for(n, 4)), dim3(kThreads), 0, s, (double *)x, g[0], g[1], g[2]);
  return hipGetLastError();
}

hipError_t layout_convert(int degree, const int nc[3], int lanes, const double *in, double *out, bool to_canonical,
                          hipStream_t s)
{
  const int n = degree + 1;
  const long long total = (long long)nc[0] * nc[1] * nc[2] * n * n * n;
  hipLaunchKernelGGL(layout_kernel, dim3(grid_for(total, 4)), dim3(kThreads), 0, s, n * n * n, nc[0], nc[1], nc[2],
                     lanes, in, out, to_canonical ? 1 : 0);
  return hipGetLastError();
}

template <int N, typename T>
static hipError_t upload_devtab(const Tables1D &t, void **dev)
{
  const DevTab<N, T> h = make_devtab<N, T>(t);
  hipError_t e = hipMalloc(dev, sizeof(h));
  if (e != hipSuccess) {
    *dev = nullptr;
    return e;
  }
  return hipMemcpy(*dev, &h, sizeof(h), hipMemcpyHostToDevice);
}

hipError_t upload_device_tables(const Tables1D &t, bool fp32, void **dev)
{
  switch (t.n) {
#define SFK_CASE(NN) \
  case NN: return fp32 ? upload_devtab<NN, float>(t, dev) : upload_devtab<NN, double>(t, dev);
    SFK_CASE(2) SFK_CASE(3) SFK_CASE(4) SFK_CASE(5) SFK_CASE(6) SFK_CASE(7) SFK_CASE(8) SFK_CASE(9) \
    SFK_CASE(10) SFK_CASE(11) SFK_CASE(12) SFK_CASE(13) SFK_CASE(14) SFK_CASE(15) SFK_CASE(16)
#undef SFK_CASE
  default: return hipErrorInvalidValue;
  }
}

hipError_t apply_1d(int nn, const std::vector<double> &M, int sym, bool forward, long long n_lines,
                    long long line_stride, int stride, const double *in, double *out, hipStream_t s)
{
  const unsigned g = grid_for(n_lines);
  switch (nn) {
#define SFK_CASE(NN)                                                                                          \
  case NN:                                                                                                    \
    hipLaunchKernelGGL(apply1d_kernel<NN>, dim3(g), dim3(kThreads), 0, s, make_eo<NN, double>(M, forward), sym, \
                       n_lines, line_stride, stride, in, out);                                                \
    break;
    SFK_CASE(2) SFK_CASE(3) SFK_CASE(4) SFK_CASE(5) SFK_CASE(6) SFK_CASE(7) SFK_CASE(8) SFK_CASE(9) \
    SFK_CASE(10) SFK_CASE(11) SFK_CASE(12) SFK_CASE(13) SFK_CASE(14) SFK_CASE(15) SFK_CASE(16)
#undef SFK_CASE
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

} // namespace sfk

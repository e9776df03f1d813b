// Allocation-free inner product (sfk_vec_dot_noalloc, include/sfk.h).
//
// sfk_vec_dot allocates its partial sums with hipMalloc on every call, which
// callers under HIP graph capture or in allocation-free loops cannot use.
// Here the partial sums live in a file-scope device array `scratch` of
// kNoallocPartials doubles (32 KiB, held by every device the library runs on
// until the process ends).  Host code never passes `scratch` itself (in host
// code the name is only a stand-in for the symbol): the device address of
// each device's copy is looked up once per device with hipGetSymbolAddress
// on the first call made with that device current, and that address is what
// the kernels receive.
//
// One workgroup sums one chunk of kChunk consecutive elements into slot
// blockIdx.x, so the summation order depends on n only (the same inputs give
// the same bits on every call); every slot the final pass reads is written
// by the same call first, so nothing has to be cleared between calls.  Calls
// that overlap on one device race on the array.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "sfk.h"
#include "sfk_internal.h"

namespace sfk {
namespace {

constexpr int kNoallocPartials = 4096;
constexpr long long kChunk = 65536;
constexpr int kThreads = 256;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoallocPartials];

template <typename T>
__global__ void __launch_bounds__(kThreads) chunk_dot_kernel(const T *x, const T *y, long long n, double *partials)
{
  __shared__ double red[kThreads];
  const long long b0 = (long long)blockIdx.x * kChunk;
  const long long b1 = b0 + kChunk < n ? b0 + kChunk : n;
  double s = 0.;
  for (long long i = b0 + threadIdx.x; i < b1; i += kThreads) s += double(x[i]) * double(y[i]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(kThreads) chunk_final_kernel(const double *partials, int g, double *result)
{
  __shared__ double red[kThreads];
  double s = 0.;
  for (int i = threadIdx.x; i < g; i += kThreads) s += partials[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *result = red[0];
}

std::mutex g_mutex;
double *g_scratch[kMaxDevices] = {};

// device address of this device's `scratch` (looked up once per device)
hipError_t scratch_of_current_device(double **ptr)
{
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = static_cast<double *>(p);
  }
  *ptr = g_scratch[dev];
  return hipSuccess;
}

} // namespace
} // namespace sfk

extern "C" sfk_status sfk_vec_dot_noalloc(int precision, const void *x, const void *y, long long n,
                                          double *result_dev, void *stream)
{
  using namespace sfk;
  if (!x || !y || !result_dev || n < 0 || (precision != SFK_FP64 && precision != SFK_FP32)) {
    set_error("sfk_vec_dot_noalloc: invalid argument");
    return SFK_ERR_INVALID_ARG;
  }
  if (n > kChunk * kNoallocPartials) {
    set_error("sfk_vec_dot_noalloc: n = " + std::to_string(n) + " exceeds the maximum " +
              std::to_string(kChunk * kNoallocPartials) + " (" + std::to_string(kNoallocPartials) +
              " partial sums of " + std::to_string(kChunk) + " elements in the static scratch array)");
    return SFK_ERR_INVALID_ARG;
  }
  double *partials = nullptr;
  hipError_t e = scratch_of_current_device(&partials);
  if (e != hipSuccess) {
    set_error(std::string("sfk_vec_dot_noalloc: scratch address: ") + hipGetErrorString(e));
    return SFK_ERR_HIP;
  }
  hipStream_t s = (hipStream_t)stream;
  const int g = (int)((n + kChunk - 1) / kChunk);
  if (g > 0) {
    if (precision == SFK_FP32)
      hipLaunchKernelGGL(chunk_dot_kernel<float>, dim3(g), dim3(kThreads), 0, s, (const float *)x, (const float *)y,
                         n, partials);
    else
      hipLaunchKernelGGL(chunk_dot_kernel<double>, dim3(g), dim3(kThreads), 0, s, (const double *)x,
                         (const double *)y, n, partials);
  }
  // g == 0: the empty sum (the final pass reads no slot)
  hipLaunchKernelGGL(chunk_final_kernel, dim3(1), dim3(kThreads), 0, s, (const double *)partials, g, result_dev);
  e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("sfk_vec_dot_noalloc: ") + hipGetErrorString(e));
    return SFK_ERR_HIP;
  }
  return SFK_OK;
}

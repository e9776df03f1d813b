// Chebyshev smoother sweep and HIP graph capture of sfk call sequences.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "sfk.h"
#include "sfk_internal.h"
#include "sfk_kernels.h"
#include "sfk_op.h"

using namespace sfk;

struct sfk_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

#define GR_FAIL(code, msg) \
  do {                     \
    set_error(msg);        \
    return code;           \
  } while (0)
#define GR_HIP(expr)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) GR_FAIL(SFK_ERR_HIP, std::string("HIP error: ") + hipGetErrorString(e_) + " in " #expr); \
  } while (0)

} // namespace

extern "C" {

sfk_status sfk_cheby_sweep(sfk_op *op, const void *rhs, void *x, void *spare, const void *diag_inv, void *tmp,
                           int degree, double lambda_min, double lambda_max, void *stream)
{
  if (!op || !rhs || !x || !spare || !diag_inv || !tmp) GR_FAIL(SFK_ERR_INVALID_ARG, "sfk_cheby_sweep: null argument");
  if (x == spare) GR_FAIL(SFK_ERR_INVALID_ARG, "sfk_cheby_sweep: x and spare must differ");
  if (degree < 1 || degree > 64) GR_FAIL(SFK_ERR_INVALID_ARG, "sfk_cheby_sweep: degree must be in 1..64");
  if (!(lambda_min > 0.) || !(lambda_max > lambda_min) || !std::isfinite(lambda_max))
    GR_FAIL(SFK_ERR_INVALID_ARG, "sfk_cheby_sweep: need 0 < lambda_min < lambda_max");
  if (op->d.kind == SFK_DG_ADVECT) GR_FAIL(SFK_ERR_UNSUPPORTED, "sfk_cheby_sweep: not for advection");
  const double theta = 0.5 * (lambda_max + lambda_min), delta = 0.5 * (lambda_max - lambda_min);
  const double sigma = theta / delta;
  // ping-pong so that the last step writes x: an odd number of steps starts
  // from a copy of x in spare
  void *src = x, *dst = spare;
  if (degree % 2 == 1) {
    const size_t bytes = size_t(op->nvec) * op->d.n_rhs * elem_size(op);
    GR_HIP(hipMemcpyAsync(spare, x, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    src = spare;
    dst = x;
  }
  double rho_old = 1. / sigma;
  for (int j = 0; j < degree; ++j) {
    double c_np, c_tm;
    if (j == 0) {
      c_np = 1. / theta;
      c_tm = 0.;
    } else {
      const double rho = 1. / (2. * sigma - rho_old);
      c_np = 2. * rho / delta;
      c_tm = rho * rho_old;
      rho_old = rho;
    }
    const sfk_status st = sfk_cheby_step(op, rhs, src, diag_inv, dst, tmp, c_np, c_tm, stream);
    if (st != SFK_OK) return st;
    void *t = src;
    src = dst;
    dst = t;
  }
  return SFK_OK;
}

sfk_status sfk_stream_create(void **stream)
{
  if (!stream) GR_FAIL(SFK_ERR_INVALID_ARG, "sfk_stream_create: null argument");
  hipStream_t s = nullptr;
  GR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return SFK_OK;
}

sfk_status sfk_stream_destroy(void *stream)
{
  if (stream) GR_HIP(hipStreamDestroy((hipStream_t)stream));
  return SFK_OK;
}

sfk_status sfk_graph_begin(void *stream)
{
  if (!stream) GR_FAIL(SFK_ERR_INVALID_ARG, "sfk_graph_begin: capture needs a non-default stream");
  GR_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeRelaxed));
  return SFK_OK;
}

sfk_status sfk_graph_end(void *stream, sfk_graph **graph)
{
  if (!stream || !graph) GR_FAIL(SFK_ERR_INVALID_ARG, "sfk_graph_end: null argument");
  *graph = nullptr;
  sfk_graph *g = new sfk_graph();
  hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g->graph);
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    GR_FAIL(SFK_ERR_HIP, std::string("sfk_graph_end: ") + hipGetErrorString(e));
  }
  *graph = g;
  return SFK_OK;
}

sfk_status sfk_graph_launch(const sfk_graph *graph, void *stream)
{
  if (!graph || !graph->exec) GR_FAIL(SFK_ERR_INVALID_ARG, "sfk_graph_launch: null graph");
  GR_HIP(hipGraphLaunch(graph->exec, (hipStream_t)stream));
  return SFK_OK;
}

sfk_status sfk_graph_destroy(sfk_graph *graph)
{
  if (!graph) return SFK_OK;
  if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
  if (graph->graph) (void)hipGraphDestroy(graph->graph);
  delete graph;
  return SFK_OK;
}

} // extern "C"

// Temporary launchers for the entry points implemented in later files.
#include "sfk_kernels.h"
namespace sfk {
hipError_t launch_sipg_general(const SipgLaunch &, hipStream_t) { return hipErrorNotSupported; }
hipError_t launch_coef_gradient(const CoefGradLaunch &, hipStream_t) { return hipErrorNotSupported; }
hipError_t launch_sine_metric(const Tables1D &, bool, const int *, double, void *, void *, hipStream_t)
{
  return hipErrorNotSupported;
}
} // namespace sfk
extern "C" sfk_status sfk_solve_cg_cheby(sfk_op *, sfk_op *, const double *, double *, double, int, int, int,
                                         sfk_solve_stats *, void *)
{
  sfk::set_error("solver not built");
  return SFK_ERR_UNSUPPORTED;
}

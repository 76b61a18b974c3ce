// FIRE / Minimizer and DDFT entry points (filled in below).
#include <hip/hip_runtime.h>

#include "api_guard.hpp"
#include "context.hpp"

using namespace cdft;

extern "C" {

int cdft_minimize_grand_potential(cdft_ctx* ctx, const cdft_fire*, const cdft_parametrization*, int, const double*,
                                  const double*, double, double*, cdft_min_result*, double*) {
  return guarded([&] {
    check_ctx(ctx);
    throw Unsupported("not built yet");
  });
}

int cdft_ddft_run(cdft_ctx* ctx, const cdft_step_config*, int, const double*, double, double*, cdft_sim_result*,
                  double*) {
  return guarded([&] {
    check_ctx(ctx);
    throw Unsupported("not built yet");
  });
}

}  // extern "C"

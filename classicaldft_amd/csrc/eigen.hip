// C ABI of the smallest-eigenpair solver (EigenvalueSolver::solve,
// include/dft/algorithms/saddle_point.hpp:136-270); the solver is in eigen.hpp.
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>
#include <vector>

#include "api_guard.hpp"
#include "context.hpp"
#include "eigen.hpp"

using namespace cdft;

extern "C" {

int cdft_smallest_eigenpair(cdft_ctx* ctx, const cdft_eigen_config* config, const double* rho, const double* vext,
                            double mu, const double* initial_guess, unsigned long long seed, double* eigenvector,
                            cdft_eigen_result* result) {
  return guarded([&] {
    check_ctx(ctx);
    if (!config || !rho || !eigenvector || !result) throw std::invalid_argument("null argument");
    if (config->max_iterations < 0 || !(config->tolerance >= 0.0) || !(config->hessian_eps >= 0.0))
      throw std::invalid_argument("bad eigenvalue solver configuration");
    Context* c = ctx->impl;
    if (config->hessian_eps == 0.0) require_fp64(c);
    CDFT_HIP(hipSetDevice(c->device));
    const long N = c->g.N;
    if (vext) CDFT_HIP(hipMemcpyAsync(c->d_vext, vext, 8 * N, hipMemcpyHostToDevice, c->stream));
    EigenRun run(c, vext ? c->d_vext : nullptr, mu, config->hessian_eps, config->fixed_boundary ? 1 : 0);
    std::unique_ptr<Hessian> hs;
    if (config->hessian_eps == 0.0) {
      hs.reset(new Hessian(c));
      run.exact = hs.get();
    }
    CDFT_HIP(hipMemcpyAsync(run.rho.p, rho, 8 * N, hipMemcpyHostToDevice, c->stream));
    // start vector, normalised on the host (saddle_point.hpp:184-185)
    std::vector<double> v0 = initial_guess ? std::vector<double>(initial_guess, initial_guess + N)
                                           : seeded_normals(static_cast<uint64_t>(seed), N);
    long double nn = 0.0L;
    for (double x : v0) nn += static_cast<long double>(x) * x;
    const double inv = 1.0 / std::sqrt(static_cast<double>(nn));
    if (!(inv > 0.0) || !std::isfinite(inv)) throw std::invalid_argument("initial guess has zero norm");
    for (double& x : v0) x *= inv;
    CDFT_HIP(hipMemcpyAsync(run.v.p, v0.data(), 8 * N, hipMemcpyHostToDevice, c->stream));
    run.base_forces();
    double lambda = 0.0;
    int iters = 0, conv = 0;
    run.solve(config->tolerance, config->max_iterations, &lambda, &iters, &conv);
    CDFT_HIP(hipMemcpyAsync(eigenvector, run.v.p, 8 * N, hipMemcpyDeviceToHost, c->stream));
    CDFT_HIP(hipStreamSynchronize(c->stream));
    result->eigenvalue = lambda;
    result->iterations = iters;
    result->converged = conv;
    result->hessian_products = run.products;
  });
}

}  // extern "C"

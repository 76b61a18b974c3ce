// The reference's other DDFT steppers and its Picard solver, on device-resident
// FP64 state:
//   split_operator_step   (include/dft/algorithms/dynamics.hpp:405-462)
//   crank_nicholson_step  (include/dft/algorithms/dynamics.hpp:464-538)
//   Picard::solve         (include/dft/algorithms/picard.hpp:63-124) with the
//                         optional fixed_mass_constraint (minimization.hpp:157-180)
// The force callback is the functional's grand-potential gradient
// (cdft_total's forces at mu, optional external field) or a host callback
// (cdft_hooks::force); Picard's constraint may also be a host callback
// (cdft_hooks::project).  Transforms are the
// FP64 general-schedule r2c / c2r; k^2 is the continuous |k|^2 of
// compute_k_squared (dynamics.hpp:310-318, Nyquist bin positive).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "api_guard.hpp"
#include "context.hpp"
#include "ddft.hpp"
#include "device_common.hpp"
#include "hessian.hpp"

namespace cdft {

namespace {

// Spectrum S (layout kz, kx, ky) times sqrt(exp(-D k^2 dt)): half a step of exact diffusion
// (split_operator_step's half_prop = sqrt(propagator), diffusion_propagator at dynamics.hpp:320-322).
__global__ void __launch_bounds__(256) k_half_diffusion(cx<double>* __restrict__ S, Geom g, double D, double dt) {
  const long nxy = static_cast<long>(g.nx) * g.ny;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < g.M;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    const long kz = i / nxy, r = i - kz * nxy;
    const int x = static_cast<int>(r / g.ny), y = static_cast<int>(r - static_cast<long>(x) * g.ny);
    const double kx = kvec(x, g.nx, g.dkx), ky = kvec(y, g.ny, g.dky), kzv = g.dkz * static_cast<double>(kz);
    const double h = sqrt(exp(-D * (kx * kx + ky * ky + kzv * kzv) * dt));
    S[i] = {S[i].re * h, S[i].im * h};
  }
}

// Crank-Nicolson combination (dynamics.hpp:515-518): W = ((1 - a) rho0_k + dt/2 (R0_k + R1_k)) / (1 + a),
// a = D k^2 dt / 2
__global__ void __launch_bounds__(256) k_cn_combine(const cx<double>* __restrict__ P0, const cx<double>* __restrict__ R0,
                                                    const cx<double>* __restrict__ R1, cx<double>* __restrict__ W,
                                                    Geom g, double D, double dt) {
  const long nxy = static_cast<long>(g.nx) * g.ny;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < g.M;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    const long kz = i / nxy, r = i - kz * nxy;
    const int x = static_cast<int>(r / g.ny), y = static_cast<int>(r - static_cast<long>(x) * g.ny);
    const double kx = kvec(x, g.nx, g.dkx), ky = kvec(y, g.ny, g.dky), kzv = g.dkz * static_cast<double>(kz);
    const double a = D * (kx * kx + ky * ky + kzv * kzv) * dt * 0.5;
    const double cm = 1.0 - a, cpi = 1.0 / (1.0 + a), h = dt * 0.5;
    W[i] = {(cm * P0[i].re + h * (R0[i].re + R1[i].re)) * cpi, (cm * P0[i].im + h * (R0[i].im + R1[i].im)) * cpi};
  }
}

// y = max(a + s b, lo) (b nullable)
__global__ void __launch_bounds__(256) k_clamp_axpy(double* __restrict__ y, const double* __restrict__ a, double s,
                                                    const double* __restrict__ b, double lo, long n) {
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * blockDim.x)
    y[i] = fmax(b ? a[i] + s * b[i] : a[i], lo);
}

// d1 <- max(raw, lo); block maxima of |d1_old - d1_new| in partials[b*8] (k_finalize_maxmin reads them)
__global__ void __launch_bounds__(256) k_clamp_dev(const double* __restrict__ raw, double* __restrict__ d1, double lo,
                                                   long n, double* partials) {
  double mx = 0.0;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    const double nv = fmax(raw[i], lo);
    mx = fmax(mx, fabs(nv - d1[i]));
    d1[i] = nv;
  }
  __shared__ double sm[256];
  sm[threadIdx.x] = mx;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[static_cast<long>(blockIdx.x) * 8 + 0] = sm[0];
    partials[static_cast<long>(blockIdx.x) * 8 + 1] = 0.0;
  }
}

// sums of f^2, rho, rho - bg (bg nullable) for the Picard residual and the mass constraint
__global__ void __launch_bounds__(256) k_picard_sums(const double* __restrict__ f, const double* __restrict__ rho,
                                                     const double* __restrict__ bg, long n, double* partials) {
  double s0 = 0, c0 = 0, s1 = 0, c1 = 0, s2 = 0, c2 = 0;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    acc2(s0, c0, f[i] * f[i]);
    acc2(s1, c1, rho[i]);
    acc2(s2, c2, bg ? rho[i] - bg[i] : rho[i]);
  }
  const double vals[3] = {s0 + c0, s1 + c1, s2 + c2};
  block_sums(vals, 3, partials);
}

// picard.hpp:107-110: rho *= exp(-mixing f / dV), clamped
__global__ void __launch_bounds__(256) k_picard_update(double* __restrict__ rho, const double* __restrict__ f,
                                                       double mixing, double idv, double lo, long n) {
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * blockDim.x)
    rho[i] = fmax(rho[i] * exp(-mixing * f[i] * idv), lo);
}

// fixed_mass_constraint (minimization.hpp:165-178): rho = bg + (target / current) (rho - bg), clamped,
// when |current| = |sum (rho - bg) dV| > 1e-30
__global__ void __launch_bounds__(256) k_mass_constraint(double* __restrict__ rho, const double* __restrict__ bg,
                                                         const double* __restrict__ sums, double target, double dv,
                                                         double lo, long n) {
  const double cur = sums[2] * dv;
  if (!(fabs(cur) > 1e-30)) return;
  const double s = target / cur;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    const double b = bg ? bg[i] : 0.0;
    rho[i] = fmax(b + s * (rho[i] - b), lo);
  }
}

struct Stepper {
  Context* c;
  const double* d_vext;
  double mu;
  long N;
  DevBuf rho, d1, f, rhs, raw, S0, SR0, SR1, W, part, red;
  long evals = 0;
  HostHooks* hooks = nullptr;  // a host force callback in place of the context's functional

  Stepper(Context* c_, const double* vx, double m)
      : c(c_), d_vext(vx), mu(m), N(c_->g.N), rho(8 * N), d1(8 * N), f(8 * N), rhs(8 * N), raw(8 * N),
        S0(16 * c_->g.M), SR0(16 * c_->g.M), SR1(16 * c_->g.M), W(16 * c_->g.M), part(8 * 8 * kVecBlocks),
        red(8 * 8) {}

  dim3 vb() const { return dim3(vec_blocks(N)); }
  dim3 sb() const { return dim3(vec_blocks(c->g.M)); }

  double compute(const double* d, double* forces) {
    ++evals;
    if (hooks && hooks->force()) return hooks->eval(c, d, forces);
    c->eval(d, d_vext, mu, forces, AliasMap{0, 0.0, 0.0});
    cdft_result r{};
    c->fetch_result(mu, &r);
    return r.grand_potential;
  }
  // excess_rhs_real (dynamics.hpp:233-241) into rhs
  void excess_rhs(const double* d, const double* forces) {
    launch_ddft_rhs(d, forces, rhs.as<double>(), c->g, part.as<double>(), c->stream);
  }
  double max_from_partials() {
    hipLaunchKernelGGL(k_finalize_maxmin, dim3(1), dim3(256), 0, c->stream, part.as<double>(),
                       static_cast<int>(vec_blocks(N)), red.as<double>());
    double h[2];
    CDFT_HIP(hipMemcpyAsync(h, red.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    CDFT_HIP(hipStreamSynchronize(c->stream));
    return h[0];
  }

  // split_operator_step (dynamics.hpp:410-462): rho <- one step; returns the final energy
  double split_step(double D, double dt, double lo) {
    cx<double>* S = S0.as<cx<double>>();
    forward_full(c, rho.as<double>(), S);
    hipLaunchKernelGGL(k_half_diffusion, sb(), dim3(256), 0, c->stream, S, c->g, D, dt);
    inverse_full(c, S, raw.as<double>());
    hipLaunchKernelGGL(k_clamp_axpy, vb(), dim3(256), 0, c->stream, d1.as<double>(), raw.as<double>(), 0.0,
                       static_cast<const double*>(nullptr), lo, N);  // rho*
    compute(d1.as<double>(), f.as<double>());
    excess_rhs(d1.as<double>(), f.as<double>());
    hipLaunchKernelGGL(k_clamp_axpy, vb(), dim3(256), 0, c->stream, raw.as<double>(), d1.as<double>(), dt,
                       static_cast<const double*>(rhs.as<double>()), lo, N);  // rho**
    forward_full(c, raw.as<double>(), S);
    hipLaunchKernelGGL(k_half_diffusion, sb(), dim3(256), 0, c->stream, S, c->g, D, dt);
    inverse_full(c, S, raw.as<double>());
    hipLaunchKernelGGL(k_clamp_axpy, vb(), dim3(256), 0, c->stream, rho.as<double>(), raw.as<double>(), 0.0,
                       static_cast<const double*>(nullptr), lo, N);
    CDFT_HIP(hipGetLastError());
    return compute(rho.as<double>(), f.as<double>());
  }

  // crank_nicholson_step (dynamics.hpp:472-538)
  double cn_step(double D, double dt, double lo, double tol, int max_it, int* iters) {
    forward_full(c, rho.as<double>(), S0.as<cx<double>>());
    compute(rho.as<double>(), f.as<double>());
    excess_rhs(rho.as<double>(), f.as<double>());
    forward_full(c, rhs.as<double>(), SR0.as<cx<double>>());
    CDFT_HIP(hipMemcpyAsync(d1.p, rho.p, 8 * N, hipMemcpyDeviceToDevice, c->stream));
    int it = 0;
    for (; it < max_it; ++it) {
      compute(d1.as<double>(), f.as<double>());
      excess_rhs(d1.as<double>(), f.as<double>());
      forward_full(c, rhs.as<double>(), SR1.as<cx<double>>());
      hipLaunchKernelGGL(k_cn_combine, sb(), dim3(256), 0, c->stream, S0.as<cx<double>>(), SR0.as<cx<double>>(),
                         SR1.as<cx<double>>(), W.as<cx<double>>(), c->g, D, dt);
      inverse_full(c, W.as<cx<double>>(), raw.as<double>());
      hipLaunchKernelGGL(k_clamp_dev, vb(), dim3(256), 0, c->stream, raw.as<double>(), d1.as<double>(), lo, N,
                         part.as<double>());
      CDFT_HIP(hipGetLastError());
      if (max_from_partials() < tol) {
        ++it;
        break;
      }
    }
    *iters = it;
    CDFT_HIP(hipMemcpyAsync(rho.p, d1.p, 8 * N, hipMemcpyDeviceToDevice, c->stream));
    return compute(rho.as<double>(), f.as<double>());
  }
};

void check_step_config(const cdft_step_config* cfg, int n_steps) {
  if (!cfg || n_steps < 0) throw std::invalid_argument("bad argument");
  if (!(cfg->dt > 0.0)) throw std::invalid_argument("step config: dt must be positive");
}

// The split-operator and Crank-Nicolson steps take a force callback only.
void check_force_hooks_only(const cdft_hooks* h) {
  if (h && (h->project || h->stop || h->snapshot))
    throw std::invalid_argument("this step takes a force callback only (no boundary, stop or snapshot callbacks)");
}

}  // namespace

}  // namespace cdft

using namespace cdft;

extern "C" {

int cdft_ddft_split_operator(cdft_ctx* ctx, const cdft_step_config* config, int n_steps, const double* vext,
                             double mu, const cdft_hooks* hooks, double* rho, double* energy) {
  return guarded([&] {
    check_ctx(ctx);
    check_step_config(config, n_steps);
    check_force_hooks_only(hooks);
    if (!rho) throw std::invalid_argument("null density");
    Context* c = ctx->impl;
    CDFT_HIP(hipSetDevice(c->device));
    const long N = c->g.N;
    if (vext) CDFT_HIP(hipMemcpyAsync(c->d_vext, vext, 8 * N, hipMemcpyHostToDevice, c->stream));
    HostHooks hk{hooks};
    Stepper st(c, vext ? c->d_vext : nullptr, mu);
    st.hooks = &hk;
    CDFT_HIP(hipMemcpyAsync(st.rho.p, rho, 8 * N, hipMemcpyHostToDevice, c->stream));
    double e = 0.0;
    for (int s = 0; s < n_steps; ++s)
      e = st.split_step(config->diffusion_coefficient, config->dt, config->min_density);
    CDFT_HIP(hipMemcpyAsync(rho, st.rho.p, 8 * N, hipMemcpyDeviceToHost, c->stream));
    CDFT_HIP(hipStreamSynchronize(c->stream));
    if (energy) *energy = e;
  });
}

int cdft_ddft_crank_nicolson(cdft_ctx* ctx, const cdft_step_config* config, int n_steps, const double* vext,
                             double mu, const cdft_hooks* hooks, double* rho, double* energy, int* iterations) {
  return guarded([&] {
    check_ctx(ctx);
    check_step_config(config, n_steps);
    check_force_hooks_only(hooks);
    if (!rho) throw std::invalid_argument("null density");
    Context* c = ctx->impl;
    CDFT_HIP(hipSetDevice(c->device));
    const long N = c->g.N;
    if (vext) CDFT_HIP(hipMemcpyAsync(c->d_vext, vext, 8 * N, hipMemcpyHostToDevice, c->stream));
    HostHooks hk{hooks};
    Stepper st(c, vext ? c->d_vext : nullptr, mu);
    st.hooks = &hk;
    CDFT_HIP(hipMemcpyAsync(st.rho.p, rho, 8 * N, hipMemcpyHostToDevice, c->stream));
    double e = 0.0;
    int it = 0;
    for (int s = 0; s < n_steps; ++s)
      e = st.cn_step(config->diffusion_coefficient, config->dt, config->min_density, config->fp_tolerance,
                     config->fp_max_iterations, &it);
    CDFT_HIP(hipMemcpyAsync(rho, st.rho.p, 8 * N, hipMemcpyDeviceToHost, c->stream));
    CDFT_HIP(hipStreamSynchronize(c->stream));
    if (energy) *energy = e;
    if (iterations) *iterations = it;
  });
}

int cdft_picard_solve(cdft_ctx* ctx, const cdft_picard* config, const double* vext, double mu,
                      const double* target, const double* background, const cdft_hooks* hooks, double* rho,
                      cdft_picard_result* result) {
  return guarded([&] {
    check_ctx(ctx);
    if (!config || !rho || !result || config->max_iterations < 0) throw std::invalid_argument("bad argument");
    if (hooks && (hooks->stop || hooks->snapshot))
      throw std::invalid_argument("Picard takes force and constraint callbacks only");
    if (target && !std::isfinite(*target)) throw std::invalid_argument("Picard: the target mass is not finite");
    Context* c = ctx->impl;
    CDFT_HIP(hipSetDevice(c->device));
    const Geom& g = c->g;
    const long N = g.N;
    if (vext) CDFT_HIP(hipMemcpyAsync(c->d_vext, vext, 8 * N, hipMemcpyHostToDevice, c->stream));
    HostHooks hk{hooks};
    Stepper st(c, vext ? c->d_vext : nullptr, mu);
    st.hooks = &hk;
    CDFT_HIP(hipMemcpyAsync(st.rho.p, rho, 8 * N, hipMemcpyHostToDevice, c->stream));
    const bool constrained = target != nullptr;
    const double target_mass = constrained ? *target : 0.0;
    if (constrained && background) CDFT_HIP(hipMemcpyAsync(st.d1.p, background, 8 * N, hipMemcpyHostToDevice, c->stream));
    const double* bg = (constrained && background) ? st.d1.as<double>() : nullptr;
    double omega = 0.0, omega_prev = 0.0, residual = 0.0;
    bool have_prev = false;
    int it = 0;
    double* r = st.rho.as<double>();
    for (; it < config->max_iterations; ++it) {
      omega = st.compute(r, st.f.as<double>());
      hipLaunchKernelGGL(k_picard_sums, st.vb(), dim3(256), 0, c->stream, st.f.as<double>(), r,
                         static_cast<const double*>(nullptr), N, st.part.as<double>());
      launch_finalize(st.part.as<double>(), static_cast<int>(vec_blocks(N)), st.red.as<double>(), c->stream);
      double h[8];
      CDFT_HIP(hipMemcpyAsync(h, st.red.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
      CDFT_HIP(hipStreamSynchronize(c->stream));
      residual = std::sqrt(h[0] / static_cast<double>(N));
      // picard.hpp:97-105
      if (residual < config->tolerance || (have_prev && std::fabs(omega - omega_prev) < config->tolerance)) break;
      omega_prev = omega;
      have_prev = true;
      hipLaunchKernelGGL(k_picard_update, st.vb(), dim3(256), 0, c->stream, r, st.f.as<double>(), config->mixing,
                         1.0 / g.dv, config->min_density, N);
      if (constrained) {
        hipLaunchKernelGGL(k_picard_sums, st.vb(), dim3(256), 0, c->stream, st.f.as<double>(), r, bg, N,
                           st.part.as<double>());
        launch_finalize(st.part.as<double>(), static_cast<int>(vec_blocks(N)), st.red.as<double>(), c->stream);
        hipLaunchKernelGGL(k_mass_constraint, st.vb(), dim3(256), 0, c->stream, r, bg, st.red.as<double>(),
                           target_mass, g.dv, config->min_density, N);
      }
      CDFT_HIP(hipGetLastError());
      if (hk.project()) hk.apply_project(c, r);  // picard.hpp:113-115
    }
    CDFT_HIP(hipMemcpyAsync(rho, r, 8 * N, hipMemcpyDeviceToHost, c->stream));
    CDFT_HIP(hipStreamSynchronize(c->stream));
    result->grand_potential = omega;
    result->residual = residual;
    result->iterations = it;
    result->converged = (residual < config->tolerance || (it > 0 && std::fabs(omega - omega_prev) < config->tolerance))
                            ? 1
                            : 0;
  });
}

}  // extern "C"

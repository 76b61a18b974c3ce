// Device-resident FIRE2 minimisation in the alias variable
// (Minimizer::grand_potential, include/dft/algorithms/minimization.hpp:280-353,
// fire::Fire, include/dft/algorithms/fire.hpp:60-178) and the DDFT
// integrating-factor scheme (integrating_factor_step + Simulation::run,
// include/dft/algorithms/dynamics.hpp:79-401,582-693).
//
// State vectors (x, v, f, rho, d1, forces) stay in HBM in FP64; the host
// keeps only the scalar control logic of the reference (FIRE's branch on the
// power, dt/alpha adaptation, the uphill counter, the EMA convergence
// monitor; DDFT's restarts and dt doubling) and reads back a handful of
// FP64 reductions per step.
//
// Evaluations of the functional on an input identical to one already
// evaluated are not repeated (FIRE's second initial compute, DDFT's first
// compute of a step and its first fixed-point iterate): the evaluation is
// deterministic, so reusing the stored result changes no output.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "api_guard.hpp"
#include "context.hpp"
#include "ddft.hpp"

namespace cdft {

namespace {

__device__ __forceinline__ bool on_face(long i, const Geom& g) {
  const long iz = i % g.nz;
  const long t = i / g.nz;
  const long iy = t % g.ny;
  const long ix = t / g.ny;
  return ix == 0 || ix == g.nx - 1 || iy == 0 || iy == g.ny - 1 || iz == 0 || iz == g.nz - 1;
}

// from_density (minimization.hpp:63-77)
__global__ void k_from_density(const double* __restrict__ rho, double* __restrict__ x, long n, AliasMap a) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = rho[i] - a.rho_min;
  if (a.kind == 1) {
    x[i] = sqrt(fmax(d, 0.0));
  } else {
    const double dd = fmin(fmax(d, 0.0), a.range * (1.0 - 1e-14));
    x[i] = sqrt(dd / (a.range - dd));
  }
}

__global__ void k_to_density(const double* __restrict__ x, double* __restrict__ rho, long n, AliasMap a) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) rho[i] = a.apply(x[i]);
}

// Sum of g over the boundary faces (grid.hpp:67-82), compact enumeration:
// x faces, then y faces with interior x, then z faces with interior x, y.
__global__ void __launch_bounds__(256) k_face_sum(const double* __restrict__ g, Geom G, long nb, double* partials) {
  double s = 0.0, c = 0.0;
  const long nx = G.nx, ny = G.ny, nz = G.nz;
  const long fx = (nx > 1 ? 2 : 1) * ny * nz;
  const long ixn = nx > 2 ? nx - 2 : 0;
  const long fy = (ny > 1 ? 2 : 1) * ixn * nz;
  const long iyn = ny > 2 ? ny - 2 : 0;
  for (long j = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; j < nb;
       j += static_cast<long>(gridDim.x) * blockDim.x) {
    long ix, iy, iz;
    if (j < fx) {
      const long f = j / (ny * nz), r = j % (ny * nz);
      ix = f == 0 ? 0 : nx - 1;
      iy = r / nz;
      iz = r % nz;
    } else if (j < fx + fy) {
      const long k = j - fx;
      const long f = k / (ixn * nz), r = k % (ixn * nz);
      iy = f == 0 ? 0 : ny - 1;
      ix = 1 + r / nz;
      iz = r % nz;
    } else {
      const long k = j - fx - fy;
      const long f = k / (ixn * iyn), r = k % (ixn * iyn);
      iz = f == 0 ? 0 : nz - 1;
      ix = 1 + r / iyn;
      iy = 1 + r % iyn;
    }
    acc2(s, c, g[iz + nz * (iy + ny * ix)]);
  }
  const double v[1] = {s + c};
  block_sums(v, 1, partials);
}

// FIRE force in the alias variable (minimization.hpp:288-302): homogeneous
// boundary average (grid.hpp:113-131), chain rule, negation.  Accumulates
// sum f^2 and the power sum v.f for the next step.
__global__ void __launch_bounds__(256) k_fire_force(const double* __restrict__ g, const double* __restrict__ x,
                                                    const double* __restrict__ v, double* __restrict__ f, Geom G,
                                                    int homog, const double* __restrict__ face, long nb, AliasMap a,
                                                    double* partials) {
  double s2 = 0, c2 = 0, sp = 0, cp = 0;
  const double avg = (homog && nb > 0) ? face[0] / static_cast<double>(nb) : 0.0;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < G.N;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    double gi = g[i];
    if (homog && on_face(i, G)) gi = avg;
    const double fi = -(gi * a.deriv(x[i]));
    f[i] = fi;
    acc2(s2, c2, fi * fi);
    acc2(sp, cp, v[i] * fi);
  }
  const double vals[2] = {s2 + c2, sp + cp};
  block_sums(vals, 2, partials);
}

// fire.hpp:110-126: optional backtrack (x -= dt_old v / 2, v = 0), then v += dt f; accumulates |v|^2
__global__ void __launch_bounds__(256) k_fire_velocity(double* __restrict__ x, double* __restrict__ v,
                                                       const double* __restrict__ f, long n, double dt_old,
                                                       double dt, int backtrack, double* partials) {
  double s = 0, c = 0;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    double vi = v[i];
    if (backtrack) {
      x[i] -= 0.5 * dt_old * vi;
      vi = 0.0;
    }
    vi += dt * f[i];
    v[i] = vi;
    acc2(s, c, vi * vi);
  }
  const double vals[1] = {s + c};
  block_sums(vals, 1, partials);
}

// fire.hpp:128-148: v = (1 - alpha) v + alpha |v|/|f| f (when |f| > 0), x += dt v
__global__ void __launch_bounds__(256) k_fire_mix(double* __restrict__ x, double* __restrict__ v,
                                                  const double* __restrict__ f, long n, double dt, double alpha,
                                                  const double* __restrict__ vsum, const double* __restrict__ fsum) {
  const double vn = sqrt(vsum[0]), fn = sqrt(fsum[0]);
  const bool mix = fn > 0.0;
  const double ratio = mix ? vn / fn : 0.0;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    double vi = v[i];
    if (mix) vi = (1.0 - alpha) * vi + alpha * ratio * f[i];
    v[i] = vi;
    x[i] += dt * vi;
  }
}

long face_count(const Geom& g) {
  const long nx = g.nx, ny = g.ny, nz = g.nz;
  const long inner = std::max(nx - 2, 0L) * std::max(ny - 2, 0L) * std::max(nz - 2, 0L);
  return nx * ny * nz - inner;
}

}  // namespace

// ------------------------------------------------------------------ Minimizer
// Device FIRE2 in the alias variable (Minimizer::grand_potential,
// minimization.hpp:280-353; fire::Fire, fire.hpp:60-165) as a state machine:
// launch_*() queue a step's device work on the context's stream without
// waiting, after_*() read its reductions once the stream has drained.  One
// host thread can so advance several independent minimisations (contexts
// with their own streams) in lockstep, the device working on all at once.
struct FireRun {
  Context* c;
  AliasMap amap;
  int homog;
  const double* d_vext;
  double mu;
  cdft_fire fire;
  DevBuf x, v, f, part, res;  // res: [0..7] force sums, [8..15] |v|^2, [16..23] face sum
  long nb;
  double* host = nullptr;     // pinned: [0..7] energy sums, [8..15] force sums
  double* energies = nullptr;  // optional record: after initialisation and every step
  // FIRE state (fire.hpp:46-58) and the convergence monitor (minimization.hpp:105-110)
  double energy = 0.0, old_energy = 0.0, sf2 = 0.0, power = 0.0, dt = 0.0, alpha = 0.0, rms = 0.0, vv = 1.0;
  int n_pos = 0, n_neg = 0, iteration = 0;
  bool converged = false;
  cdft_result last_result{};

  FireRun(Context* c_, AliasMap a, int h, const double* vx, double m, const cdft_fire& fr)
      : c(c_), amap(a), homog(h), d_vext(vx), mu(m), fire(fr), x(8 * c_->g.N), v(8 * c_->g.N), f(8 * c_->g.N),
        part(8 * 8 * kVecBlocks), res(8 * 24) {
    nb = face_count(c->g);
    CDFT_HIP(hipHostMalloc(reinterpret_cast<void**>(&host), 16 * sizeof(double), hipHostMallocDefault));
  }
  ~FireRun() {
    if (host) (void)hipHostFree(host);
  }
  FireRun(const FireRun&) = delete;
  FireRun& operator=(const FireRun&) = delete;

  // compute(x): energy Omega and f in d_f; device sums f^2 and v.f in res[0], res[1]
  void launch_compute() {
    const Geom& g = c->g;
    c->eval(x.as<double>(), d_vext, mu, c->d_force, amap);
    if (homog) {
      hipLaunchKernelGGL(k_face_sum, dim3(vec_blocks(nb)), dim3(256), 0, c->stream, c->d_force, g, nb,
                         part.as<double>());
      launch_finalize(part.as<double>(), static_cast<int>(vec_blocks(nb)), res.as<double>() + 16, c->stream);
    }
    hipLaunchKernelGGL(k_fire_force, dim3(vec_blocks(g.N)), dim3(256), 0, c->stream, c->d_force, x.as<double>(),
                       v.as<double>(), f.as<double>(), g, homog, res.as<double>() + 16, nb, amap, part.as<double>());
    launch_finalize(part.as<double>(), static_cast<int>(vec_blocks(g.N)), res.as<double>(), c->stream);
    CDFT_HIP(hipGetLastError());
    CDFT_HIP(hipMemcpyAsync(host, c->d_res, 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CDFT_HIP(hipMemcpyAsync(host + 8, res.p, 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  void read_compute() {
    cdft_result r{};
    c->result_from(host, mu, &r);
    last_result = r;
    energy = r.grand_potential;
    sf2 = host[8];
    power = host[9];
    rms = std::sqrt(sf2 / static_cast<double>(c->g.N));
  }
  void sync() { CDFT_HIP(hipStreamSynchronize(c->stream)); }

  // Fire::initialize (fire.hpp:60-84) and the extra compute of minimization.hpp:312-313
  void launch_start(const double* rho0) {
    const Geom& g = c->g;
    CDFT_HIP(hipSetDevice(c->device));
    CDFT_HIP(hipMemcpyAsync(c->d_rho, rho0, 8 * g.N, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_from_density, dim3((g.N + 255) / 256), dim3(256), 0, c->stream, c->d_rho,
                       x.as<double>(), g.N, amap);
    CDFT_HIP(hipMemsetAsync(v.p, 0, 8 * g.N, c->stream));
    launch_compute();
  }
  void after_start() {
    read_compute();
    dt = fire.dt;
    alpha = fire.alpha_start;
    converged = false;  // minimization.hpp:321: the loop starts unconverged
    if (energies) energies[0] = energy;
  }
  bool active() const { return !converged && iteration < fire.max_steps; }

  // Fire::step (fire.hpp:86-165); `power` is v.f of the current forces
  void launch_step() {
    const Geom& g = c->g;
    old_energy = energy;
    bool backtrack = false;
    const double dt_old = dt;
    if (power > 0.0 || iteration == 0) {
      ++n_pos;
      n_neg = 0;
      if (n_pos > fire.n_delay) {
        dt = std::min(dt * fire.f_inc, fire.dt_max);
        alpha *= fire.f_alpha;
      }
    } else {
      n_pos = 0;
      ++n_neg;
      if (n_neg > fire.max_uphill) throw std::runtime_error("fire::step: exceeded max uphill steps");
      backtrack = true;
      if (iteration > fire.n_delay) {
        dt = std::max(dt * fire.f_dec, fire.dt_min);
        alpha = fire.alpha_start;
      }
    }
    CDFT_HIP(hipSetDevice(c->device));
    const unsigned nbk = vec_blocks(g.N);
    hipLaunchKernelGGL(k_fire_velocity, dim3(nbk), dim3(256), 0, c->stream, x.as<double>(), v.as<double>(),
                       f.as<double>(), g.N, dt_old, dt, backtrack ? 1 : 0, part.as<double>());
    launch_finalize(part.as<double>(), static_cast<int>(nbk), res.as<double>() + 8, c->stream);
    hipLaunchKernelGGL(k_fire_mix, dim3(nbk), dim3(256), 0, c->stream, x.as<double>(), v.as<double>(),
                       f.as<double>(), g.N, dt, alpha, res.as<double>() + 8, res.as<double>());
    launch_compute();
  }
  void after_step() {
    read_compute();
    ++iteration;
    // convergence_monitor (minimization.hpp:105-110,329-332)
    vv = vv * 0.9 + 0.1 * (std::fabs(energy - old_energy) / dt);
    const double volume = static_cast<double>(c->g.N) * c->g.dv;
    converged = std::fabs(vv / volume) < fire.force_tolerance;
    if (energies) energies[iteration] = energy;
  }

  // the final total() (minimization.hpp:339-343) is the last compute's evaluation
  void finish(double* rho_out, cdft_min_result* result) {
    const Geom& g = c->g;
    CDFT_HIP(hipSetDevice(c->device));
    hipLaunchKernelGGL(k_to_density, dim3((g.N + 255) / 256), dim3(256), 0, c->stream, x.as<double>(), c->d_rho,
                       g.N, amap);
    CDFT_HIP(hipMemcpyAsync(rho_out, c->d_rho, 8 * g.N, hipMemcpyDeviceToHost, c->stream));
    sync();
    result->free_energy = last_result.free_energy;
    result->grand_potential = last_result.grand_potential;
    result->rms_force = rms;
    result->iterations = iteration;
    result->converged = converged ? 1 : 0;
    result->dt = dt;
    result->alpha = alpha;
  }
};

}  // namespace cdft

using namespace cdft;

extern "C" {

int cdft_minimize_grand_potential(cdft_ctx* ctx, const cdft_fire* fire, const cdft_parametrization* param,
                                  int use_homogeneous_boundary, const double* rho0, const double* vext, double mu,
                                  double* rho_out, cdft_min_result* result, double* energies) {
  return guarded([&] {
    check_ctx(ctx);
    if (!fire || !param || !rho0 || !rho_out || !result) throw std::invalid_argument("null argument");
    if (param->kind != 0 && param->kind != 1) throw std::invalid_argument("unknown parametrization");
    Context* c = ctx->impl;
    CDFT_HIP(hipSetDevice(c->device));
    const Geom& g = c->g;
    const AliasMap amap{param->kind == 0 ? 1 : 2, param->rho_min, param->rho_max - param->rho_min};
    if (vext) CDFT_HIP(hipMemcpyAsync(c->d_vext, vext, 8 * g.N, hipMemcpyHostToDevice, c->stream));
    FireRun run(c, amap, use_homogeneous_boundary ? 1 : 0, vext ? c->d_vext : nullptr, mu, *fire);
    run.energies = energies;
    run.launch_start(rho0);
    run.sync();
    run.after_start();
    while (run.active()) {
      run.launch_step();
      run.sync();
      run.after_step();
    }
    run.finish(rho_out, result);
  });
}

int cdft_minimize_grand_potential_batch(int n, cdft_ctx* const* ctxs, const cdft_fire* fire,
                                        const cdft_parametrization* param, int use_homogeneous_boundary,
                                        const double* const* rho0, const double* mu, double* const* rho_out,
                                        cdft_min_result* results, int* status) {
  return guarded([&] {
    if (n < 0 || (n > 0 && (!ctxs || !fire || !param || !rho0 || !mu || !rho_out || !results || !status)))
      throw std::invalid_argument("null argument");
    if (param->kind != 0 && param->kind != 1) throw std::invalid_argument("unknown parametrization");
    for (int i = 0; i < n; ++i) {
      check_ctx(ctxs[i]);
      if (!rho0[i] || !rho_out[i]) throw std::invalid_argument("null density");
      for (int j = 0; j < i; ++j)
        if (ctxs[j] == ctxs[i]) throw std::invalid_argument("the runs of a batch need distinct contexts");
      if (ctxs[i]->impl->device != ctxs[0]->impl->device)
        throw std::invalid_argument("the contexts of a batch must share a device");
    }
    const AliasMap amap{param->kind == 0 ? 1 : 2, param->rho_min, param->rho_max - param->rho_min};
    std::vector<std::unique_ptr<FireRun>> runs;
    for (int i = 0; i < n; ++i) {
      Context* c = ctxs[i]->impl;
      CDFT_HIP(hipSetDevice(c->device));
      runs.emplace_back(new FireRun(c, amap, use_homogeneous_boundary ? 1 : 0, nullptr, mu[i], *fire));
      status[i] = CDFT_OK;
    }
    for (auto& r : runs) r->launch_start(rho0[&r - &runs[0]]);
    for (auto& r : runs) {
      r->sync();
      r->after_start();
    }
    // lockstep: queue every active run's step, then wait for and read each
    std::vector<char> launched(static_cast<size_t>(n), 0);
    for (;;) {
      bool any = false;
      for (int i = 0; i < n; ++i) {
        launched[static_cast<size_t>(i)] = 0;
        if (status[i] != CDFT_OK || !runs[static_cast<size_t>(i)]->active()) continue;
        try {
          runs[static_cast<size_t>(i)]->launch_step();
          launched[static_cast<size_t>(i)] = 1;
          any = true;
        } catch (const std::runtime_error& e) {  // FIRE's uphill limit ends that run only
          if (dynamic_cast<const HipError*>(&e)) throw;
          status[i] = CDFT_ERUNTIME;
        }
      }
      if (!any) break;
      for (int i = 0; i < n; ++i) {
        if (!launched[static_cast<size_t>(i)]) continue;
        runs[static_cast<size_t>(i)]->sync();
        runs[static_cast<size_t>(i)]->after_step();
      }
    }
    for (int i = 0; i < n; ++i) runs[static_cast<size_t>(i)]->finish(rho_out[i], &results[i]);
  });
}

int cdft_ddft_run(cdft_ctx* ctx, const cdft_step_config* config, int n_steps, const double* vext, double mu,
                  double* rho, cdft_sim_result* result, double* energies, double* times) {
  return guarded([&] {
    check_ctx(ctx);
    if (!config || !rho || !result || n_steps < 0) throw std::invalid_argument("bad argument");
    Context* c = ctx->impl;
    CDFT_HIP(hipSetDevice(c->device));
    const Geom& g = c->g;
    if (vext) CDFT_HIP(hipMemcpyAsync(c->d_vext, vext, 8 * g.N, hipMemcpyHostToDevice, c->stream));
    DdftRun run(c, vext ? c->d_vext : nullptr, mu);
    CDFT_HIP(hipMemcpyAsync(run.rho0.p, rho, 8 * g.N, hipMemcpyHostToDevice, c->stream));
    // mass (dynamics.hpp:591-596)
    const AliasMap id{0, 0.0, 0.0};
    auto mass_of = [&](const double* d) {
      launch_sum(d, id, run.part.as<double>(), g, c->stream);
      launch_finalize(run.part.as<double>(), sum_blocks(g), run.sums.as<double>() + 32, c->stream);
      double s;
      CDFT_HIP(hipMemcpyAsync(&s, run.sums.as<double>() + 32, 8, hipMemcpyDeviceToHost, c->stream));
      CDFT_HIP(hipStreamSynchronize(c->stream));
      return s * g.dv;
    };
    const double m0 = mass_of(run.rho0.as<double>());
    double e = run.compute(run.rho0.as<double>(), run.f0.as<double>());
    if (energies) energies[0] = e;
    if (times) times[0] = 0.0;
    double dt = config->dt;
    int successes = 0;
    double time = 0.0;
    for (int s = 1; s <= n_steps; ++s) {
      const double before = dt;
      e = c->precision == CDFT_PRECISION_MIXED ? run.step<float>(&dt, config->fp_tolerance, config->fp_max_iterations)
                                               : run.step<double>(&dt, config->fp_tolerance, config->fp_max_iterations);
      time += dt;
      // adaptive timestep (dynamics.hpp:640-649)
      if (dt < before) successes = 0;
      else ++successes;
      if (successes >= 5 && dt < config->dt_max) {
        dt = std::min(2.0 * dt, config->dt_max);
        successes = 0;
      }
      if (energies) energies[s] = e;
      if (times) times[s] = time;
    }
    CDFT_HIP(hipMemcpyAsync(rho, run.rho0.p, 8 * g.N, hipMemcpyDeviceToHost, c->stream));
    CDFT_HIP(hipStreamSynchronize(c->stream));
    result->mass_initial = m0;
    result->mass_final = mass_of(run.rho0.as<double>());
    result->time = time;
    result->energy = e;
    result->dt = dt;
    result->evaluations = run.evals;
  });
}

}  // extern "C"

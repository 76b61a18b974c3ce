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
#include "eigen.hpp"

namespace cdft {

namespace {

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
// With fixed mass (minimization.hpp:207-211) the Lagrange multiplier
// lambda = lam[0] / target (lam[0] = sum g rho after the boundary average)
// shifts every component by -lambda dV first; lam is null otherwise.
// With mpart set, the block maxima of |g| / dV (the density-space force after
// the boundary average and the multiplier, without dV) go to mpart[b*8] for
// the optional max-force stop.
__global__ void __launch_bounds__(256) k_fire_force(const double* __restrict__ g, const double* __restrict__ x,
                                                    const double* __restrict__ v, double* __restrict__ f, Geom G,
                                                    int homog, const double* __restrict__ face, long nb, AliasMap a,
                                                    const double* __restrict__ lam, double target, double* partials,
                                                    double* mpart) {
  double s2 = 0, c2 = 0, sp = 0, cp = 0, gmax = 0.0;
  const double avg = (homog && nb > 0) ? face[0] / static_cast<double>(nb) : 0.0;
  const double shift = lam ? (lam[0] / target) * G.dv : 0.0;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < G.N;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    double gi = g[i];
    if (homog && on_face(i, G)) gi = avg;
    gi -= shift;
    // fmax drops a NaN operand: a non-finite force counts as infinite
    gmax = fmax(gmax, isfinite(gi) ? fabs(gi) : INFINITY);
    const double fi = -(gi * a.deriv(x[i]));
    f[i] = fi;
    acc2(s2, c2, fi * fi);
    acc2(sp, cp, v[i] * fi);
  }
  const double vals[2] = {s2 + c2, sp + cp};
  block_sums(vals, 2, partials);
  if (mpart) {
    __shared__ double smx[256];
    smx[threadIdx.x] = gmax;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) smx[threadIdx.x] = fmax(smx[threadIdx.x], smx[threadIdx.x + w]);
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      mpart[static_cast<long>(blockIdx.x) * 8 + 0] = smx[0] / G.dv;
      mpart[static_cast<long>(blockIdx.x) * 8 + 1] = 0.0;
    }
  }
}

// Fixed mass (minimization.hpp:196-201): rho *= target / (sum rho dV) when the mass is positive.
__global__ void __launch_bounds__(256) k_scale_to_mass(double* __restrict__ rho, long n,
                                                       const double* __restrict__ sum, double target, double dv) {
  const double mass = sum[0] * dv;
  if (!(mass > 0.0)) return;
  const double s = target / mass;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * blockDim.x)
    rho[i] *= s;
}

// sum g rho with the boundary components of g replaced by their average (minimization.hpp:203-210)
__global__ void __launch_bounds__(256) k_lambda_dot(const double* __restrict__ g, const double* __restrict__ rho,
                                                    Geom G, int homog, const double* __restrict__ face, long nb,
                                                    double* partials) {
  double s = 0, c = 0;
  const double avg = (homog && nb > 0) ? face[0] / static_cast<double>(nb) : 0.0;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < G.N;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    const double gi = (homog && on_face(i, G)) ? avg : g[i];
    acc2(s, c, gi * rho[i]);
  }
  const double vals[1] = {s + c};
  block_sums(vals, 1, partials);
}

// ---- Search::find (saddle_point.hpp:367-589): interior/boundary sums, the
// interior mass rescale, and the force of the eigenvector-following FIRE run.

// sum over interior (face = 0) or boundary (face = 1) points of a (times b when given)
__global__ void __launch_bounds__(256) k_region_sum(const double* __restrict__ a, const double* __restrict__ b,
                                                    Geom G, int face, double* partials) {
  double s = 0, c = 0;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < G.N;
       i += static_cast<long>(gridDim.x) * blockDim.x)
    if (on_face(i, G) == (face != 0)) acc2(s, c, b ? a[i] * b[i] : a[i]);
  const double vals[1] = {s + c};
  block_sums(vals, 1, partials);
}

// saddle_point.hpp:451-461: interior densities scaled to the interior target mass
__global__ void __launch_bounds__(256) k_scale_interior(double* __restrict__ rho, const double* __restrict__ sum,
                                                        double mt, Geom G) {
  const double m = sum[0] * G.dv;
  if (!(m > 0.0 && mt > 0.0)) return;
  const double s = mt / m;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < G.N;
       i += static_cast<long>(gridDim.x) * blockDim.x)
    if (!on_face(i, G)) rho[i] *= s;
}

// saddle_point.hpp:463-484: fixed boundary, interior Lagrange multiplier
// mu_L = sum_int g rho / mt, alias transform; f = the alias force, and the
// partial sums of dir . f
__global__ void __launch_bounds__(256) k_saddle_alias(const double* __restrict__ g, const double* __restrict__ x,
                                                      AliasMap a, const double* __restrict__ lsum, double mt,
                                                      const double* __restrict__ dir, double* __restrict__ f,
                                                      Geom G, double* partials) {
  const double muL = mt > 0.0 ? lsum[0] / mt : lsum[0];
  double s = 0, c = 0;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < G.N;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    const double gi = on_face(i, G) ? 0.0 : g[i] - muL * G.dv;
    const double fa = gi * a.deriv(x[i]);
    f[i] = fa;
    acc2(s, c, dir[i] * fa);
  }
  const double vals[1] = {s + c};
  block_sums(vals, 1, partials);
}

// saddle_point.hpp:486-489: f <- -(f - (dir . f) dir); sums f^2 and v.f as k_fire_force does
__global__ void __launch_bounds__(256) k_saddle_project(double* __restrict__ f, const double* __restrict__ dir,
                                                        const double* __restrict__ psum, const double* __restrict__ v,
                                                        long n, double* partials) {
  const double pr = psum[0];
  double s2 = 0, c2 = 0, sp = 0, cp = 0;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    const double fi = -(f[i] - pr * dir[i]);
    f[i] = fi;
    acc2(s2, c2, fi * fi);
    acc2(sp, cp, v[i] * fi);
  }
  const double vals[2] = {s2 + c2, sp + cp};
  block_sums(vals, 2, partials);
}

// dir = e * d rho / d x (the alias transform of the eigenvector, saddle_point.hpp:429-431)
__global__ void __launch_bounds__(256) k_alias_dir(const double* __restrict__ e, const double* __restrict__ x,
                                                   AliasMap a, double* __restrict__ dir, long n) {
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * blockDim.x)
    dir[i] = e[i] * a.deriv(x[i]);
}

__global__ void __launch_bounds__(256) k_scale(double* __restrict__ y, double s, long n) {
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * blockDim.x)
    y[i] *= s;
}

// FIRE2 state (fire.hpp:46-58) and the convergence monitor
// (minimization.hpp:105-110), kept on the device: whole steps replay from a
// graph with no host decision between them.
struct FireCtl {
  double energy, old_energy, dt, dt_old, alpha, vv;
  int n_pos, n_neg, iteration, backtrack;
  int skip;   // this step changes nothing: the run has ended
  int state;  // FireState
};
enum FireState : int { FS_RUN = 0, FS_CONVERGED = 1, FS_MAX_STEPS = 2, FS_UPHILL = 3 };

struct FireConst {
  cdft_fire p;
  double mu, volume;
  int rms_stop;  // stop on FIRE's own rms-force test (fire.hpp:101-103) instead of the energy monitor
  double ndof;   // degrees of freedom of the rms force
};

// Omega from the evaluation sums, in the operation order of Context::result_from
__device__ __forceinline__ double fire_omega(const double* res, double mu) {
#pragma clang fp contract(off)
  const double F = res[RS_IDEAL] + res[RS_EXT] + res[RS_HS] + res[RS_MF];
  return F - mu * res[RS_MASS];
}

// Fire::initialize (fire.hpp:60-84) after the first compute; fres[0] = sum f^2
__global__ void k_fire_init(FireCtl* ctl, const double* __restrict__ res, const double* __restrict__ fres,
                            FireConst k, double* energies) {
  FireCtl c{};
  c.energy = fire_omega(res, k.mu);
  c.dt = k.p.dt;
  c.alpha = k.p.alpha_start;
  c.vv = 1.0;
  c.skip = 1;
  c.state = k.p.max_steps > 0 ? FS_RUN : FS_MAX_STEPS;
  if (k.rms_stop && sqrt(fres[0] / k.ndof) < k.p.force_tolerance) c.state = FS_CONVERGED;
  *ctl = c;
  if (energies) energies[0] = c.energy;
}

// Fire::step (fire.hpp:86-108): the branch on the power v.f of the current forces (fres[1])
__global__ void k_fire_begin(FireCtl* ctl, const double* __restrict__ fres, FireConst k) {
#pragma clang fp contract(off)
  FireCtl c = *ctl;
  if (c.state != FS_RUN) {
    ctl->skip = 1;
    return;
  }
  c.old_energy = c.energy;
  c.backtrack = 0;
  c.dt_old = c.dt;
  if (fres[1] > 0.0 || c.iteration == 0) {
    ++c.n_pos;
    c.n_neg = 0;
    if (c.n_pos > k.p.n_delay) {
      c.dt = fmin(c.dt * k.p.f_inc, k.p.dt_max);
      c.alpha *= k.p.f_alpha;
    }
  } else {
    c.n_pos = 0;
    ++c.n_neg;
    if (c.n_neg > k.p.max_uphill) {  // the reference throws here
      c.state = FS_UPHILL;
      c.skip = 1;
      *ctl = c;
      return;
    }
    c.backtrack = 1;
    if (c.iteration > k.p.n_delay) {
      c.dt = fmax(c.dt * k.p.f_dec, k.p.dt_min);
      c.alpha = k.p.alpha_start;
    }
  }
  c.skip = 0;
  *ctl = c;
}

// After the step's compute: energy, iteration count, convergence_monitor (minimization.hpp:105-110,329-332)
__global__ void k_fire_end(FireCtl* ctl, const double* __restrict__ res, const double* __restrict__ fres,
                           FireConst k, double* energies) {
#pragma clang fp contract(off)
  FireCtl c = *ctl;
  if (c.skip) return;
  c.energy = fire_omega(res, k.mu);
  ++c.iteration;
  c.vv = c.vv * 0.9 + 0.1 * (fabs(c.energy - c.old_energy) / c.dt);
  bool conv = k.rms_stop ? sqrt(fres[0] / k.ndof) < k.p.force_tolerance
                         : fabs(c.vv / k.volume) < k.p.force_tolerance;
  if (k.p.force_max_tolerance > 0.0 && fres[48] < k.p.force_max_tolerance) conv = true;  // the max-force stop
  if (conv) c.state = FS_CONVERGED;
  else if (c.iteration >= k.p.max_steps) c.state = FS_MAX_STEPS;
  if (energies) energies[c.iteration] = c.energy;
  *ctl = c;
}

// fire.hpp:110-126: optional backtrack (x -= dt_old v / 2, v = 0), then v += dt f; accumulates |v|^2
__global__ void __launch_bounds__(256) k_fire_velocity(double* __restrict__ x, double* __restrict__ v,
                                                       const double* __restrict__ f, long n,
                                                       const FireCtl* __restrict__ ctl, double* partials) {
  if (ctl->skip) return;
  const double dt_old = ctl->dt_old, dt = ctl->dt;
  const bool backtrack = ctl->backtrack != 0;
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
                                                  const double* __restrict__ f, long n,
                                                  const FireCtl* __restrict__ ctl, const double* __restrict__ vsum,
                                                  const double* __restrict__ fsum) {
  if (ctl->skip) return;
  const double dt = ctl->dt, alpha = ctl->alpha;
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
// minimization.hpp:280-353; fire::Fire, fire.hpp:60-165).  The whole step,
// control included, runs on the device and is captured once as a graph; the
// host replays it in blocks of kFireBlock steps and reads the 64-byte
// control state between blocks.  A step after the run has ended (converged,
// step limit, uphill limit) changes nothing, so a block may overrun the end.
// Several runs (contexts with their own streams) advance together: one
// graph holds a block of steps of every running run (BatchGraph), or, where
// graphs are off, every run's block is launched before any is waited for.
constexpr int kFireBlock = 8;

struct FireRun {
  Context* c;
  AliasMap amap;
  int homog;
  const double* d_vext;
  FireConst k;
  // res: [0..7] force sums, [8..15] |v|^2, [16..23] face sum, [24..31] mass sum, [32..39] sum g rho,
  // [40..47] dir . f, [48..55] max |force| / dV
  DevBuf x, v, f, part, mpart, res, ctl, en;
  long nb;
  bool fixed_mass = false;  // Minimizer::fixed_mass: rescale rho to `target`, mu = 0 in the loop
  double target = 0.0;
  // Search::find's FIRE run: interior mass `target`, fixed boundary, forces projected off `dir`
  bool saddle = false;
  const double* dir = nullptr;
  FireCtl* host = nullptr;  // pinned copy of the control state
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  FireRun(Context* c_, AliasMap a, int h, const double* vx, double m, const cdft_fire& fr, bool record)
      : c(c_), amap(a), homog(h), d_vext(vx), x(8 * c_->g.N), v(8 * c_->g.N), f(8 * c_->g.N),
        part(8 * 8 * kVecBlocks), mpart(fr.force_max_tolerance > 0.0 ? 8 * 8 * kVecBlocks : 0), res(8 * 56),
        ctl(sizeof(FireCtl)),
        en(record ? 8 * (static_cast<size_t>(std::max(fr.max_steps, 0)) + 1) : 0) {
    k.p = fr;
    k.mu = m;
    k.volume = static_cast<double>(c->g.N) * c->g.dv;
    k.rms_stop = 0;
    k.ndof = static_cast<double>(c->g.N);
    nb = face_count(c->g);
    CDFT_HIP(hipHostMalloc(reinterpret_cast<void**>(&host), sizeof(FireCtl), hipHostMallocDefault));
  }
  ~FireRun() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (host) (void)hipHostFree(host);
  }
  FireRun(const FireRun&) = delete;
  FireRun& operator=(const FireRun&) = delete;

  // compute(x): Omega sums in d_res, forces f; device sums f^2 and v.f in res[0], res[1]
  void launch_compute() {
    const Geom& g = c->g;
    const unsigned nbk = vec_blocks(g.N);
    if (saddle) {  // saddle_point.hpp:444-491
      const AliasMap id{0, 0.0, 0.0};
      double* r = res.as<double>();
      hipLaunchKernelGGL(k_to_density, dim3((g.N + 255) / 256), dim3(256), 0, c->stream, x.as<double>(), c->d_rho,
                         g.N, amap);
      hipLaunchKernelGGL(k_region_sum, dim3(nbk), dim3(256), 0, c->stream, c->d_rho,
                         static_cast<const double*>(nullptr), g, 0, part.as<double>());
      launch_finalize(part.as<double>(), static_cast<int>(nbk), r + 24, c->stream);
      hipLaunchKernelGGL(k_scale_interior, dim3(nbk), dim3(256), 0, c->stream, c->d_rho, r + 24, target, g);
      c->eval_direct(c->d_rho, d_vext, k.mu, c->d_force, id);
      hipLaunchKernelGGL(k_region_sum, dim3(nbk), dim3(256), 0, c->stream, c->d_force,
                         static_cast<const double*>(c->d_rho), g, 0, part.as<double>());
      launch_finalize(part.as<double>(), static_cast<int>(nbk), r + 32, c->stream);
      hipLaunchKernelGGL(k_saddle_alias, dim3(nbk), dim3(256), 0, c->stream, c->d_force, x.as<double>(), amap,
                         r + 32, target, dir, f.as<double>(), g, part.as<double>());
      launch_finalize(part.as<double>(), static_cast<int>(nbk), r + 40, c->stream);
      hipLaunchKernelGGL(k_saddle_project, dim3(nbk), dim3(256), 0, c->stream, f.as<double>(), dir, r + 40,
                         v.as<double>(), g.N, part.as<double>());
      launch_finalize(part.as<double>(), static_cast<int>(nbk), r, c->stream);
      return;
    }
    if (fixed_mass) {
      // rho = to_density(x) rescaled to the target mass, evaluated at mu = 0 (minimization.hpp:193-206)
      const AliasMap id{0, 0.0, 0.0};
      hipLaunchKernelGGL(k_to_density, dim3((g.N + 255) / 256), dim3(256), 0, c->stream, x.as<double>(), c->d_rho,
                         g.N, amap);
      launch_sum(c->d_rho, id, c->d_sum_part, g, c->stream);
      launch_finalize(c->d_sum_part, c->nblk_sum, res.as<double>() + 24, c->stream);
      hipLaunchKernelGGL(k_scale_to_mass, dim3(nbk), dim3(256), 0, c->stream, c->d_rho, g.N,
                         res.as<double>() + 24, target, g.dv);
      c->eval_direct(c->d_rho, d_vext, 0.0, c->d_force, id);
    } else {
      c->eval_direct(x.as<double>(), d_vext, k.mu, c->d_force, amap);
    }
    if (homog) {
      hipLaunchKernelGGL(k_face_sum, dim3(vec_blocks(nb)), dim3(256), 0, c->stream, c->d_force, g, nb,
                         part.as<double>());
      launch_finalize(part.as<double>(), static_cast<int>(vec_blocks(nb)), res.as<double>() + 16, c->stream);
    }
    if (fixed_mass) {
      hipLaunchKernelGGL(k_lambda_dot, dim3(nbk), dim3(256), 0, c->stream, c->d_force, c->d_rho, g, homog,
                         res.as<double>() + 16, nb, part.as<double>());
      launch_finalize(part.as<double>(), static_cast<int>(nbk), res.as<double>() + 32, c->stream);
    }
    hipLaunchKernelGGL(k_fire_force, dim3(nbk), dim3(256), 0, c->stream, c->d_force, x.as<double>(),
                       v.as<double>(), f.as<double>(), g, homog, res.as<double>() + 16, nb, amap,
                       fixed_mass ? static_cast<const double*>(res.as<double>() + 32) : nullptr, target,
                       part.as<double>(), mpart.as<double>());
    launch_finalize(part.as<double>(), static_cast<int>(nbk), res.as<double>(), c->stream);
    if (mpart.p)
      hipLaunchKernelGGL(k_finalize_maxmin, dim3(1), dim3(256), 0, c->stream, mpart.as<double>(),
                         static_cast<int>(nbk), res.as<double>() + 48);
  }

  // one Fire::step (fire.hpp:86-165) with its compute and the monitor
  void launch_step_kernels() {
    const Geom& g = c->g;
    const unsigned nbk = vec_blocks(g.N);
    auto* cp = ctl.as<FireCtl>();
    hipLaunchKernelGGL(k_fire_begin, dim3(1), dim3(1), 0, c->stream, cp, res.as<double>(), k);
    hipLaunchKernelGGL(k_fire_velocity, dim3(nbk), dim3(256), 0, c->stream, x.as<double>(), v.as<double>(),
                       f.as<double>(), g.N, cp, part.as<double>());
    launch_finalize(part.as<double>(), static_cast<int>(nbk), res.as<double>() + 8, c->stream);
    hipLaunchKernelGGL(k_fire_mix, dim3(nbk), dim3(256), 0, c->stream, x.as<double>(), v.as<double>(),
                       f.as<double>(), g.N, cp, res.as<double>() + 8, res.as<double>());
    launch_compute();
    hipLaunchKernelGGL(k_fire_end, dim3(1), dim3(1), 0, c->stream, cp, c->d_res, res.as<double>(), k,
                       en.as<double>());
  }

  // Fire::initialize (fire.hpp:60-84) and the extra compute of minimization.hpp:312-313
  void launch_start(const double* rho0, bool on_device = false) {
    const Geom& g = c->g;
    CDFT_HIP(hipSetDevice(c->device));
    CDFT_HIP(hipMemcpyAsync(c->d_rho, rho0, 8 * g.N, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                            c->stream));
    hipLaunchKernelGGL(k_from_density, dim3((g.N + 255) / 256), dim3(256), 0, c->stream, c->d_rho,
                       x.as<double>(), g.N, amap);
    CDFT_HIP(hipMemsetAsync(v.p, 0, 8 * g.N, c->stream));
    launch_compute();
    hipLaunchKernelGGL(k_fire_init, dim3(1), dim3(1), 0, c->stream, ctl.as<FireCtl>(), c->d_res,
                       static_cast<const double*>(res.as<double>()), k, en.as<double>());
    CDFT_HIP(hipGetLastError());
    if (c->use_graphs) {
      CDFT_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      try {
        launch_step_kernels();
      } catch (...) {
        hipGraph_t g0 = nullptr;
        (void)hipStreamEndCapture(c->stream, &g0);
        if (g0) (void)hipGraphDestroy(g0);
        throw;
      }
      CDFT_HIP(hipStreamEndCapture(c->stream, &graph));
      CDFT_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    poll();
  }

  // queue up to kFireBlock steps (no more than the step limit leaves), then a read of the control state
  void launch_block() {
    CDFT_HIP(hipSetDevice(c->device));
    const int left = std::max(1, k.p.max_steps - host->iteration);
    for (int i = 0; i < std::min(kFireBlock, left); ++i) {
      if (exec) CDFT_HIP(hipGraphLaunch(exec, c->stream));
      else launch_step_kernels();
    }
    CDFT_HIP(hipGetLastError());
    poll();
  }
  void poll() { CDFT_HIP(hipMemcpyAsync(host, ctl.p, sizeof(FireCtl), hipMemcpyDeviceToHost, c->stream)); }
  void sync() { CDFT_HIP(hipStreamSynchronize(c->stream)); }
  bool running() const { return host->state == FS_RUN; }
  bool uphill() const { return host->state == FS_UPHILL; }

  // The final total() (minimization.hpp:339-343) is the last compute's evaluation.  With fixed
  // mass the final density is the last compute's rescaled one, evaluated at the chemical potential
  // (minimization.hpp:256-266).
  void finish(double* rho_out, cdft_min_result* result, double* energies, double mu_final = 0.0) {
    const Geom& g = c->g;
    CDFT_HIP(hipSetDevice(c->device));
    if (fixed_mass) {
      const AliasMap id{0, 0.0, 0.0};
      c->eval_direct(c->d_rho, d_vext, mu_final, nullptr, id);
    } else {
      hipLaunchKernelGGL(k_to_density, dim3((g.N + 255) / 256), dim3(256), 0, c->stream, x.as<double>(),
                         c->d_rho, g.N, amap);
    }
    CDFT_HIP(hipMemcpyAsync(rho_out, c->d_rho, 8 * g.N, hipMemcpyDeviceToHost, c->stream));
    double er[8], fs[8];
    CDFT_HIP(hipMemcpyAsync(er, c->d_res, sizeof(er), hipMemcpyDeviceToHost, c->stream));
    CDFT_HIP(hipMemcpyAsync(fs, res.p, sizeof(fs), hipMemcpyDeviceToHost, c->stream));
    const FireCtl* h = host;
    if (energies && en.p)
      CDFT_HIP(hipMemcpyAsync(energies, en.p, 8 * (static_cast<size_t>(h->iteration) + 1), hipMemcpyDeviceToHost,
                              c->stream));
    sync();
    cdft_result r{};
    c->result_from(er, fixed_mass ? mu_final : k.mu, &r);
    result->free_energy = r.free_energy;
    result->grand_potential = r.grand_potential;
    result->rms_force = std::sqrt(fs[0] / static_cast<double>(g.N));
    result->iterations = h->iteration;
    result->converged = h->state == FS_CONVERGED ? 1 : 0;
    result->dt = h->dt;
    result->alpha = h->alpha;
  }
};


// A stream of its own for a batch's graph launches and control reads
// (default priority; created and destroyed with the batch call).
struct BatchStream {
  hipStream_t s = nullptr;
  BatchStream() { CDFT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
  ~BatchStream() {
    if (s) (void)hipStreamDestroy(s);
  }
  BatchStream(const BatchStream&) = delete;
  BatchStream& operator=(const BatchStream&) = delete;
};

// The batch graph: `layers` steps of every run in `live`, captured in one
// stream capture.  The capture forks from the batch stream onto every run's
// context stream, records each run's FIRE step kernels (launch_step_kernels)
// `layers` times on that stream, and joins back: a flat graph of kernel nodes
// whose only edges chain one run's own kernels, so the runtime executes the
// runs' chains concurrently.
struct BatchGraph {
  hipGraph_t g = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<hipEvent_t> ev;  // [0] the fork, [1 + i] run i's join
  ~BatchGraph() {
    clear();
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
  void clear() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (g) (void)hipGraphDestroy(g);
    exec = nullptr;
    g = nullptr;
  }
  void build(const std::vector<FireRun*>& live, int layers, hipStream_t bs) {
    clear();
    while (ev.size() < live.size() + 1) {
      hipEvent_t e = nullptr;
      CDFT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      ev.push_back(e);
    }
    CDFT_HIP(hipStreamBeginCapture(bs, hipStreamCaptureModeThreadLocal));
    try {
      CDFT_HIP(hipEventRecord(ev[0], bs));
      for (FireRun* r : live) CDFT_HIP(hipStreamWaitEvent(r->c->stream, ev[0], 0));
      for (int s = 0; s < layers; ++s)
        for (FireRun* r : live) r->launch_step_kernels();
      for (size_t i = 0; i < live.size(); ++i) {
        CDFT_HIP(hipEventRecord(ev[i + 1], live[i]->c->stream));
        CDFT_HIP(hipStreamWaitEvent(bs, ev[i + 1], 0));
      }
    } catch (...) {
      hipGraph_t g0 = nullptr;
      (void)hipStreamEndCapture(bs, &g0);
      if (g0) (void)hipGraphDestroy(g0);
      throw;
    }
    CDFT_HIP(hipStreamEndCapture(bs, &g));
    CDFT_HIP(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
  }
  BatchGraph() = default;
  BatchGraph(const BatchGraph&) = delete;
  BatchGraph& operator=(const BatchGraph&) = delete;
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
    FireRun run(c, amap, use_homogeneous_boundary ? 1 : 0, vext ? c->d_vext : nullptr, mu, *fire,
                energies != nullptr);
    run.launch_start(rho0);
    run.sync();
    while (run.running()) {
      run.launch_block();
      run.sync();
    }
    if (run.uphill()) throw std::runtime_error("fire::step: exceeded max uphill steps");
    run.finish(rho_out, result, energies);
  });
}

int cdft_minimize_fixed_mass(cdft_ctx* ctx, const cdft_fire* fire, const cdft_parametrization* param,
                             int use_homogeneous_boundary, const double* rho0, const double* vext, double mu,
                             double target_mass, double* rho_out, cdft_min_result* result, double* energies) {
  return guarded([&] {
    check_ctx(ctx);
    if (!fire || !param || !rho0 || !rho_out || !result) throw std::invalid_argument("null argument");
    if (param->kind != 0 && param->kind != 1) throw std::invalid_argument("unknown parametrization");
    if (!(target_mass > 0.0)) throw std::invalid_argument("fixed_mass: target mass must be positive");
    Context* c = ctx->impl;
    CDFT_HIP(hipSetDevice(c->device));
    const Geom& g = c->g;
    const AliasMap amap{param->kind == 0 ? 1 : 2, param->rho_min, param->rho_max - param->rho_min};
    if (vext) CDFT_HIP(hipMemcpyAsync(c->d_vext, vext, 8 * g.N, hipMemcpyHostToDevice, c->stream));
    // the loop's energy is F (mu = 0); the chemical potential enters only the final result
    FireRun run(c, amap, use_homogeneous_boundary ? 1 : 0, vext ? c->d_vext : nullptr, 0.0, *fire,
                energies != nullptr);
    run.fixed_mass = true;
    run.target = target_mass;
    run.launch_start(rho0);
    run.sync();
    while (run.running()) {
      run.launch_block();
      run.sync();
    }
    if (run.uphill()) throw std::runtime_error("fire::step: exceeded max uphill steps");
    run.finish(rho_out, result, energies, mu);
  });
}

int cdft_saddle_search(cdft_ctx* ctx, const cdft_saddle_config* config, const double* rho0, const double* vext,
                       double mu, double target_mass, unsigned long long seed, double* rho_out, double* eigenvector,
                       cdft_saddle_result* result) {
  return guarded([&] {
    check_ctx(ctx);
    if (!config || !rho0 || !rho_out || !result) throw std::invalid_argument("null argument");
    const cdft_parametrization& pr = config->param;
    if (pr.kind != 0 && pr.kind != 1) throw std::invalid_argument("unknown parametrization");
    if (config->eigen.hessian_eps < 0.0 || config->max_outer_iterations < 0)
      throw std::invalid_argument("bad saddle search configuration");
    Context* c = ctx->impl;
    if (config->eigen.hessian_eps == 0.0) require_fp64(c);
    CDFT_HIP(hipSetDevice(c->device));
    const Geom& g = c->g;
    const long N = g.N;
    const size_t nb = 8 * static_cast<size_t>(N);
    const AliasMap amap{pr.kind == 0 ? 1 : 2, pr.rho_min, pr.rho_max - pr.rho_min};
    if (vext) CDFT_HIP(hipMemcpyAsync(c->d_vext, vext, nb, hipMemcpyHostToDevice, c->stream));
    const double* dvx = vext ? c->d_vext : nullptr;
    DevBuf rho(nb), eig(nb), xcur(nb), dir(nb);
    CDFT_HIP(hipMemcpyAsync(rho.p, rho0, nb, hipMemcpyHostToDevice, c->stream));
    EigenRun er(c, dvx, mu, config->eigen.hessian_eps, 1);
    std::unique_ptr<Hessian> hs;
    if (config->eigen.hessian_eps == 0.0) {
      hs.reset(new Hessian(c));
      er.exact = hs.get();
    }
    const dim3 vb(vec_blocks(N));
    auto host_sum = [&](const double* a, const double* b, int face) {
      hipLaunchKernelGGL(k_region_sum, vb, dim3(256), 0, c->stream, a, b, g, face, er.part.as<double>());
      launch_finalize(er.part.as<double>(), static_cast<int>(vec_blocks(N)), er.red.as<double>(), c->stream);
      double h = 0.0;
      CDFT_HIP(hipMemcpyAsync(&h, er.red.p, 8, hipMemcpyDeviceToHost, c->stream));
      CDFT_HIP(hipStreamSynchronize(c->stream));
      return h;
    };
    const long nbd = face_count(g);  // boundary points
    double omega = 0.0, lambda = 0.0, ftol = config->tolerance_start;
    bool converged = false, have_eig = false;
    long evals = 0;
    int outer = 0;
    for (; outer < config->max_outer_iterations; ++outer) {
      const double omega_old = omega;
      // 1. smallest eigenpair at rho (saddle_point.hpp:411-420)
      CDFT_HIP(hipMemcpyAsync(er.rho.p, rho.p, nb, hipMemcpyDeviceToDevice, c->stream));
      if (outer >= config->eigenvalue_clear_count && have_eig) {
        CDFT_HIP(hipMemcpyAsync(er.v.p, eig.p, nb, hipMemcpyDeviceToDevice, c->stream));
      } else {
        std::vector<double> v0 = seeded_normals(static_cast<uint64_t>(seed) + static_cast<uint64_t>(outer), N);
        CDFT_HIP(hipMemcpyAsync(er.v.p, v0.data(), nb, hipMemcpyHostToDevice, c->stream));
        CDFT_HIP(hipStreamSynchronize(c->stream));
      }
      const double vn = std::sqrt(er.dots({{er.v.as<double>(), er.v.as<double>()}})[0]);
      er.lincomb(er.v.as<double>(), 1.0 / vn, er.v.as<double>());
      er.base_forces();
      int it = 0, conv = 0;
      er.solve(config->eigen.tolerance, config->eigen.max_iterations, &lambda, &it, &conv);
      CDFT_HIP(hipMemcpyAsync(eig.p, er.v.p, nb, hipMemcpyDeviceToDevice, c->stream));
      have_eig = true;
      // 2. the alias-space direction (saddle_point.hpp:428-435) and the boundary mass (:437-447)
      hipLaunchKernelGGL(k_from_density, dim3((N + 255) / 256), dim3(256), 0, c->stream, rho.as<double>(),
                         xcur.as<double>(), N, amap);
      hipLaunchKernelGGL(k_alias_dir, vb, dim3(256), 0, c->stream, eig.as<double>(), xcur.as<double>(), amap,
                         dir.as<double>(), N);
      const double dn = std::sqrt(er.dots({{dir.as<double>(), dir.as<double>()}})[0]);
      if (dn > 1e-30) hipLaunchKernelGGL(k_scale, vb, dim3(256), 0, c->stream, dir.as<double>(), 1.0 / dn, N);
      const double sb = host_sum(rho.as<double>(), nullptr, 1);
      const double avg = nbd > 0 ? sb / static_cast<double>(nbd) : 0.0;
      const double m_boundary = avg * g.dv * static_cast<double>(nbd);
      // 3. FIRE at fixed interior mass with the eigen-direction projection (saddle_point.hpp:449-520)
      cdft_fire fc = config->fire;
      fc.force_tolerance = std::max(ftol, config->fire.force_tolerance);
      fc.force_max_tolerance = 0.0;  // the search stops on FIRE's rms force only
      ftol /= config->tolerance_factor;
      FireRun run(c, amap, 0, dvx, mu, fc, false);
      run.saddle = true;
      run.target = target_mass - m_boundary;
      run.dir = dir.as<double>();
      run.k.rms_stop = 1;
      run.launch_start(rho.as<double>(), true);
      run.sync();
      while (run.running()) {
        run.launch_block();
        run.sync();
      }
      if (run.uphill()) throw std::runtime_error("fire::step: exceeded max uphill steps");
      evals += 1 + run.host->iteration;
      // 4. the final density is the last compute's; Omega at mu (saddle_point.hpp:523-545)
      CDFT_HIP(hipMemcpyAsync(rho.p, c->d_rho, nb, hipMemcpyDeviceToDevice, c->stream));
      c->eval_direct(rho.as<double>(), dvx, mu, nullptr, AliasMap{0, 0.0, 0.0});
      cdft_result r{};
      c->fetch_result(mu, &r);
      omega = r.grand_potential;
      if (outer > 0 && std::fabs(omega - omega_old) < config->omega_tolerance) {
        converged = true;
        break;
      }
    }
    CDFT_HIP(hipMemcpyAsync(rho_out, rho.p, nb, hipMemcpyDeviceToHost, c->stream));
    if (eigenvector) CDFT_HIP(hipMemcpyAsync(eigenvector, eig.p, nb, hipMemcpyDeviceToHost, c->stream));
    CDFT_HIP(hipStreamSynchronize(c->stream));
    result->grand_potential = omega;
    result->eigenvalue = lambda;
    result->outer_iterations = outer + 1;  // as the reference counts it (saddle_point.hpp:584)
    result->converged = converged ? 1 : 0;
    result->evaluations = evals;
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
      runs.emplace_back(new FireRun(c, amap, use_homogeneous_boundary ? 1 : 0, nullptr, mu[i], *fire, false));
    }
    for (int i = 0; i < n; ++i) runs[static_cast<size_t>(i)]->launch_start(rho0[i]);
    for (auto& r : runs) r->sync();
    bool graphs = true;
    for (auto& r : runs) graphs = graphs && r->exec != nullptr;
    std::vector<FireRun*> live;
    if (graphs && n > 0) {
      // One graph advances every running state point by up to kFireBlock
      // steps (BatchGraph); it is rebuilt when the set of running state
      // points or the block length changes.
      CDFT_HIP(hipSetDevice(ctxs[0]->impl->device));
      BatchStream bs;
      BatchGraph bg;
      std::vector<FireRun*> built;
      int built_layers = 0;
      for (;;) {
        live.clear();
        int layers = kFireBlock;
        for (auto& r : runs)
          if (r->running()) {
            live.push_back(r.get());
            layers = std::min(layers, std::max(1, r->k.p.max_steps - r->host->iteration));
          }
        if (live.empty()) break;
        if (live != built || layers != built_layers) {
          bg.build(live, layers, bs.s);
          built = live;
          built_layers = layers;
        }
        CDFT_HIP(hipGraphLaunch(bg.exec, bs.s));
        for (FireRun* r : live)
          CDFT_HIP(hipMemcpyAsync(r->host, r->ctl.p, sizeof(FireCtl), hipMemcpyDeviceToHost, bs.s));
        CDFT_HIP(hipStreamSynchronize(bs.s));
      }
    } else {
      // every running run's block is queued before any is waited for
      for (;;) {
        live.clear();
        for (auto& r : runs)
          if (r->running()) live.push_back(r.get());
        if (live.empty()) break;
        for (FireRun* r : live) r->launch_block();
        for (FireRun* r : live) r->sync();
      }
    }
    for (int i = 0; i < n; ++i) {
      FireRun& r = *runs[static_cast<size_t>(i)];
      status[i] = r.uphill() ? CDFT_ERUNTIME : CDFT_OK;  // FIRE's uphill limit ends that run only
      r.finish(rho_out[i], &results[i], nullptr);
    }
  });
}

int cdft_ddft_run(cdft_ctx* ctx, const cdft_step_config* config, int n_steps, const double* vext, double mu,
                  const cdft_hooks* hooks, double* rho, cdft_sim_result* result, double* energies, double* times) {
  return guarded([&] {
    check_ctx(ctx);
    if (!config || !rho || !result || n_steps < 0) throw std::invalid_argument("bad argument");
    Context* c = ctx->impl;
    CDFT_HIP(hipSetDevice(c->device));
    const Geom& g = c->g;
    if (vext) CDFT_HIP(hipMemcpyAsync(c->d_vext, vext, 8 * g.N, hipMemcpyHostToDevice, c->stream));
    HostHooks hk{hooks};
    DdftRun run(c, vext ? c->d_vext : nullptr, mu);
    run.hooks = &hk;
    CDFT_HIP(hipMemcpyAsync(run.rho0.p, rho, 8 * g.N, hipMemcpyHostToDevice, c->stream));
    // the boundary condition on the initial state (dynamics.hpp:586-588)
    if (hk.project()) hk.apply_project(c, run.rho0.as<double>());
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
    if (hk.snapshot()) hk.take_snapshot(c, 0, 0.0, e, run.rho0.as<double>());  // dynamics.hpp:603-608
    double dt = config->dt;
    int successes = 0, steps = 0;
    double time = 0.0;
    bool forces_stale = false;  // the boundary condition changed the density after its forces were taken
    // Without stop or snapshot callbacks (and with the context's functional) the
    // step energies are needed only at the end: each step's raw sums stay on
    // the device and are read back once.
    const bool defer = !hk.stop() && !hk.snapshot() && !hk.force();
    DevBuf esums(defer ? 8 * sizeof(double) * static_cast<size_t>(n_steps + 1) : 0);
    for (int s = 1; s <= n_steps; ++s) {
      if (forces_stale) {
        run.compute(run.rho0.as<double>(), run.f0.as<double>());  // the step's first compute (dynamics.hpp:345)
        forces_stale = false;
      }
      const double before = dt;
      double* sums_to = defer ? esums.as<double>() + 8 * s : nullptr;
      e = c->precision == CDFT_PRECISION_MIXED
              ? run.step<float>(&dt, config->fp_tolerance, config->fp_max_iterations, sums_to)
              : run.step<double>(&dt, config->fp_tolerance, config->fp_max_iterations, sums_to);
      time += dt;
      steps = s;
      if (hk.project()) {  // dynamics.hpp:629-631
        hk.apply_project(c, run.rho0.as<double>());
        forces_stale = true;
      }
      // adaptive timestep (dynamics.hpp:640-649)
      if (dt < before) successes = 0;
      else ++successes;
      if (successes >= 5 && dt < config->dt_max) {
        dt = std::min(2.0 * dt, config->dt_max);
        successes = 0;
      }
      if (energies) energies[s] = e;
      if (times) times[s] = time;
      if (hk.snapshot() && hooks->snapshot_interval > 0 && s % hooks->snapshot_interval == 0)  // :668-675
        hk.take_snapshot(c, s, time, e, run.rho0.as<double>());
      if (hk.stop() && hk.should_stop(s, time, e)) break;  // dynamics.hpp:677-682
    }
    if (defer && steps > 0) {
      std::vector<double> es(8 * static_cast<size_t>(steps + 1));
      CDFT_HIP(hipMemcpyAsync(es.data(), esums.p, sizeof(double) * es.size(), hipMemcpyDeviceToHost, c->stream));
      CDFT_HIP(hipStreamSynchronize(c->stream));
      for (int s = 1; s <= steps; ++s) {
        cdft_result r{};
        c->result_from(es.data() + 8 * s, mu, &r);
        if (energies) energies[s] = r.grand_potential;
        e = r.grand_potential;
      }
    }
    CDFT_HIP(hipMemcpyAsync(rho, run.rho0.p, 8 * g.N, hipMemcpyDeviceToHost, c->stream));
    CDFT_HIP(hipStreamSynchronize(c->stream));
    result->mass_initial = m0;
    result->mass_final = mass_of(run.rho0.as<double>());
    result->time = time;
    result->energy = e;
    result->dt = dt;
    result->evaluations = run.evals;
    result->steps = steps;
  });
}

}  // extern "C"

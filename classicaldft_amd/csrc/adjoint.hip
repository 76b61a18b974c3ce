// Reverse-mode adjoint of the DDFT integrating-factor scheme, FP64.
//
// The forward map is integrating_factor_step (include/dft/algorithms/
// dynamics.hpp:333-401) with a fixed dt and a fixed count P of fixed-point
// iterations: from rho0 with forces f0 = F(rho0), g = f / dV,
//   R(d, f)   = div(d grad(f / dV)) - Lap(d)          (dynamics.hpp:135-229)
//   d^(0)     = rho0
//   d^(i+1)   = P_U0 rho0 + P_(U1-U2) R(rho0, f0) + P_U2 R(d^(i), F(d^(i)))
//   rho1      = d^(P)
// where P_a x = c2r(a(k) r2c(x)) / N and U0, U1, U2 are the propagator
// factors of apply_propagator (dynamics.hpp:247-289).  Every P_a has a real
// multiplier that is even in k, so it is symmetric; the Jacobian H of the
// forces is the Hessian of beta Omega times dV, also symmetric.  The backward
// pass of one step therefore needs, per fixed-point iterate, one full
// transform pair, the transposed stencil and one Hessian-vector product:
//   R_bar_i        = P_U2 d_bar^(i+1)
//   rho0_bar      += P_U0 d_bar^(i+1),  R0_bar += P_(U1-U2) d_bar^(i+1)
//   (d_bar, g_bar) = stencil^T(d^(i), g_i; R_bar_i),  d_bar += H(d^(i)) g_bar / dV
// and the external field collects V_bar += g_bar at every evaluation
// (the forces hold V dV).
//
// Checkpointing: the forward pass keeps rho every C steps (C = ceil(sqrt K)
// by default).  The backward pass recomputes one segment of C steps from its
// checkpoint, keeping each step's iterates and forces on the device (2P
// fields per step), and runs the segment's steps in reverse.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "api_guard.hpp"
#include "context.hpp"
#include "ddft.hpp"
#include "hessian.hpp"

namespace cdft {

namespace {

__device__ __forceinline__ double min_image(double d, double L) { return d - L * rint(d / L); }

// Sum of Gaussian wells; prm holds (A, w, cx, cy, cz) per well.
__global__ void __launch_bounds__(256) k_wells(const double* __restrict__ prm, int nw, Geom g,
                                               double* __restrict__ V) {
  const double Lx = g.nx * g.dx, Ly = g.ny * g.dx, Lz = g.nz * g.dx;
  for (long p = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; p < g.N;
       p += static_cast<long>(gridDim.x) * blockDim.x) {
    const long iz = p % g.nz, iy = (p / g.nz) % g.ny, ix = p / (static_cast<long>(g.ny) * g.nz);
    double v = 0.0;
    for (int j = 0; j < nw; ++j) {
      const double* q = prm + 5 * j;
      const double ex = min_image(ix * g.dx - q[2], Lx), ey = min_image(iy * g.dx - q[3], Ly),
                   ez = min_image(iz * g.dx - q[4], Lz);
      const double d2 = ex * ex + ey * ey + ez * ez;
      v += q[0] * exp(-d2 / (2.0 * q[1] * q[1]));
    }
    V[p] = v;
  }
}

// dL/d(A, w, cx, cy, cz) of every well from V_bar: per-block sums into
// partials[blk * 5 nw + 5 j + c].
__global__ void __launch_bounds__(256) k_wells_grad(const double* __restrict__ prm, int nw, Geom g,
                                                    const double* __restrict__ vbar, double* __restrict__ partials) {
  __shared__ double sh[5][256];
  const double Lx = g.nx * g.dx, Ly = g.ny * g.dx, Lz = g.nz * g.dx;
  for (int j = 0; j < nw; ++j) {
    const double* q = prm + 5 * j;
    const double A = q[0], w = q[1], iw2 = 1.0 / (w * w);
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (long p = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; p < g.N;
         p += static_cast<long>(gridDim.x) * blockDim.x) {
      const long iz = p % g.nz, iy = (p / g.nz) % g.ny, ix = p / (static_cast<long>(g.ny) * g.nz);
      const double ex = min_image(ix * g.dx - q[2], Lx), ey = min_image(iy * g.dx - q[3], Ly),
                   ez = min_image(iz * g.dx - q[4], Lz);
      const double d2 = ex * ex + ey * ey + ez * ez;
      const double G = exp(-0.5 * d2 * iw2);
      const double vb = vbar[p], avg = vb * A * G * iw2;
      s[0] += vb * G;          // dV/dA = G
      s[1] += avg * d2 / w;    // dV/dw = A G d^2 / w^3
      s[2] += avg * ex;        // dV/dc = A G (r - c) / w^2
      s[3] += avg * ey;
      s[4] += avg * ez;
    }
    for (int c = 0; c < 5; ++c) sh[c][threadIdx.x] = s[c];
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
      if (threadIdx.x < h)
        for (int c = 0; c < 5; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x < 5) partials[static_cast<long>(blockIdx.x) * 5 * nw + 5 * j + threadIdx.x] = sh[threadIdx.x][0];
    __syncthreads();
  }
}

// Transposed right-hand side stencil.  For R = div(rho grad g) - Lap(rho)
// with g = force / dV and the face fluxes of k_ddft_rhs, given R_bar:
//   rho_bar = -D sum_axes [(g+ - g)(Rb+ - Rb) + (g - g-)(Rb - Rb-)] - Lap(Rb)
//   g_bar   =  D sum_axes [(rho + rho+)(Rb+ - Rb) - (rho- + rho)(Rb - Rb-)]
// Writes rs = rho_bar, fbar = g_bar / dV, and adds g_bar to vbar.
__global__ void __launch_bounds__(256) k_rhs_adjoint(const double* __restrict__ rho, const double* __restrict__ force,
                                                     const double* __restrict__ Rb, double* __restrict__ rs,
                                                     double* __restrict__ fbar, double* __restrict__ vbar, Geom G) {
  const long nx = G.nx, ny = G.ny, nz = G.nz, nyz = ny * nz;
  const double D = 0.5 / (G.dx * G.dx), D2 = 1.0 / (G.dx * G.dx), idv = 1.0 / G.dv;
  for (long p = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; p < G.N;
       p += static_cast<long>(gridDim.x) * blockDim.x) {
    const long iz = p % nz, iy = (p / nz) % ny, ix = p / nyz;
    const long nb[3][2] = {{((ix + 1) % nx) * nyz + iy * nz + iz, ((ix - 1 + nx) % nx) * nyz + iy * nz + iz},
                           {ix * nyz + ((iy + 1) % ny) * nz + iz, ix * nyz + ((iy - 1 + ny) % ny) * nz + iz},
                           {ix * nyz + iy * nz + (iz + 1) % nz, ix * nyz + iy * nz + (iz - 1 + nz) % nz}};
    const double r0 = rho[p], g0 = force[p] * idv, b0 = Rb[p];
    double rsv = 0.0, gb = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const long P = nb[a][0], M = nb[a][1];
      const double rp = rho[P], rm = rho[M], gp = force[P] * idv, gm = force[M] * idv, bp = Rb[P], bm = Rb[M];
      rsv -= D * ((gp - g0) * (bp - b0) + (g0 - gm) * (b0 - bm));
      rsv -= D2 * (bp + bm - 2.0 * b0);
      gb += D * ((r0 + rp) * (bp - b0) - (rm + r0) * (b0 - bm));
    }
    rs[p] = rsv;
    fbar[p] = gb * idv;
    vbar[p] += gb;
  }
}

// Propagator factors applied to the spectrum D of d_bar (layout kz, kx, ky):
// Arho += U0 D; first iterate: AR0 += U1 D; otherwise AR0 += (U1 - U2) D and
// Rout = U2 D.  Same factors as k_xprop (dynamics.hpp:247-289).
__global__ void __launch_bounds__(256) k_prop_adjoint(const cx<double>* __restrict__ Dk, cx<double>* __restrict__ Arho,
                                                      cx<double>* __restrict__ AR0, cx<double>* __restrict__ Rout,
                                                      const double* __restrict__ lam, const double* __restrict__ fac,
                                                      double dt, Geom g, int first) {
  const long nxy = static_cast<long>(g.nx) * g.ny;
  for (long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; idx < g.M;
       idx += static_cast<long>(gridDim.x) * blockDim.x) {
    const int kz = static_cast<int>(idx / nxy);
    const long r = idx - kz * nxy;
    const int x = static_cast<int>(r / g.ny), y = static_cast<int>(r - static_cast<long>(x) * g.ny);
    const double L = lam[x] + lam[g.nx + y] + lam[g.nx + g.ny + kz];
    const double ex = fac[x] * fac[g.nx + y] * fac[g.nx + g.ny + kz];
    const bool zero = idx == 0;
    const double U0 = ex;
    const double U1 = zero ? dt : (ex - 1.0) / L;
    const double U2 = zero ? dt / 2.0 : (ex - 1.0 - dt * L) / (L * L * dt);
    const cx<double> d = Dk[idx];
    Arho[idx] = {Arho[idx].re + U0 * d.re, Arho[idx].im + U0 * d.im};
    const double c0 = first ? U1 : U1 - U2;
    AR0[idx] = {AR0[idx].re + c0 * d.re, AR0[idx].im + c0 * d.im};
    if (!first) Rout[idx] = {U2 * d.re, U2 * d.im};
  }
}

__global__ void __launch_bounds__(256) k_add3(double* __restrict__ out, const double* __restrict__ a,
                                              const double* __restrict__ b, const double* __restrict__ c, long n) {
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * blockDim.x)
    out[i] = a[i] + b[i] + (c ? c[i] : 0.0);
}

// loss partial sums of (rho - t)^2 and rho_bar = 2 (rho - t) dV
__global__ void __launch_bounds__(256) k_loss(const double* __restrict__ rho, const double* __restrict__ t,
                                              double* __restrict__ rbar, long n, double dv, double* partials) {
  double s = 0.0, c = 0.0;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    const double d = rho[i] - t[i];
    acc2(s, c, d * d);
    rbar[i] = 2.0 * d * dv;
  }
  const double v[1] = {s + c};
  block_sums(v, 1, partials);
}

void check_config(const cdft_adjoint_config* cfg) {
  if (!cfg) throw std::invalid_argument("null adjoint config");
  if (!(cfg->dt > 0.0)) throw std::invalid_argument("adjoint config: dt must be positive");
  if (cfg->fp_iterations < 1) throw std::invalid_argument("adjoint config: fp_iterations must be >= 1");
  if (cfg->n_steps < 1) throw std::invalid_argument("adjoint config: n_steps must be >= 1");
}

// The forward map (fixed dt, P iterations) on a DdftRun whose rho0 holds the
// density and f0 its forces.  When `st` is given it receives the step's
// states: st[0] = rho0, st[1] = f0, st[2i] = d^(i), st[2i+1] = F(d^(i)) for
// i = 1..P-1.
void fixed_step(DdftRun& run, double dt, int P, double* const* st, double* min_d) {
  Context* c = run.c;
  const long N = c->g.N;
  const size_t bytes = 8 * static_cast<size_t>(N);
  if (st) {
    CDFT_HIP(hipMemcpyAsync(st[0], run.rho0.p, bytes, hipMemcpyDeviceToDevice, c->stream));
    CDFT_HIP(hipMemcpyAsync(st[1], run.f0.p, bytes, hipMemcpyDeviceToDevice, c->stream));
  }
  run.rhs_forward<double>(run.rho0.as<double>(), run.f0.as<double>(), run.SR0.as<cx<double>>(), 1, true);
  run.forward<double>(run.rho0.as<double>(), run.Srho0.as<cx<double>>(), 0, true);
  run.update_timestep(dt);
  CDFT_HIP(hipMemcpyAsync(run.d1.p, run.rho0.p, bytes, hipMemcpyDeviceToDevice, c->stream));
  for (int it = 0; it < P; ++it) {
    const bool first = it == 0;  // d^(0) = rho0: its forces are f0
    if (!first) {
      run.compute(run.d1.as<double>(), run.f1.as<double>());
      if (st) {
        CDFT_HIP(hipMemcpyAsync(st[2 * it], run.d1.p, bytes, hipMemcpyDeviceToDevice, c->stream));
        CDFT_HIP(hipMemcpyAsync(st[2 * it + 1], run.f1.p, bytes, hipMemcpyDeviceToDevice, c->stream));
      }
      run.rhs_forward<double>(run.d1.as<double>(), run.f1.as<double>(), run.SR1.as<cx<double>>(), 2, false);
    }
    double dev = 0.0, mn = 0.0;
    run.propagate<double>(first, &dev, &mn);
    *min_d = std::min(*min_d, mn);
  }
  run.compute(run.d1.as<double>(), run.f1.as<double>());
  CDFT_HIP(hipMemcpyAsync(run.rho0.p, run.d1.p, bytes, hipMemcpyDeviceToDevice, c->stream));
  CDFT_HIP(hipMemcpyAsync(run.f0.p, run.f1.p, bytes, hipMemcpyDeviceToDevice, c->stream));
}

struct Adjoint {
  Context* c;
  DdftRun& run;
  Hessian hess;
  int P;
  double dt;
  DevBuf Dk, Arho, AR0, Rk;                  // spectra (kz, kx, ky), FP64
  DevBuf db, Rb, rs, fbar, h, rb, abar, vbar;  // real fields

  Adjoint(Context* c_, DdftRun& r, int P_, double dt_)
      : c(c_), run(r), hess(c_), P(P_), dt(dt_), Dk(16 * c_->g.M), Arho(16 * c_->g.M), AR0(16 * c_->g.M),
        Rk(16 * c_->g.M), db(8 * c_->g.N), Rb(8 * c_->g.N), rs(8 * c_->g.N), fbar(8 * c_->g.N), h(8 * c_->g.N),
        rb(8 * c_->g.N), abar(8 * c_->g.N), vbar(8 * c_->g.N) {
    CDFT_HIP(hipMemsetAsync(vbar.p, 0, 8 * c->g.N, c->stream));
  }

  // abar: rho_bar of the step's output on entry, of its input on return.
  void step_back(double* const* st) {
    const Geom& g = c->g;
    const size_t nb = 8 * static_cast<size_t>(g.N);
    const dim3 vb(vec_blocks(g.N)), vs(vec_blocks(g.M));
    run.update_timestep(dt);
    CDFT_HIP(hipMemsetAsync(Arho.p, 0, 16 * g.M, c->stream));
    CDFT_HIP(hipMemsetAsync(AR0.p, 0, 16 * g.M, c->stream));
    CDFT_HIP(hipMemcpyAsync(db.p, abar.p, nb, hipMemcpyDeviceToDevice, c->stream));
    for (int i = P - 1; i >= 0; --i) {
      forward_full(c, db.as<double>(), Dk.as<cx<double>>());
      hipLaunchKernelGGL(k_prop_adjoint, vs, dim3(256), 0, c->stream, Dk.as<cx<double>>(), Arho.as<cx<double>>(),
                         AR0.as<cx<double>>(), Rk.as<cx<double>>(), run.lam.as<double>(), run.fac.as<double>(), dt,
                         g, i == 0 ? 1 : 0);
      if (i > 0) {
        inverse_full(c, Rk.as<cx<double>>(), Rb.as<double>());
        hipLaunchKernelGGL(k_rhs_adjoint, vb, dim3(256), 0, c->stream, st[2 * i], st[2 * i + 1], Rb.as<double>(),
                           rs.as<double>(), fbar.as<double>(), vbar.as<double>(), g);
        hess.apply(st[2 * i], fbar.as<double>(), h.as<double>());
        hipLaunchKernelGGL(k_add3, vb, dim3(256), 0, c->stream, db.as<double>(), rs.as<double>(), h.as<double>(),
                           static_cast<const double*>(nullptr), g.N);
      }
    }
    inverse_full(c, AR0.as<cx<double>>(), Rb.as<double>());
    hipLaunchKernelGGL(k_rhs_adjoint, vb, dim3(256), 0, c->stream, st[0], st[1], Rb.as<double>(), rs.as<double>(),
                       fbar.as<double>(), vbar.as<double>(), g);
    hess.apply(st[0], fbar.as<double>(), h.as<double>());
    inverse_full(c, Arho.as<cx<double>>(), rb.as<double>());
    hipLaunchKernelGGL(k_add3, vb, dim3(256), 0, c->stream, abar.as<double>(), rb.as<double>(), rs.as<double>(),
                       static_cast<const double*>(h.as<double>()), g.N);
    CDFT_HIP(hipGetLastError());
  }
};

void upload(Context* c, void* dst, const double* src, long n) {
  CDFT_HIP(hipMemcpyAsync(dst, src, 8 * static_cast<size_t>(n), hipMemcpyHostToDevice, c->stream));
}

void download(Context* c, double* dst, const void* src, long n) {
  CDFT_HIP(hipMemcpyAsync(dst, src, 8 * static_cast<size_t>(n), hipMemcpyDeviceToHost, c->stream));
}

void wells_field(Context* c, int nw, const double* params, DevBuf& prm, double* V) {
  if (nw < 0 || (nw > 0 && !params)) throw std::invalid_argument("bad well parameters");
  for (int j = 0; j < nw; ++j)
    if (!(params[5 * j + 1] > 0.0)) throw std::invalid_argument("well width must be positive");
  if (nw > 0) upload(c, prm.p, params, 5L * nw);
  hipLaunchKernelGGL(k_wells, dim3(vec_blocks(c->g.N)), dim3(256), 0, c->stream, prm.as<double>(), nw, c->g, V);
  CDFT_HIP(hipGetLastError());
}

}  // namespace

}  // namespace cdft

using namespace cdft;

extern "C" {

int cdft_hessian_vector(cdft_ctx* ctx, const double* rho, const double* u, double* out) {
  return guarded([&] {
    check_ctx(ctx);
    if (!rho || !u || !out) throw std::invalid_argument("null array");
    Context* c = ctx->impl;
    require_fp64(c);
    CDFT_HIP(hipSetDevice(c->device));
    const long N = c->g.N;
    DevBuf dr(8 * N), du(8 * N), dout(8 * N);
    upload(c, dr.p, rho, N);
    upload(c, du.p, u, N);
    Hessian hs(c);
    hs.apply(dr.as<double>(), du.as<double>(), dout.as<double>());
    download(c, out, dout.p, N);
    CDFT_HIP(hipStreamSynchronize(c->stream));
  });
}

int cdft_wells_potential(cdft_ctx* ctx, int n_wells, const double* params, double* vext) {
  return guarded([&] {
    check_ctx(ctx);
    if (!vext) throw std::invalid_argument("null array");
    Context* c = ctx->impl;
    CDFT_HIP(hipSetDevice(c->device));
    DevBuf prm(8 * 5 * static_cast<size_t>(std::max(n_wells, 1))), V(8 * c->g.N);
    wells_field(c, n_wells, params, prm, V.as<double>());
    download(c, vext, V.p, c->g.N);
    CDFT_HIP(hipStreamSynchronize(c->stream));
  });
}

int cdft_ddft_fixed_run(cdft_ctx* ctx, const cdft_adjoint_config* config, const double* vext, double mu, double* rho,
                        double* min_density) {
  return guarded([&] {
    check_ctx(ctx);
    check_config(config);
    if (!rho) throw std::invalid_argument("null array");
    Context* c = ctx->impl;
    require_fp64(c);
    CDFT_HIP(hipSetDevice(c->device));
    const long N = c->g.N;
    if (vext) upload(c, c->d_vext, vext, N);
    DdftRun run(c, vext ? c->d_vext : nullptr, mu);
    upload(c, run.rho0.p, rho, N);
    run.compute(run.rho0.as<double>(), run.f0.as<double>());
    double mn = 1e300;
    for (int s = 0; s < config->n_steps; ++s) fixed_step(run, config->dt, config->fp_iterations, nullptr, &mn);
    download(c, rho, run.rho0.p, N);
    CDFT_HIP(hipStreamSynchronize(c->stream));
    if (min_density) *min_density = mn;
  });
}

int cdft_ddft_adjoint_wells(cdft_ctx* ctx, const cdft_adjoint_config* config, int n_wells, const double* params,
                            double mu, const double* rho0, const double* target, cdft_adjoint_result* result,
                            double* grad, double* grad_rho0, double* grad_vext, double* rho_final) {
  return guarded([&] {
    check_ctx(ctx);
    check_config(config);
    if (!rho0 || !target || !result || (n_wells > 0 && !grad)) throw std::invalid_argument("null array");
    Context* c = ctx->impl;
    require_fp64(c);
    CDFT_HIP(hipSetDevice(c->device));
    const Geom& g = c->g;
    const long N = g.N;
    const size_t nb = 8 * static_cast<size_t>(N);
    const int K = config->n_steps, P = config->fp_iterations;
    const int C = config->checkpoint_interval > 0
                      ? std::min(config->checkpoint_interval, K)
                      : std::max(1, static_cast<int>(std::ceil(std::sqrt(static_cast<double>(K)))));
    const int ncp = (K + C - 1) / C;

    DevBuf prm(8 * 5 * static_cast<size_t>(std::max(n_wells, 1))), V(nb), tgt(nb);
    wells_field(c, n_wells, params, prm, V.as<double>());
    upload(c, tgt.p, target, N);
    DdftRun run(c, V.as<double>(), mu);
    upload(c, run.rho0.p, rho0, N);

    // forward pass: checkpoints every C steps
    DevBuf cps(nb * ncp);
    double mn = 1e300;
    run.compute(run.rho0.as<double>(), run.f0.as<double>());
    for (int s = 0; s < K; ++s) {
      if (s % C == 0)
        CDFT_HIP(hipMemcpyAsync(cps.as<double>() + static_cast<size_t>(s / C) * N, run.rho0.p, nb,
                                hipMemcpyDeviceToDevice, c->stream));
      fixed_step(run, config->dt, P, nullptr, &mn);
    }
    if (rho_final) download(c, rho_final, run.rho0.p, N);

    Adjoint adj(c, run, P, config->dt);
    hipLaunchKernelGGL(k_loss, dim3(vec_blocks(N)), dim3(256), 0, c->stream, run.rho0.as<double>(),
                       tgt.as<double>(), adj.abar.as<double>(), N, g.dv, run.part.as<double>());
    launch_finalize(run.part.as<double>(), static_cast<int>(vec_blocks(N)), run.sums.as<double>() + 32, c->stream);
    double loss = 0.0;
    download(c, &loss, run.sums.as<double>() + 32, 1);
    CDFT_HIP(hipStreamSynchronize(c->stream));

    // backward pass, one recomputed segment at a time
    DevBuf store(nb * 2 * static_cast<size_t>(P) * C);
    std::vector<double*> ptrs(2 * static_cast<size_t>(P) * C, nullptr);
    for (size_t i = 0; i < ptrs.size(); ++i) ptrs[i] = store.as<double>() + i * N;
    for (int m = ncp - 1; m >= 0; --m) {
      const int s0 = m * C, s1 = std::min(K, s0 + C);
      CDFT_HIP(hipMemcpyAsync(run.rho0.p, cps.as<double>() + static_cast<size_t>(m) * N, nb, hipMemcpyDeviceToDevice,
                              c->stream));
      run.compute(run.rho0.as<double>(), run.f0.as<double>());
      for (int s = s0; s < s1; ++s) fixed_step(run, config->dt, P, ptrs.data() + 2 * P * (s - s0), &mn);
      for (int s = s1 - 1; s >= s0; --s) adj.step_back(ptrs.data() + 2 * P * (s - s0));
    }

    if (n_wells > 0) {
      const unsigned nblk = vec_blocks(N);
      DevBuf part(8 * 5 * static_cast<size_t>(n_wells) * nblk);
      hipLaunchKernelGGL(k_wells_grad, dim3(nblk), dim3(256), 0, c->stream, prm.as<double>(), n_wells, g,
                         adj.vbar.as<double>(), part.as<double>());
      CDFT_HIP(hipGetLastError());
      std::vector<double> hp(5 * static_cast<size_t>(n_wells) * nblk);
      download(c, hp.data(), part.p, static_cast<long>(hp.size()));
      CDFT_HIP(hipStreamSynchronize(c->stream));
      for (int k = 0; k < 5 * n_wells; ++k) {
        long double s = 0.0L;
        for (unsigned b = 0; b < nblk; ++b) s += hp[static_cast<size_t>(b) * 5 * n_wells + k];
        grad[k] = static_cast<double>(s);
      }
    }
    if (grad_rho0) download(c, grad_rho0, adj.abar.p, N);
    if (grad_vext) download(c, grad_vext, adj.vbar.p, N);
    CDFT_HIP(hipStreamSynchronize(c->stream));
    result->loss = loss * g.dv;
    result->min_density = mn;
    result->evaluations = run.evals;
    result->hessian_products = adj.hess.products;
    result->checkpoints = ncp;
    result->store_bytes = static_cast<double>(nb) * (ncp + 2.0 * P * C);
  });
}

}  // extern "C"

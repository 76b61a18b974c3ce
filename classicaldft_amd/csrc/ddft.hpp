// DDFT integrating-factor machinery shared by the forward scheme
// (cdft_ddft_run) and its adjoint: the right-hand side stencil, the
// propagator fused with the x stage of the transforms, the FP64 update of
// the iterate, and DdftRun, which owns the device state of one run
// (integrating_factor_step, include/dft/algorithms/dynamics.hpp:79-401).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <utility>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "context.hpp"
#include "device_common.hpp"
#include "line_fft.hpp"

namespace cdft {

// Kernels, helpers and DdftRun have internal linkage: each translation unit
// that includes this header gets its own copy.
namespace {

// Boundary mask of the grid (grid.hpp:67-82): a point on any face of the box.
__device__ __forceinline__ bool on_face(long i, const Geom& g) {
  const long iz = i % g.nz;
  const long t = i / g.nz;
  const long iy = t % g.ny;
  const long ix = t / g.ny;
  return ix == 0 || ix == g.nx - 1 || iy == 0 || iy == g.ny - 1 || iz == 0 || iz == g.nz - 1;
}

__device__ __forceinline__ void acc2(double& s, double& c, double x) {
  const double t = s + x;
  c += (fabs(s) >= fabs(x)) ? ((s - t) + x) : ((x - t) + s);
  s = t;
}

// Block-reduce up to 4 sums into partials[blockIdx.x*8 + i] (slots 4..7 zero).
__device__ __forceinline__ void block_sums(const double* v, int nv, double* partials, long row = blockIdx.x) {
  __shared__ double sh[4][256];
  for (int i = 0; i < 4; ++i) sh[i][threadIdx.x] = i < nv ? v[i] : 0.0;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int i = 0; i < 4; ++i) sh[i][threadIdx.x] += sh[i][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 8) partials[row * 8 + threadIdx.x] = threadIdx.x < 4 ? sh[threadIdx.x][0] : 0.0;
}

// ---------------------------------------------------------------- DDFT kernels
// R = div(rho grad(F/dV)) - Laplacian(rho), periodic 7-point stencils (dynamics.hpp:135-229).
// A thread owns one (y, z) column and marches over an x chunk of xc planes,
// keeping the x-1, x, x+1 values in registers: the x neighbours are read
// once, the y and z neighbours come from the same plane.  Block row
// blockIdx.y * gridDim.x + blockIdx.x of partials gets the block's sum of R.
__global__ void __launch_bounds__(256) k_ddft_rhs(const double* __restrict__ rho, const double* __restrict__ force,
                                                  double* __restrict__ rhs, Geom G, int xc, double* partials) {
  double s = 0, c = 0;
  const int nx = G.nx, ny = G.ny, nz = G.nz, nyz = ny * nz;
  const double D = 0.5 / (G.dx * G.dx), D2 = 1.0 / (G.dx * G.dx), idv = 1.0 / G.dv;
  const int x0 = blockIdx.y * xc, x1 = min(x0 + xc, nx);
  for (int col = blockIdx.x * blockDim.x + threadIdx.x; col < nyz; col += gridDim.x * blockDim.x) {
    const int iy = col / nz, iz = col - iy * nz;
    const int cyp = iy + 1 == ny ? col - (ny - 1) * nz : col + nz;
    const int cym = iy == 0 ? col + (ny - 1) * nz : col - nz;
    const int czp = iz + 1 == nz ? col - (nz - 1) : col + 1;
    const int czm = iz == 0 ? col + (nz - 1) : col - 1;
    const long xm0 = x0 == 0 ? nx - 1 : x0 - 1;
    double rmx = rho[xm0 * nyz + col], xmx = force[xm0 * nyz + col] * idv;
    double d0 = rho[static_cast<long>(x0) * nyz + col], x0v = force[static_cast<long>(x0) * nyz + col] * idv;
    for (int x = x0; x < x1; ++x) {
      const long b = static_cast<long>(x) * nyz, bp = static_cast<long>(x + 1 == nx ? 0 : x + 1) * nyz;
      const double rpx = rho[bp + col], xpx = force[bp + col] * idv;
      const double rpy = rho[b + cyp], rmy = rho[b + cym], rpz = rho[b + czp], rmz = rho[b + czm];
      const double xpy = force[b + cyp] * idv, xmy = force[b + cym] * idv;
      const double xpz = force[b + czp] * idv, xmz = force[b + czm] * idv;
      double r = D * ((rpx + d0) * (xpx - x0v) - (d0 + rmx) * (x0v - xmx)) +
                 D * ((rpy + d0) * (xpy - x0v) - (d0 + rmy) * (x0v - xmy)) +
                 D * ((rpz + d0) * (xpz - x0v) - (d0 + rmz) * (x0v - xmz));
      r -= D2 * (rpx + rmx - 2.0 * d0) + D2 * (rpy + rmy - 2.0 * d0) + D2 * (rpz + rmz - 2.0 * d0);
      rhs[b + col] = r;
      acc2(s, c, r);
      rmx = d0;
      xmx = x0v;
      d0 = rpx;
      x0v = xpx;
    }
  }
  const double vals[1] = {s + c};
  block_sums(vals, 1, partials, blockIdx.y * gridDim.x + blockIdx.x);
}

// Integrating-factor propagator fused with the x stage of the transforms
// (dynamics.hpp:247-289): x-FFT of R1 (z, y done), combine
// w = U0 rho0_k + U1 R0_k + U2 (R1_k - R0_k) on the mean-free spectra (the
// k = 0 mode is propagated separately in FP64), inverse x-FFT.
template <typename S, typename C, int EPT, int NTH>
__global__ void __launch_bounds__(NTH) k_xprop(const cx<S>* __restrict__ SR1, const cx<S>* __restrict__ Srho0,
                                               const cx<S>* __restrict__ SR0, cx<S>* __restrict__ Sout, Geom g,
                                               FftPlan px, const cx<C>* __restrict__ twx, int ty,
                                               const double* __restrict__ lam, const double* __restrict__ fac,
                                               double dt, int first_iter) {
  extern __shared__ unsigned char smem_raw[];
  cx<C>* buf = reinterpret_cast<cx<C>*>(smem_raw);
  const int nx = g.nx, ny = g.ny, ls = nx + 1;
  const int y0 = blockIdx.x * ty;
  const int kz = blockIdx.y;
  const long off = static_cast<long>(kz) * nx * ny;
  if (!first_iter) {
    for (int e = threadIdx.x; e < ty * nx; e += blockDim.x) {
      const int x = e / ty, yl = e - x * ty, y = y0 + yl;
      const cx<S> v = (y < ny) ? SR1[off + x * ny + y] : cx<S>{S(0), S(0)};
      buf[yl * ls + x] = {static_cast<C>(v.re), static_cast<C>(v.im)};
    }
    __syncthreads();
    fft_lines<C, EPT>(buf, ty, ls, px, -1, twx);
  }
  // lam/fac layout: [0, nx) x axis, [nx, nx+ny) y axis, [nx+ny, nx+ny+nzh) z axis
  cx<C> wv[EPT];
#pragma unroll
  for (int m = 0; m < EPT; ++m) {
    const int e = threadIdx.x + m * blockDim.x;
    wv[m] = {C(0), C(0)};
    if (e < ty * nx) {
      const int x = e / ty, yl = e - x * ty, y = min(y0 + yl, ny - 1);
      const double L = lam[x] + lam[nx + y] + lam[nx + ny + kz];
      const double ex = fac[x] * fac[nx + y] * fac[nx + ny + kz];
      const bool zero = (x == 0 && y == 0 && kz == 0);
      const double U0 = ex;
      const double U1 = zero ? dt : (ex - 1.0) / L;
      const double U2 = zero ? dt / 2.0 : (ex - 1.0 - dt * L) / (L * L * dt);
      const long idx = off + x * ny + y;
      const cx<S> r0 = SR0[idx], p0 = Srho0[idx];
      double r1re = static_cast<double>(r0.re), r1im = static_cast<double>(r0.im);
      if (!first_iter) {
        r1re = static_cast<double>(buf[yl * ls + x].re);
        r1im = static_cast<double>(buf[yl * ls + x].im);
      }
      double wr = U0 * static_cast<double>(p0.re) + U1 * static_cast<double>(r0.re) +
                  U2 * (r1re - static_cast<double>(r0.re));
      double wi = U0 * static_cast<double>(p0.im) + U1 * static_cast<double>(r0.im) +
                  U2 * (r1im - static_cast<double>(r0.im));
      if (zero) wr = wi = 0.0;  // mean mode handled in FP64 by k_ddft_update
      wv[m] = {static_cast<C>(wr), static_cast<C>(wi)};
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < EPT; ++m) {
    const int e = threadIdx.x + m * blockDim.x;
    if (e < ty * nx) {
      const int x = e / ty, yl = e - x * ty;
      buf[yl * ls + x] = wv[m];
    }
  }
  __syncthreads();
  fft_lines<C, EPT>(buf, ty, ls, px, +1, twx);
  for (int e = threadIdx.x; e < ty * nx; e += blockDim.x) {
    const int x = e / ty, yl = e - x * ty, y = y0 + yl;
    if (y < ny) Sout[off + x * ny + y] = {static_cast<S>(buf[yl * ls + x].re), static_cast<S>(buf[yl * ls + x].im)};
  }
}

// d1_new = mean_new + raw - mean(raw); max |d1_new - d1_old|; min d1_new (apply_propagator, dynamics.hpp:289-303).
// sums: [0] sum rho0, [8] sum R0, [16] sum R1; the sum of raw is reduced by
// every block from launch_sum's raw_rows (column 0 of raw_n rows).  d1_old is
// `old` (rho0 on an attempt's first iteration, else d1 itself).
__global__ void __launch_bounds__(256) k_ddft_update(const double* __restrict__ raw, const double* old,
                                                     double* d1, long n, const double* __restrict__ sums,
                                                     const double* __restrict__ raw_rows, int raw_n, double dt,
                                                     int first_iter, double* partials) {
  const double inv_n = 1.0 / static_cast<double>(n);
  const double s0 = sums[0], r0 = sums[8], r1 = first_iter ? sums[8] : sums[16];
  const double mean_new = (s0 + dt * r0 + 0.5 * dt * (r1 - r0)) * inv_n;
  const double shift = mean_new - block_sum_col0<256>(raw_rows, raw_n) * inv_n;
  double mx = 0.0, mn = 1e300;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    const double nv = raw[i] + shift;
    // fmax drops a NaN operand: a non-finite iterate is reported as an infinite deviation
    mx = fmax(mx, isfinite(nv) ? fabs(old[i] - nv) : INFINITY);
    mn = fmin(mn, nv);
    d1[i] = nv;
  }
  __shared__ double smx[256], smn[256];
  smx[threadIdx.x] = mx;
  smn[threadIdx.x] = mn;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      smx[threadIdx.x] = fmax(smx[threadIdx.x], smx[threadIdx.x + w]);
      smn[threadIdx.x] = fmin(smn[threadIdx.x], smn[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[static_cast<long>(blockIdx.x) * 8 + 0] = smx[0];
    partials[static_cast<long>(blockIdx.x) * 8 + 1] = smn[0];
  }
}

__global__ void k_finalize_maxmin(const double* __restrict__ partials, int nblk, double* __restrict__ res) {
  __shared__ double smx[256], smn[256];
  double mx = 0.0, mn = 1e300;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    mx = fmax(mx, partials[static_cast<long>(b) * 8 + 0]);
    mn = fmin(mn, partials[static_cast<long>(b) * 8 + 1]);
  }
  smx[threadIdx.x] = mx;
  smn[threadIdx.x] = mn;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      smx[threadIdx.x] = fmax(smx[threadIdx.x], smx[threadIdx.x + w]);
      smn[threadIdx.x] = fmin(smn[threadIdx.x], smn[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    res[0] = smx[0];
    res[1] = smn[0];
  }
}

// the 8 raw sums of an evaluation, device to device (one launch, no copy engine)
__global__ void k_copy_sums(const double* __restrict__ src, double* __restrict__ dst) {
  if (threadIdx.x < 8) dst[threadIdx.x] = src[threadIdx.x];
}

constexpr int kVecBlocks = 1024;

unsigned vec_blocks(long n) {
  long b = (n + 255) / 256;
  return static_cast<unsigned>(std::min<long>(std::max<long>(b, 1), kVecBlocks));
}

// k_ddft_rhs over the grid with at most kVecBlocks blocks; returns the number
// of block sums it writes to partials.
inline int launch_ddft_rhs(const double* rho, const double* force, double* rhs, const Geom& g, double* partials,
                           hipStream_t st) {
  if (static_cast<long>(g.ny) * g.nz >= (1L << 31)) throw Unsupported("y-z plane too large for the DDFT stencil");
  const int nyz = g.ny * g.nz;
  const int bx = std::min((nyz + 255) / 256, kVecBlocks);
  const int nxc = std::max(1, std::min(g.nx, kVecBlocks / bx));
  const int xc = (g.nx + nxc - 1) / nxc;
  const int by = (g.nx + xc - 1) / xc;
  hipLaunchKernelGGL(k_ddft_rhs, dim3(bx, by), dim3(256), 0, st, rho, force, rhs, g, xc, partials);
  return bx * by;
}

struct DevBuf {
  void* p = nullptr;
  DevBuf() = default;
  explicit DevBuf(size_t bytes) {
    if (bytes) CDFT_HIP(hipMalloc(&p, bytes));
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// ---------------------------------------------------------------- host callbacks
// Runs the callbacks of a cdft_hooks (any force callable, boundary / stop
// conditions, snapshots and constraints, as the reference's algorithms take
// them: dynamics.hpp:21-22,545-567, picard.hpp:16-21) on host copies of
// device fields.  A nonzero callback status aborts the entry point.
struct HostHooks {
  const cdft_hooks* h = nullptr;
  std::vector<double> a, b;

  bool force() const { return h && h->force; }
  bool project() const { return h && h->project; }
  bool stop() const { return h && h->stop; }
  bool snapshot() const { return h && h->snapshot; }

  void fetch(Context* c, const double* d) {
    a.resize(static_cast<size_t>(c->g.N));
    CDFT_HIP(hipMemcpyAsync(a.data(), d, 8 * c->g.N, hipMemcpyDeviceToHost, c->stream));
    CDFT_HIP(hipStreamSynchronize(c->stream));
  }
  void store(Context* c, double* d, const std::vector<double>& v) {
    CDFT_HIP(hipMemcpyAsync(d, v.data(), 8 * c->g.N, hipMemcpyHostToDevice, c->stream));
    CDFT_HIP(hipStreamSynchronize(c->stream));
  }
  // energy and forces of the force callback at d_rho (forces to d_forces)
  double eval(Context* c, const double* d_rho, double* d_forces) {
    fetch(c, d_rho);
    b.assign(static_cast<size_t>(c->g.N), 0.0);
    double e = 0.0;
    if (h->force(h->user, a.data(), &e, b.data()) != 0) throw std::runtime_error("the force callback failed");
    if (d_forces) store(c, d_forces, b);
    return e;
  }
  // the boundary condition / constraint on d_rho, in place
  void apply_project(Context* c, double* d_rho) {
    fetch(c, d_rho);
    if (h->project(h->user, a.data()) != 0) throw std::runtime_error("the boundary / constraint callback failed");
    store(c, d_rho, a);
  }
  bool should_stop(int step, double time, double e) const {
    int flag = 0;
    if (h->stop(h->user, step, time, e, &flag) != 0) throw std::runtime_error("the stop callback failed");
    return flag != 0;
  }
  void take_snapshot(Context* c, int step, double time, double e, const double* d_rho) {
    fetch(c, d_rho);
    if (h->snapshot(h->user, step, time, e, a.data()) != 0) throw std::runtime_error("the snapshot callback failed");
  }
};

// Restart limit of one integrating-factor step: the reference divides dt by
// 10 on every restart with no floor (dynamics.hpp:374-392); here a step that
// restarts more than this many times fails instead of running on with a
// vanishing dt.
constexpr int kMaxRestarts = 30;

// ------------------------------------------------------------------ DDFT
struct DdftRun {
  Context* c;
  const double* d_vext;
  double mu;
  DevBuf rho0, d1, f0, f1, rhs, raw;
  DevBuf Srho0, SR0, SR1, Sw;
  DevBuf lam, fac, part, sums;
  DevBuf rpart;  // block sums of the propagated field (read by k_ddft_update while it writes part)
  std::vector<double> lam_h;
  double st_dt = 0.0;
  long evals = 0;
  HostHooks* hooks = nullptr;  // a host force callback in place of the context's functional
  double* mm_h = nullptr;      // pinned, coherent host memory: {max deviation, min d1} of the last propagation
  double* mm_d = nullptr;      // mm_h's device address
  hipEvent_t mm_ready = nullptr;

  DdftRun(Context* c_, const double* vx, double m)
      : c(c_), d_vext(vx), mu(m), rho0(8 * c_->g.N), d1(8 * c_->g.N), f0(8 * c_->g.N), f1(8 * c_->g.N),
        rhs(8 * c_->g.N), raw(8 * c_->g.N), Srho0(c_->csize() * c_->g.M), SR0(c_->csize() * c_->g.M),
        SR1(c_->csize() * c_->g.M), Sw(c_->csize() * c_->g.M),
        lam(8 * (c_->g.nx + c_->g.ny + c_->g.nzh)), fac(8 * (c_->g.nx + c_->g.ny + c_->g.nzh)),
        part(8 * 8 * kVecBlocks), sums(8 * 40), rpart(8 * 8 * kVecBlocks) {
    CDFT_HIP(hipHostMalloc(reinterpret_cast<void**>(&mm_h), 2 * sizeof(double), hipHostMallocCoherent));
    CDFT_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&mm_d), mm_h, 0));
    CDFT_HIP(hipEventCreateWithFlags(&mm_ready, hipEventDisableTiming));
    const Geom& g = c->g;
    // make_if_state (dynamics.hpp:79-117): Lambda_d = 2 (cos(2 pi i / N_d) - 1) / dx^2
    const double PI = 3.14159265358979323846, Dx = 1.0 / (g.dx * g.dx);
    lam_h.resize(g.nx + g.ny + g.nzh);
    for (int i = 0; i < g.nx; ++i) lam_h[i] = 2.0 * Dx * (std::cos(2.0 * PI * i / g.nx) - 1.0);
    for (int i = 0; i < g.ny; ++i) lam_h[g.nx + i] = 2.0 * Dx * (std::cos(2.0 * PI * i / g.ny) - 1.0);
    for (int i = 0; i < g.nzh; ++i) lam_h[g.nx + g.ny + i] = 2.0 * Dx * (std::cos(2.0 * PI * i / g.nz) - 1.0);
    CDFT_HIP(hipMemcpy(lam.p, lam_h.data(), 8 * lam_h.size(), hipMemcpyHostToDevice));
  }

  ~DdftRun() {
    if (mm_h) (void)hipHostFree(mm_h);
    if (mm_ready) (void)hipEventDestroy(mm_ready);
  }
  DdftRun(const DdftRun&) = delete;
  DdftRun& operator=(const DdftRun&) = delete;

  // update_timestep (dynamics.hpp:121-129)
  void update_timestep(double dt) {
    if (std::fabs(st_dt - dt) < 1e-30) return;
    st_dt = dt;
    std::vector<double> f(lam_h.size());
    for (size_t i = 0; i < f.size(); ++i) f[i] = std::exp(lam_h[i] * dt);
    CDFT_HIP(hipMemcpyAsync(fac.p, f.data(), 8 * f.size(), hipMemcpyHostToDevice, c->stream));
    CDFT_HIP(hipStreamSynchronize(c->stream));
  }

  // With `sums_to` the evaluation's raw sums (8 doubles) are copied there on
  // the device and the energy is not read back (NaN returned): a run that
  // needs the energies only at its end keeps the host out of the step.
  double compute(const double* rho, double* forces, double* sums_to = nullptr) {
    if (hooks && hooks->force()) {
      ++evals;
      return hooks->eval(c, rho, forces);
    }
    enqueue_eval(rho, forces, sums_to);
    return finish_eval(sums_to);
  }

  // The context's evaluation, enqueued on the stream (no host wait); with
  // `sums_to` its raw sums are copied there on the device.
  void enqueue_eval(const double* rho, double* forces, double* sums_to) {
    ++evals;
    c->eval(rho, d_vext, mu, forces, AliasMap{0, 0.0, 0.0});
    if (sums_to) {
      hipLaunchKernelGGL(k_copy_sums, dim3(1), dim3(64), 0, c->stream, c->d_res, sums_to);
      CDFT_HIP(hipGetLastError());
    }
  }

  // The grand potential of the last enqueued evaluation (NaN with `sums_to`).
  double finish_eval(double* sums_to) {
    if (sums_to) return std::numeric_limits<double>::quiet_NaN();
    double er[8];
    CDFT_HIP(hipMemcpyAsync(er, c->d_res, sizeof(er), hipMemcpyDeviceToHost, c->stream));
    CDFT_HIP(hipStreamSynchronize(c->stream));
    cdft_result r{};
    c->result_from(er, mu, &r);
    return r.grand_potential;
  }

  // Full forward transform of a real field into S (mean removed; its FP64 sum
  // to sums[slot]).  `rows` > 0: part already holds that many block sums of
  // the field (its producer reduced them), else launch_sum forms them.
  template <typename R>
  void forward(const double* field, cx<R>* S, int slot, bool x_stage, int rows = 0) {
    const Geom& g = c->g;
    const AliasMap id{0, 0.0, 0.0};
    double* s = sums.as<double>() + 8 * slot;
    if (rows <= 0) {
      launch_sum(field, id, part.as<double>(), g, c->stream);
      rows = sum_blocks(g);
    }
    // the z-R2C reduces the block sums itself and stores the sum in s
    launch_zr2c<R, R>(field, S, g, c->plan[2], c->tw<R>(2), id, nullptr, c->stream, part.as<double>(), rows, s);
    launch_lines<R, R>(S, static_cast<long>(g.nzh) * g.nx, c->plan[1], c->tw<R>(1), -1, c->stream);
    if (x_stage) launch_xlines<R, R>(S, 1, g, c->plan[0], c->tw<R>(0), -1, c->stream);
  }

  // excess_rhs_fourier (dynamics.hpp:212-229) into S (x stage optional)
  template <typename R>
  void rhs_forward(const double* rho, const double* forces, cx<R>* S, int slot, bool x_stage) {
    const Geom& g = c->g;
    const int rows = launch_ddft_rhs(rho, forces, rhs.as<double>(), g, part.as<double>(), c->stream);
    forward<R>(rhs.as<double>(), S, slot, x_stage, rows);
  }

  // one propagator application; returns {max deviation, min d1}
  template <typename R>
  void propagate(bool first_iter, double* dev, double* mn) {
    propagate_async<R>(first_iter);
    propagate_wait(dev, mn);
  }

  // {max deviation, min d1} of the last propagate_async, once it has run
  void propagate_wait(double* dev, double* mn) {
    CDFT_HIP(hipEventSynchronize(mm_ready));
    *dev = mm_h[0];
    *mn = mm_h[1];
  }

  // one propagator application, enqueued; its {max deviation, min d1} reach
  // the pinned mm_h when mm_ready completes
  template <typename R>
  void propagate_async(bool first_iter) {
    const Geom& g = c->g;
    const int ty = 16;
    const size_t sh = sizeof(cx<R>) * ty * (g.nx + 1);
    const dim3 grid((g.ny + ty - 1) / ty, g.nzh);
    // up to 8 elements per thread (512 threads for the larger tiles of the
    // LDS-pass lengths), else 16
    if (ty * g.nx <= 8 * 256)
      hipLaunchKernelGGL((k_xprop<R, R, 8, 256>), grid, dim3(256), sh, c->stream, SR1.as<cx<R>>(), Srho0.as<cx<R>>(),
                         SR0.as<cx<R>>(), Sw.as<cx<R>>(), g, c->plan[0], c->tw<R>(0), ty, lam.as<double>(),
                         fac.as<double>(), st_dt, first_iter ? 1 : 0);
    else if (ty * g.nx <= 16 * 256 && slot_lanes(g.nx) == 0)
      hipLaunchKernelGGL((k_xprop<R, R, 8, 512>), grid, dim3(512), sh, c->stream, SR1.as<cx<R>>(), Srho0.as<cx<R>>(),
                         SR0.as<cx<R>>(), Sw.as<cx<R>>(), g, c->plan[0], c->tw<R>(0), ty, lam.as<double>(),
                         fac.as<double>(), st_dt, first_iter ? 1 : 0);
    else if (ty * g.nx <= 16 * 256)
      hipLaunchKernelGGL((k_xprop<R, R, 16, 256>), grid, dim3(256), sh, c->stream, SR1.as<cx<R>>(), Srho0.as<cx<R>>(),
                         SR0.as<cx<R>>(), Sw.as<cx<R>>(), g, c->plan[0], c->tw<R>(0), ty, lam.as<double>(),
                         fac.as<double>(), st_dt, first_iter ? 1 : 0);
    else
      throw Unsupported("x extent too large for the propagator kernel");
    launch_lines<R, R>(Sw.as<cx<R>>(), static_cast<long>(g.nzh) * g.nx, c->plan[1], c->tw<R>(1), +1, c->stream);
    launch_zc2r<R, R>(Sw.as<cx<R>>(), 1, raw.as<double>(), nullptr, g.inv_n, g, c->plan[2], c->tw<R>(2), nullptr,
                      DcTable{}, c->stream);
    const AliasMap id{0, 0.0, 0.0};
    launch_sum(raw.as<double>(), id, rpart.as<double>(), g, c->stream);
    hipLaunchKernelGGL(k_ddft_update, dim3(vec_blocks(g.N)), dim3(256), 0, c->stream, raw.as<double>(),
                       first_iter ? static_cast<const double*>(rho0.as<double>()) : d1.as<double>(), d1.as<double>(),
                       g.N, sums.as<double>(), rpart.as<double>(), sum_blocks(g), st_dt, first_iter ? 1 : 0,
                       part.as<double>());
    // the reduction writes straight into the pinned, coherent mm_h
    hipLaunchKernelGGL(k_finalize_maxmin, dim3(1), dim3(256), 0, c->stream, part.as<double>(),
                       static_cast<int>(vec_blocks(g.N)), mm_d);
    CDFT_HIP(hipGetLastError());
    CDFT_HIP(hipEventRecord(mm_ready, c->stream));
  }

  // integrating_factor_step (dynamics.hpp:333-401); rho0 holds the densities and f0 their forces.
  template <typename R>
  double step(double* cfg_dt, double tol, int max_it, double* sums_to = nullptr) {
    rhs_forward<R>(rho0.as<double>(), f0.as<double>(), SR0.as<cx<R>>(), 1, true);
    forward<R>(rho0.as<double>(), Srho0.as<cx<R>>(), 0, true);
    // Every propagation is followed by an evaluation at the new iterate
    // whatever its test decides (the step's last, or the next iteration's):
    // with the context's functional it is enqueued before the host waits for
    // the test, so the GPU evaluates while the host reads the flag.  A restart
    // discards it.
    const bool ahead = !(hooks && hooks->force());
    bool restart;
    int restarts = 0;
    do {
      restart = false;
      update_timestep(*cfg_dt);
      // d1 starts as rho0: the first propagate reads rho0 as the old iterate and writes all of d1
      double deviation = 1e30;
      for (int it = 0; it < max_it; ++it) {
        const bool first = it == 0;  // d1 == rho0: compute({d1}) equals the step's first compute
        if (!first) {
          if (!ahead) compute(d1.as<double>(), f1.as<double>());
          rhs_forward<R>(d1.as<double>(), f1.as<double>(), SR1.as<cx<R>>(), 2, false);
        }
        const double old = deviation;
        double mn = 0.0;
        propagate_async<R>(first);
        if (ahead) enqueue_eval(d1.as<double>(), f1.as<double>(), sums_to);
        propagate_wait(&deviation, &mn);
        if (!std::isfinite(deviation) || !std::isfinite(mn))
          throw std::runtime_error("integrating_factor_step: the iterate is not finite (dt = " +
                                   std::to_string(*cfg_dt) + ")");
        if (mn < 0.0 || (it > 0 && deviation > old)) {
          restart = true;
          break;
        }
        if (deviation < tol) break;
      }
      if (!restart && deviation > tol) restart = true;
      if (restart) {
        if (ahead && max_it > 0) --evals;  // the evaluation enqueued after the last propagation is discarded
        if (++restarts > kMaxRestarts)
          throw std::runtime_error("integrating_factor_step: no convergence after " + std::to_string(kMaxRestarts) +
                                   " restarts (dt = " + std::to_string(*cfg_dt) + ")");
        *cfg_dt /= 10.0;
      }
    } while (restart);
    const double e = ahead ? finish_eval(sums_to) : compute(d1.as<double>(), f1.as<double>(), sums_to);
    // the new density and its forces start the next step: swap the buffers
    // instead of copying them (d1 and f1 are rewritten before they are read)
    std::swap(rho0.p, d1.p);
    std::swap(f0.p, f1.p);
    return e;
  }
};


}  // namespace

}  // namespace cdft

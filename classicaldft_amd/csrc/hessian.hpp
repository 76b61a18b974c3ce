// Exact Hessian-vector products of the functional, FP64: H(rho) u with H the
// Jacobian of the forces delta(beta Omega)/delta rho * dV (symmetric: the
// Hessian of beta Omega times dV).  Shared by the DDFT adjoint and the
// smallest-eigenpair solver.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "context.hpp"
#include "ddft.hpp"

namespace cdft {

// Kernels and helpers with internal linkage: each translation unit that
// includes this header gets its own copy.
namespace {

// Pointwise part of H(rho) u: dg = the u-directional derivatives of the 11
// FMT derivative channels (dual numbers through fmt_phi_grad_t), local =
// dV (u / rho + W_mf u).  n, dn: weighted densities of rho and of u.
__global__ void __launch_bounds__(256) k_hvp_point(const double* __restrict__ n, const double* __restrict__ dn,
                                                   const double* __restrict__ rho, const double* __restrict__ u,
                                                   double* __restrict__ dg, double* __restrict__ local, Geom g,
                                                   FmtParams fp, ChannelSet cs) {
  const long N = g.N;
  for (long p = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; p < N;
       p += static_cast<long>(gridDim.x) * blockDim.x) {
    const double r = rho[p];
    double lf = r > 1e-300 ? u[p] / r : 0.0;  // d/drho of ln max(rho, 1e-300) (ideal_gas.hpp:18-38)
    if (cs.mf) lf += dn[(cs.hs ? 11 : 0) * N + p];
    local[p] = lf * g.dv;
    if (cs.hs) {
      Dual w[11], gg[11];
#pragma unroll
      for (int i = 0; i < 11; ++i) w[i] = Dual(n[i * N + p], dn[i * N + p]);
      fmt_phi_grad_t<Dual>(fp, w, gg);
#pragma unroll
      for (int i = 0; i < 11; ++i) dg[i * N + p] = gg[i].d;
    }
  }
}

void require_fp64(const Context* c) {
  if (c->precision != CDFT_PRECISION_FP64) throw Unsupported("the adjoint needs a context of CDFT_PRECISION_FP64");
}

// Full FP64 transforms between a real field and a spectrum (kz, kx, ky).
void forward_full(Context* c, const double* f, cx<double>* S) {
  const Geom& g = c->g;
  const AliasMap id{0, 0.0, 0.0};
  launch_zr2c<double, double>(f, S, g, c->plan[2], c->tw<double>(2), id, nullptr, c->stream);
  launch_lines<double, double>(S, static_cast<long>(g.nzh) * g.nx, c->plan[1], c->tw<double>(1), -1, c->stream);
  launch_xlines<double, double>(S, 1, g, c->plan[0], c->tw<double>(0), -1, c->stream);
}

// out = c2r(S) / N; S is overwritten.
void inverse_full(Context* c, cx<double>* S, double* out) {
  const Geom& g = c->g;
  launch_xlines<double, double>(S, 1, g, c->plan[0], c->tw<double>(0), +1, c->stream);
  launch_lines<double, double>(S, static_cast<long>(g.nzh) * g.nx, c->plan[1], c->tw<double>(1), +1, c->stream);
  launch_zc2r<double, double>(S, 1, out, nullptr, g.inv_n, g, c->plan[2], c->tw<double>(2), nullptr, DcTable{},
                              c->stream);
}

// H(rho) u on the device (general-schedule transforms, FP64).
struct Hessian {
  Context* c;
  DevBuf n, dn, dg, local;
  long products = 0;

  explicit Hessian(Context* c_)
      : c(c_), n(8 * std::max(c_->cs.nf, 1) * c_->g.N), dn(8 * std::max(c_->cs.nf, 1) * c_->g.N),
        dg(8 * 11 * c_->g.N * c_->cs.hs), local(8 * c_->g.N) {}

  void apply(const double* rho, const double* u, double* out) {
    const Geom& g = c->g;
    c->weighted_densities(rho, n.as<double>());
    c->weighted_densities(u, dn.as<double>());
    hipLaunchKernelGGL(k_hvp_point, dim3(vec_blocks(g.N)), dim3(256), 0, c->stream, n.as<double>(), dn.as<double>(),
                       rho, u, dg.as<double>(), local.as<double>(), g, c->fp, c->cs);
    if (c->cs.nb > 0) {
      const AliasMap id{0, 0.0, 0.0};
      auto* SF = static_cast<cx<double>*>(c->d_SF);
      auto* SO = static_cast<cx<double>*>(c->d_SO);
      const long ylines = static_cast<long>(g.nzh) * g.nx;
      for (int ch = 0; ch < c->cs.nb; ++ch)
        launch_zr2c<double, double>(dg.as<double>() + ch * g.N, SF + ch * g.M, g, c->plan[2], c->tw<double>(2), id,
                                    nullptr, c->stream);
      launch_lines<double, double>(SF, ylines * c->cs.nb, c->plan[1], c->tw<double>(1), -1, c->stream);
      launch_xbwd<double, double>(SF, SO, static_cast<const cx<double>*>(c->d_WTg), g, c->plan[0],
                                  c->tw<double>(0), c->cs.nb, c->stream);
      launch_lines<double, double>(SO, ylines, c->plan[1], c->tw<double>(1), +1, c->stream);
      launch_zc2r<double, double>(SO, 1, out, local.as<double>(), g.dv, g, c->plan[2], c->tw<double>(2), nullptr,
                                  DcTable{}, c->stream);
    } else {
      CDFT_HIP(hipMemcpyAsync(out, local.p, 8 * g.N, hipMemcpyDeviceToDevice, c->stream));
    }
    CDFT_HIP(hipGetLastError());
    ++products;
  }
};

}  // namespace

}  // namespace cdft

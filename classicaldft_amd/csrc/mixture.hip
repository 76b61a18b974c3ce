// Multi-species functionals::total (functionals.hpp:64-95) on the device, in
// FP64.  Its 3-D transforms run on the fused kernels where the grid has them
// (power-of-two extents 64 and 128, x also 256; spectra in the sub-plane
// order), on the general schedule otherwise (plane-major spectra).  In mixed
// precision (fused grids only) the real-space measure, derivative and
// mean-field fields are stored in FP32; every transform, spectrum, pointwise
// term and sum stays FP64.
//
// Hard spheres (hard_sphere.hpp:174-254): the 16 measures eta, n2, n1, n0,
// n_v2[3], n_v1[3], T[6] are sums over species of each species' weighted
// density, with n1, n0 and n_v1 scaled by that species' radius
// (hard_sphere.hpp:80-86,186-198).  Being linear in the densities, the sums
// are formed in k-space before one inverse transform per channel.  The
// pointwise kernel returns Phi and its 16 partial derivatives; each species'
// force is the k-space combination of those with its own weights
// (w2 <- d n2 + d n1 / (4 pi R) + d n0 / (4 pi R^2), the vector weight,
// conjugated, <- d n_v2 + d n_v1 / (4 pi R); hard_sphere.hpp:118-134,236-254).
//
// Mean field (mean_field.hpp:233-281): per species s, the potential
// U_s = sum over self pairs (s, s) of w * rho_s plus half of w * rho_t over
// cross pairs (s, t); the force gets U_s dV and F_mf = 1/2 sum_s rho_s U_s dV,
// which equals the reference's 1/2 rho_i (w * rho_j) dV per listed pair.
//
// Ideal gas and external field per species (ideal_gas.hpp:18-38,
// external_field.hpp:16-31); Omega = F - sum_s mu_s N_s.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "api_guard.hpp"
#include "context.hpp"
#include "device_common.hpp"

namespace cdft {
namespace {

constexpr int kMixFmt = 16;  // eta, n2, n1, n0, n_v2[3], n_v1[3], T[6]
// Transformed FMT channels: all but T_zz, which follows from the trace
// identity T_zz = n2 - T_xx - T_yy of every species' weighted densities (the
// diagonal tensor weights sum to w2), so its derivative folds into those three.
constexpr int kMixFmtT = 15;
constexpr int kMixThreads = 256;

struct MixParams {
  int S, ninter, hs;
  int hp, nxh, q;  // spectra in the fused sub-plane order (and its half extents)
  double R[kMaxSpecies], invR[kMaxSpecies], ifr[kMaxSpecies], ifr2[kMaxSpecies];  // 1/(4 pi R), 1/(4 pi R^2)
  int si[kMaxMixInteractions], sj[kMaxMixInteractions];
};

struct MixPoint {
  int S, hs, nfm, nu;  // species, FMT on, FMT channel count, mean-field potential count
  double mu[kMaxSpecies];
  double dv;
  FmtParams fp;
};

// Carving of Context::d_mix (T, the channel-layout staging of the fused
// transforms, only on fused-schedule grids).
struct MixLayout {
  int nch = 0, nblk = 0;
  size_t R = 0, F = 0, X = 0, B = 0, local = 0, rho = 0, vext = 0, force = 0, part = 0, res = 0, T = 0,
         total = 0;
};

MixLayout mix_layout(const Context* c) {
  MixLayout L;
  const size_t S = static_cast<size_t>(c->n_species), N = static_cast<size_t>(c->g.N),
               M = static_cast<size_t>(c->g.M);
  const int nfm = c->cs.hs ? kMixFmtT : 0, nu = c->inters.empty() ? 0 : c->n_species;
  L.nch = std::max(nfm + nu, 1);
  L.nblk = static_cast<int>(std::min<long>(1024, (c->g.N + kMixThreads - 1) / kMixThreads));
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += (bytes + 255) / 256 * 256;
    return at;
  };
  L.R = take(16 * S * M);
  L.F = take(16 * static_cast<size_t>(L.nch) * M);
  L.X = take((c->precision == CDFT_PRECISION_MIXED ? 4 : 8) * static_cast<size_t>(L.nch) * N);
  L.B = take(16 * S * M);
  L.local = take(8 * S * N);
  L.rho = take(8 * S * N);
  L.vext = take(8 * S * N);
  L.force = take(8 * S * N);
  L.part = take(8 * 8 * static_cast<size_t>(L.nblk));
  L.res = take(8 * 8);
  if (c->fast) L.T = take(16 * std::max<size_t>(static_cast<size_t>(L.nch), S) * M);
  L.total = off;
  return L;
}

template <typename T>
T* at(const Context* c, size_t off) {
  return reinterpret_cast<T*>(static_cast<char*>(c->d_mix) + off);
}

__device__ __forceinline__ cx<double> cmul(cx<double> a, cx<double> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cx<double> cscale(cx<double> a, double s) { return {a.re * s, a.im * s}; }
__device__ __forceinline__ cx<double> cadd(cx<double> a, cx<double> b) { return {a.re + b.re, a.im + b.im}; }

// k-vector of spectral element i: plane-major [kz][kx][ky], or the fused
// sub-plane order [kz][r][p][s][k] (kx = 2 k + p, ky = r + 2 s)
__device__ __forceinline__ void kvec3(long i, const Geom& g, const MixParams& p, double* kx, double* ky, double* kz) {
  if (p.hp) {
    const int k = static_cast<int>(i % p.nxh);
    const long t = i / p.nxh;
    const int s = static_cast<int>(t % p.q);
    const long u = t / p.q;
    *kx = kvec(2 * k + static_cast<int>(u & 1), g.nx, g.dkx);
    *ky = kvec(static_cast<int>((u >> 1) & 1) + 2 * s, g.ny, g.dky);
    *kz = g.dkz * static_cast<double>(u >> 2);
    return;
  }
  const long nxy = static_cast<long>(g.nx) * g.ny;
  const long z = i / nxy, r = i - z * nxy;
  const int x = static_cast<int>(r / g.ny), y = static_cast<int>(r - static_cast<long>(x) * g.ny);
  *kx = kvec(x, g.nx, g.dkx);
  *ky = kvec(y, g.ny, g.dky);
  *kz = g.dkz * static_cast<double>(z);
}

// Forward channels at each spectral element: the 16 summed FMT measures
// (when hs) and the S mean-field potentials (when there are interactions).
__global__ __launch_bounds__(kMixThreads) void k_mix_fwd(const cx<double>* __restrict__ rhoK,
                                                         const double* __restrict__ mfw, cx<double>* __restrict__ F,
                                                         Geom g, MixParams p) {
  const long M = g.M;
  // (species values are read where used: a register array indexed by a
  // runtime species number would live in scratch memory)
  for (long i = static_cast<long>(blockIdx.x) * kMixThreads + threadIdx.x; i < M;
       i += static_cast<long>(gridDim.x) * kMixThreads) {
    int ch = 0;
    if (p.hs) {
      double kx, ky, kz;
      kvec3(i, g, p, &kx, &ky, &kz);
      const double k2 = kx * kx + ky * ky + kz * kz;
      const double k = sqrt(k2);
      cx<double> acc[kMixFmtT];
#pragma unroll
      for (int c = 0; c < kMixFmtT; ++c) acc[c] = {0.0, 0.0};
      for (int s = 0; s < p.S; ++s) {
        double w3, w2;
        radial(k, p.R[s], g.inv_n, &w3, &w2);
        const WRad wr = wrad({w3, w2}, k2, p.invR[s]);
        const cx<double> r = rhoK[s * M + i];
        acc[0] = cadd(acc[0], cscale(r, w3));
        acc[1] = cadd(acc[1], cscale(r, w2));
        acc[2] = cadd(acc[2], cscale(r, w2 * p.ifr[s]));
        acc[3] = cadd(acc[3], cscale(r, w2 * p.ifr2[s]));
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const cx<double> wv = cmul(channel_weight_t(2 + a, false, kx, ky, kz, wr, 0.0, p.invR[s]), r);
          acc[4 + a] = cadd(acc[4 + a], wv);
          acc[7 + a] = cadd(acc[7 + a], cscale(wv, p.ifr[s]));
        }
#pragma unroll
        for (int q = 0; q < 5; ++q) {  // T_xx .. T_yz (T_zz by the trace identity)
          const double wq = channel_weight_t(5 + q, false, kx, ky, kz, wr, 0.0, p.invR[s]).re;
          acc[10 + q] = cadd(acc[10 + q], cscale(r, wq));
        }
      }
#pragma unroll
      for (int c = 0; c < kMixFmtT; ++c) F[c * M + i] = acc[c];
      ch = kMixFmtT;
    }
    if (p.ninter > 0) {
      for (int s = 0; s < p.S; ++s) {
        cx<double> u = {0.0, 0.0};
        for (int q = 0; q < p.ninter; ++q) {
          const int a = p.si[q], b = p.sj[q];
          if (a != s && b != s) continue;
          const double w = mfw[q * M + i];
          if (a == b)
            u = cadd(u, cscale(rhoK[s * M + i], w));  // self pair
          else
            u = cadd(u, cscale(rhoK[(a == s ? b : a) * M + i], 0.5 * w));  // half of a cross pair
        }
        F[(ch + s) * M + i] = u;
      }
    }
  }
}

// Pointwise: Phi and its 16 partials (written over the measure fields, which
// are FP64 or FP32 (XT); the arithmetic is FP64), the local force of every
// species and the block sums [F_id, F_ext, F_hs, F_mf, sum mu rho, sum rho]
// (each still without dV).
template <typename XT>
__global__ __launch_bounds__(kMixThreads) void k_mix_point(XT* __restrict__ X, const double* __restrict__ rho,
                                                           const double* __restrict__ vext,
                                                           double* __restrict__ local, double* __restrict__ partials,
                                                           long N, MixPoint mp) {
  double sums[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (long p = static_cast<long>(blockIdx.x) * kMixThreads + threadIdx.x; p < N;
       p += static_cast<long>(gridDim.x) * kMixThreads) {
    if (mp.hs) {
      double m[kMixFmt], d[kMixFmt];
#pragma unroll
      for (int c = 0; c < kMixFmtT; ++c) m[c] = static_cast<double>(X[c * N + p]);
      m[15] = m[1] - m[10] - m[13];  // T_zz
      sums[2] += fmt_phi_grad_mix_t<double>(mp.fp, m, d);
      d[1] += d[15];  // adjoint of the trace identity
      d[10] -= d[15];
      d[13] -= d[15];
#pragma unroll
      for (int c = 0; c < kMixFmtT; ++c) X[c * N + p] = static_cast<XT>(d[c]);
    }
    for (int s = 0; s < mp.S; ++s) {
      const double r = rho[s * N + p];
      const double lr = flog(fmax(r, 1e-300));  // ideal_gas.hpp:27
      double f = lr - mp.mu[s];
      sums[0] += r * (lr - 1.0);
      sums[4] += mp.mu[s] * r;
      sums[5] += r;
      if (vext) {
        const double v = vext[s * N + p];
        f += v;
        sums[1] += r * v;
      }
      if (mp.nu) {
        const double u = static_cast<double>(X[(mp.nfm + s) * N + p]);
        f += u;
        sums[3] += 0.5 * r * u;
      }
      local[s * N + p] = f * mp.dv;
    }
  }
  block_reduce<8>(sums, partials + 8 * static_cast<long>(blockIdx.x));
}

// Backward: B_s = sum over the 16 derivative spectra times species s' weights.
__global__ __launch_bounds__(kMixThreads) void k_mix_bwd(const cx<double>* __restrict__ D, cx<double>* __restrict__ B,
                                                         Geom g, MixParams p) {
  const long M = g.M;
  for (long i = static_cast<long>(blockIdx.x) * kMixThreads + threadIdx.x; i < M;
       i += static_cast<long>(gridDim.x) * kMixThreads) {
    double kx, ky, kz;
    kvec3(i, g, p, &kx, &ky, &kz);
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double k = sqrt(k2);
    cx<double> d[kMixFmtT];  // the T_zz derivative is folded into n2, T_xx and T_yy
#pragma unroll
    for (int c = 0; c < kMixFmtT; ++c) d[c] = D[c * M + i];
    for (int s = 0; s < p.S; ++s) {
      double w3, w2;
      radial(k, p.R[s], g.inv_n, &w3, &w2);
      const WRad wr = wrad({w3, w2}, k2, p.invR[s]);
      cx<double> b = cscale(d[0], w3);
      const cx<double> scal = cadd(cadd(d[1], cscale(d[2], p.ifr[s])), cscale(d[3], p.ifr2[s]));
      b = cadd(b, cscale(scal, w2));
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const cx<double> dv = cadd(d[4 + a], cscale(d[7 + a], p.ifr[s]));
        b = cadd(b, cmul(channel_weight_t(2 + a, true, kx, ky, kz, wr, 0.0, p.invR[s]), dv));
      }
#pragma unroll
      for (int q = 0; q < 5; ++q)
        b = cadd(b, cscale(d[10 + q], channel_weight_t(5 + q, false, kx, ky, kz, wr, 0.0, p.invR[s]).re));
      B[s * M + i] = b;
    }
  }
}

// Per-term split of the forces of one species (functionals.hpp:64-95): the
// ideal (ideal_gas.hpp:29-33), external (external_field.hpp:24-28) and
// mean-field (conv dV, mean_field.hpp:259-277) forces are pointwise; the
// hard-sphere force is the total force less those three.
template <typename CT>
__global__ __launch_bounds__(kMixThreads) void k_split_terms(const double* __restrict__ rho,
                                                             const double* __restrict__ vext,
                                                             const CT* __restrict__ conv,
                                                             const double* __restrict__ total, double mu, double dv,
                                                             long n, int hs, double* __restrict__ f_id,
                                                             double* __restrict__ f_ext, double* __restrict__ f_mf,
                                                             double* __restrict__ f_hs) {
  for (long p = static_cast<long>(blockIdx.x) * kMixThreads + threadIdx.x; p < n;
       p += static_cast<long>(gridDim.x) * kMixThreads) {
    const double id = (flog(fmax(rho[p], 1e-300)) - mu) * dv;
    const double ext = vext ? vext[p] * dv : 0.0;
    const double mf = conv ? static_cast<double>(conv[p]) * dv : 0.0;
    if (f_id) f_id[p] = id;
    if (f_ext) f_ext[p] = ext;
    if (f_mf) f_mf[p] = mf;
    if (f_hs) f_hs[p] = hs ? total[p] - id - ext - mf : 0.0;
  }
}

unsigned spectral_blocks(long M) {
  return static_cast<unsigned>(std::min<long>(2048, (M + kMixThreads - 1) / kMixThreads));
}

MixParams mix_params(const Context* c) {
  const double PI = 3.14159265358979323846;
  MixParams p{};
  p.S = c->n_species;
  p.ninter = static_cast<int>(c->inters.size());
  p.hs = c->cs.hs;
  for (int s = 0; s < p.S; ++s) {
    const double R = 0.5 * c->diam[s];
    p.R[s] = R;
    p.invR[s] = 1.0 / R;
    p.ifr[s] = 1.0 / (4.0 * PI * R);
    p.ifr2[s] = p.ifr[s] / R;
  }
  for (int q = 0; q < p.ninter; ++q) {
    p.si[q] = c->inters[q].si;
    p.sj[q] = c->inters[q].sj;
  }
  p.hp = c->fast ? 1 : 0;
  p.nxh = c->g.nx / 2;
  p.q = c->g.ny / 2;
  return p;
}

// Forward transform of `count` real fields (N apart) into spectra (M apart);
// FP32 fields (I = float) only on the fused path.
template <typename I>
void r2c_fields(Context* c, const I* in, cx<double>* out, int count) {
  const Geom& g = c->g;
  if (c->fast) {
    launch_fields_forward(in, out, at<cx<double>>(c, mix_layout(c).T), count, c->fg, c->ftw64, c->stream);
    return;
  }
  if constexpr (!std::is_same<I, double>::value) throw std::logic_error("FP32 fields need the fused transforms");
  else launch_zr2c_fields<double, double>(in, count, out, g, c->plan[2], c->tw<double>(2), c->stream);
  launch_lines<double, double>(out, static_cast<long>(g.nzh) * g.nx * count, c->plan[1], c->tw<double>(1), -1,
                               c->stream);
  launch_xlines<double, double>(out, count, g, c->plan[0], c->tw<double>(0), -1, c->stream);
}

// Inverse x and y transforms of `count` spectra in place (the z C2R follows).
void inverse_xy(Context* c, cx<double>* sp, int count) {
  const Geom& g = c->g;
  launch_xlines<double, double>(sp, count, g, c->plan[0], c->tw<double>(0), +1, c->stream);
  launch_lines<double, double>(sp, static_cast<long>(g.nzh) * g.nx * count, c->plan[1], c->tw<double>(1), +1,
                               c->stream);
}

// Inverse transform of `count` spectra (M apart) into real fields (N apart):
// out = scale * c2r(spec) + add (add nullable).  The general path transforms
// the spectra in place; FP32 output (O = float) only on the fused path.
template <typename O>
void c2r_fields(Context* c, cx<double>* spec, O* out, const double* add, double scale, int count) {
  const Geom& g = c->g;
  if (c->fast) {
    launch_fields_inverse(spec, out, add, scale, at<cx<double>>(c, mix_layout(c).T), count, c->fg, c->ftw64,
                          c->stream);
    return;
  }
  if constexpr (!std::is_same<O, double>::value) {
    throw std::logic_error("FP32 fields need the fused transforms");
  } else {
    inverse_xy(c, spec, count);
    launch_zc2r<double, double>(spec, count, out, add, scale, g, c->plan[2], c->tw<double>(2), nullptr, DcTable{},
                                c->stream);
  }
}

// One evaluation on device arrays (S * N each; vext and force nullable);
// the sums land in the layout's res block.  XT: the real-space field type.
template <typename XT>
void mixture_eval_t(Context* c, const double* rho, const double* vext, const double* mu, double* force) {
  const Geom& g = c->g;
  const MixLayout L = mix_layout(c);
  const MixParams p = mix_params(c);
  const int S = c->n_species;
  const int nfm = c->cs.hs ? kMixFmtT : 0, nu = c->inters.empty() ? 0 : S;
  auto* R = at<cx<double>>(c, L.R);
  auto* F = at<cx<double>>(c, L.F);
  auto* X = at<XT>(c, L.X);
  auto* B = at<cx<double>>(c, L.B);
  auto* local = at<double>(c, L.local);
  auto* part = at<double>(c, L.part);
  auto* res = at<double>(c, L.res);

  if (nfm + nu > 0) {
    r2c_fields(c, rho, R, S);
    k_mix_fwd<<<spectral_blocks(g.M), kMixThreads, 0, c->stream>>>(R, c->d_mix_mf, F, g, p);
    c2r_fields(c, F, X, nullptr, 1.0, nfm + nu);
  }
  MixPoint mp{};
  mp.S = S;
  mp.hs = c->cs.hs;
  mp.nfm = nfm;
  mp.nu = nu;
  for (int s = 0; s < S; ++s) mp.mu[s] = mu[s];
  mp.dv = g.dv;
  mp.fp = c->fp;
  k_mix_point<XT><<<L.nblk, kMixThreads, 0, c->stream>>>(X, rho, vext, local, part, g.N, mp);
  if (force) {
    if (nfm > 0) {
      r2c_fields(c, X, F, kMixFmtT);
      k_mix_bwd<<<spectral_blocks(g.M), kMixThreads, 0, c->stream>>>(F, B, g, p);
      c2r_fields(c, B, force, local, g.dv, S);
    } else {
      CDFT_HIP(hipMemcpyAsync(force, local, sizeof(double) * S * g.N, hipMemcpyDeviceToDevice, c->stream));
    }
  }
  launch_finalize(part, L.nblk, res, c->stream);
  CDFT_HIP(hipGetLastError());
}

void mixture_eval(Context* c, const double* rho, const double* vext, const double* mu, double* force) {
  if (c->precision == CDFT_PRECISION_MIXED) mixture_eval_t<float>(c, rho, vext, mu, force);
  else mixture_eval_t<double>(c, rho, vext, mu, force);
}

void mixture_result(Context* c, cdft_result* out) {
  const MixLayout L = mix_layout(c);
  double res[8];
  CDFT_HIP(hipMemcpyAsync(res, at<double>(c, L.res), sizeof(res), hipMemcpyDeviceToHost, c->stream));
  CDFT_HIP(hipStreamSynchronize(c->stream));
  const double dv = c->g.dv;
  out->ideal = res[0] * dv;
  out->external = res[1] * dv;
  out->hard_sphere = res[2] * dv;
  out->mean_field = res[3] * dv;
  out->mass = res[5] * dv;
  // functionals.hpp:79-92: F = F_id + F_ext + F_hs + F_mf, Omega = F - sum_s mu_s N_s
  out->free_energy = out->ideal + out->external + out->hard_sphere + out->mean_field;
  out->grand_potential = out->free_energy - res[4] * dv;
}

// Bulk (uniform) thermodynamics, bulk/thermodynamics.hpp:17-255, for host
// calls and for the batched device kernel alike: the model in a plain struct.
struct BulkParams {
  int S, ninter;
  FmtParams fp;
  double diam[kMaxSpecies];
  int si[kMaxMixInteractions], sj[kMaxMixInteractions];
  double a_vdw[kMaxMixInteractions];
};

BulkParams bulk_params(const Context* c) {
  BulkParams b{};
  b.S = c->n_species;
  b.ninter = static_cast<int>(c->inters.size());
  b.fp = c->fp;
  for (int s = 0; s < b.S; ++s) b.diam[s] = c->diam[s];
  for (int q = 0; q < b.ninter; ++q) {
    b.si[q] = c->inters[q].si;
    b.sj[q] = c->inters[q].sj;
    b.a_vdw[q] = c->a_vdw[q];
  }
  return b;
}

// The summed uniform measures of fmt::make_uniform_measures (measures.hpp:49-65)
// in the mixture channel order.
__host__ __device__ void uniform_measures(const BulkParams& b, const double* rho, double (&m)[kMixFmt]) {
  const double PI = 3.14159265358979323846;
  for (int c = 0; c < kMixFmt; ++c) m[c] = 0.0;
  for (int s = 0; s < b.S; ++s) {
    const double d = b.diam[s], r = rho[s];
    const double n2 = PI * r * d * d;
    m[0] += PI / 6.0 * r * d * d * d;
    m[1] += n2;
    m[2] += 0.5 * d * r;
    m[3] += r;
    m[10] += n2 / 3.0;  // T_xx, T_yy, T_zz
    m[13] += n2 / 3.0;
    m[15] += n2 / 3.0;
  }
}

void check_bulk_rho(const double* rho, int S) {
  if (!rho) throw std::invalid_argument("densities are required");
  for (int s = 0; s < S; ++s)
    if (!(rho[s] > 0.0) || !std::isfinite(rho[s])) throw std::invalid_argument("bulk densities must be positive");
}

// BulkThermodynamics::free_energy_density (thermodynamics.hpp:206-220).  As
// in the reference, a cross pair counts a rho_i rho_j here while the
// functional and the chemical potential give it half of that.
__host__ __device__ double bulk_free_energy_density(const BulkParams& b, const double* rho) {
  double f = 0.0;
  for (int s = 0; s < b.S; ++s) f += rho[s] * (log(rho[s]) - 1.0);  // ideal, :19-27
  if (b.fp.model != FMT_NONE) {  // hard spheres, :38-52
    double m[kMixFmt], g[kMixFmt];
    uniform_measures(b, rho, m);
    f += fmt_phi_grad_mix_t<double>(b.fp, m, g);
  }
  for (int q = 0; q < b.ninter; ++q) f += (b.si[q] == b.sj[q] ? 0.5 : 1.0) * b.a_vdw[q] * rho[b.si[q]] * rho[b.sj[q]];
  return f;
}

// BulkThermodynamics::chemical_potential (thermodynamics.hpp:222-237) of every species.
__host__ __device__ void bulk_chemical_potentials(const BulkParams& b, const double* rho, double* mu) {
  const double PI = 3.14159265358979323846;
  double m[kMixFmt], g[kMixFmt];
  const bool hs = b.fp.model != FMT_NONE;
  if (hs) {
    uniform_measures(b, rho, m);
    fmt_phi_grad_mix_t<double>(b.fp, m, g);
  }
  for (int s = 0; s < b.S; ++s) {
    double v = log(rho[s]);
    if (hs) {  // excess_chemical_potential, :54-89 (the tensor partials are zero for scalar models)
      const double d = b.diam[s];
      v += g[0] * PI / 6.0 * d * d * d + g[1] * PI * d * d + g[2] * 0.5 * d + g[3];
      v += (g[10] + g[13] + g[15]) * PI * d * d / 3.0;
    }
    for (int q = 0; q < b.ninter; ++q) {
      const int i = b.si[q], j = b.sj[q];
      if (i == s && j == s)
        v += b.a_vdw[q] * rho[j];
      else if (i == s)
        v += 0.5 * b.a_vdw[q] * rho[j];
      else if (j == s)
        v += 0.5 * b.a_vdw[q] * rho[i];
    }
    mu[s] = v;
  }
}

// Many bulk states at once, one thread each (phase-diagram and coexistence
// scans): f, the chemical potentials and P = sum rho mu - f
// (thermodynamics.hpp:206-255); each output nullable.
__global__ __launch_bounds__(kMixThreads) void k_bulk_batch(BulkParams b, long n, const double* __restrict__ rho,
                                                            double* __restrict__ f, double* __restrict__ mu,
                                                            double* __restrict__ p) {
  const long i = static_cast<long>(blockIdx.x) * kMixThreads + threadIdx.x;
  if (i >= n) return;
  double r[kMaxSpecies], m[kMaxSpecies];
  for (int s = 0; s < b.S; ++s) r[s] = rho[i * b.S + s];
  const double fe = bulk_free_energy_density(b, r);
  bulk_chemical_potentials(b, r, m);
  if (f) f[i] = fe;
  if (mu)
    for (int s = 0; s < b.S; ++s) mu[i * b.S + s] = m[s];
  if (p) {
    double v = -fe;
    for (int s = 0; s < b.S; ++s) v += r[s] * m[s];
    p[i] = v;
  }
}

void check_mu(const double* mu, int S) {
  if (!mu) throw std::invalid_argument("chemical potentials are required");
  for (int s = 0; s < S; ++s)
    if (!std::isfinite(mu[s])) throw std::invalid_argument("chemical potential is not finite");
}

// A smooth positive density for every species (autotuning input).
__global__ __launch_bounds__(kMixThreads) void k_mix_tune_density(double* rho, Geom g, int S) {
  const long n = g.N * S;
  for (long i = static_cast<long>(blockIdx.x) * kMixThreads + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * kMixThreads) {
    const long s = i / g.N, p = i - s * g.N;
    const long iz = p % g.nz, iy = (p / g.nz) % g.ny, ix = p / (static_cast<long>(g.ny) * g.nz);
    const double a = 6.283185307179586 * (static_cast<double>(ix) / g.nx + static_cast<double>(iy) / g.ny);
    rho[i] = (0.3 + 0.2 * sin(a) * cos(6.283185307179586 * static_cast<double>(iz) / g.nz)) / (1.0 + s);
  }
}

}  // namespace

double mixture_tune_time(Context* c, int reps, double* omega) {
  const MixLayout L = mix_layout(c);
  double* rho = at<double>(c, L.rho);
  double* force = at<double>(c, L.force);
  const Geom& g = c->g;
  k_mix_tune_density<<<spectral_blocks(g.N * c->n_species), kMixThreads, 0, c->stream>>>(rho, g, c->n_species);
  CDFT_HIP(hipGetLastError());
  const std::vector<double> mu(static_cast<size_t>(c->n_species), 0.0);
  mixture_eval(c, rho, nullptr, mu.data(), force);
  hipEvent_t a = nullptr, b = nullptr;
  CDFT_HIP(hipEventCreate(&a));
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    throw HipError("hipEventCreate failed");
  }
  float ms = 0.0f;
  try {
    CDFT_HIP(hipEventRecord(a, c->stream));
    for (int i = 0; i < reps; ++i) mixture_eval(c, rho, nullptr, mu.data(), force);
    CDFT_HIP(hipEventRecord(b, c->stream));
    cdft_result r{};
    mixture_result(c, &r);
    CDFT_HIP(hipEventElapsedTime(&ms, a, b));
    *omega = r.grand_potential;
  } catch (...) {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    throw;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return 1000.0 * ms / reps;
}

void init_mixture(Context* c) {
  const Geom& g = c->g;
  const MixLayout L = mix_layout(c);
  CDFT_HIP(hipMalloc(&c->d_mix, L.total));
  const int ni = static_cast<int>(c->inters.size());
  if (ni == 0) return;
  CDFT_HIP(hipMalloc(reinterpret_cast<void**>(&c->d_mix_mf), sizeof(double) * g.M * ni));
  // the B block (S * M complex, S >= 2) is free scratch at set-up
  auto* spec = at<cx<double>>(c, L.B);
  for (int q = 0; q < ni; ++q) {
    c->a_vdw.push_back(mf_weight_spectrum(c, c->inters[q], spec));
    if (c->fast)  // the fused transforms' sub-plane order
      launch_fast_mf_layout(spec, c->d_mix_mf + static_cast<long>(q) * g.M, g, c->stream);
    else
      store_real_spectrum(spec, c->d_mix_mf + static_cast<long>(q) * g.M, g.M, c->stream);
  }
  CDFT_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace cdft

using namespace cdft;

extern "C" {

int cdft_species_count(const cdft_ctx* ctx, int* n_species) {
  return guarded([&] {
    check_ctx_any(ctx);
    if (!n_species) throw std::invalid_argument("null argument");
    *n_species = ctx->impl->n_species;
  });
}

// One-species contexts take the one-species path with mu[0].
int cdft_total_species_device(cdft_ctx* ctx, const double* d_rho, const double* d_vext, const double* mu,
                              cdft_result* result, double* d_forces) {
  int n = 0;
  const int rc = cdft_species_count(ctx, &n);
  if (rc != CDFT_OK) return rc;
  if (!mu) return guarded([] { throw std::invalid_argument("chemical potentials are required"); });
  if (n == 1) return cdft_total_device(ctx, d_rho, d_vext, mu[0], result, d_forces);
  return guarded([&] {
    Context* c = ctx->impl;
    if (!d_rho || !result) throw std::invalid_argument("cdft_total_species_device: rho and result are required");
    check_mu(mu, c->n_species);
    CDFT_HIP(hipSetDevice(c->device));
    mixture_eval(c, d_rho, d_vext, mu, d_forces);
    mixture_result(c, result);
  });
}

int cdft_total_species(cdft_ctx* ctx, const double* rho, const double* vext, const double* mu, cdft_result* result,
                       double* forces) {
  int n = 0;
  const int rc = cdft_species_count(ctx, &n);
  if (rc != CDFT_OK) return rc;
  if (!mu) return guarded([] { throw std::invalid_argument("chemical potentials are required"); });
  if (n == 1) return cdft_total(ctx, rho, vext, mu[0], result, forces);
  return guarded([&] {
    Context* c = ctx->impl;
    if (!rho || !result) throw std::invalid_argument("cdft_total_species: rho and result are required");
    check_mu(mu, c->n_species);
    CDFT_HIP(hipSetDevice(c->device));
    const MixLayout L = mix_layout(c);
    const size_t bytes = sizeof(double) * c->n_species * c->g.N;
    double* d_rho = at<double>(c, L.rho);
    double* d_vext = vext ? at<double>(c, L.vext) : nullptr;
    double* d_force = forces ? at<double>(c, L.force) : nullptr;
    CDFT_HIP(hipMemcpyAsync(d_rho, rho, bytes, hipMemcpyHostToDevice, c->stream));
    if (vext) CDFT_HIP(hipMemcpyAsync(d_vext, vext, bytes, hipMemcpyHostToDevice, c->stream));
    mixture_eval(c, d_rho, d_vext, mu, d_force);
    mixture_result(c, result);
    if (forces) {
      CDFT_HIP(hipMemcpyAsync(forces, d_force, bytes, hipMemcpyDeviceToHost, c->stream));
      CDFT_HIP(hipStreamSynchronize(c->stream));
    }
  });
}

int cdft_contributions(cdft_ctx* ctx, const double* rho, const double* vext, const double* mu, cdft_result* result,
                       double* f_ideal, double* f_external, double* f_hard_sphere, double* f_mean_field) {
  return guarded([&] {
    check_ctx_any(ctx);
    Context* c = ctx->impl;
    if (!rho || !result) throw std::invalid_argument("cdft_contributions: rho and result are required");
    const int S = c->n_species;
    check_mu(mu, S);
    CDFT_HIP(hipSetDevice(c->device));
    const long N = c->g.N;
    const size_t field = sizeof(double) * N;
    double* outs[4] = {f_ideal, f_external, f_hard_sphere, f_mean_field};
    // device block: [4 S fields of output] [S*N rho] [S*N vext] [S*N total] [nf*N weighted densities]
    const bool one = S == 1;
    const long nwd = (one && c->cs.mf && (f_mean_field || f_hard_sphere)) ? c->cs.nf : 0;
    double* blk = nullptr;
    CDFT_HIP(hipMalloc(&blk, field * (4 * S + 3 * S + nwd)));
    try {
      double* d_out = blk;
      double* d_rho = blk + 4L * S * N;
      double* d_vext = vext ? d_rho + static_cast<long>(S) * N : nullptr;
      double* d_tot = d_rho + 2L * S * N;
      double* d_wd = d_rho + 3L * S * N;
      CDFT_HIP(hipMemcpyAsync(d_rho, rho, field * S, hipMemcpyHostToDevice, c->stream));
      if (vext) CDFT_HIP(hipMemcpyAsync(d_vext, vext, field * S, hipMemcpyHostToDevice, c->stream));
      const double* conv[kMaxSpecies] = {};
      const float* conv32[kMaxSpecies] = {};  // a mixed-precision mixture's FP32 potentials
      if (one) {  // (no cached graph: this call's buffers are its own)
        c->eval_direct(d_rho, d_vext, mu[0], d_tot, AliasMap{0, 0.0, 0.0});
        c->fetch_result(mu[0], result);
        if (nwd > 0) {
          c->weighted_densities(d_rho, d_wd);
          conv[0] = d_wd + (nwd - 1) * N;  // the mean-field convolution is the last channel
        }
      } else {
        mixture_eval(c, d_rho, d_vext, mu, d_tot);
        mixture_result(c, result);
        if (!c->inters.empty()) {  // the per-species potentials U_s stay in the channel fields
          const MixLayout L = mix_layout(c);
          const int nfm = c->cs.hs ? kMixFmtT : 0;
          for (int s = 0; s < S; ++s) {
            if (c->precision == CDFT_PRECISION_MIXED)
              conv32[s] = at<float>(c, L.X) + static_cast<long>(nfm + s) * N;
            else
              conv[s] = at<double>(c, L.X) + static_cast<long>(nfm + s) * N;
          }
        }
      }
      const unsigned nb = static_cast<unsigned>(std::min<long>(2048, (N + kMixThreads - 1) / kMixThreads));
      for (int s = 0; s < S; ++s) {
        const long o = static_cast<long>(s) * N;
        double* dst[4];
        for (int t = 0; t < 4; ++t) dst[t] = outs[t] ? d_out + (static_cast<long>(t) * S * N + o) : nullptr;
        if (conv32[s])
          k_split_terms<float><<<nb, kMixThreads, 0, c->stream>>>(d_rho + o, d_vext ? d_vext + o : nullptr,
                                                                  conv32[s], d_tot + o, mu[s], c->g.dv, N,
                                                                  c->cs.hs, dst[0], dst[1], dst[3], dst[2]);
        else
          k_split_terms<double><<<nb, kMixThreads, 0, c->stream>>>(d_rho + o, d_vext ? d_vext + o : nullptr,
                                                                   conv[s], d_tot + o, mu[s], c->g.dv, N,
                                                                   c->cs.hs, dst[0], dst[1], dst[3], dst[2]);
      }
      CDFT_HIP(hipGetLastError());
      for (int t = 0; t < 4; ++t)
        if (outs[t])
          CDFT_HIP(hipMemcpyAsync(outs[t], d_out + static_cast<long>(t) * S * N, field * S, hipMemcpyDeviceToHost,
                                  c->stream));
      CDFT_HIP(hipStreamSynchronize(c->stream));
    } catch (...) {
      (void)hipFree(blk);
      throw;
    }
    CDFT_HIP(hipFree(blk));
  });
}

int cdft_bulk_free_energy_density(const cdft_ctx* ctx, const double* rho, double* f) {
  return guarded([&] {
    check_ctx_any(ctx);
    if (!f) throw std::invalid_argument("null argument");
    check_bulk_rho(rho, ctx->impl->n_species);
    *f = bulk_free_energy_density(bulk_params(ctx->impl), rho);
  });
}

int cdft_bulk_chemical_potentials(const cdft_ctx* ctx, const double* rho, double* mu) {
  return guarded([&] {
    check_ctx_any(ctx);
    if (!mu) throw std::invalid_argument("null argument");
    check_bulk_rho(rho, ctx->impl->n_species);
    bulk_chemical_potentials(bulk_params(ctx->impl), rho, mu);
  });
}

int cdft_bulk_pressure(const cdft_ctx* ctx, const double* rho, double* p) {
  return guarded([&] {
    check_ctx_any(ctx);
    if (!p) throw std::invalid_argument("null argument");
    const Context* c = ctx->impl;
    check_bulk_rho(rho, c->n_species);
    std::vector<double> mu(static_cast<size_t>(c->n_species));
    const BulkParams b = bulk_params(c);
    bulk_chemical_potentials(b, rho, mu.data());
    double v = -bulk_free_energy_density(b, rho);  // P = sum rho mu - f (thermodynamics.hpp:239-246)
    for (int s = 0; s < c->n_species; ++s) v += rho[s] * mu[s];
    *p = v;
  });
}

int cdft_bulk_batch(const cdft_ctx* ctx, long n_states, const double* rho, double* f, double* mu, double* p) {
  return guarded([&] {
    check_ctx_any(ctx);
    const Context* c = ctx->impl;
    const int S = c->n_species;
    if (n_states < 0 || (n_states > 0 && !rho)) throw std::invalid_argument("cdft_bulk_batch: bad argument");
    if (n_states == 0) return;
    for (long i = 0; i < n_states * S; ++i)
      if (!(rho[i] > 0.0) || !std::isfinite(rho[i])) throw std::invalid_argument("bulk densities must be positive");
    CDFT_HIP(hipSetDevice(c->device));
    const size_t nr = sizeof(double) * n_states * S, n1 = sizeof(double) * n_states;
    double* blk = nullptr;
    CDFT_HIP(hipMalloc(&blk, 2 * nr + 2 * n1));
    try {
      double* d_rho = blk;
      double* d_mu = blk + n_states * S;
      double* d_f = d_mu + n_states * S;
      double* d_p = d_f + n_states;
      CDFT_HIP(hipMemcpyAsync(d_rho, rho, nr, hipMemcpyHostToDevice, c->stream));
      const unsigned nb = static_cast<unsigned>((n_states + kMixThreads - 1) / kMixThreads);
      k_bulk_batch<<<nb, kMixThreads, 0, c->stream>>>(bulk_params(c), n_states, d_rho, f ? d_f : nullptr,
                                                      mu ? d_mu : nullptr, p ? d_p : nullptr);
      CDFT_HIP(hipGetLastError());
      if (f) CDFT_HIP(hipMemcpyAsync(f, d_f, n1, hipMemcpyDeviceToHost, c->stream));
      if (mu) CDFT_HIP(hipMemcpyAsync(mu, d_mu, nr, hipMemcpyDeviceToHost, c->stream));
      if (p) CDFT_HIP(hipMemcpyAsync(p, d_p, n1, hipMemcpyDeviceToHost, c->stream));
      CDFT_HIP(hipStreamSynchronize(c->stream));
    } catch (...) {
      (void)hipFree(blk);
      throw;
    }
    CDFT_HIP(hipFree(blk));
  });
}

int cdft_time_total_species_device(cdft_ctx* ctx, const double* d_rho, const double* d_vext, const double* mu,
                                   int count, double* d_forces, double* ms) {
  return guarded([&] {
    check_ctx_any(ctx);
    Context* c = ctx->impl;
    if (c->n_species == 1) throw std::invalid_argument("one-species context: use cdft_time_total_device");
    if (!d_rho || !ms || count < 1) throw std::invalid_argument("cdft_time_total_species_device: bad argument");
    check_mu(mu, c->n_species);
    CDFT_HIP(hipSetDevice(c->device));
    hipEvent_t e0, e1;
    CDFT_HIP(hipEventCreate(&e0));
    CDFT_HIP(hipEventCreate(&e1));
    CDFT_HIP(hipEventRecord(e0, c->stream));
    for (int i = 0; i < count; ++i) mixture_eval(c, d_rho, d_vext, mu, d_forces);
    CDFT_HIP(hipEventRecord(e1, c->stream));
    CDFT_HIP(hipEventSynchronize(e1));
    float t = 0.0f;
    CDFT_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  });
}

}  // extern "C"

// HIP kernels of the free-energy + gradient evaluation for gfx950.
//
// Data layout in HBM
//   real fields (rho, V, forces, x, v):  double, idx = iz + Nz*(iy + Ny*ix)
//                                        (grid.hpp:35-37 of the reference)
//   spectral work arrays:                cx<R>, "plane-major": idx = iy + Ny*(ix + Nx*kz),
//                                        kz in [0, Nz/2], one array of M per channel
// The plane-major order makes the y lines contiguous, the x lines
// Ny-strided inside one contiguous kz plane, and lets a z-line kernel that
// owns consecutive (ix,iy) lines write every kz plane with unit stride.
//
// Evaluation schedule (one species, FMT + optional mean field):
//   k_zr2c      z-R2C of rho                                   (real -> S0)
//   k_lines     y-FFT of S0
//   k_xfwd      x-FFT, multiply by all weights, 12 inverse x-FFTs (S0 -> SF[12])
//   k_lines     inverse y-FFT of the 12 channels
//   k_zpoint    per z-line: 12 C2R, pointwise Phi/dPhi + local terms + partial
//               sums, 11 R2C of the combined derivatives      (SF -> SB[11], in place)
//   k_lines     y-FFT of the 11 channels
//   k_xbwd      x-FFT, conj-weighted channel sum, inverse x-FFT (SB -> SO)
//   k_lines     inverse y-FFT of SO
//   k_zc2r      z-C2R, * dV, + local force                      (SO -> force)
//   k_finalize  deterministic compensated sum of the block partials
#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>

#include "fft_lds.hpp"
#include "fmt_point.hpp"
#include "kernels.hpp"

namespace cdft {

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ void two_sum(double& s, double& c, double x) {
  // Neumaier compensated accumulation of x into (s, c)
  const double t = s + x;
  c += (fabs(s) >= fabs(x)) ? ((s - t) + x) : ((x - t) + s);
  s = t;
}

template <typename R>
__device__ __forceinline__ R kvec(int i, int n, R dk) {
  return dk * R(i <= n / 2 ? i : i - n);
}

// Radial FMT functions w3(k) = 4 pi (sin kR - kR cos kR)/k^3, w2(k) = 4 pi R sin(kR)/k
// (weights.hpp:174-191), each times 1/N.  The vector channel prefactor of
// weights.hpp:195-201 equals w3 and the tensor coefficients of
// weights.hpp:206-222 are A = w2 - 3 w3/R, B = w3/R (identities of the closed forms).
__device__ __forceinline__ void radial(double k, double R, double inv_n, double* w3, double* w2) {
  const double PI = 3.14159265358979323846;
  if (k < 1e-12) {
    *w3 = (4.0 * PI / 3.0) * R * R * R * inv_n;
    *w2 = 4.0 * PI * R * R * inv_n;
    return;
  }
  const double x = k * R;
  double s, c;
  sincos(x, &s, &c);
  *w3 = 4.0 * PI * (s - x * c) / (k * k * k) * inv_n;
  *w2 = 4.0 * PI * R * s / k * inv_n;
}

__device__ __forceinline__ void radial(float k, float R, float inv_n, float* w3, float* w2) {
  const float PI = 3.14159265358979323846f;
  const float x = k * R;
  if (x < 1.0f) {
    // series of (sin x - x cos x)/x^3 and sin x / x: no cancellation at small kR in FP32
    const float x2 = x * x;
    const float p3 = 1.0f / 3.0f -
                     x2 * (1.0f / 30.0f - x2 * (1.0f / 840.0f - x2 * (1.0f / 45360.0f - x2 * (1.0f / 3991680.0f))));
    const float p2 = 1.0f - x2 * (1.0f / 6.0f - x2 * (1.0f / 120.0f - x2 * (1.0f / 5040.0f - x2 * (1.0f / 362880.0f))));
    *w3 = 4.0f * PI * R * R * R * p3 * inv_n;
    *w2 = 4.0f * PI * R * R * p2 * inv_n;
    return;
  }
  float s, c;
  sincosf(x, &s, &c);
  *w3 = 4.0f * PI * (s - x * c) / (k * k * k) * inv_n;
  *w2 = 4.0f * PI * R * s / k * inv_n;
}

// Weight of channel ch at k: ch 0 w3, 1 w2, 2..4 i k_a w3 (conj when `adj`),
// 5..10 tensor xx,xy,xz,yy,yz,zz, MF (=11) the mean-field table value.
template <typename R>
__device__ __forceinline__ cx<R> channel_weight(int ch, bool adj, R kx, R ky, R kz, R w3, R w2, R mf, R invR) {
  if (ch == 0) return {w3, R(0)};
  if (ch == 1) return {w2, R(0)};
  if (ch <= 4) {
    const R ka = ch == 2 ? kx : (ch == 3 ? ky : kz);
    return {R(0), adj ? -ka * w3 : ka * w3};
  }
  if (ch <= 10) {
    const R k2 = kx * kx + ky * ky + kz * kz;
    const R ik2 = k2 > R(0) ? R(1) / k2 : R(0);
    const R A = w2 - R(3) * w3 * invR;
    const R B = w3 * invR;
    const int q = ch - 5;
    const R ki = q < 3 ? kx : (q < 5 ? ky : kz);
    const R kj = (q == 0) ? kx : (q == 1 || q == 3) ? ky : kz;
    const bool diag = (q == 0 || q == 3 || q == 5);
    return {A * ki * kj * ik2 + (diag ? B : R(0)), R(0)};
  }
  return {mf, R(0)};
}

// ---------------------------------------------------------------- z-line kernels
// Lines l = ix*Ny + iy of the real field are contiguous; a block owns `nlb`
// consecutive lines (nlb even) and packs line pairs into one complex FFT.

template <typename R>
__global__ void k_zr2c(const double* __restrict__ in, cx<R>* __restrict__ out, Geom g, FftPlan pz,
                       const cx<R>* __restrict__ twz, int nlb, AliasMap amap, const double* __restrict__ shift_sum) {
  extern __shared__ unsigned char smem_raw[];
  cx<R>* buf0 = reinterpret_cast<cx<R>*>(smem_raw);
  const int nz = g.nz, nzh = g.nzh;
  cx<R>* buf1 = buf0 + nlb * nzh;
  const long l0 = static_cast<long>(blockIdx.x) * nlb;
  const int ls = nz;
  const double shift = shift_sum ? shift_sum[0] * g.inv_n : 0.0;
  for (int e = threadIdx.x; e < nlb * nz; e += blockDim.x) {
    const int ll = e / nz, z = e - ll * nz;
    const long l = l0 + ll;
    R val = R(0);
    if (l < g.nxy) val = static_cast<R>(amap.apply(in[l * nz + z]) - shift);
    R* dst = reinterpret_cast<R*>(&buf0[(ll >> 1) * ls + z]);
    dst[ll & 1] = val;
  }
  __syncthreads();
  cx<R>* res = fft_lines(buf0, buf1, nlb >> 1, ls, pz, -1, twz);
  __syncthreads();
  for (int e = threadIdx.x; e < nlb * nzh; e += blockDim.x) {
    const int kz = e / nlb, ll = e - kz * nlb;
    const long l = l0 + ll;
    if (l >= g.nxy) continue;
    const cx<R>* line = res + (ll >> 1) * ls;
    const cx<R> zk = line[kz];
    const cx<R> zm = cconj(line[kz == 0 ? 0 : nz - kz]);
    cx<R> v;
    if ((ll & 1) == 0) v = {R(0.5) * (zk.re + zm.re), R(0.5) * (zk.im + zm.im)};
    else v = {R(0.5) * (zk.im - zm.im), R(-0.5) * (zk.re - zm.re)};
    out[static_cast<long>(kz) * g.nxy + l] = v;
  }
}

// Build the packed complex z-spectrum of line pairs from a staged half
// spectrum (stage[ll*nzh + kz]) with Hermitian projection of kz=0 and Nyquist.
template <typename R>
__device__ __forceinline__ void build_c2r_lines(const cx<R>* __restrict__ stage, cx<R>* __restrict__ dst, int nlb,
                                                int nz, int nzh) {
  const int np = nlb >> 1;
  for (int e = threadIdx.x; e < np * nz; e += blockDim.x) {
    const int p = e / nz, k = e - p * nz;
    const bool mirror = k >= nzh;
    const int kk = mirror ? nz - k : k;
    cx<R> A = stage[(2 * p) * nzh + kk];
    cx<R> B = stage[(2 * p + 1) * nzh + kk];
    if (kk == 0 || (2 * kk == nz)) {
      A.im = R(0);
      B.im = R(0);
    }
    if (mirror) {
      A.im = -A.im;
      B.im = -B.im;
    }
    dst[p * nz + k] = {A.re - B.im, A.im + B.re};
  }
}

// Generic z-C2R of `nch` channels: out[c*N + p] = scale * c2r(S_c)[p] + (add ? add[p] : 0)
template <typename R>
__global__ void k_zc2r(const cx<R>* __restrict__ S, double* __restrict__ out, const double* __restrict__ add,
                       double scale, Geom g, FftPlan pz, const cx<R>* __restrict__ twz, int nlb,
                       const double* __restrict__ shift_sum, DcTable dc) {
  extern __shared__ unsigned char smem_raw[];
  cx<R>* buf0 = reinterpret_cast<cx<R>*>(smem_raw);
  const int nz = g.nz, nzh = g.nzh;
  cx<R>* buf1 = buf0 + nlb * nzh;
  const int ch = blockIdx.y;
  const cx<R>* Sc = S + static_cast<long>(ch) * g.M;
  const long l0 = static_cast<long>(blockIdx.x) * nlb;
  for (int e = threadIdx.x; e < nlb * nzh; e += blockDim.x) {
    const int kz = e / nlb, ll = e - kz * nlb;
    const long l = l0 + ll;
    buf1[ll * nzh + kz] = (l < g.nxy) ? Sc[static_cast<long>(kz) * g.nxy + l] : cx<R>{R(0), R(0)};
  }
  __syncthreads();
  build_c2r_lines(buf1, buf0, nlb, nz, nzh);
  __syncthreads();
  cx<R>* res = fft_lines(buf0, buf1, nlb >> 1, nz, pz, +1, twz);
  __syncthreads();
  double* oc = out + static_cast<long>(ch) * g.N;
  const double cst = shift_sum ? dc.v[ch] * shift_sum[0] * g.inv_n : 0.0;
  for (int e = threadIdx.x; e < nlb * nz; e += blockDim.x) {
    const int ll = e / nz, z = e - ll * nz;
    const long l = l0 + ll;
    if (l >= g.nxy) continue;
    const cx<R> v = res[(ll >> 1) * nz + z];
    const long p = l * nz + z;
    const double val = static_cast<double>((ll & 1) ? v.im : v.re) * scale + cst;
    oc[p] = add ? val + add[p] : val;
  }
}

// FP64 block sums of amap(in): partials[b*8 + 0]; slots 1..7 zeroed.
__global__ void __launch_bounds__(256) k_sum(const double* __restrict__ in, AliasMap amap, long n,
                                             double* __restrict__ partials) {
  __shared__ double sh[256];
  double s = 0.0, c = 0.0;
  const long stride = static_cast<long>(gridDim.x) * blockDim.x;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    two_sum(s, c, amap.apply(in[i]));
  sh[threadIdx.x] = s + c;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 8) partials[static_cast<long>(blockIdx.x) * 8 + threadIdx.x] = threadIdx.x == 0 ? sh[0] : 0.0;
}

// ---------------------------------------------------------------- y lines (contiguous)
template <typename R>
__global__ void k_lines(cx<R>* __restrict__ S, long nlines, FftPlan p, const cx<R>* __restrict__ tw, int nlb, int sg) {
  extern __shared__ unsigned char smem_raw[];
  cx<R>* buf0 = reinterpret_cast<cx<R>*>(smem_raw);
  const int L = p.L;
  cx<R>* buf1 = buf0 + nlb * L;
  const long line0 = static_cast<long>(blockIdx.x) * nlb;
  const int nl = static_cast<int>(min(static_cast<long>(nlb), nlines - line0));
  cx<R>* base = S + line0 * L;
  for (int e = threadIdx.x; e < nl * L; e += blockDim.x) buf0[e] = base[e];
  __syncthreads();
  cx<R>* res = fft_lines(buf0, buf1, nl, L, p, sg, tw);
  __syncthreads();
  for (int e = threadIdx.x; e < nl * L; e += blockDim.x) base[e] = res[e];
}

// ---------------------------------------------------------------- x lines (strided in a kz plane)
// Block (y-tile, kz, channel): TY consecutive y columns, all x.  LDS line stride nx+1.
template <typename R>
__global__ void k_xlines(cx<R>* __restrict__ S, Geom g, FftPlan px, const cx<R>* __restrict__ twx, int ty, int sg) {
  extern __shared__ unsigned char smem_raw[];
  cx<R>* buf0 = reinterpret_cast<cx<R>*>(smem_raw);
  const int nx = g.nx, ny = g.ny, ls = nx + 1;
  cx<R>* buf1 = buf0 + ty * ls;
  const int y0 = blockIdx.x * ty;
  const int kz = blockIdx.y;
  cx<R>* plane = S + static_cast<long>(blockIdx.z) * g.M + static_cast<long>(kz) * nx * ny;
  for (int e = threadIdx.x; e < ty * nx; e += blockDim.x) {
    const int x = e / ty, yl = e - x * ty, y = y0 + yl;
    buf0[yl * ls + x] = (y < ny) ? plane[x * ny + y] : cx<R>{R(0), R(0)};
  }
  __syncthreads();
  cx<R>* res = fft_lines(buf0, buf1, ty, ls, px, sg, twx);
  __syncthreads();
  for (int e = threadIdx.x; e < ty * nx; e += blockDim.x) {
    const int x = e / ty, yl = e - x * ty, y = y0 + yl;
    if (y < ny) plane[x * ny + y] = res[yl * ls + x];
  }
}

// Fused forward k-space stage: x-FFT of rho_hat, multiply by every channel
// weight (analytic FMT weights, mean-field table), inverse x-FFT per channel.
template <typename R, int EPT>
__global__ void k_xfwd(const cx<R>* __restrict__ S0, cx<R>* __restrict__ SF, const R* __restrict__ mf, Geom g,
                       FftPlan px, const cx<R>* __restrict__ twx, int ty, ChannelSet cs) {
  extern __shared__ unsigned char smem_raw[];
  cx<R>* buf0 = reinterpret_cast<cx<R>*>(smem_raw);
  const int nx = g.nx, ny = g.ny, ls = nx + 1;
  cx<R>* buf1 = buf0 + ty * ls;
  const int y0 = blockIdx.x * ty;
  const int kz = blockIdx.y;
  const long plane_off = static_cast<long>(kz) * nx * ny;
  const cx<R>* plane = S0 + plane_off;
  for (int e = threadIdx.x; e < ty * nx; e += blockDim.x) {
    const int x = e / ty, yl = e - x * ty, y = y0 + yl;
    buf0[yl * ls + x] = (y < ny) ? plane[x * ny + y] : cx<R>{R(0), R(0)};
  }
  __syncthreads();
  const cx<R>* res = fft_lines(buf0, buf1, ty, ls, px, -1, twx);
  __syncthreads();
  cx<R> rh[EPT];
  R w3[EPT], w2[EPT], wm[EPT];
  const R kzv = R(g.dkz) * R(kz);
  const R Rr = R(g.R);
#pragma unroll
  for (int m = 0; m < EPT; ++m) {
    const int e = threadIdx.x + m * blockDim.x;
    rh[m] = {R(0), R(0)};
    w3[m] = w2[m] = wm[m] = R(0);
    if (e < ty * nx) {
      const int x = e / ty, yl = e - x * ty, y = y0 + yl;
      rh[m] = res[yl * ls + x];
      if (y < ny) {
        const R kx = kvec<R>(x, nx, R(g.dkx)), ky = kvec<R>(y, ny, R(g.dky));
        const R k = sqrt(kx * kx + ky * ky + kzv * kzv);
        if (cs.hs) radial(k, Rr, R(g.inv_n), &w3[m], &w2[m]);
        if (cs.mf) wm[m] = mf[plane_off + x * ny + y];
      }
    }
  }
  __syncthreads();
  const R invR = R(1) / Rr;
  for (int c = 0; c < cs.nf; ++c) {
    const int ch = cs.hs ? c : 11;
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
      const int e = threadIdx.x + m * blockDim.x;
      if (e < ty * nx) {
        const int x = e / ty, yl = e - x * ty, y = y0 + yl;
        const R kx = kvec<R>(x, nx, R(g.dkx)), ky = kvec<R>(min(y, ny - 1), ny, R(g.dky));
        const cx<R> w = channel_weight<R>(ch, false, kx, ky, kzv, w3[m], w2[m], wm[m], invR);
        buf0[yl * ls + x] = cmul(w, rh[m]);
      }
    }
    __syncthreads();
    const cx<R>* r = fft_lines(buf0, buf1, ty, ls, px, +1, twx);
    __syncthreads();
    cx<R>* outp = SF + static_cast<long>(c) * g.M + plane_off;
    for (int e = threadIdx.x; e < ty * nx; e += blockDim.x) {
      const int x = e / ty, yl = e - x * ty, y = y0 + yl;
      if (y < ny) outp[x * ny + y] = r[yl * ls + x];
    }
    __syncthreads();
  }
}

// Fused backward k-space stage: x-FFT of every derivative channel,
// sum of weight * channel (conj weight on the odd vector channels,
// hard_sphere.hpp:118-134), inverse x-FFT of the sum.
template <typename R, int EPT>
__global__ void k_xbwd(const cx<R>* __restrict__ SB, cx<R>* __restrict__ SO, Geom g, FftPlan px,
                       const cx<R>* __restrict__ twx, int ty, int nb) {
  extern __shared__ unsigned char smem_raw[];
  cx<R>* buf0 = reinterpret_cast<cx<R>*>(smem_raw);
  const int nx = g.nx, ny = g.ny, ls = nx + 1;
  cx<R>* buf1 = buf0 + ty * ls;
  const int y0 = blockIdx.x * ty;
  const int kz = blockIdx.y;
  const long plane_off = static_cast<long>(kz) * nx * ny;
  const R kzv = R(g.dkz) * R(kz);
  const R Rr = R(g.R);
  const R invR = R(1) / Rr;
  cx<R> acc[EPT];
  R w3[EPT], w2[EPT];
#pragma unroll
  for (int m = 0; m < EPT; ++m) {
    acc[m] = {R(0), R(0)};
    w3[m] = w2[m] = R(0);
    const int e = threadIdx.x + m * blockDim.x;
    if (e < ty * nx) {
      const int x = e / ty, yl = e - x * ty, y = y0 + yl;
      if (y < ny) {
        const R kx = kvec<R>(x, nx, R(g.dkx)), ky = kvec<R>(y, ny, R(g.dky));
        radial(sqrt(kx * kx + ky * ky + kzv * kzv), Rr, R(g.inv_n), &w3[m], &w2[m]);
      }
    }
  }
  for (int c = 0; c < nb; ++c) {
    const cx<R>* plane = SB + static_cast<long>(c) * g.M + plane_off;
    for (int e = threadIdx.x; e < ty * nx; e += blockDim.x) {
      const int x = e / ty, yl = e - x * ty, y = y0 + yl;
      buf0[yl * ls + x] = (y < ny) ? plane[x * ny + y] : cx<R>{R(0), R(0)};
    }
    __syncthreads();
    const cx<R>* r = fft_lines(buf0, buf1, ty, ls, px, -1, twx);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
      const int e = threadIdx.x + m * blockDim.x;
      if (e < ty * nx) {
        const int x = e / ty, yl = e - x * ty, y = y0 + yl;
        const R kx = kvec<R>(x, nx, R(g.dkx)), ky = kvec<R>(min(y, ny - 1), ny, R(g.dky));
        const cx<R> w = channel_weight<R>(c, true, kx, ky, kzv, w3[m], w2[m], R(0), invR);
        const cx<R> t = cmul(w, r[yl * ls + x]);
        acc[m] = cadd(acc[m], t);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int m = 0; m < EPT; ++m) {
    const int e = threadIdx.x + m * blockDim.x;
    if (e < ty * nx) {
      const int x = e / ty, yl = e - x * ty;
      buf0[yl * ls + x] = acc[m];
    }
  }
  __syncthreads();
  const cx<R>* r = fft_lines(buf0, buf1, ty, ls, px, +1, twx);
  __syncthreads();
  cx<R>* outp = SO + plane_off;
  for (int e = threadIdx.x; e < ty * nx; e += blockDim.x) {
    const int x = e / ty, yl = e - x * ty, y = y0 + yl;
    if (y < ny) outp[x * ny + y] = r[yl * ls + x];
  }
}

// ---------------------------------------------------------------- fused z-line pointwise kernel
// Per block of nlb consecutive z-lines: C2R of every forward channel, the
// pointwise FMT Phi/dPhi (FP64), the ideal-gas, external-field and
// mean-field local terms, compensated partial sums, and R2C of the 11
// combined derivative channels (written in place over SF's first channels).
template <typename R, int PPT>
__global__ void __launch_bounds__(512) k_zpoint(cx<R>* SF, const double* __restrict__ rho_in,
                                                const double* __restrict__ vext, double* __restrict__ local_force,
                                                double* __restrict__ partials, Geom g, FftPlan pz,
                                                const cx<R>* __restrict__ twz, int nlb, ChannelSet cs, FmtParams fp,
                                                double mu, AliasMap amap, const double* __restrict__ shift_sum,
                                                DcTable dc) {
  extern __shared__ unsigned char smem_raw[];
  cx<R>* buf0 = reinterpret_cast<cx<R>*>(smem_raw);
  const int nz = g.nz, nzh = g.nzh;
  cx<R>* buf1 = buf0 + nlb * nzh;
  const long l0 = static_cast<long>(blockIdx.x) * nlb;
  const int npts = nlb * nz;
  R wv[12][PPT];

  for (int c = 0; c < cs.nf; ++c) {
    const cx<R>* Sc = SF + static_cast<long>(c) * g.M;
    for (int e = threadIdx.x; e < nlb * nzh; e += blockDim.x) {
      const int kz = e / nlb, ll = e - kz * nlb;
      const long l = l0 + ll;
      buf1[ll * nzh + kz] = (l < g.nxy) ? Sc[static_cast<long>(kz) * g.nxy + l] : cx<R>{R(0), R(0)};
    }
    __syncthreads();
    build_c2r_lines(buf1, buf0, nlb, nz, nzh);
    __syncthreads();
    const cx<R>* res = fft_lines(buf0, buf1, nlb >> 1, nz, pz, +1, twz);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      const int q = threadIdx.x + m * blockDim.x;
      R val = R(0);
      if (q < npts) {
        const int ll = q / nz, z = q - ll * nz;
        const cx<R> v = res[(ll >> 1) * nz + z];
        val = (ll & 1) ? v.im : v.re;
      }
      wv[c][m] = val;
    }
    __syncthreads();
  }

  // pointwise; the combined derivatives overwrite wv[0..10] once consumed
  double s_id = 0, c_id = 0, s_ext = 0, c_ext = 0, s_hs = 0, c_hs = 0, s_mf = 0, c_mf = 0, s_n = 0, c_n = 0;
  R (&gv)[12][PPT] = wv;
  const double dv = g.dv;
  const int mfc = cs.hs ? 11 : 0;
  const double shift = shift_sum ? shift_sum[0] * g.inv_n : 0.0;
#pragma unroll
  for (int m = 0; m < PPT; ++m) {
    const int q = threadIdx.x + m * blockDim.x;
    const long l = l0 + q / nz;
    if (q < npts && l < g.nxy) {
      const long p = l0 * nz + q;
      const double rho = amap.apply(rho_in[p]);
      const double lr = log(fmax(rho, 1e-300));
      double lf = (lr - mu) * dv;                // ideal_gas.hpp:312-318
      two_sum(s_id, c_id, rho * (lr - 1.0));
      two_sum(s_n, c_n, rho);
      if (vext) {                                 // external_field.hpp:16-31
        const double V = vext[p];
        lf += V * dv;
        two_sum(s_ext, c_ext, rho * V);
      }
      if (cs.mf) {                                // mean_field.hpp:259-266 (self pair)
        const double conv = static_cast<double>(wv[mfc][m]) + dc.v[mfc] * shift;
        lf += conv * dv;
        two_sum(s_mf, c_mf, 0.5 * rho * conv);
      }
      if (cs.hs) {
        double w[11], gg[11];
#pragma unroll
        for (int i = 0; i < 11; ++i) w[i] = static_cast<double>(wv[i][m]) + dc.v[i] * shift;
        const double phi = fmt_phi_grad(fp, w, gg);
        two_sum(s_hs, c_hs, phi);
#pragma unroll
        for (int i = 0; i < 11; ++i) gv[i][m] = static_cast<R>(gg[i]);
      }
      local_force[p] = lf;
    } else {
#pragma unroll
      for (int i = 0; i < 11; ++i) gv[i][m] = R(0);
    }
  }

  // block reduction of the five compensated sums
  __shared__ double red[5][512];
  {
    const double v5[5] = {s_id + c_id, s_ext + c_ext, s_hs + c_hs, s_mf + c_mf, s_n + c_n};
    for (int i = 0; i < 5; ++i) red[i][threadIdx.x] = v5[i];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s)
        for (int i = 0; i < 5; ++i) red[i][threadIdx.x] += red[i][threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x < 8)
      partials[static_cast<long>(blockIdx.x) * 8 + threadIdx.x] = threadIdx.x < 5 ? red[threadIdx.x][0] * dv : 0.0;
  }

  // forward z R2C of the combined derivatives
  for (int c = 0; c < cs.nb; ++c) {
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      const int q = threadIdx.x + m * blockDim.x;
      if (q < npts) {
        const int ll = q / nz, z = q - ll * nz;
        R* dst = reinterpret_cast<R*>(&buf0[(ll >> 1) * nz + z]);
        dst[ll & 1] = gv[c][m];
      }
    }
    __syncthreads();
    const cx<R>* res = fft_lines(buf0, buf1, nlb >> 1, nz, pz, -1, twz);
    __syncthreads();
    cx<R>* out = SF + static_cast<long>(c) * g.M;
    for (int e = threadIdx.x; e < nlb * nzh; e += blockDim.x) {
      const int kz = e / nlb, ll = e - kz * nlb;
      const long l = l0 + ll;
      if (l >= g.nxy) continue;
      const cx<R>* line = res + (ll >> 1) * nz;
      const cx<R> zk = line[kz];
      const cx<R> zm = cconj(line[kz == 0 ? 0 : nz - kz]);
      cx<R> v;
      if ((ll & 1) == 0) v = {R(0.5) * (zk.re + zm.re), R(0.5) * (zk.im + zm.im)};
      else v = {R(0.5) * (zk.im - zm.im), R(-0.5) * (zk.re - zm.re)};
      out[static_cast<long>(kz) * g.nxy + l] = v;
    }
    __syncthreads();
  }
}

// Deterministic compensated sum of per-block partials (nblk x 8 doubles) into res[0..7].
__global__ void k_finalize(const double* __restrict__ partials, int nblk, double* __restrict__ res) {
  __shared__ double sh[8][256];
  __shared__ double shc[8][256];
  for (int i = 0; i < 8; ++i) {
    double s = 0, c = 0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) two_sum(s, c, partials[static_cast<long>(b) * 8 + i]);
    sh[i][threadIdx.x] = s;
    shc[i][threadIdx.x] = c;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    const int i = threadIdx.x;
    double s = 0, c = 0;
    for (int t = 0; t < static_cast<int>(blockDim.x); ++t) {
      two_sum(s, c, sh[i][t]);
      c += shc[i][t];
    }
    res[i] = s + c;
  }
}

// ---------------------------------------------------------------- launchers
template <typename R>
static int zlines_per_block(int nz) {
  int nlb = 2048 / nz;
  if (nlb < 2) nlb = 2;
  if (nlb & 1) --nlb;
  if (nlb > 32) nlb = 32;
  return nlb;
}

template <typename R>
void launch_zr2c(const double* in, cx<R>* out, const Geom& g, const FftPlan& pz, const cx<R>* twz, AliasMap amap,
                 const double* shift_sum, hipStream_t st) {
  const int nlb = zlines_per_block<R>(g.nz);
  const unsigned nblk = static_cast<unsigned>((g.nxy + nlb - 1) / nlb);
  const size_t sh = 2 * sizeof(cx<R>) * nlb * g.nzh;
  hipLaunchKernelGGL(k_zr2c<R>, dim3(nblk), dim3(256), sh, st, in, out, g, pz, twz, nlb, amap, shift_sum);
}

template <typename R>
void launch_zc2r(const cx<R>* S, int nch, double* out, const double* add, double scale, const Geom& g,
                 const FftPlan& pz, const cx<R>* twz, const double* shift_sum, DcTable dc, hipStream_t st) {
  const int nlb = zlines_per_block<R>(g.nz);
  const unsigned nblk = static_cast<unsigned>((g.nxy + nlb - 1) / nlb);
  const size_t sh = 2 * sizeof(cx<R>) * nlb * g.nzh;
  hipLaunchKernelGGL(k_zc2r<R>, dim3(nblk, nch), dim3(256), sh, st, S, out, add, scale, g, pz, twz, nlb,
                     shift_sum, dc);
}

int sum_blocks(const Geom& g) {
  long b = (g.N + 4095) / 4096;
  if (b > 1024) b = 1024;
  if (b < 1) b = 1;
  return static_cast<int>(b);
}

void launch_sum(const double* in, AliasMap amap, double* partials, const Geom& g, hipStream_t st) {
  hipLaunchKernelGGL(k_sum, dim3(sum_blocks(g)), dim3(256), 0, st, in, amap, g.N, partials);
}

template <typename R>
void launch_lines(cx<R>* S, long nlines, const FftPlan& p, const cx<R>* tw, int sg, hipStream_t st) {
  int nlb = 2048 / p.L;
  if (nlb < 1) nlb = 1;
  const unsigned nblk = static_cast<unsigned>((nlines + nlb - 1) / nlb);
  const size_t sh = 2 * sizeof(cx<R>) * nlb * p.L;
  hipLaunchKernelGGL(k_lines<R>, dim3(nblk), dim3(256), sh, st, S, nlines, p, tw, nlb, sg);
}

template <typename R>
static int x_tile(const Geom& g) {
  return sizeof(R) == 4 ? 16 : 8;
}

template <typename R>
void launch_xlines(cx<R>* S, int nch, const Geom& g, const FftPlan& px, const cx<R>* twx, int sg, hipStream_t st) {
  const int ty = x_tile<R>(g);
  const size_t sh = 2 * sizeof(cx<R>) * ty * (g.nx + 1);
  hipLaunchKernelGGL(k_xlines<R>, dim3((g.ny + ty - 1) / ty, g.nzh, nch), dim3(256), sh, st, S, g, px, twx, ty, sg);
}

template <typename R>
void launch_xfwd(const cx<R>* S0, cx<R>* SF, const R* mf, const Geom& g, const FftPlan& px, const cx<R>* twx,
                 ChannelSet cs, hipStream_t st) {
  const int ty = x_tile<R>(g);
  const size_t sh = 2 * sizeof(cx<R>) * ty * (g.nx + 1);
  const int ept = (ty * g.nx + 255) / 256;
  if (ept > 16) throw std::runtime_error("x extent too large for the fused k-space kernel");
  dim3 grid((g.ny + ty - 1) / ty, g.nzh);
  if (ept <= 4) hipLaunchKernelGGL((k_xfwd<R, 4>), grid, dim3(256), sh, st, S0, SF, mf, g, px, twx, ty, cs);
  else if (ept <= 8) hipLaunchKernelGGL((k_xfwd<R, 8>), grid, dim3(256), sh, st, S0, SF, mf, g, px, twx, ty, cs);
  else hipLaunchKernelGGL((k_xfwd<R, 16>), grid, dim3(256), sh, st, S0, SF, mf, g, px, twx, ty, cs);
}

template <typename R>
void launch_xbwd(const cx<R>* SB, cx<R>* SO, const Geom& g, const FftPlan& px, const cx<R>* twx, int nb,
                 hipStream_t st) {
  const int ty = x_tile<R>(g);
  const size_t sh = 2 * sizeof(cx<R>) * ty * (g.nx + 1);
  const int ept = (ty * g.nx + 255) / 256;
  if (ept > 16) throw std::runtime_error("x extent too large for the fused k-space kernel");
  dim3 grid((g.ny + ty - 1) / ty, g.nzh);
  if (ept <= 4) hipLaunchKernelGGL((k_xbwd<R, 4>), grid, dim3(256), sh, st, SB, SO, g, px, twx, ty, nb);
  else if (ept <= 8) hipLaunchKernelGGL((k_xbwd<R, 8>), grid, dim3(256), sh, st, SB, SO, g, px, twx, ty, nb);
  else hipLaunchKernelGGL((k_xbwd<R, 16>), grid, dim3(256), sh, st, SB, SO, g, px, twx, ty, nb);
}

// z-lines per block of the fused pointwise kernel: 2048 points (FP32) or
// 1024 points (FP64) per 512-thread block, so each thread keeps 4 or 2
// points x 12 channels in registers.
template <typename R>
static int zpoint_nlb(int nz) {
  int nlb = (sizeof(R) == 4 ? 2048 : 1024) / nz;
  if (nlb < 2) nlb = 2;
  if (nlb & 1) --nlb;
  if (nlb > 32) nlb = 32;
  return nlb;
}

template <typename R>
int zpoint_blocks(const Geom& g) {
  const int nlb = zpoint_nlb<R>(g.nz);
  return static_cast<int>((g.nxy + nlb - 1) / nlb);
}
template int zpoint_blocks<float>(const Geom&);
template int zpoint_blocks<double>(const Geom&);

template <typename R>
void launch_zpoint(cx<R>* SF, const double* rho, const double* vext, double* local_force, double* partials,
                   const Geom& g, const FftPlan& pz, const cx<R>* twz, ChannelSet cs, FmtParams fp, double mu,
                   AliasMap amap, const double* shift_sum, DcTable dc, hipStream_t st) {
  const int nlb = zpoint_nlb<R>(g.nz);
  const unsigned nblk = static_cast<unsigned>((g.nxy + nlb - 1) / nlb);
  const size_t sh = 2 * sizeof(cx<R>) * nlb * g.nzh;
  const int ppt = (nlb * g.nz + 511) / 512;
  if (ppt > 4) throw std::runtime_error("z extent too large for the fused pointwise kernel");
  if (ppt <= 2)
    hipLaunchKernelGGL((k_zpoint<R, 2>), dim3(nblk), dim3(512), sh, st, SF, rho, vext, local_force, partials, g, pz,
                       twz, nlb, cs, fp, mu, amap, shift_sum, dc);
  else
    hipLaunchKernelGGL((k_zpoint<R, 4>), dim3(nblk), dim3(512), sh, st, SF, rho, vext, local_force, partials, g, pz,
                       twz, nlb, cs, fp, mu, amap, shift_sum, dc);
}

void launch_finalize(const double* partials, int nblk, double* res, hipStream_t st) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, partials, nblk, res);
}

#define CDFT_INST(R)                                                                                              \
  template void launch_zr2c<R>(const double*, cx<R>*, const Geom&, const FftPlan&, const cx<R>*, AliasMap,       \
                               const double*, hipStream_t);                                                       \
  template void launch_zc2r<R>(const cx<R>*, int, double*, const double*, double, const Geom&, const FftPlan&,   \
                               const cx<R>*, const double*, DcTable, hipStream_t);                                \
  template void launch_lines<R>(cx<R>*, long, const FftPlan&, const cx<R>*, int, hipStream_t);                   \
  template void launch_xlines<R>(cx<R>*, int, const Geom&, const FftPlan&, const cx<R>*, int, hipStream_t);      \
  template void launch_xfwd<R>(const cx<R>*, cx<R>*, const R*, const Geom&, const FftPlan&, const cx<R>*,         \
                               ChannelSet, hipStream_t);                                                          \
  template void launch_xbwd<R>(const cx<R>*, cx<R>*, const Geom&, const FftPlan&, const cx<R>*, int, hipStream_t); \
  template void launch_zpoint<R>(cx<R>*, const double*, const double*, double*, double*, const Geom&,            \
                                 const FftPlan&, const cx<R>*, ChannelSet, FmtParams, double, AliasMap,            \
                                 const double*, DcTable, hipStream_t);

CDFT_INST(float)
CDFT_INST(double)

}  // namespace cdft

// HIP kernels of the free-energy + gradient evaluation for gfx950.
//
// Data layout in HBM
//   real fields (rho, V, forces, x, v):  double, idx = iz + Nz*(iy + Ny*ix)
//                                        (grid.hpp:35-37 of the reference)
//   spectral work arrays:                cx<S>, "plane-major": idx = iy + Ny*(ix + Nx*kz),
//                                        kz in [0, Nz/2], one array of M per channel
// The plane-major order makes the y lines contiguous, the x lines
// Ny-strided inside one contiguous kz plane, and lets a z-line kernel that
// owns consecutive (ix,iy) lines write every kz plane with unit stride.
//
// Every transform kernel is templated on the storage type S of the spectral
// arrays in HBM and on the arithmetic type C used in LDS and registers.  The
// mixed-precision evaluation stores FP32 everywhere (half the traffic) and
// computes the scalar channels' (eta, n2, mean field) inverse transforms and
// the x stage of every channel in FP64: a smooth density concentrates its
// spectrum in a few large low-k values, whose FP32 arithmetic errors would
// otherwise add up coherently in Omega.  The vector and tensor channels' y
// and z stages and the backward path run in FP32.
//
// Evaluation schedule (one species, FMT + optional mean field):
//   k_sum       FP64 block sums of rho (mean-density split)
//   k_zr2c      z-R2C of rho - mean (each block reduces the sums) (real -> S0)
//   k_lines     y-FFT of S0
//   k_xfwd      x-FFT, 7 weighted inverse x-FFTs -> channels   (S0 -> SF[nf])
//   k_lines     inverse y-FFT of the nf channels
//   k_zpoint    per z-line: nf C2R, pointwise Phi/dPhi + local terms + partial
//               sums, 11 R2C of the combined derivatives        (SF -> SF[11], in place)
//   k_lines     y-FFT of the 11 channels
//   k_xbwd      6 grouped x-FFTs, conj-weighted sum, inverse x-FFT (SF -> SO)
// For the z extents of kg_zpoint (48, 80, 96, 160) SF holds reduced channel
// sets instead: nf - 2 forward (v_z and T_zz are formed in the point kernel)
// and 9 backward (their derivatives folded into the others).
//   k_lines     inverse y-FFT of SO
//   k_zc2r      z-C2R, * dV, + local force                      (SO -> force);
//               its block 0 first reduces the point kernel's partial sums
//   k_finalize  the same reduction when no forces are evaluated
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>
#include <cmath>
#include <stdexcept>
#include <type_traits>

#include "device_common.hpp"
#include "fft_lds.hpp"
#include "line_fft.hpp"
#include "fmt_point.hpp"
#include "kernels.hpp"

namespace cdft {

constexpr int NT = 256;  // threads per block of every transform kernel

// v * w for an FP64 weight: in FP32 the weight is applied as the pair
// (FP32 value, FP32 residual), like the FFT constants.
template <typename C>
__device__ __forceinline__ cx<C> wmul(cx<C> v, cx<double> w) {
  if constexpr (sizeof(C) == 4) {
    const cx<float> hi = {static_cast<float>(w.re), static_cast<float>(w.im)};
    const cx<float> lo = {static_cast<float>(w.re - static_cast<double>(hi.re)),
                          static_cast<float>(w.im - static_cast<double>(hi.im))};
    return cadd(cmul(v, hi), cmul(v, lo));
  } else {
    return cmul(v, w);
  }
}

// R2C unpack of a packed line pair: line spectrum Z of (a + i b); returns A[k] or B[k].
template <typename C>
__device__ __forceinline__ cx<C> unpack_pair(const cx<C>* line, int kz, int nz, bool odd) {
  const cx<C> zk = line[kz];
  const cx<C> zm = cconj(line[kz == 0 ? 0 : nz - kz]);
  if (!odd) return {C(0.5) * (zk.re + zm.re), C(0.5) * (zk.im + zm.im)};
  return {C(0.5) * (zk.im - zm.im), C(-0.5) * (zk.re - zm.re)};
}

// C2R pack of line pairs from staged half spectra (stage[ll*nzh + kz]) with
// Hermitian projection of the kz = 0 and Nyquist bins; writes Z[k] for all k.
template <typename C, typename T>
__device__ __forceinline__ void build_c2r_lines(const cx<T>* __restrict__ stage, cx<C>* __restrict__ dst, int npairs,
                                                int nz, int nzh) {
  const UDiv dnz(nz);  // block-uniform divisors
  for (int e = threadIdx.x; e < npairs * nz; e += blockDim.x) {
    const int p = dnz.div(e), k = e - p * nz;
    const bool mirror = k >= nzh;
    const int kk = mirror ? nz - k : k;
    cx<C> A = cvt<C>(stage[(2 * p) * nzh + kk]);
    cx<C> B = cvt<C>(stage[(2 * p + 1) * nzh + kk]);
    if (kk == 0 || (2 * kk == nz)) {
      A.im = C(0);
      B.im = C(0);
    }
    if (mirror) {
      A.im = -A.im;
      B.im = -B.im;
    }
    dst[p * nz + k] = {A.re - B.im, A.im + B.re};
  }
}

// C2R of staged half spectra of ng groups of np line pairs (group gi at
// stage + gi * hs, line ll of it at ll * nzh): the lines are built (as
// build_c2r_lines does) straight into the registers of a slot of T lanes,
// transformed there, and written to buf[(gi * np + p) * L ...].  Ends with a
// block barrier; the caller synchronises before (stage visible).
template <typename C, typename T_, int L, int T>
__device__ __forceinline__ void slot_c2r_lines(const cx<T_>* __restrict__ stage, int hs, int nzh, cx<C>* buf, int ng,
                                               int np, const cx<C>* __restrict__ tw) {
  constexpr int E = L / T;
  const int nl = ng * np;
  const int nslot = static_cast<int>(blockDim.x) / T;
  const int slot = static_cast<int>(threadIdx.x) / T, lane = static_cast<int>(threadIdx.x) % T;
  for (int j = slot; j < nl; j += nslot) {
    const int gi = j / np, pr = j - gi * np;
    const cx<T_>* sa = stage + gi * hs + (2 * pr) * nzh;  // line a; line b follows at + nzh
    cx<C> v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int k = lane + T * q;
      const bool mirror = k >= nzh;
      const int kk = mirror ? L - k : k;
      cx<C> A = cvt<C>(sa[kk]), B = cvt<C>(sa[nzh + kk]);
      if (kk == 0 || 2 * kk == L) {
        A.im = C(0);
        B.im = C(0);
      }
      if (mirror) {
        A.im = -A.im;
        B.im = -B.im;
      }
      v[q] = {A.re - B.im, A.im + B.re};
    }
    cx<C>* line = buf + static_cast<long>(j) * L;
    slot_fft<C, L, T, +1, false, false>(v, line, lane, tw);
    wave_sync();
#pragma unroll
    for (int e = 0; e < E; ++e) line[lane + T * e] = v[e];
    wave_sync();
  }
  __syncthreads();
}

// build_c2r_lines for every group, then the inverse transform of all ng * np
// lines; the slot lengths of fft_lines build in registers (slot_c2r_lines).
template <typename C, typename T_, int MAXE>
__device__ __forceinline__ void c2r_lines(const cx<T_>* __restrict__ stage, int hs, cx<C>* buf, int ng, int np, int nz,
                                          int nzh, const FftPlan& pz, const cx<C>* __restrict__ tw) {
  switch (pz.L) {
    case 48: slot_c2r_lines<C, T_, 48, 8>(stage, hs, nzh, buf, ng, np, tw); return;
    case 64: slot_c2r_lines<C, T_, 64, 8>(stage, hs, nzh, buf, ng, np, tw); return;
    case 80: slot_c2r_lines<C, T_, 80, 8>(stage, hs, nzh, buf, ng, np, tw); return;
    case 96: slot_c2r_lines<C, T_, 96, 16>(stage, hs, nzh, buf, ng, np, tw); return;
    case 128: slot_c2r_lines<C, T_, 128, 16>(stage, hs, nzh, buf, ng, np, tw); return;
    case 160: slot_c2r_lines<C, T_, 160, 16>(stage, hs, nzh, buf, ng, np, tw); return;
    default:
      for (int gi = 0; gi < ng; ++gi) build_c2r_lines(stage + gi * hs, buf + gi * np * nz, np, nz, nzh);
      __syncthreads();
      fft_inplace<C, MAXE>(buf, ng * np, nz, pz, +1, tw);
      return;
  }
}

// ---------------------------------------------------------------- z-line kernels
// Lines l = ix*Ny + iy of the real field are contiguous; a block owns `nlb`
// consecutive lines (nlb even) and packs line pairs into one complex FFT.

template <typename S, typename C>
__global__ void __launch_bounds__(NT) k_zr2c(const double* __restrict__ in, cx<S>* __restrict__ out, Geom g,
                                             FftPlan pz, const cx<C>* __restrict__ twz, int nlb, AliasMap amap,
                                             const double* __restrict__ shift_sum, const double* __restrict__ sum_rows,
                                             int sum_n, double* __restrict__ sum_out) {
  const UDiv dnlb(nlb), dnz(g.nz);  // block-uniform divisors
  extern __shared__ unsigned char smem_raw[];
  cx<C>* buf = reinterpret_cast<cx<C>*>(smem_raw);
  const int nz = g.nz, nzh = g.nzh;
  const long l0 = static_cast<long>(blockIdx.x) * nlb;
  double shift = shift_sum ? shift_sum[0] * g.inv_n : 0.0;
  if (sum_rows) {
    // the density sum from k_sum's block rows, reduced by every block (no
    // reduction launch comes between); block 0 stores it for the later kernels
    const double total = block_sum_col0<NT>(sum_rows, sum_n);
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 8) sum_out[threadIdx.x] = threadIdx.x == 0 ? total : 0.0;
    shift = total * g.inv_n;
  }
  in += static_cast<long>(blockIdx.y) * g.N;  // field blockIdx.y of a multi-field launch
  out += static_cast<long>(blockIdx.y) * g.M;
  for (int e = threadIdx.x; e < nlb * nz; e += blockDim.x) {
    const int ll = dnz.div(e), z = e - ll * nz;
    const long l = l0 + ll;
    C val = C(0);
    if (l < g.nxy) val = static_cast<C>(amap.apply(in[l * nz + z]) - shift);
    C* dst = reinterpret_cast<C*>(&buf[(ll >> 1) * nz + z]);
    dst[ll & 1] = val;
  }
  __syncthreads();
  fft_lines<C, 8>(buf, nlb >> 1, nz, pz, -1, twz);
  for (int e = threadIdx.x; e < nlb * nzh; e += blockDim.x) {
    const int kz = dnlb.div(e), ll = e - kz * nlb;
    const long l = l0 + ll;
    if (l >= g.nxy) continue;
    out[static_cast<long>(kz) * g.nxy + l] = cvt<S>(unpack_pair(buf + (ll >> 1) * nz, kz, nz, ll & 1));
  }
}

// Generic z-C2R of `nch` channels: out[c*N + p] = scale * c2r(S_c)[p] + cst_c + (add ? add[c*N + p] : 0)
// with cst_c = dc.v[c] * shift_sum[0] / N (the mean-density split) when shift_sum is given.
template <typename S, typename C>
__global__ void __launch_bounds__(NT) k_zc2r(const cx<S>* __restrict__ Sp, double* __restrict__ out,
                                             const double* __restrict__ add, double scale, Geom g, FftPlan pz,
                                             const cx<C>* __restrict__ twz, int nlb,
                                             const double* __restrict__ shift_sum, DcTable dc,
                                             const double* __restrict__ fin_rows, int fin_n, double* __restrict__ fin_res) {
  if (fin_rows && blockIdx.x == 0 && blockIdx.y == 0) {
    // the partial sums of an earlier kernel, reduced by block 0 before its own
    // lines (while the other blocks run) instead of by a launch of their own
    __shared__ double sh[NT], shc[NT];
    reduce_rows<NT>(fin_rows, fin_n, fin_res, sh, shc);
    __syncthreads();
  }
  const UDiv dnlb(nlb), dnz(g.nz);  // block-uniform divisors
  extern __shared__ unsigned char smem_raw[];
  const int nz = g.nz, nzh = g.nzh;
  cx<C>* buf = reinterpret_cast<cx<C>*>(smem_raw);
  cx<S>* stage = reinterpret_cast<cx<S>*>(buf + (nlb >> 1) * nz);
  const int ch = blockIdx.y;
  const cx<S>* Sc = Sp + static_cast<long>(ch) * g.M;
  const long l0 = static_cast<long>(blockIdx.x) * nlb;
  for (int e = threadIdx.x; e < nlb * nzh; e += blockDim.x) {
    const int kz = dnlb.div(e), ll = e - kz * nlb;
    const long l = l0 + ll;
    stage[ll * nzh + kz] = (l < g.nxy) ? Sc[static_cast<long>(kz) * g.nxy + l] : cx<S>{S(0), S(0)};
  }
  __syncthreads();
  c2r_lines<C, S, 8>(stage, 0, buf, 1, nlb >> 1, nz, nzh, pz, twz);
  double* oc = out + static_cast<long>(ch) * g.N;
  const double* ac = add ? add + static_cast<long>(ch) * g.N : nullptr;
  const double cst = shift_sum ? dc.v[ch] * shift_sum[0] * g.inv_n : 0.0;
  for (int e = threadIdx.x; e < nlb * nz; e += blockDim.x) {
    const int ll = dnz.div(e), z = e - ll * nz;
    const long l = l0 + ll;
    if (l >= g.nxy) continue;
    const cx<C> v = buf[(ll >> 1) * nz + z];
    const long p = l * nz + z;
    const double val = static_cast<double>((ll & 1) ? v.im : v.re) * scale + cst;
    oc[p] = ac ? val + ac[p] : val;
  }
}

// FP64 block sums of amap(in): partials[b*8 + 0]; slots 1..7 zeroed.
// Per-block sums of rho.  Each thread issues kSumUnroll loads (stride: the
// whole grid) before adding them, so their latencies overlap.
constexpr int kSumUnroll = 8;
__global__ void __launch_bounds__(NT) k_sum(const double* __restrict__ in, AliasMap amap, long n,
                                            double* __restrict__ partials) {
  __shared__ double sh[NT / 64];
  double s = 0.0, c = 0.0;
  const long stride = static_cast<long>(gridDim.x) * blockDim.x;
  for (long base = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; base < n; base += kSumUnroll * stride) {
    double v[kSumUnroll];
#pragma unroll
    for (int u = 0; u < kSumUnroll; ++u) {
      const long i = base + u * stride;
      v[u] = i < n ? amap.apply(in[i]) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kSumUnroll; ++u) two_sum(s, c, v[u]);
  }
  double x = s + c;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x < 8) {
    double t = 0.0;
    if (threadIdx.x == 0)
#pragma unroll
      for (int w = 0; w < NT / 64; ++w) t += sh[w];
    partials[static_cast<long>(blockIdx.x) * 8 + threadIdx.x] = t;
  }
}

// ---------------------------------------------------------------- y lines (contiguous)
// Lines line0 .. line0 + nl - 1 by one block.
template <typename S, typename C>
__device__ __forceinline__ void lines_block(cx<S>* __restrict__ Sp, long line0, int nl, const FftPlan& p,
                                            const cx<C>* __restrict__ tw, int sg) {
  extern __shared__ unsigned char smem_raw[];
  cx<C>* buf = reinterpret_cast<cx<C>*>(smem_raw);
  const int L = p.L;
  cx<S>* base = Sp + line0 * L;
  for (int e = threadIdx.x; e < nl * L; e += blockDim.x) buf[e] = cvt<C>(base[e]);
  __syncthreads();
  fft_lines<C, 8>(buf, nl, L, p, sg, tw);
  for (int e = threadIdx.x; e < nl * L; e += blockDim.x) base[e] = cvt<S>(buf[e]);
}

template <typename S, typename C>
__global__ void __launch_bounds__(NT, 4) k_lines(cx<S>* __restrict__ Sp, long nlines, FftPlan p,
                                                 const cx<C>* __restrict__ tw, int nlb, int sg) {
  const long line0 = static_cast<long>(blockIdx.x) * nlb;
  lines_block<S, C>(Sp, line0, static_cast<int>(min(static_cast<long>(nlb), nlines - line0)), p, tw, sg);
}

// Lines of two arithmetic precisions in one launch: lines [lo0, lo1) in CL,
// nlb per block, the others in CH, hb per block (blocks [0, na) take lines
// [0, lo0), the next nb blocks [lo0, lo1), the rest [lo1, nlines)), so both
// kinds of block fit the same LDS.
template <typename S, typename CH, typename CL>
__global__ void __launch_bounds__(NT, 4) k_lines_split(cx<S>* __restrict__ Sp, long nlines, FftPlan p,
                                                       const cx<CH>* __restrict__ twh,
                                                       const cx<CL>* __restrict__ twl, int nlb, int hb, int sg,
                                                       long lo0, long lo1, int na, int nb) {
  const long b = blockIdx.x;
  if (b < na) {
    const long line0 = b * hb;
    lines_block<S, CH>(Sp, line0, static_cast<int>(min(static_cast<long>(hb), lo0 - line0)), p, twh, sg);
  } else if (b < na + nb) {
    const long line0 = lo0 + (b - na) * nlb;
    lines_block<S, CL>(Sp, line0, static_cast<int>(min(static_cast<long>(nlb), lo1 - line0)), p, twl, sg);
  } else {
    const long line0 = lo1 + (b - na - nb) * hb;
    lines_block<S, CH>(Sp, line0, static_cast<int>(min(static_cast<long>(hb), nlines - line0)), p, twh, sg);
  }
}

// ---------------------------------------------------------------- x lines (strided in a kz plane)
// Block (y-tile, kz, channel): TY consecutive y columns, all x.  LDS line stride nx+1.
template <typename S, typename C, int MAXE, int NTH = NT>
__global__ void __launch_bounds__(NTH) k_xlines(cx<S>* __restrict__ Sp, Geom g, FftPlan px,
                                               const cx<C>* __restrict__ twx, int ty, int sg) {
  const UDiv dty(ty);  // block-uniform divisors
  extern __shared__ unsigned char smem_raw[];
  cx<C>* buf = reinterpret_cast<cx<C>*>(smem_raw);
  const int nx = g.nx, ny = g.ny, ls = nx + 1;
  const int y0 = blockIdx.x * ty;
  const int kz = blockIdx.y;
  cx<S>* plane = Sp + static_cast<long>(blockIdx.z) * g.M + static_cast<long>(kz) * nx * ny;
  for (int e = threadIdx.x; e < ty * nx; e += blockDim.x) {
    const int x = dty.div(e), yl = e - x * ty, y = y0 + yl;
    buf[yl * ls + x] = (y < ny) ? cvt<C>(plane[x * ny + y]) : cx<C>{C(0), C(0)};
  }
  __syncthreads();
  fft_lines<C, MAXE>(buf, ty, ls, px, sg, twx);
  for (int e = threadIdx.x; e < ty * nx; e += blockDim.x) {
    const int x = dty.div(e), yl = e - x * ty, y = y0 + yl;
    if (y < ny) plane[x * ny + y] = cvt<S>(buf[yl * ls + x]);
  }
}

// Channel order of kg_zpoint with the hard-sphere channels.  Forward round
// positions hold the channels (index as in the 12-channel set: 0 eta, 1 n2,
// 2..4 v, 5..10 T, 11 mean field) eta, n2, [mf,] v_z, v_x, v_y, T_xx, T_xy,
// T_xz, T_yy, T_yz; T_zz comes from the trace identity.  SF holds the reduced
// set: eta, n2, v_x, v_y, T_xx, T_xy, T_xz, T_yy, T_yz[, mf] (v_z is formed from
// eta's spectrum); backward: the same nine, eta carrying the v_z term and n2,
// T_xx, T_yy the T_zz derivative.
__host__ __device__ constexpr int kg_fwd_channel(int nf, int pos) {
  return nf == 12 ? (pos < 2 ? pos : pos == 2 ? 11 : pos == 3 ? 4 : pos < 6 ? pos - 2 : pos - 1)
                  : (pos < 2 ? pos : pos == 2 ? 4 : pos < 5 ? pos - 1 : pos);
}
__host__ __device__ constexpr int kg_fwd_slot(int ch) {  // -1: v_z, T_zz (not stored)
  return ch < 4 ? ch : ch == 4 || ch == 10 ? -1 : ch == 11 ? 9 : ch - 1;
}
__host__ __device__ constexpr int kg_bwd_channel(int pos) { return pos == 0 ? 0 : pos == 1 ? 4 : pos < 5 ? pos - 1 : pos; }
__host__ __device__ constexpr int kg_bwd_slot(int ch) { return kg_fwd_slot(ch); }

// Fused forward k-space stage: x-FFT of rho_hat, x-inverse transforms of rho_hat
// times the channel weights, every channel stored (S0 -> SF).
//
// ky and kz are constant along an x line, so a weight's factors in ky and kz
// commute with the x transforms: with the closed forms of weights.hpp:91-184
// (w3, w2, aik2 = (w2 - 3 w3 / R) / k^2 radial) the 12 channels come from 7
// x-inverse transforms ("sources"):
//   0: w3 rho        -> eta, v_y = i ky (.), v_z = i kz (.)   (eta kept in LDS)
//   1: w2 rho        -> n2
//   2: mf rho        -> the mean-field channel
//   3: i kx w3 rho   -> v_x
//   4: aik2 kx^2 rho -> T_xx = (.) + eta / R
//   5: aik2 kx rho   -> T_xy = ky (.), T_xz = kz (.)
//   6: aik2 rho      -> T_yy = ky^2 (.) + eta / R, T_yz = ky kz (.), T_zz = kz^2 (.) + eta / R
// The LDS holds G slabs of C: G sources per batched transform in C, as many
// more in a narrower CV (the vector and tensor sources 3..6 use CV, the scalar
// ones C); with G = 2 the x extent has slot transforms (the host checks), and
// only those are compiled in.  With `red` SF receives the reduced set of the
// kg_zpoint point kernel.  EPT >= ty*nx/blockDim.x elements owned per thread.
template <typename S, typename C, typename CV, int EPT, int G>
__global__ void __launch_bounds__(2 * NT) k_xfwd(const cx<S>* __restrict__ S0, cx<S>* __restrict__ SF,
                                             const S* __restrict__ mf, const cx<double>* __restrict__ wt, Geom g,
                                             FftPlan px, const cx<C>* __restrict__ twx,
                                             const cx<CV>* __restrict__ twv, int ty, ChannelSet cs, int red) {
  const UDiv dty(ty);  // block-uniform divisors
  extern __shared__ unsigned char smem_raw[];
  cx<C>* buf = reinterpret_cast<cx<C>*>(smem_raw);
  const int nx = g.nx, ny = g.ny, ls = nx + 1;
  const int slab = ty * ls;
  cx<C>* keep = buf + G * slab;  // eta's x-inverse, for T_xx, T_yy and T_zz
  const int y0 = blockIdx.x * ty;
  const int kz = blockIdx.y;
  const long plane_off = static_cast<long>(kz) * nx * ny;
  const cx<S>* plane = S0 + plane_off;
  for (int e = threadIdx.x; e < ty * nx; e += blockDim.x) {
    const int x = dty.div(e), yl = e - x * ty, y = y0 + yl;
    buf[yl * ls + x] = (y < ny) ? cvt<C>(plane[x * ny + y]) : cx<C>{C(0), C(0)};
  }
  __syncthreads();
  if constexpr (G == 2) slot_lines_only<C>(buf, ty, ls, px, -1, twx);
  else fft_lines<C, EPT>(buf, ty, ls, px, -1, twx);
  cx<C> rh[EPT];
  WRad wr[EPT];
  const double kzv = g.dkz * kz;
  const double invR = 1.0 / g.R;
#pragma unroll
  for (int m = 0; m < EPT; ++m) {
    const int e = threadIdx.x + m * blockDim.x;
    rh[m] = {C(0), C(0)};
    wr[m] = {0.0, 0.0, 0.0};
    if (e < ty * nx) {
      const int x = dty.div(e), yl = e - x * ty, y = y0 + yl;
      rh[m] = buf[yl * ls + x];
      if (y < ny && cs.hs) {
        const double kx = kvec(x, nx, g.dkx), ky = kvec(y, ny, g.dky);
        wr[m] = wrad(wt[plane_off + x * ny + y], kx * kx + ky * ky + kzv * kzv, invR);
      }
    }
  }
  __syncthreads();
  // the block-uniform source list: 0..6, without 2 when there is no mean
  // field, only 2 when there is no hard-sphere term
  const int nsrc = cs.hs ? 6 + (cs.mf ? 1 : 0) : (cs.mf ? 1 : 0);
  auto src = [&](int i) { return cs.hs ? (cs.mf || i < 2 ? i : i + 1) : 2; };
  // channel c's array in SF: the 12-channel order, or with `red` the reduced
  // set of the kg_zpoint point kernel (no v_z, T_zz: oslot < 0)
  auto oslot = [&](int c) { return red ? kg_fwd_slot(c) : c; };
  const int mf_slot = cs.hs ? oslot(11) : 0;
  auto out_ch = [&](int c) { return SF + static_cast<long>(oslot(c)) * g.M + plane_off; };
  // A batch of ng sources from list position c0 in arithmetic CR: the LDS
  // buffer holds GR slabs of CR (G slabs of C).
  auto run_batch = [&](auto zero, const auto* __restrict__ twr, int c0, int ng) {
    using CR = decltype(zero);
    constexpr int GR = G * static_cast<int>(sizeof(cx<C>) / sizeof(cx<CR>));
    cx<CR>* bufr = reinterpret_cast<cx<CR>*>(smem_raw);
    for (int gi = 0; gi < ng; ++gi) {
      const int sc = src(c0 + gi);
#pragma unroll
      for (int m = 0; m < EPT; ++m) {
        const int e = threadIdx.x + m * blockDim.x;
        if (e < ty * nx) {
          const int x = dty.div(e), yl = e - x * ty, y = y0 + yl;
          const double kx = kvec(x, nx, g.dkx);
          const WRad w = wr[m];
          cx<double> wc;
          switch (sc) {
            case 0: wc = {w.w3, 0.0}; break;
            case 1: wc = {w.w2, 0.0}; break;
            case 2: {
              // mean-field table: FP64 values, or FP32 (value, residual) planes
              const long o = plane_off + x * ny + min(y, ny - 1);
              wc = {sizeof(S) == 8 ? static_cast<double>(mf[o])
                                   : static_cast<double>(mf[o]) + static_cast<double>(mf[g.M + o]),
                    0.0};
              break;
            }
            case 3: wc = {0.0, kx * w.w3}; break;
            case 4: wc = {w.aik2 * kx * kx, 0.0}; break;
            case 5: wc = {w.aik2 * kx, 0.0}; break;
            default: wc = {w.aik2, 0.0}; break;
          }
          bufr[gi * slab + yl * ls + x] = cvt<CR>(wmul(rh[m], wc));
        }
      }
    }
    __syncthreads();
    // register-resident slot transforms where the length has them (with 12
    // channel transforms per block they measured slower than the LDS passes,
    // 160^3: 502 against 308 us; with the 7 sources they are faster)
    if constexpr (G == 2) slot_lines_only<CR>(bufr, ng * ty, ls, px, +1, twr);  // x extent with slots (host)
    else fft_lines<CR, EPT * GR>(bufr, ng * ty, ls, px, +1, twr);
    for (int gi = 0; gi < ng; ++gi) {
      const int sc = src(c0 + gi);
      const cx<CR>* vb = bufr + gi * slab;
      for (int e = threadIdx.x; e < ty * nx; e += blockDim.x) {
        const int x = dty.div(e), yl = e - x * ty, y = y0 + yl;
        if (y >= ny) continue;
        const long o = static_cast<long>(x) * ny + y;
        const cx<C> v = cvt<C>(vb[yl * ls + x]);
        const C ky = static_cast<C>(kvec(y, ny, g.dky)), kzc = static_cast<C>(kzv), ir = static_cast<C>(invR);
        switch (sc) {
          case 0:
            keep[yl * ls + x] = v;
            out_ch(0)[o] = cvt<S>(v);
            out_ch(3)[o] = cvt<S>(cx<C>{-ky * v.im, ky * v.re});
            if (!red) out_ch(4)[o] = cvt<S>(cx<C>{-kzc * v.im, kzc * v.re});
            break;
          case 1: out_ch(1)[o] = cvt<S>(v); break;
          case 2: (SF + static_cast<long>(mf_slot) * g.M + plane_off)[o] = cvt<S>(v); break;
          case 3: out_ch(2)[o] = cvt<S>(v); break;
          case 4: {
            const cx<C> k = keep[yl * ls + x];
            out_ch(5)[o] = cvt<S>(cx<C>{v.re + ir * k.re, v.im + ir * k.im});
            break;
          }
          case 5:
            out_ch(6)[o] = cvt<S>(cx<C>{ky * v.re, ky * v.im});
            out_ch(7)[o] = cvt<S>(cx<C>{kzc * v.re, kzc * v.im});
            break;
          default: {
            const cx<C> k = keep[yl * ls + x];
            const C yy = ky * ky, yz = ky * kzc, zz = kzc * kzc;
            out_ch(8)[o] = cvt<S>(cx<C>{yy * v.re + ir * k.re, yy * v.im + ir * k.im});
            out_ch(9)[o] = cvt<S>(cx<C>{yz * v.re, yz * v.im});
            if (!red) out_ch(10)[o] = cvt<S>(cx<C>{zz * v.re + ir * k.re, zz * v.im + ir * k.im});
            break;
          }
        }
      }
    }
    __syncthreads();
  };
  // the scalar sources (eta, n2, mean field) in C; with CV narrower than C the
  // vector and tensor sources 3..6 in CV, more of them per batch
  const int nscal = (cs.hs && !std::is_same<C, CV>::value) ? (cs.mf ? 3 : 2) : nsrc;
  for (int c0 = 0; c0 < nscal; c0 += G) run_batch(C(0), twx, c0, min(G, nscal - c0));
  if constexpr (!std::is_same<C, CV>::value) {
    constexpr int GV = G * static_cast<int>(sizeof(cx<C>) / sizeof(cx<CV>));
    for (int c0 = nscal; c0 < nsrc; c0 += GV) run_batch(CV(0), twv, c0, min(GV, nsrc - c0));
  }
}

// Fused backward k-space stage: x-FFT of the derivative channels, sum of
// weight * channel with conj weights on the odd vector channels
// (hard_sphere.hpp:118-134), inverse x-FFT of the sum (SB -> SO).
//
// The conj weights are linear in (w3, w2, aik2) with factors in k, and ky, kz
// are constant along an x line, so the 11 channels are combined into 6 groups
// before the x transforms, and the sum is
//   S = w3 (X0 - i kx X1) + w2 X2 + aik2 (kx^2 X3 + kx X4 + X5),   Xq = x-FFT of
//   P0 = G_eta - i ky G_vy - i kz G_vz + (G_Txx + G_Tyy + G_Tzz) / R,  P1 = G_vx,  P2 = G_n2,
//   P3 = G_Txx,  P4 = ky G_Txy + kz G_Txz,  P5 = ky^2 G_Tyy + ky kz G_Tyz + kz^2 G_Tzz,
// accumulated as A = X0 - i kx X1, B = X2, C = kx^2 X3 + kx X4 + X5 and
// combined once per element with the radial factors.  G groups per batched
// transform; a group's loads combine its channels in registers.  With `red`
// SB holds kg_zpoint's nine channels (no v_z, T_zz: their terms are folded
// into eta, n2, T_xx and T_yy).
template <typename S, typename C, int EPT, int G>
__global__ void __launch_bounds__(2 * NT, (sizeof(C) == 4 ? 3 : 2)) k_xbwd(const cx<S>* __restrict__ SB, cx<S>* __restrict__ SO,
                                             const cx<double>* __restrict__ wt, Geom g, FftPlan px,
                                             const cx<C>* __restrict__ twx, int ty, int red) {
  constexpr int NG = 6;
  const UDiv dty(ty);  // block-uniform divisors
  extern __shared__ unsigned char smem_raw[];
  cx<C>* buf = reinterpret_cast<cx<C>*>(smem_raw);
  const int nx = g.nx, ny = g.ny, ls = nx + 1;
  const int slab = ty * ls;
  const int y0 = blockIdx.x * ty;
  const int kz = blockIdx.y;
  const long plane_off = static_cast<long>(kz) * nx * ny;
  const double kzv = g.dkz * kz;
  const double invR = 1.0 / g.R;
  cx<C> accA[EPT], accB[EPT], accC[EPT];
#pragma unroll
  for (int m = 0; m < EPT; ++m) accA[m] = accB[m] = accC[m] = {C(0), C(0)};
  auto in_ch = [&](int c, long o) { return cvt<C>(SB[static_cast<long>(c) * g.M + plane_off + o]); };
  for (int c0 = 0; c0 < NG; c0 += G) {
    const int ng = min(G, NG - c0);
    for (int e = threadIdx.x; e < ty * nx; e += blockDim.x) {
      const int x = dty.div(e), yl = e - x * ty, y = y0 + yl;
      const long o = static_cast<long>(x) * ny + y;
      const C ky = static_cast<C>(kvec(min(y, ny - 1), ny, g.dky)), kzc = static_cast<C>(kzv);
      const C ir = static_cast<C>(invR);
      for (int gi = 0; gi < ng; ++gi) {
        cx<C> p = {C(0), C(0)};
        if (y < ny && red) {
          // the nine channels of kg_zpoint (eta with the v_z term, n2, T_xx and
          // T_yy with the T_zz derivative): 0 eta, 1 n2, 2 v_x, 3 v_y, 4 T_xx,
          // 5 T_xy, 6 T_xz, 7 T_yy, 8 T_yz
          switch (c0 + gi) {  // block-uniform
            case 0: {
              const cx<C> ge = in_ch(0, o), gy = in_ch(3, o);
              const cx<C> t = cadd(in_ch(4, o), in_ch(7, o));
              p = {ge.re + ky * gy.im + ir * t.re, ge.im - ky * gy.re + ir * t.im};
              break;
            }
            case 1: p = in_ch(2, o); break;
            case 2: p = in_ch(1, o); break;
            case 3: p = in_ch(4, o); break;
            case 4: {
              const cx<C> a = in_ch(5, o), b = in_ch(6, o);
              p = {ky * a.re + kzc * b.re, ky * a.im + kzc * b.im};
              break;
            }
            default: {
              const cx<C> a = in_ch(7, o), b = in_ch(8, o);
              const C yy = ky * ky, yz = ky * kzc;
              p = {yy * a.re + yz * b.re, yy * a.im + yz * b.im};
              break;
            }
          }
        } else if (y < ny) {
          switch (c0 + gi) {  // block-uniform
            case 0: {
              const cx<C> ge = in_ch(0, o), gy = in_ch(3, o), gz = in_ch(4, o);
              const cx<C> t = cadd(cadd(in_ch(5, o), in_ch(8, o)), in_ch(10, o));
              p = {ge.re + ky * gy.im + kzc * gz.im + ir * t.re, ge.im - ky * gy.re - kzc * gz.re + ir * t.im};
              break;
            }
            case 1: p = in_ch(2, o); break;
            case 2: p = in_ch(1, o); break;
            case 3: p = in_ch(5, o); break;
            case 4: {
              const cx<C> a = in_ch(6, o), b = in_ch(7, o);
              p = {ky * a.re + kzc * b.re, ky * a.im + kzc * b.im};
              break;
            }
            default: {
              const cx<C> a = in_ch(8, o), b = in_ch(9, o), c = in_ch(10, o);
              const C yy = ky * ky, yz = ky * kzc, zz = kzc * kzc;
              p = {yy * a.re + yz * b.re + zz * c.re, yy * a.im + yz * b.im + zz * c.im};
              break;
            }
          }
        }
        buf[gi * slab + yl * ls + x] = p;
      }
    }
    __syncthreads();
    // the wider batches are launched for x extents with slot transforms only
    if constexpr (G > (sizeof(C) == 4 ? 2 : 1)) slot_lines_only<C>(buf, ng * ty, ls, px, -1, twx);
    else fft_lines<C, EPT * G>(buf, ng * ty, ls, px, -1, twx);
    for (int gi = 0; gi < ng; ++gi) {
      const int grp = c0 + gi;  // block-uniform
#pragma unroll
      for (int m = 0; m < EPT; ++m) {
        const int e = threadIdx.x + m * blockDim.x;
        if (e < ty * nx) {
          const int x = dty.div(e), yl = e - x * ty;
          const cx<C> X = buf[gi * slab + yl * ls + x];
          const C kx = static_cast<C>(kvec(x, nx, g.dkx));
          cx<C>& A = accA[m];
          cx<C>& Cs = accC[m];
          switch (grp) {
            case 0: A = cadd(A, X); break;
            case 1: A = {A.re + kx * X.im, A.im - kx * X.re}; break;  // -i kx X1
            case 2: accB[m] = cadd(accB[m], X); break;
            case 3: Cs = {Cs.re + (kx * kx) * X.re, Cs.im + (kx * kx) * X.im}; break;
            case 4: Cs = {Cs.re + kx * X.re, Cs.im + kx * X.im}; break;
            default: Cs = cadd(Cs, X); break;
          }
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int m = 0; m < EPT; ++m) {
    const int e = threadIdx.x + m * blockDim.x;
    if (e < ty * nx) {
      const int x = dty.div(e), yl = e - x * ty, y = y0 + yl;
      cx<C> s = {C(0), C(0)};
      if (y < ny) {
        // the radial factors of the element (FP64 table), in the path's precision
        const double kx = kvec(x, nx, g.dkx), ky = kvec(y, ny, g.dky);
        const WRad w = wrad(wt[plane_off + x * ny + y], kx * kx + ky * ky + kzv * kzv, invR);
        const C w3 = static_cast<C>(w.w3), w2 = static_cast<C>(w.w2), a2 = static_cast<C>(w.aik2);
        const cx<C> a = accA[m], b = accB[m], c = accC[m];
        s = {w3 * a.re + w2 * b.re + a2 * c.re, w3 * a.im + w2 * b.im + a2 * c.im};
      }
      buf[yl * ls + x] = s;
    }
  }
  __syncthreads();
  if constexpr (G > (sizeof(C) == 4 ? 2 : 1)) slot_lines_only<C>(buf, ty, ls, px, +1, twx);
  else fft_lines<C, EPT>(buf, ty, ls, px, +1, twx);
  cx<S>* outp = SO + plane_off;
  for (int e = threadIdx.x; e < ty * nx; e += blockDim.x) {
    const int x = dty.div(e), yl = e - x * ty, y = y0 + yl;
    if (y < ny) outp[x * ny + y] = cvt<S>(buf[yl * ls + x]);
  }
}

// ---------------------------------------------------------------- fused z-line pointwise kernel
// Per block of nlb consecutive z-lines (PPT * NT points): C2R of every forward
// channel (arithmetic CF, ZP_G channels per batched transform), the pointwise
// FMT Phi/dPhi (FP64), the ideal-gas, external-field and mean-field local
// terms, compensated partial sums, and R2C of the 11 combined derivative
// channels (arithmetic CB) written in place over SF's first channels.  NF
// forward channels (12: FMT + mean field, 11: FMT, 1: mean field, 0) and NB
// backward channels (11 or 0) are compile-time so the per-point values stay
// in registers.

constexpr int ZP_G = 4;  // channels per batched z transform in k_zpoint

template <typename S, typename CF, typename CB, int PPT, int NF, int NB>
__global__ void __launch_bounds__(NT) k_zpoint(cx<S>* SF, const double* __restrict__ rho_in,
                                               const double* __restrict__ vext, double* __restrict__ local_force,
                                               double* __restrict__ partials, Geom g, FftPlan pz,
                                               const cx<CF>* __restrict__ twf, const cx<CB>* __restrict__ twb,
                                               int nlb, FmtParams fp, double mu, AliasMap amap,
                                               const double* __restrict__ shift_sum, DcTable dc) {
  const UDiv dnlb(nlb), dnz(g.nz);  // block-uniform divisors
  constexpr bool HS = NF >= 11;
  constexpr bool MF = (NF == 12 || NF == 1);
  constexpr int MFC = HS ? 11 : 0;
  constexpr int NFA = NF > 0 ? NF : 1;
  extern __shared__ unsigned char smem_raw[];
  const int nz = g.nz, nzh = g.nzh;
  const int np = nlb >> 1;                 // packed line pairs per channel
  const int zs = np * nz;                  // complex elements per channel in the transform buffer
  constexpr size_t CMAX = sizeof(cx<CF>) > sizeof(cx<CB>) ? sizeof(cx<CF>) : sizeof(cx<CB>);
  cx<CF>* fbuf = reinterpret_cast<cx<CF>*>(smem_raw);
  cx<CB>* bbuf = reinterpret_cast<cx<CB>*>(smem_raw);
  cx<S>* stage = reinterpret_cast<cx<S>*>(smem_raw + CMAX * ZP_G * zs);
  const long l0 = static_cast<long>(blockIdx.x) * nlb;
  const int npts = nlb * nz;
  const int hs = nlb * nzh;                // staged half-spectrum elements per channel
  S wv[NFA][PPT];

// fully unrolled: wv is indexed by the channel, which must stay a constant
#pragma clang loop unroll(full)
  for (int c0 = 0; c0 < NF; c0 += ZP_G) {
    const int ng = min(ZP_G, NF - c0);
    for (int gi = 0; gi < ng; ++gi) {
      const cx<S>* Sc = SF + static_cast<long>(c0 + gi) * g.M;
      for (int e = threadIdx.x; e < hs; e += blockDim.x) {
        const int kz = dnlb.div(e), ll = e - kz * nlb;
        const long l = l0 + ll;
        stage[gi * hs + ll * nzh + kz] = (l < g.nxy) ? Sc[static_cast<long>(kz) * g.nxy + l] : cx<S>{S(0), S(0)};
      }
    }
    __syncthreads();
    c2r_lines<CF, S, ZP_G * PPT / 2>(stage, hs, fbuf, ng, np, nz, nzh, pz, twf);
#pragma unroll
    for (int gi = 0; gi < ZP_G; ++gi) {
      const int c = c0 + gi;
      if (c < NF) {
#pragma unroll
        for (int m = 0; m < PPT; ++m) {
          const int q = threadIdx.x + m * blockDim.x;
          S val = S(0);
          if (q < npts) {
            const int ll = dnz.div(q), z = q - ll * nz;
            const cx<CF> v = fbuf[gi * zs + (ll >> 1) * nz + z];
            val = static_cast<S>((ll & 1) ? v.im : v.re);
          }
          wv[c][m] = val;
        }
      }
    }
    __syncthreads();
  }

  // pointwise; one point at a time (rolled loop) so that only one point's FP64
  // temporaries are live; channel values move in and out with unrolled selects.
  double acc[5] = {0, 0, 0, 0, 0}, cmp[5] = {0, 0, 0, 0, 0};  // ideal, external, hs, mf, mass
  const double dv = g.dv;
  const double shift = shift_sum ? shift_sum[0] * g.inv_n : 0.0;
#pragma unroll 1
  for (int m = 0; m < PPT; ++m) {
    const int q = threadIdx.x + m * blockDim.x;
    const long l = l0 + dnz.div(q);
    S cur[NFA];
#pragma unroll
    for (int i = 0; i < NF; ++i) {
      S t = wv[i][0];
#pragma unroll
      for (int j = 1; j < PPT; ++j) t = (m == j) ? wv[i][j] : t;
      cur[i] = t;
    }
    S out[11];
#pragma unroll
    for (int i = 0; i < 11; ++i) out[i] = S(0);
    if (q < npts && l < g.nxy) {
      const long p = l0 * nz + q;
      const double rho = amap.apply(rho_in[p]);
      const double lr = flog(fmax(rho, 1e-300));
      double lf = (lr - mu) * dv;                // ideal_gas.hpp:27-33
      two_sum(acc[0], cmp[0], rho * (lr - 1.0));
      two_sum(acc[4], cmp[4], rho);
      if (vext) {                                 // external_field.hpp:16-31
        const double V = vext[p];
        lf += V * dv;
        two_sum(acc[1], cmp[1], rho * V);
      }
      if constexpr (MF) {                         // mean_field.hpp:259-266 (self pair)
        const double conv = static_cast<double>(cur[MFC]) + dc.v[MFC] * shift;
        lf += conv * dv;
        two_sum(acc[3], cmp[3], 0.5 * rho * conv);
      }
      if constexpr (HS) {
        double w[11], gg[11];
#pragma unroll
        for (int i = 0; i < 11; ++i) w[i] = static_cast<double>(cur[i]) + dc.v[i] * shift;
        const double phi = fmt_phi_grad(fp, w, gg);
        two_sum(acc[2], cmp[2], phi);
#pragma unroll
        for (int i = 0; i < 11; ++i) out[i] = static_cast<S>(gg[i]);
      }
      local_force[p] = lf;
    }
    if constexpr (HS) {
#pragma unroll
      for (int i = 0; i < 11; ++i)
#pragma unroll
        for (int j = 0; j < PPT; ++j) wv[i][j] = (m == j) ? out[i] : wv[i][j];
    }
  }
  {
    double tot[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) tot[i] = acc[i] + cmp[i];
    __shared__ double out5[5];
    block_reduce<5>(tot, out5);
    __syncthreads();
    if (threadIdx.x < 8)
      partials[static_cast<long>(blockIdx.x) * 8 + threadIdx.x] = threadIdx.x < 5 ? out5[threadIdx.x] * dv : 0.0;
  }

  // forward z R2C of the combined derivatives, ZP_G channels per batched transform
// fully unrolled: wv is indexed by the channel, which must stay a constant
#pragma clang loop unroll(full)
  for (int c0 = 0; c0 < NB; c0 += ZP_G) {
    const int ng = min(ZP_G, NB - c0);
#pragma unroll
    for (int gi = 0; gi < ZP_G; ++gi) {
      const int c = c0 + gi;
      if (c < NB) {
#pragma unroll
        for (int m = 0; m < PPT; ++m) {
          const int q = threadIdx.x + m * blockDim.x;
          if (q < npts) {
            const int ll = dnz.div(q), z = q - ll * nz;
            CB* dst = reinterpret_cast<CB*>(&bbuf[gi * zs + (ll >> 1) * nz + z]);
            dst[ll & 1] = static_cast<CB>(wv[c][m]);
          }
        }
      }
    }
    __syncthreads();
    fft_lines<CB, ZP_G * PPT / 2>(bbuf, ng * np, nz, pz, -1, twb);
    for (int gi = 0; gi < ng; ++gi) {
      cx<S>* outc = SF + static_cast<long>(c0 + gi) * g.M;
      for (int e = threadIdx.x; e < hs; e += blockDim.x) {
        const int kz = dnlb.div(e), ll = e - kz * nlb;
        const long l = l0 + ll;
        if (l >= g.nxy) continue;
        outc[static_cast<long>(kz) * g.nxy + l] =
            cvt<S>(unpack_pair(bbuf + gi * zs + (ll >> 1) * nz, kz, nz, ll & 1));
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- kg_zpoint
// The point kernel of the general schedule for z lengths 48, 80, 96 and 160,
// with register-resident z transforms in place of k_zpoint's block-wide LDS
// passes.  A block of W waves owns LINES consecutive z lines; wave w owns the
// LPW line pairs 2 (w LPW + lp) + {0, 1}.  A round transforms 4 channels: slot
// s = cr * LPW + lp of a wave (T lanes, E = L / T elements per lane) runs
// channel 4 R + cr of line pair lp, so every wave works on its own lines and
// the block synchronises only around the staging of each round:
//   forward round: the block loads the 4 channels' half spectra of its lines
//     (coalesced over the lines) into LDS; each slot builds its packed line
//     pair in registers, runs the C2R and leaves the lines in its exchange
//     buffer; each lane then reads its PPL points (line position f = lid + 64 j
//     of the wave's 2 LPW L values) of every channel of the round;
//   pointwise: Phi/dPhi, local terms and sums, as k_zpoint;
//   backward round: each lane writes its points of 4 derivative channels into
//     the slots' buffers; each slot runs the R2C of its line pair and unpacks
//     the two half spectra into LDS; the block stores them coalesced.
template <int L, typename S>
struct KgCfg {
  static constexpr int T = (L == 96 || L == 160) ? 16 : 8;
  static constexpr int E = L / T;
  static constexpr int NS = 64 / T;   // slots per wave
  static constexpr int LPW = NS / 4;  // line pairs per wave (4 channels per round)
  static constexpr int W = 4;
  static constexpr int NTH = 64 * W;
  static constexpr int LINES = 2 * LPW * W;
  static constexpr int NZH = L / 2 + 1;
  static constexpr int PPL = 2 * LPW * L / 64;  // points per lane
  static_assert(PPL * 64 == 2 * LPW * L, "the wave's points must divide over its lanes");
  static constexpr int stage_bytes = (4 * NZH * LINES * static_cast<int>(sizeof(cx<S>)) + 15) / 16 * 16;
  static constexpr int xb_bytes(int cmax) { return W * NS * L * cmax; }
};

template <typename S, typename CF, typename CB, int L, int NF, int NB>
__global__ void __launch_bounds__((KgCfg<L, S>::NTH), (sizeof(CF) > sizeof(CB) ? 3 : 2)) kg_zpoint(cx<S>* SF, const double* __restrict__ rho_in,
                                                              const double* __restrict__ vext,
                                                              double* __restrict__ local_force,
                                                              double* __restrict__ partials, Geom g,
                                                              const cx<CF>* __restrict__ twf,
                                                              const cx<CB>* __restrict__ twb, FmtParams fp, double mu,
                                                              AliasMap amap, const double* __restrict__ shift_sum,
                                                              DcTable dc) {
  using K = KgCfg<L, S>;
  constexpr int T = K::T, E = K::E, LPW = K::LPW, LINES = K::LINES, NZH = K::NZH, PPL = K::PPL;
  constexpr bool HS = NF >= 11;
  constexpr bool MF = (NF == 12 || NF == 1);
  constexpr int MFC = HS ? 11 : 0;
  constexpr int NFA = NF > 0 ? NF : 1;
  // Slot buffers hold CB values.  With CF wider (mixed precision) round 0's CF
  // transforms use the staging area and the slot buffers together: wave w's
  // slot s at (w NS + s) L sizeof(cx<CF>), once every wave has read its round-0
  // inputs from the staging area (three 4-wave blocks per CU instead of two).
  static_assert(sizeof(CB) <= sizeof(CF), "the backward arithmetic is not wider than the forward one");
  constexpr int XBE = sizeof(cx<CB>);
  static_assert(K::W * K::NS * L * static_cast<int>(sizeof(cx<CF>)) <= K::stage_bytes + K::xb_bytes(XBE),
                "round 0's buffers fit the staging area and the slot buffers");
  // With the hard-sphere channels SF holds the reduced set (kg_fwd_slot,
  // kg_bwd_slot): T_zz follows from the trace identity at the points, n_v,z
  // from eta's staged spectrum, and their derivatives fold into the others.
  constexpr int NPF = HS ? NF - 1 : NF, NPB = HS && NB > 0 ? 10 : NB;
  constexpr int NRF = (NPF + 3) / 4, NRB = (NPB + 3) / 4;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<S>* stage = reinterpret_cast<cx<S>*>(smem_raw);  // [cr][kz][line]
  const int tid = static_cast<int>(threadIdx.x);
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lid = tid & 63;
  const int slot = lid / T, lane = lid % T;
  const int cr = slot / LPW, lp = slot % LPW;
  unsigned char* xwave = smem_raw + K::stage_bytes + wave * K::NS * L * XBE;
  // slot buffer of slot s of this wave (L complex CB values), and its round-0 CF buffer
  auto sbuf = [&](int s) { return xwave + s * L * XBE; };
  auto wbuf = [&](int s) { return smem_raw + (wave * K::NS + s) * L * static_cast<int>(sizeof(cx<CF>)); };
  const long l0 = static_cast<long>(blockIdx.x) * LINES;
  const long nxy = g.nxy;
  const int la = 2 * (wave * LPW + lp);  // the slot's line pair, block-local
  S wv[NFA][PPL];

  // ---- forward rounds.  Round R + 1's half spectra are loaded into registers
  // before round R's transforms, so their memory latency overlaps that work.
  // Round 0 holds the scalar channels eta, n2 (and the mean field) and runs in
  // CF; the vector and tensor channels of the later rounds run in CB (FP32 in
  // mixed precision, as in the fused schedule).
  auto fch = [](int pos) { return HS ? kg_fwd_channel(NF, pos) : pos; };               // channel (wv index)
  auto fsf = [](int pos) { return HS ? kg_fwd_slot(kg_fwd_channel(NF, pos)) : pos; };  // SF slot, -1: from eta
  constexpr int NSTG = 4 * NZH * LINES, KPF = (NSTG + K::NTH - 1) / K::NTH;
  cx<S> pf[KPF];
  auto load_round = [&](int R) {
#pragma unroll
    for (int k = 0; k < KPF; ++k) {
      const int idx = tid + K::NTH * k;
      const int c4 = idx / (NZH * LINES), rem = idx - c4 * (NZH * LINES);
      const int kz = rem / LINES, ln = rem - kz * LINES;
      const int pos = 4 * R + c4;
      pf[k] = {S(0), S(0)};
      if (idx < NSTG && pos < NPF && fsf(pos) >= 0 && l0 + ln < nxy)
        pf[k] = SF[static_cast<long>(fsf(pos)) * g.M + static_cast<long>(kz) * nxy + l0 + ln];
    }
  };
  auto store_round = [&]() {
#pragma unroll
    for (int k = 0; k < KPF; ++k) {
      const int idx = tid + K::NTH * k;
      if (idx < NSTG) stage[idx] = pf[k];
    }
  };
  // one round's C2R in arithmetic CR: slot (cr, lp) transforms position 4 R + cr
  // of its line pair; then every lane reads its points of the round's channels
  auto fwd_round = [&](auto zero, const auto* __restrict__ tw, int R) {
    using CR = decltype(zero);
    constexpr bool WIDE = sizeof(cx<CR>) > XBE;  // round 0 in CF, buffers over the staging area
    auto rbuf = [&](int s) { return WIDE ? wbuf(s) : sbuf(s); };
    const int pos = 4 * R + cr;
    cx<CR>* xb = reinterpret_cast<cx<CR>*>(rbuf(slot));
    cx<CR> v[E];
    if (pos < NPF) {
      // n_v,z = i kz eta: eta's staged spectrum (position 0 of round 0) times i kz
      const bool vz = fsf(pos) < 0;
      const cx<S>* st = stage + (vz ? 0 : cr) * NZH * LINES;
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const int k = lane + T * q;
        const bool mirror = k >= NZH;
        const int kk = mirror ? L - k : k;
        cx<CR> A = cvt<CR>(st[kk * LINES + la]), B = cvt<CR>(st[kk * LINES + la + 1]);
        if (vz) {
          const CR kzp = static_cast<CR>(g.dkz * kk);
          A = {-kzp * A.im, kzp * A.re};
          B = {-kzp * B.im, kzp * B.re};
        }
        if (kk == 0 || 2 * kk == L) {
          A.im = CR(0);
          B.im = CR(0);
        }
        if (mirror) {
          A.im = -A.im;
          B.im = -B.im;
        }
        v[q] = {A.re - B.im, A.im + B.re};
      }
    }
    if constexpr (WIDE) __syncthreads();  // every wave has read its inputs from the staging area
    if (pos < NPF) {
      slot_fft<CR, L, T, +1, false, false>(v, xb, lane, tw);
      wave_sync();
#pragma unroll
      for (int e = 0; e < E; ++e) xb[lane + T * e] = v[e];
    }
    wave_sync();
    // the lane's points of the round's channels: position f = lid + 64 j of the
    // wave's line pairs (line pair f / 2L, z = (f % 2L) / 2, line (f % 2L) & 1)
#pragma unroll
    for (int c4 = 0; c4 < 4; ++c4) {
      if (4 * R + c4 < NPF) {
        const int ch = fch(4 * R + c4);
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
          const int f = lid + 64 * j, pl = f / (2 * L), r2 = f - pl * 2 * L;
          const CR* src = reinterpret_cast<const CR*>(rbuf(c4 * LPW + pl));
          wv[ch][j] = static_cast<S>(src[r2]);
        }
      }
    }
    wave_sync();  // the buffers are rewritten by the next round
  };
  if constexpr (NRF > 0) {
    load_round(0);
    store_round();
    __syncthreads();
  }
#pragma unroll
  for (int R = 0; R < NRF; ++R) {
    if (R + 1 < NRF) load_round(R + 1);
    if (R == 0) fwd_round(CF(0), twf, R);
    else fwd_round(CB(0), twb, R);
    if (R + 1 < NRF) {
      __syncthreads();  // every slot is done with this round's staging
      store_round();
      __syncthreads();
    }
  }

  // ---- pointwise (as k_zpoint)
  double acc[5] = {0, 0, 0, 0, 0}, cmp[5] = {0, 0, 0, 0, 0};  // ideal, external, hs, mf, mass
  const double dv = g.dv;
  const double shift = shift_sum ? shift_sum[0] * g.inv_n : 0.0;
#pragma unroll
  for (int j = 0; j < PPL; ++j) {
    const int f = lid + 64 * j, pl = f / (2 * L), r2 = f - pl * 2 * L;
    const int z = r2 >> 1, line = 2 * (wave * LPW + pl) + (r2 & 1);
    const long l = l0 + line;
    S out[11];
#pragma unroll
    for (int i = 0; i < 11; ++i) out[i] = S(0);
    if (l < nxy) {
      const long p = l * L + z;
      const double rho = amap.apply(rho_in[p]);
      const double lr = flog(fmax(rho, 1e-300));
      double lfv = (lr - mu) * dv;  // ideal_gas.hpp:27-33
      two_sum(acc[0], cmp[0], rho * (lr - 1.0));
      two_sum(acc[4], cmp[4], rho);
      if (vext) {  // external_field.hpp:16-31
        const double V = vext[p];
        lfv += V * dv;
        two_sum(acc[1], cmp[1], rho * V);
      }
      if constexpr (MF) {  // mean_field.hpp:259-266 (self pair)
        const double conv = static_cast<double>(wv[MFC][j]) + dc.v[MFC] * shift;
        lfv += conv * dv;
        two_sum(acc[3], cmp[3], 0.5 * rho * conv);
      }
      if constexpr (HS) {
        double w[11], gg[11];
#pragma unroll
        for (int i = 0; i < 10; ++i) w[i] = static_cast<double>(wv[i][j]) + dc.v[i] * shift;
        w[10] = w[1] - w[5] - w[8];  // T_zz = n2 - T_xx - T_yy (weights.hpp:121-137,170-178)
        const double phi = fmt_phi_grad(fp, w, gg);
        two_sum(acc[2], cmp[2], phi);
        // its derivative folds into n2, T_xx and T_yy
        gg[1] += gg[10];
        gg[5] -= gg[10];
        gg[8] -= gg[10];
#pragma unroll
        for (int i = 0; i < 10; ++i) out[i] = static_cast<S>(gg[i]);
      }
      local_force[p] = lfv;
    }
    if constexpr (HS) {
#pragma unroll
      for (int i = 0; i < 11; ++i) wv[i][j] = out[i];
    }
  }
  {
    double tot[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) tot[i] = acc[i] + cmp[i];
    __shared__ double out5[5];
    block_reduce<5>(tot, out5);
    __syncthreads();
    if (tid < 8) partials[static_cast<long>(blockIdx.x) * 8 + tid] = tid < 5 ? out5[tid] * dv : 0.0;
  }

  // ---- backward rounds: R2C of the combined derivatives, written in place over SF's first
  // channels.  Positions: eta, n_v,z, n2, v_x, v_y, T_xx, T_xy, T_xz, T_yy, T_yz (kg_bwd_slot);
  // n_v,z's spectrum times -i kz joins eta's when they are stored.
  auto bch = [](int pos) { return kg_bwd_channel(pos); };
#pragma unroll
  for (int R = 0; R < NRB; ++R) {
#pragma unroll
    for (int c4 = 0; c4 < 4; ++c4) {
      if (4 * R + c4 < NPB) {
        const int ch = bch(4 * R + c4);
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
          const int f = lid + 64 * j, pl = f / (2 * L), r2 = f - pl * 2 * L;
          CB* dst = reinterpret_cast<CB*>(sbuf(c4 * LPW + pl));
          dst[r2] = static_cast<CB>(wv[ch][j]);
        }
      }
    }
    wave_sync();
    if (R > 0) __syncthreads();  // the previous round's store is done with the staging buffer
    const int c = 4 * R + cr;
    cx<CB>* xb = reinterpret_cast<cx<CB>*>(sbuf(slot));
    if (c < NPB) {
      cx<CB> v[E];
#pragma unroll
      for (int q = 0; q < E; ++q) v[q] = xb[lane + T * q];
      slot_fft<CB, L, T, -1, false, false>(v, xb, lane, twb);
      wave_sync();
#pragma unroll
      for (int e = 0; e < E; ++e) xb[lane + T * e] = v[e];
      wave_sync();
      // unpack the two real lines' half spectra (unpack_pair) into the staging buffer
      cx<S>* st = stage + cr * NZH * LINES;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int k = lane + T * e;
        if (k < NZH) {
          const cx<CB> zk = xb[k], zm0 = xb[k == 0 ? 0 : L - k];
          const cx<CB> zm = {zm0.re, -zm0.im};
          st[k * LINES + la] = cvt<S>(cx<CB>{CB(0.5) * (zk.re + zm.re), CB(0.5) * (zk.im + zm.im)});
          st[k * LINES + la + 1] = cvt<S>(cx<CB>{CB(0.5) * (zk.im - zm.im), CB(-0.5) * (zk.re - zm.re)});
        }
      }
    }
    __syncthreads();
    const int nc = min(4, NPB - 4 * R);
    for (int idx = tid; idx < nc * NZH * LINES; idx += K::NTH) {
      const int c4 = idx / (NZH * LINES), rem = idx - c4 * (NZH * LINES);
      const int kz = rem / LINES, ln = rem - kz * LINES;
      const int sl = kg_bwd_slot(bch(4 * R + c4));
      if (l0 + ln < nxy && sl >= 0) {
        cx<S> v = stage[idx];
        if (R == 0 && c4 == 0) {  // eta + (-i kz) n_v,z (position 1)
          const cx<S> z = stage[idx + NZH * LINES];
          const S kzp = static_cast<S>(g.dkz * kz);
          v = {v.re + kzp * z.im, v.im - kzp * z.re};
        }
        SF[static_cast<long>(sl) * g.M + static_cast<long>(kz) * nxy + l0 + ln] = v;
      }
    }
  }
}

// Deterministic compensated sum of per-block partials (nblk x 8 doubles) into res[0..7]:
// strided Neumaier sums per thread, then a fixed-shape tree of (sum, compensation) pairs.
// Deterministic compensated column sums of `nrows` rows of 8 partial sums by
// one block (reduce_rows).
constexpr int kFinThreads = 1024;
__global__ void __launch_bounds__(kFinThreads) k_finalize(const double* __restrict__ partials, int nrows,
                                                          double* __restrict__ res) {
  __shared__ double sh[kFinThreads];
  __shared__ double shc[kFinThreads];
  reduce_rows<kFinThreads>(partials, nrows, res, sh, shc);
}

// ---------------------------------------------------------------- launchers
// z-lines per block of the plain z kernels: up to 2048 complex elements (two
// real lines per element) per transform buffer.
template <typename C>
static int zlines_per_block(const FftPlan& pz) {
  if (pz.lines > 0) return pz.lines;
  int nlb = (sizeof(C) == 4 ? 4096 : 2048) / pz.L;
  if (nlb < 2) nlb = 2;
  if (nlb & 1) --nlb;
  if (nlb > 64) nlb = 64;
  return nlb;
}

template <typename S, typename C>
void launch_zr2c(const double* in, cx<S>* out, const Geom& g, const FftPlan& pz, const cx<C>* twz, AliasMap amap,
                 const double* shift_sum, hipStream_t st, const double* sum_rows, int sum_n, double* sum_out) {
  const int nlb = zlines_per_block<C>(pz);
  if ((nlb >> 1) * g.nz > 8 * NT) throw std::runtime_error("z extent too large for the z transform kernel");
  const unsigned nblk = static_cast<unsigned>((g.nxy + nlb - 1) / nlb);
  const size_t sh = sizeof(cx<C>) * (nlb >> 1) * g.nz;
  hipLaunchKernelGGL((k_zr2c<S, C>), dim3(nblk), dim3(NT), sh, st, in, out, g, pz, twz, nlb, amap, shift_sum,
                     sum_rows, sum_n, sum_out);
}

template <typename S, typename C>
void launch_zr2c_fields(const double* in, int nfields, cx<S>* out, const Geom& g, const FftPlan& pz,
                        const cx<C>* twz, hipStream_t st) {
  const int nlb = zlines_per_block<C>(pz);
  if ((nlb >> 1) * g.nz > 8 * NT) throw std::runtime_error("z extent too large for the z transform kernel");
  if (nfields < 1 || nfields > 65535) throw std::runtime_error("field count out of range");
  const unsigned nblk = static_cast<unsigned>((g.nxy + nlb - 1) / nlb);
  const size_t sh = sizeof(cx<C>) * (nlb >> 1) * g.nz;
  const AliasMap id{0, 0.0, 0.0};
  hipLaunchKernelGGL((k_zr2c<S, C>), dim3(nblk, nfields), dim3(NT), sh, st, in, out, g, pz, twz, nlb, id,
                     static_cast<const double*>(nullptr), static_cast<const double*>(nullptr), 0,
                     static_cast<double*>(nullptr));
}

template <typename S, typename C>
void launch_zc2r(const cx<S>* Sp, int nch, double* out, const double* add, double scale, const Geom& g,
                 const FftPlan& pz, const cx<C>* twz, const double* shift_sum, DcTable dc, hipStream_t st,
                 const double* fin_rows, int fin_n, double* fin_res) {
  const int nlb = zlines_per_block<C>(pz);
  if ((nlb >> 1) * g.nz > 8 * NT) throw std::runtime_error("z extent too large for the z transform kernel");
  const unsigned nblk = static_cast<unsigned>((g.nxy + nlb - 1) / nlb);
  const size_t sh = sizeof(cx<C>) * (nlb >> 1) * g.nz + sizeof(cx<S>) * nlb * g.nzh;
  hipLaunchKernelGGL((k_zc2r<S, C>), dim3(nblk, nch), dim3(NT), sh, st, Sp, out, add, scale, g, pz, twz, nlb,
                     shift_sum, dc, fin_rows, fin_n, fin_res);
}

int sum_blocks(const Geom& g) {
  long b = (g.N + NT * kSumUnroll - 1) / (NT * kSumUnroll);
  if (b > 1024) b = 1024;
  if (b < 1) b = 1;
  return static_cast<int>(b);
}

void launch_sum(const double* in, AliasMap amap, double* partials, const Geom& g, hipStream_t st) {
  hipLaunchKernelGGL(k_sum, dim3(sum_blocks(g)), dim3(NT), 0, st, in, amap, g.N, partials);
}

template <typename S, typename C>
void launch_lines(cx<S>* Sp, long nlines, const FftPlan& p, const cx<C>* tw, int sg, hipStream_t st) {
  // one line per slot by default (every slot busy); the LDS-pass lengths take
  // at most 8 * NT elements per block
  const int T = slot_lanes(p.L);
  const int nlb = p.lines > 0 ? p.lines : (T ? NT / T : 2048 / p.L);
  if (nlb < 1 || (!T && nlb * p.L > 8 * NT))
    throw std::runtime_error("line length too large for the line transform kernel");
  const unsigned nblk = static_cast<unsigned>((nlines + nlb - 1) / nlb);
  const size_t sh = sizeof(cx<C>) * nlb * p.L;
  hipLaunchKernelGGL((k_lines<S, C>), dim3(nblk), dim3(NT), sh, st, Sp, nlines, p, tw, nlb, sg);
}

void launch_lines_split(cx<float>* Sp, long nlines, const FftPlan& p, const cx<double>* twh, const cx<float>* twl,
                        long lo0, long lo1, int sg, hipStream_t st) {
  const int T = slot_lanes(p.L);
  const int nlb = p.lines > 0 ? p.lines : (T ? NT / T : 2048 / p.L);
  if (nlb < 1 || (!T && nlb * p.L > 8 * NT))
    throw std::runtime_error("line length too large for the line transform kernel");
  // FP64 blocks take half as many lines, in the same LDS
  const int hb = std::max(1, nlb / 2);
  const long na = (lo0 + hb - 1) / hb, nb = (lo1 - lo0 + nlb - 1) / nlb, nc = (nlines - lo1 + hb - 1) / hb;
  const unsigned nblk = static_cast<unsigned>(na + nb + nc);
  const size_t sh = std::max(sizeof(cx<float>) * nlb, sizeof(cx<double>) * hb) * p.L;
  hipLaunchKernelGGL((k_lines_split<float, double, float>), dim3(nblk), dim3(NT), sh, st, Sp, nlines, p, twh, twl,
                     nlb, hb, sg, lo0, lo1, static_cast<int>(na), static_cast<int>(nb));
}

// Dynamic LDS above 64 KiB is enabled per kernel instantiation, once per process.
template <typename K>
static void kg_smem(K kernel, int bytes) {
  static std::mutex mu;
  static std::vector<const void*> done;
  const void* key = reinterpret_cast<const void*>(kernel);
  std::lock_guard<std::mutex> lk(mu);
  if (std::find(done.begin(), done.end(), key) != done.end()) return;
  if (hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
    throw std::runtime_error("hipFuncSetAttribute(MaxDynamicSharedMemorySize) failed");
  done.push_back(key);
}

static constexpr int kXTile = 16;  // y columns per x-line block (128-byte rows in FP32)
static constexpr size_t kLdsBytes = 160 * 1024;  // LDS of one workgroup on gfx950

template <typename S, typename C>
void launch_xlines(cx<S>* Sp, int nch, const Geom& g, const FftPlan& px, const cx<C>* twx, int sg, hipStream_t st) {
  const int ty = px.tile > 0 ? px.tile : kXTile;
  const size_t sh = sizeof(cx<C>) * ty * (g.nx + 1);
  dim3 grid((g.ny + ty - 1) / ty, g.nzh, nch);
  // up to 8 elements per thread: 256 threads per block, or 512 for larger
  // tiles of the LDS-pass lengths (twice the waves in flight per CU)
  if (ty * g.nx <= 8 * NT)
    hipLaunchKernelGGL((k_xlines<S, C, 8>), grid, dim3(NT), sh, st, Sp, g, px, twx, ty, sg);
  else if (ty * g.nx <= 16 * NT && slot_lanes(g.nx) == 0)
    hipLaunchKernelGGL((k_xlines<S, C, 8, 2 * NT>), grid, dim3(2 * NT), sh, st, Sp, g, px, twx, ty, sg);
  else if (ty * g.nx <= 16 * NT)
    hipLaunchKernelGGL((k_xlines<S, C, 16>), grid, dim3(NT), sh, st, Sp, g, px, twx, ty, sg);
  else
    throw std::runtime_error("x extent too large for the x transform kernel");
}

template <typename S, typename C, typename CV>
void launch_xfwd(const cx<S>* S0, cx<S>* SF, const S* mf, const cx<double>* wt, const Geom& g, const FftPlan& px,
                 const cx<C>* twx, const cx<CV>* twv, ChannelSet cs, bool red, hipStream_t st) {
  const int rd = red && cs.hs ? 1 : 0;
  const int ty = px.tile > 0 ? px.tile : kXTile;
  // sources per batched transform: two slabs of C (the vector and tensor
  // sources of a narrower CV four) where they fit the LDS, else one
  auto lds = [&](int G) { return sizeof(cx<C>) * (G + (cs.hs ? 1 : 0)) * ty * (g.nx + 1); };
  const bool g2 = lds(2) <= kLdsBytes && slot_lanes(g.nx) != 0;  // the two-slab form has slot transforms only
  const size_t sh = lds(g2 ? 2 : 1);
  if (sh > kLdsBytes) throw std::runtime_error("x tile too large for the LDS of the forward k-space kernel");
  dim3 grid((g.ny + ty - 1) / ty, g.nzh);
  // up to 8 elements per thread: 256 threads per block, or 512 for larger tiles
  if (ty * g.nx > 16 * NT) throw std::runtime_error("x extent too large for the fused k-space kernel");
  const int nth = ty * g.nx <= 8 * NT ? NT : 2 * NT;
  if (g2) {
    kg_smem(k_xfwd<S, C, CV, 8, 2>, static_cast<int>(kLdsBytes));
    hipLaunchKernelGGL((k_xfwd<S, C, CV, 8, 2>), grid, dim3(nth), sh, st, S0, SF, mf, wt, g, px, twx, twv, ty, cs,
                       rd);
  } else {
    kg_smem(k_xfwd<S, C, CV, 8, 1>, static_cast<int>(kLdsBytes));
    hipLaunchKernelGGL((k_xfwd<S, C, CV, 8, 1>), grid, dim3(nth), sh, st, S0, SF, mf, wt, g, px, twx, twv, ty, cs,
                       rd);
  }
}

template <typename S, typename C>
void launch_xbwd(const cx<S>* SB, cx<S>* SO, const cx<double>* wt, const Geom& g, const FftPlan& px,
                 const cx<C>* twx, int nb, hipStream_t st) {
  if (nb != 11 && nb != 9)
    throw std::runtime_error("the backward x kernel takes the 11 hard-sphere derivative channels or kg_zpoint's 9");
  const int red = nb == 9 ? 1 : 0;
  const int ty = px.tile > 0 ? px.tile : kXTile;
  // groups per batched transform: GW (3 of the 6 in FP32, 2 in FP64) for x
  // extents with slot transforms where the LDS allows, else GN
  constexpr int GN = sizeof(C) == 4 ? 2 : 1, GW = sizeof(C) == 4 ? 3 : 2;
  auto lds = [&](int G) { return sizeof(cx<C>) * G * ty * (g.nx + 1); };
  const bool wide = slot_lanes(g.nx) != 0 && lds(GW) <= kLdsBytes;
  const size_t sh = lds(wide ? GW : GN);
  if (sh > kLdsBytes) throw std::runtime_error("x tile too large for the LDS of the backward k-space kernel");
  dim3 grid((g.ny + ty - 1) / ty, g.nzh);
  // up to 8 elements per thread: 256 threads per block, or 512 for larger tiles
  if (ty * g.nx > 16 * NT) throw std::runtime_error("x extent too large for the fused k-space kernel");
  const int nth = ty * g.nx <= 8 * NT ? NT : 2 * NT;
  if (wide) {
    kg_smem(k_xbwd<S, C, 8, GW>, static_cast<int>(kLdsBytes));
    hipLaunchKernelGGL((k_xbwd<S, C, 8, GW>), grid, dim3(nth), sh, st, SB, SO, wt, g, px, twx, ty, red);
  } else {
    kg_smem(k_xbwd<S, C, 8, GN>, static_cast<int>(kLdsBytes));
    hipLaunchKernelGGL((k_xbwd<S, C, 8, GN>), grid, dim3(nth), sh, st, SB, SO, wt, g, px, twx, ty, red);
  }
}

// z-lines per block of the fused pointwise kernel: PPT*NT points (4 per thread
// in FP32 storage, 2 in FP64).
template <typename S>
static int zpoint_nlb(const FftPlan& pz) {
  if (pz.tile > 0) return pz.tile;
  int nlb = (sizeof(S) == 4 ? 4 : 2) * NT / pz.L;
  if (nlb < 2) nlb = 2;
  if (nlb & 1) --nlb;
  if (nlb > 64) nlb = 64;
  return nlb;
}

// z lengths with the register-resident point kernel (kg_zpoint)
static bool kg_length(int L) { return L == 48 || L == 80 || L == 96 || L == 160; }
bool zpoint_reduced(int nz, const ChannelSet& cs) { return cs.hs && kg_length(nz); }
template <typename S>
static int kg_lines(int L) {
  switch (L) {
    case 48: return KgCfg<48, S>::LINES;
    case 80: return KgCfg<80, S>::LINES;
    case 96: return KgCfg<96, S>::LINES;
    default: return KgCfg<160, S>::LINES;
  }
}

template <typename S>
int zpoint_blocks(const Geom& g, const FftPlan& pz) {
  const int nlb = kg_length(pz.L) ? kg_lines<S>(pz.L) : zpoint_nlb<S>(pz);
  return static_cast<int>((g.nxy + nlb - 1) / nlb);
}
template int zpoint_blocks<float>(const Geom&, const FftPlan&);
template int zpoint_blocks<double>(const Geom&, const FftPlan&);

template <typename S, typename CF, typename CB, int PPT, int NF, int NB>
static void zpoint_go(unsigned nblk, size_t sh, hipStream_t st, cx<S>* SF, const double* rho, const double* vext,
                      double* lf, double* partials, const Geom& g, const FftPlan& pz, const cx<CF>* twf,
                      const cx<CB>* twb, int nlb, FmtParams fp, double mu, AliasMap amap, const double* shift_sum,
                      DcTable dc) {
  hipLaunchKernelGGL((k_zpoint<S, CF, CB, PPT, NF, NB>), dim3(nblk), dim3(NT), sh, st, SF, rho, vext, lf, partials,
                     g, pz, twf, twb, nlb, fp, mu, amap, shift_sum, dc);
}

template <typename S, typename CF, typename CB, int L, int NF, int NB>
static void kg_go(hipStream_t st, cx<S>* SF, const double* rho, const double* vext, double* lf, double* partials,
                  const Geom& g, const cx<CF>* twf, const cx<CB>* twb, FmtParams fp, double mu, AliasMap amap,
                  const double* shift_sum, DcTable dc) {
  using K = KgCfg<L, S>;
  const int sh = K::stage_bytes + K::xb_bytes(static_cast<int>(sizeof(cx<CB>)));  // slot buffers of CB values
  auto k = kg_zpoint<S, CF, CB, L, NF, NB>;
  kg_smem(k, sh);
  const unsigned nblk = static_cast<unsigned>((g.nxy + K::LINES - 1) / K::LINES);
  hipLaunchKernelGGL(k, dim3(nblk), dim3(K::NTH), sh, st, SF, rho, vext, lf, partials, g, twf, twb, fp, mu, amap,
                     shift_sum, dc);
}

template <typename S, typename CF, typename CB, int L>
static void kg_launch(cx<S>* SF, const double* rho, const double* vext, double* lf, double* partials,
                      const Geom& g, const cx<CF>* twf, const cx<CB>* twb, ChannelSet cs, FmtParams fp, double mu,
                      AliasMap amap, const double* shift_sum, DcTable dc, hipStream_t st) {
#define CDFT_KG(NF, NB) kg_go<S, CF, CB, L, NF, NB>(st, SF, rho, vext, lf, partials, g, twf, twb, fp, mu, amap, shift_sum, dc)
  if (cs.nf == 12 && cs.nb == 11) CDFT_KG(12, 11);
  else if (cs.nf == 12 && cs.nb == 0) CDFT_KG(12, 0);
  else if (cs.nf == 11 && cs.nb == 11) CDFT_KG(11, 11);
  else if (cs.nf == 11 && cs.nb == 0) CDFT_KG(11, 0);
  else if (cs.nf == 1 && cs.nb == 0) CDFT_KG(1, 0);
  else if (cs.nf == 0 && cs.nb == 0) CDFT_KG(0, 0);
  else throw std::runtime_error("unsupported channel configuration");
#undef CDFT_KG
}

template <typename S, typename CF, typename CB>
void launch_zpoint(cx<S>* SF, const double* rho, const double* vext, double* local_force, double* partials,
                   const Geom& g, const FftPlan& pz, const cx<CF>* twf, const cx<CB>* twb, ChannelSet cs,
                   FmtParams fp, double mu, AliasMap amap, const double* shift_sum, DcTable dc, hipStream_t st) {
  if (kg_length(pz.L)) {
    switch (pz.L) {
      case 48: kg_launch<S, CF, CB, 48>(SF, rho, vext, local_force, partials, g, twf, twb, cs, fp, mu, amap, shift_sum, dc, st); return;
      case 80: kg_launch<S, CF, CB, 80>(SF, rho, vext, local_force, partials, g, twf, twb, cs, fp, mu, amap, shift_sum, dc, st); return;
      case 96: kg_launch<S, CF, CB, 96>(SF, rho, vext, local_force, partials, g, twf, twb, cs, fp, mu, amap, shift_sum, dc, st); return;
      default: kg_launch<S, CF, CB, 160>(SF, rho, vext, local_force, partials, g, twf, twb, cs, fp, mu, amap, shift_sum, dc, st); return;
    }
  }
  constexpr int PPT = sizeof(S) == 4 ? 4 : 2;
  const int nlb = zpoint_nlb<S>(pz);
  if (nlb * g.nz > PPT * NT) throw std::runtime_error("z extent too large for the fused pointwise kernel");
  const unsigned nblk = static_cast<unsigned>((g.nxy + nlb - 1) / nlb);
  const int zs = (nlb >> 1) * g.nz;
  const size_t sh = std::max(sizeof(cx<CF>), sizeof(cx<CB>)) * ZP_G * zs + sizeof(cx<S>) * ZP_G * nlb * g.nzh;
#define CDFT_ZP(NF, NB)                                                                                         \
  zpoint_go<S, CF, CB, PPT, NF, NB>(nblk, sh, st, SF, rho, vext, local_force, partials, g, pz, twf, twb, nlb, fp, \
                                    mu, amap, shift_sum, dc)
  if (cs.nf == 12 && cs.nb == 11) CDFT_ZP(12, 11);
  else if (cs.nf == 12 && cs.nb == 0) CDFT_ZP(12, 0);
  else if (cs.nf == 11 && cs.nb == 11) CDFT_ZP(11, 11);
  else if (cs.nf == 11 && cs.nb == 0) CDFT_ZP(11, 0);
  else if (cs.nf == 1 && cs.nb == 0) CDFT_ZP(1, 0);
  else if (cs.nf == 0 && cs.nb == 0) CDFT_ZP(0, 0);
  else throw std::runtime_error("unsupported channel configuration");
#undef CDFT_ZP
}

// FP64 radial weights (w3, w2) of every element of the half spectrum in the
// general schedule's order kz * Nx * Ny + kx * Ny + ky; they depend only on the
// grid and the radius, so the x kernels read them instead of evaluating
// sincos per element and launch.
__global__ void __launch_bounds__(NT) k_wtable_general(cx<double>* __restrict__ wt, Geom g) {
  const long nxy = static_cast<long>(g.nx) * g.ny;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < g.M;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    const long kz = i / nxy, r = i - kz * nxy;
    const int x = static_cast<int>(r / g.ny), y = static_cast<int>(r - static_cast<long>(x) * g.ny);
    const double kx = kvec(x, g.nx, g.dkx), ky = kvec(y, g.ny, g.dky), kzv = g.dkz * static_cast<double>(kz);
    double w3, w2;
    radial(sqrt(kx * kx + ky * ky + kzv * kzv), g.R, g.inv_n, &w3, &w2);
    wt[i] = {w3, w2};
  }
}

void launch_wtable_general(cx<double>* wt, const Geom& g, hipStream_t st) {
  hipLaunchKernelGGL(k_wtable_general, dim3(1024), dim3(NT), 0, st, wt, g);
}

void launch_finalize(const double* partials, int nrows, double* res, hipStream_t st) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinThreads), 0, st, partials, nrows, res);
}

#define CDFT_INST_SC(S, C)                                                                                         \
  template void launch_zr2c<S, C>(const double*, cx<S>*, const Geom&, const FftPlan&, const cx<C>*, AliasMap,      \
                                  const double*, hipStream_t, const double*, int, double*);                        \
  template void launch_zr2c_fields<S, C>(const double*, int, cx<S>*, const Geom&, const FftPlan&, const cx<C>*,    \
                                         hipStream_t);                                                             \
  template void launch_zc2r<S, C>(const cx<S>*, int, double*, const double*, double, const Geom&, const FftPlan&, \
                                  const cx<C>*, const double*, DcTable, hipStream_t, const double*, int, double*); \
  template void launch_lines<S, C>(cx<S>*, long, const FftPlan&, const cx<C>*, int, hipStream_t);                 \
  template void launch_xlines<S, C>(cx<S>*, int, const Geom&, const FftPlan&, const cx<C>*, int, hipStream_t);    \
  template void launch_xfwd<S, C, C>(const cx<S>*, cx<S>*, const S*, const cx<double>*, const Geom&,               \
                                     const FftPlan&, const cx<C>*, const cx<C>*, ChannelSet, bool, hipStream_t);  \
  template void launch_xbwd<S, C>(const cx<S>*, cx<S>*, const cx<double>*, const Geom&, const FftPlan&,           \
                                  const cx<C>*, int, hipStream_t);

CDFT_INST_SC(float, float)
CDFT_INST_SC(float, double)
CDFT_INST_SC(double, double)
template void launch_xfwd<float, double, float>(const cx<float>*, cx<float>*, const float*, const cx<double>*,
                                                const Geom&, const FftPlan&, const cx<double>*, const cx<float>*,
                                                ChannelSet, bool, hipStream_t);

template void launch_zpoint<float, double, float>(cx<float>*, const double*, const double*, double*, double*,
                                                  const Geom&, const FftPlan&, const cx<double>*, const cx<float>*,
                                                  ChannelSet, FmtParams, double, AliasMap, const double*, DcTable,
                                                  hipStream_t);
template void launch_zpoint<float, float, float>(cx<float>*, const double*, const double*, double*, double*,
                                                 const Geom&, const FftPlan&, const cx<float>*, const cx<float>*,
                                                 ChannelSet, FmtParams, double, AliasMap, const double*, DcTable,
                                                 hipStream_t);
template void launch_zpoint<double, double, double>(cx<double>*, const double*, const double*, double*, double*,
                                                    const Geom&, const FftPlan&, const cx<double>*,
                                                    const cx<double>*, ChannelSet, FmtParams, double, AliasMap,
                                                    const double*, DcTable, hipStream_t);

}  // namespace cdft

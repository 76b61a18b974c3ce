// Device helpers shared by the evaluation kernels: compensated FP64
// accumulation, precision conversion, wave vectors, the analytic FMT weights
// and a block reduction.
#pragma once

#include <hip/hip_runtime.h>

#include "fft_lds.hpp"

namespace cdft {

__device__ __forceinline__ void two_sum(double& s, double& c, double x) {
  // Neumaier compensated accumulation of x into (s, c)
  const double t = s + x;
  c += (fabs(s) >= fabs(x)) ? ((s - t) + x) : ((x - t) + s);
  s = t;
}

// Deterministic compensated column sums res[0..7] of `nrows` rows of 8
// partial sums, by one block of NT threads (sh, shc: 2 NT doubles of LDS).
// Thread t owns column t % 8 of rows t / 8 + (NT / 8) j and issues its loads
// 16 at a time before adding them; the per-thread (sum, compensation) pairs
// of a column then merge in a fixed tree order.
template <int NT>
__device__ __forceinline__ void reduce_rows(const double* __restrict__ partials, int nrows, double* __restrict__ res,
                                            double* sh, double* shc) {
  constexpr int RS = NT / 8, U = 16;  // rows per sweep, loads in flight per thread
  const int t = threadIdx.x, comp = t & 7, row = t >> 3;
  double s = 0.0, c = 0.0;
  for (int b0 = row; b0 < nrows; b0 += U * RS) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = b0 + u * RS;
      v[u] = b < nrows ? partials[static_cast<long>(b) * 8 + comp] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) two_sum(s, c, v[u]);
  }
  sh[t] = s;
  shc[t] = c;
  __syncthreads();
  for (int w = RS / 2; w > 0; w >>= 1) {
    if (row < w) {
      double s2 = sh[t], c2 = shc[t] + shc[t + 8 * w];
      two_sum(s2, c2, sh[t + 8 * w]);
      sh[t] = s2;
      shc[t] = c2;
    }
    __syncthreads();
  }
  if (t < 8) res[t] = sh[t] + shc[t];
}

// Sum of column 0 of n rows of 8 partial sums, returned to every thread of a
// block of NT threads; every block computes it in the same fixed order
// (compensated per thread, then a tree of (sum, compensation) pairs), so a
// kernel can reduce an earlier kernel's block sums itself instead of waiting
// for a reduction launch.
template <int NT>
__device__ __forceinline__ double block_sum_col0(const double* __restrict__ rows, int n) {
  __shared__ double sh[NT], shc[NT];
  double s = 0.0, c = 0.0;
  for (int r = threadIdx.x; r < n; r += NT) two_sum(s, c, rows[static_cast<long>(r) * 8]);
  sh[threadIdx.x] = s;
  shc[threadIdx.x] = c;
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) {
      double s2 = sh[threadIdx.x], c2 = shc[threadIdx.x] + shc[threadIdx.x + w];
      two_sum(s2, c2, sh[threadIdx.x + w]);
      sh[threadIdx.x] = s2;
      shc[threadIdx.x] = c2;
    }
    __syncthreads();
  }
  const double total = sh[0] + shc[0];
  __syncthreads();  // the arrays may be reused after the call
  return total;
}

template <typename D, typename T>
__device__ __forceinline__ cx<D> cvt(cx<T> a) {
  return {static_cast<D>(a.re), static_cast<D>(a.im)};
}

__device__ __forceinline__ double kvec(int i, int n, double dk) { return dk * static_cast<double>(i <= n / 2 ? i : i - n); }

// Radial FMT functions w3(k) = 4 pi (sin kR - kR cos kR)/k^3, w2(k) = 4 pi R sin(kR)/k
// (weights.hpp:91-108,159-162), each times 1/N.  The vector channel prefactor of
// weights.hpp:110-119,164-168 equals w3 and the tensor coefficients of
// weights.hpp:121-137,170-178 are A = w2 - 3 w3/R, B = w3/R (identities of the closed forms).
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

// Radial factors of one element: w3, w2 and aik2 = (w2 - 3 w3/R) / k^2 (0 at
// k = 0), from which every FMT channel weight follows without a division.
struct WRad {
  double w3, w2, aik2;
};
__device__ __forceinline__ WRad wrad(cx<double> w, double k2, double invR) {
  return {w.re, w.im, k2 > 0.0 ? (w.im - 3.0 * w.re * invR) / k2 : 0.0};
}
// Weight of channel ch from the radial factors (the closed forms of
// weights.hpp:91-184): ch 0 w3, 1 w2, 2..4 i k_a w3 (conj when `adj`), 5..10
// tensor aik2 k_i k_j + delta_ij w3 / R, 11 the mean-field table value.
__device__ __forceinline__ cx<double> channel_weight_t(int ch, bool adj, double kx, double ky, double kz, WRad w,
                                                       double mf, double invR) {
  if (ch == 0) return {w.w3, 0.0};
  if (ch == 1) return {w.w2, 0.0};
  if (ch <= 4) {
    const double ka = ch == 2 ? kx : (ch == 3 ? ky : kz);
    return {0.0, adj ? -ka * w.w3 : ka * w.w3};
  }
  if (ch <= 10) {
    const int q = ch - 5;
    const double ki = q < 3 ? kx : (q < 5 ? ky : kz);
    const double kj = (q == 0) ? kx : (q == 1 || q == 3) ? ky : kz;
    const bool diag = (q == 0 || q == 3 || q == 5);
    return {w.aik2 * ki * kj + (diag ? w.w3 * invR : 0.0), 0.0};
  }
  return {mf, 0.0};
}

// Weight of channel ch at k (FP64): ch 0 w3, 1 w2, 2..4 i k_a w3 (conj when
// `adj`), 5..10 tensor xx,xy,xz,yy,yz,zz, 11 the mean-field table value.
__device__ __forceinline__ cx<double> channel_weight(int ch, bool adj, double kx, double ky, double kz, double w3,
                                                     double w2, double mf, double invR) {
  if (ch == 0) return {w3, 0.0};
  if (ch == 1) return {w2, 0.0};
  if (ch <= 4) {
    const double ka = ch == 2 ? kx : (ch == 3 ? ky : kz);
    return {0.0, adj ? -ka * w3 : ka * w3};
  }
  if (ch <= 10) {
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double ik2 = k2 > 0.0 ? 1.0 / k2 : 0.0;
    const double A = w2 - 3.0 * w3 * invR;
    const double B = w3 * invR;
    const int q = ch - 5;
    const double ki = q < 3 ? kx : (q < 5 ? ky : kz);
    const double kj = (q == 0) ? kx : (q == 1 || q == 3) ? ky : kz;
    const bool diag = (q == 0 || q == 3 || q == 5);
    return {A * ki * kj * ik2 + (diag ? B : 0.0), 0.0};
  }
  return {mf, 0.0};
}

// Wave-then-block sum of NV doubles; thread 0 of the block gets the totals in out[].
template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double* out) {
  __shared__ double wsum[16][NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double x = v[i];
    for (int off = warpSize / 2; off > 0; off >>= 1) x += __shfl_xor(x, off);
    v[i] = x;
  }
  const int wave = threadIdx.x / warpSize, lane = threadIdx.x % warpSize;
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) wsum[wave][i] = v[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + warpSize - 1) / warpSize;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double t = 0.0;
      for (int w = 0; w < nw; ++w) t += wsum[w][i];
      out[i] = t;
    }
  }
}

}  // namespace cdft

// LDS-resident Stockham FFT building blocks for gfx950 (CDNA4).
//
// A work-group holds `nl` lines of length L in one LDS buffer (line stride
// `ls`, complex elements) and transforms all of them cooperatively, in place:
// in each radix pass every thread gathers its butterflies into registers,
// the block synchronises, and every thread scatters its outputs.  Batching
// many lines (and several channels) into one call keeps every thread busy in
// every pass.  Radices 2, 3, 4, 5 and 8 cover every grid length of the form
// 2^a 3^b 5^c.  Twiddles come from a per-length table exp(-2 pi i m / L),
// m < L, computed in FP64 on the host.  The arithmetic type C is independent
// of the type the caller stores in HBM.
//
// In FP32 every twiddle and every radix constant is applied as an FP32 pair
// (value + residual), which removes the coherent part of the FP32 error.
//
// sign = -1 is the forward transform (e^{-ikr}), +1 the backward one; both
// are unnormalised, as FFTW's r2c / c2r are in the reference
// (src/math/fourier.cpp:41-57).
#pragma once

#include <hip/hip_runtime.h>

namespace cdft {

template <typename R>
struct cx {
  R re, im;
};

// Exact n / d for 0 <= n < 2^24 and a divisor 1 <= d < 2^20 that is the same
// for the whole block, by a multiply-high: m = ceil(2^k / d) with
// k = max(32, 24 + ceil(log2 d)), n / d = (n m) >> k (the round-up method of
// Granlund and Montgomery: m d - 2^k < d and n < 2^24 keep the floor exact).
// Costs two instructions per quotient where a runtime division costs about
// twenty; m is formed once, in FP64 (exact: 2^k / d is either an integer or
// at least 1/d from one).
struct UDiv {
  unsigned d, m;
  int sh;
  __device__ explicit UDiv(int dv) : d(static_cast<unsigned>(dv)) {
    const int l = 32 - __clz(static_cast<int>(d - 1u));  // ceil(log2 d), 0 for d = 1
    const int k = l > 8 ? 24 + l : 32;
    sh = k - 32;
    m = d > 1u ? static_cast<unsigned>(ceil(ldexp(1.0, k) / static_cast<double>(d))) : 0u;
  }
  __device__ int div(int n) const {
    return d > 1u ? static_cast<int>(__umulhi(static_cast<unsigned>(n), m) >> sh) : n;
  }
};

template <typename R>
__device__ __forceinline__ cx<R> cadd(cx<R> a, cx<R> b) {
  return {a.re + b.re, a.im + b.im};
}
template <typename R>
__device__ __forceinline__ cx<R> csub(cx<R> a, cx<R> b) {
  return {a.re - b.re, a.im - b.im};
}
template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R>
__device__ __forceinline__ cx<R> cscale(cx<R> a, R s) {
  return {a.re * s, a.im * s};
}
template <typename R>
__device__ __forceinline__ cx<R> cconj(cx<R> a) {
  return {a.re, -a.im};
}
// (sg * i) * a
template <typename R>
__device__ __forceinline__ cx<R> cmul_si(cx<R> a, int sg) {
  return sg < 0 ? cx<R>{a.im, -a.re} : cx<R>{-a.im, a.re};
}

// x * c for a real constant c.  In FP32 the constant is applied as the sum of
// its FP32 value and its FP32 residual (c = hi + lo), so the product carries
// the constant to ~2^-48: FP32 rounding of a constant is the same for every
// line of a batch and would otherwise add up coherently.
template <typename R>
__device__ __forceinline__ R kmul(R x, double c) {
  if constexpr (sizeof(R) == 4) {
    const float hi = static_cast<float>(c);
    const float lo = static_cast<float>(c - static_cast<double>(hi));
    return fmaf(x, lo, x * hi);
  } else {
    return x * c;
  }
}

// Per-axis transform plan: length and radix sequence (product == L).
struct FftPlan {
  int L;
  int np;
  int rad[12];  // radices of the passes, in order
  // Launch shape of the general-schedule kernels along this axis (0: the
  // built-in default; set by the autotuner):
  //   lines  lines per block of the axis' line transforms;
  //   tile   x axis: y columns per block of the x kernels;
  //          z axis: z lines per block of the pointwise kernel.
  int lines;
  int tile;
};

// ---- radix-r DFT kernels: v[m] <- sum_i v[i] exp(sg * 2 pi i m i / r) -------
template <typename R>
__device__ __forceinline__ void dft2(cx<R>* v) {
  cx<R> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <typename R>
__device__ __forceinline__ void dft4(cx<R>* v, int sg) {
  cx<R> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
  cx<R> b0 = cadd(v[1], v[3]), b1 = cmul_si(csub(v[1], v[3]), sg);
  v[0] = cadd(a0, b0);
  v[1] = cadd(a1, b1);
  v[2] = csub(a0, b0);
  v[3] = csub(a1, b1);
}

// PAIR = false applies 1/sqrt2 as one rounded multiply (the register-resident
// slot transforms, whose twiddles are single rounded values too).
template <typename R, bool PAIR = true>
__device__ __forceinline__ void dft8(cx<R>* v, int sg) {
  cx<R> e[4] = {v[0], v[2], v[4], v[6]};
  cx<R> o[4] = {v[1], v[3], v[5], v[7]};
  dft4(e, sg);
  dft4(o, sg);
  constexpr double h = 0.70710678118654752440084436210484903928;
  auto kh = [](R x) { return PAIR ? kmul<R>(x, h) : x * static_cast<R>(h); };
  // w8^1 = (1 + sg i)/sqrt2, w8^2 = sg i, w8^3 = (-1 + sg i)/sqrt2
  cx<R> o1 = {kh(o[1].re - R(sg) * o[1].im), kh(o[1].im + R(sg) * o[1].re)};
  cx<R> o2 = cmul_si(o[2], sg);
  cx<R> o3 = {kh(-o[3].re - R(sg) * o[3].im), kh(-o[3].im + R(sg) * o[3].re)};
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

template <typename R, bool PAIR = true>
__device__ __forceinline__ void dft3(cx<R>* v, int sg) {
  constexpr double s3 = 0.86602540378443864676372317075293618347;
  auto km = [](R x) { return PAIR ? kmul<R>(x, s3) : x * static_cast<R>(s3); };
  cx<R> t = cadd(v[1], v[2]);
  cx<R> dd = cmul_si(csub(v[1], v[2]), sg);
  cx<R> d = {km(dd.re), km(dd.im)};
  cx<R> m = {v[0].re - R(0.5) * t.re, v[0].im - R(0.5) * t.im};
  v[0] = cadd(v[0], t);
  v[1] = cadd(m, d);
  v[2] = csub(m, d);
}

template <typename R, bool PAIR = true>
__device__ __forceinline__ void dft5(cx<R>* v, int sg) {
  constexpr double c1 = 0.30901699437494742410229341718281905886;   // cos(2pi/5)
  constexpr double c2 = -0.80901699437494742410229341718281905886;  // cos(4pi/5)
  constexpr double s1 = 0.95105651629515357211643933337938214340;   // sin(2pi/5)
  constexpr double s2 = 0.58778525229247312916870595463907276860;   // sin(4pi/5)
  auto km = [](R x, double c) { return PAIR ? kmul<R>(x, c) : x * static_cast<R>(c); };
  cx<R> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  cx<R> t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  cx<R> a1 = {v[0].re + km(t1.re, c1) + km(t2.re, c2), v[0].im + km(t1.im, c1) + km(t2.im, c2)};
  cx<R> a2 = {v[0].re + km(t1.re, c2) + km(t2.re, c1), v[0].im + km(t1.im, c2) + km(t2.im, c1)};
  cx<R> b1 = cmul_si(cx<R>{km(t3.re, s1) + km(t4.re, s2), km(t3.im, s1) + km(t4.im, s2)}, sg);
  cx<R> b2 = cmul_si(cx<R>{km(t3.re, s2) - km(t4.re, s1), km(t3.im, s2) - km(t4.im, s1)}, sg);
  v[0] = cadd(v[0], cadd(t1, t2));
  v[1] = cadd(a1, b1);
  v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2);
  v[3] = csub(a2, b2);
}

// Radix 6 and 10 (the first pass of the register-resident transforms of
// lengths 3 * 2^a and 5 * 2^a): the radix-3 / radix-5 transforms of the even
// and the odd inputs, the odd ones times exp(sg 2 pi i k / r), k < r/2, then
// one radix-2 stage: X[k] = E[k] + W^k O[k], X[k + r/2] = E[k] - W^k O[k].
// The constants are single rounded values (these radices serve only the slot
// transforms, whose twiddles are single rounded values too).
template <typename R>
__device__ __forceinline__ void dft6(cx<R>* v, int sg) {
  constexpr double s3 = 0.86602540378443864676372317075293618347;  // sin(pi/3)
  cx<R> e[3] = {v[0], v[2], v[4]};
  cx<R> o[3] = {v[1], v[3], v[5]};
  dft3<R, false>(e, sg);
  dft3<R, false>(o, sg);
  const R ss = static_cast<R>(sg) * static_cast<R>(s3);
  const cx<R> o1 = cmul(o[1], cx<R>{R(0.5), ss});   // W^1 = (1/2, sg sqrt3/2)
  const cx<R> o2 = cmul(o[2], cx<R>{R(-0.5), ss});  // W^2 = (-1/2, sg sqrt3/2)
  v[0] = cadd(e[0], o[0]);
  v[3] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[4] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[5] = csub(e[2], o2);
}

template <typename R>
__device__ __forceinline__ void dft10(cx<R>* v, int sg) {
  constexpr double c36 = 0.80901699437494742410229341718281905886;  // cos(pi/5)
  constexpr double s36 = 0.58778525229247312916870595463907276860;  // sin(pi/5)
  constexpr double c72 = 0.30901699437494742410229341718281905886;  // cos(2pi/5)
  constexpr double s72 = 0.95105651629515357211643933337938214340;  // sin(2pi/5)
  cx<R> e[5] = {v[0], v[2], v[4], v[6], v[8]};
  cx<R> o[5] = {v[1], v[3], v[5], v[7], v[9]};
  dft5<R, false>(e, sg);
  dft5<R, false>(o, sg);
  const R sg1 = static_cast<R>(sg);
  // W^k = exp(sg 2 pi i k / 10), k = 1..4
  const cx<R> w[4] = {{static_cast<R>(c36), sg1 * static_cast<R>(s36)},
                      {static_cast<R>(c72), sg1 * static_cast<R>(s72)},
                      {static_cast<R>(-c72), sg1 * static_cast<R>(s72)},
                      {static_cast<R>(-c36), sg1 * static_cast<R>(s36)}};
  v[0] = cadd(e[0], o[0]);
  v[5] = csub(e[0], o[0]);
#pragma unroll
  for (int k = 1; k < 5; ++k) {
    const cx<R> t = cmul(o[k], w[k - 1]);
    v[k] = cadd(e[k], t);
    v[k + 5] = csub(e[k], t);
  }
}

template <typename R, int RAD, bool PAIR = true>
__device__ __forceinline__ void dft_r(cx<R>* v, int sg) {
  if constexpr (RAD == 2) dft2(v);
  else if constexpr (RAD == 3) dft3<R, PAIR>(v, sg);
  else if constexpr (RAD == 4) dft4(v, sg);
  else if constexpr (RAD == 5) dft5<R, PAIR>(v, sg);
  else if constexpr (RAD == 6) dft6(v, sg);
  else if constexpr (RAD == 10) dft10(v, sg);
  else dft8<R, PAIR>(v, sg);
}

// One in-place Stockham pass over `nl` lines held in one LDS buffer: every
// thread first gathers, twiddles and transforms all of its butterflies in
// registers, the block synchronises, then every thread writes its outputs.
// MAXB bounds the butterflies per thread (nl * L / RAD <= MAXB * blockDim.x).
// The FP32 twiddle table holds the rounded values in [0, L) and their
// residuals in [L, 2L); FP64 tables hold [0, L).
template <typename C, int RAD, int MAXB>
__device__ __forceinline__ void pass_ip(cx<C>* buf, int nl, int ls, int L, int ns, int sg,
                                        const cx<C>* __restrict__ tw) {
  const int nb = L / RAD;
  const int tot = nl * nb;
  const int tws = L / (ns * RAD);
  const UDiv dnb(nb), dns(ns);
  cx<C> v[MAXB][RAD];
#pragma unroll
  for (int m = 0; m < MAXB; ++m) {
    const int b = threadIdx.x + m * blockDim.x;
    if (b < tot) {
      const int line = dnb.div(b);
      const int j = b - line * nb;
      const int k = j - dns.div(j) * ns;
      const cx<C>* s = buf + line * ls;
#pragma unroll
      for (int i = 0; i < RAD; ++i) v[m][i] = s[j + i * nb];
      if (ns > 1) {
#pragma unroll
        for (int i = 1; i < RAD; ++i) {
          const int ti = i * k * tws;
          cx<C> w = tw[ti];
          if (sg > 0) w.im = -w.im;
          if constexpr (sizeof(C) == 4) {
            cx<C> wl = tw[L + ti];
            if (sg > 0) wl.im = -wl.im;
            v[m][i] = cadd(cmul(v[m][i], w), cmul(v[m][i], wl));
          } else {
            v[m][i] = cmul(v[m][i], w);
          }
        }
      }
      dft_r<C, RAD>(v[m], sg);
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < MAXB; ++m) {
    const int b = threadIdx.x + m * blockDim.x;
    if (b < tot) {
      const int line = dnb.div(b);
      const int j = b - line * nb;
      const int jq = dns.div(j);
      const int k = j - jq * ns;
      const int base = jq * ns * RAD + k;
      cx<C>* d = buf + line * ls;
#pragma unroll
      for (int i = 0; i < RAD; ++i) d[base + i * ns] = v[m][i];
    }
  }
  __syncthreads();
}

// Transform `nl` lines held in `buf` in place.  MAXE bounds the elements per
// thread (nl * L <= MAXE * blockDim.x).  The caller synchronises before the
// call (inputs visible); the call ends with a barrier (outputs visible).
template <typename C, int MAXE>
__device__ __forceinline__ void fft_inplace(cx<C>* buf, int nl, int ls, const FftPlan& p, int sg, const cx<C>* __restrict__ tw) {
  int ns = 1;
  for (int q = 0; q < p.np; ++q) {
    const int r = p.rad[q];
    switch (r) {
      case 8: pass_ip<C, 8, (MAXE + 7) / 8>(buf, nl, ls, p.L, ns, sg, tw); break;
      case 4: pass_ip<C, 4, (MAXE + 3) / 4>(buf, nl, ls, p.L, ns, sg, tw); break;
      case 2: pass_ip<C, 2, (MAXE + 1) / 2>(buf, nl, ls, p.L, ns, sg, tw); break;
      case 3: pass_ip<C, 3, (MAXE + 2) / 3>(buf, nl, ls, p.L, ns, sg, tw); break;
      default: pass_ip<C, 5, (MAXE + 4) / 5>(buf, nl, ls, p.L, ns, sg, tw); break;
    }
    ns *= r;
  }
}

}  // namespace cdft

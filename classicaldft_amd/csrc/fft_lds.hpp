// LDS-resident Stockham FFT building blocks for gfx950 (CDNA4).
//
// A work-group holds `nl` lines of length L in LDS (line stride `ls`, complex
// elements) and transforms all of them cooperatively: each radix pass reads
// one buffer and writes the other (ping-pong), so no thread ever waits on
// another inside a pass; one barrier separates passes.  Radices 2, 3, 4, 5
// and 8 cover every grid length of the form 2^a 3^b 5^c.  Twiddles come from
// a per-length table exp(-2 pi i m / L), m < L, computed in FP64 on the host.
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

// Per-axis transform plan: length and radix sequence (product == L).
struct FftPlan {
  int L;
  int np;
  int rad[12];
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

template <typename R>
__device__ __forceinline__ void dft8(cx<R>* v, int sg) {
  cx<R> e[4] = {v[0], v[2], v[4], v[6]};
  cx<R> o[4] = {v[1], v[3], v[5], v[7]};
  dft4(e, sg);
  dft4(o, sg);
  const R h = R(0.70710678118654752440084436210484903928);
  // w8^1 = (1 + sg i)/sqrt2, w8^2 = sg i, w8^3 = (-1 + sg i)/sqrt2
  cx<R> o1 = {h * (o[1].re - R(sg) * o[1].im), h * (o[1].im + R(sg) * o[1].re)};
  cx<R> o2 = cmul_si(o[2], sg);
  cx<R> o3 = {h * (-o[3].re - R(sg) * o[3].im), h * (-o[3].im + R(sg) * o[3].re)};
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

template <typename R>
__device__ __forceinline__ void dft3(cx<R>* v, int sg) {
  const R s3 = R(0.86602540378443864676372317075293618347);
  cx<R> t = cadd(v[1], v[2]);
  cx<R> d = cscale(cmul_si(csub(v[1], v[2]), sg), s3);
  cx<R> m = {v[0].re - R(0.5) * t.re, v[0].im - R(0.5) * t.im};
  v[0] = cadd(v[0], t);
  v[1] = cadd(m, d);
  v[2] = csub(m, d);
}

template <typename R>
__device__ __forceinline__ void dft5(cx<R>* v, int sg) {
  const R c1 = R(0.30901699437494742410229341718281905886);   // cos(2pi/5)
  const R c2 = R(-0.80901699437494742410229341718281905886);  // cos(4pi/5)
  const R s1 = R(0.95105651629515357211643933337938214340);   // sin(2pi/5)
  const R s2 = R(0.58778525229247312916870595463907276860);   // sin(4pi/5)
  cx<R> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  cx<R> t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  cx<R> a1 = {v[0].re + c1 * t1.re + c2 * t2.re, v[0].im + c1 * t1.im + c2 * t2.im};
  cx<R> a2 = {v[0].re + c2 * t1.re + c1 * t2.re, v[0].im + c2 * t1.im + c1 * t2.im};
  cx<R> b1 = cmul_si(cx<R>{s1 * t3.re + s2 * t4.re, s1 * t3.im + s2 * t4.im}, sg);
  cx<R> b2 = cmul_si(cx<R>{s2 * t3.re - s1 * t4.re, s2 * t3.im - s1 * t4.im}, sg);
  v[0] = cadd(v[0], cadd(t1, t2));
  v[1] = cadd(a1, b1);
  v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2);
  v[3] = csub(a2, b2);
}

template <typename R, int RAD>
__device__ __forceinline__ void dft_r(cx<R>* v, int sg) {
  if constexpr (RAD == 2) dft2(v);
  else if constexpr (RAD == 3) dft3(v, sg);
  else if constexpr (RAD == 4) dft4(v, sg);
  else if constexpr (RAD == 5) dft5(v, sg);
  else dft8(v, sg);
}

// One Stockham pass over all lines: src -> dst.
template <typename R, int RAD>
__device__ __forceinline__ void stockham_pass(const cx<R>* __restrict__ src, cx<R>* __restrict__ dst, int nl,
                                              int ls, int L, int ns, int sg, const cx<R>* __restrict__ tw) {
  const int nb = L / RAD;
  const int tot = nl * nb;
  const int tws = L / (ns * RAD);
  for (int b = threadIdx.x; b < tot; b += blockDim.x) {
    const int line = b / nb;
    const int j = b - line * nb;
    const int k = j % ns;
    const cx<R>* s = src + line * ls;
    cx<R> v[RAD];
#pragma unroll
    for (int i = 0; i < RAD; ++i) v[i] = s[j + i * nb];
    if (ns > 1) {
#pragma unroll
      for (int i = 1; i < RAD; ++i) {
        cx<R> w = tw[i * k * tws];
        if (sg > 0) w.im = -w.im;
        v[i] = cmul(v[i], w);
      }
    }
    dft_r<R, RAD>(v, sg);
    const int base = (j / ns) * ns * RAD + k;
    cx<R>* d = dst + line * ls;
#pragma unroll
    for (int i = 0; i < RAD; ++i) d[base + i * ns] = v[i];
  }
}

// Transform `nl` lines held in buf0; returns the buffer holding the result
// (buf0 or buf1).  Caller must __syncthreads() before (data in buf0 visible)
// and after (before reading the result).
template <typename R>
__device__ cx<R>* fft_lines(cx<R>* buf0, cx<R>* buf1, int nl, int ls, const FftPlan& p, int sg,
                            const cx<R>* __restrict__ tw) {
  cx<R>* a = buf0;
  cx<R>* b = buf1;
  int ns = 1;
  for (int q = 0; q < p.np; ++q) {
    const int r = p.rad[q];
    switch (r) {
      case 8: stockham_pass<R, 8>(a, b, nl, ls, p.L, ns, sg, tw); break;
      case 4: stockham_pass<R, 4>(a, b, nl, ls, p.L, ns, sg, tw); break;
      case 2: stockham_pass<R, 2>(a, b, nl, ls, p.L, ns, sg, tw); break;
      case 3: stockham_pass<R, 3>(a, b, nl, ls, p.L, ns, sg, tw); break;
      default: stockham_pass<R, 5>(a, b, nl, ls, p.L, ns, sg, tw); break;
    }
    ns *= r;
    cx<R>* t = a;
    a = b;
    b = t;
    if (q + 1 < p.np) __syncthreads();
  }
  return a;
}

}  // namespace cdft

// Register-resident line FFTs for the fused evaluation kernels (gfx950).
//
// A "slot" is a group of T consecutive lanes of one wavefront (T divides 64)
// that transforms one line of compile-time length L.  Every lane holds
// E = L / T elements in registers:
//   on entry  v[q] = x[lane + T q]   (q < E)
//   on exit   v[e] = X[lane + T e]   (e < E)
// The transform is a Stockham autosort FFT whose first pass has radix E (its
// inputs are exactly the lane's registers) and whose last pass writes its
// outputs straight back to registers in the order above; the passes between
// exchange data through the slot's private, padded LDS buffer.  Lanes of a
// slot live in one wavefront, so the exchanges need only wavefront-level
// ordering, never a work-group barrier: waves of a block run their
// transforms independently and overlap one another's memory traffic.
//
// sign SG = -1 is the forward transform exp(-i k x), +1 the backward one;
// both are unnormalised (src/math/fourier.cpp:41-57 of the reference).
// Twiddles come from one table holding, for every power-of-two length
// 4 <= L <= kTwMax, the L values exp(-2 pi i m / L) at offset L - 4 (so the
// butterflies of a pass read consecutive entries), correctly rounded to the
// arithmetic type, and so is the radix-8 constant 1/sqrt2.  (Carrying the
// FP32 rounding residual as a second product changed the Omega and force
// errors of a 128^3 FP32 evaluation by under 5% and cost 2.5x the multiplies.)
#pragma once

#include <hip/hip_runtime.h>

#include "fft_lds.hpp"

namespace cdft {

constexpr int kTwMax = 256;            // longest line with a twiddle table
constexpr int kTwLen = 2 * kTwMax - 4;  // entries of the concatenated tables (L = 4 .. kTwMax)

// Orders the LDS traffic of the lanes of one wavefront (no s_barrier).  The
// fences cover LDS only, so global loads (twiddles, the next inputs) can
// still be issued early across it.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Padded position inside a slot buffer: one spare element every 16 spreads
// the strided writes of the early passes over the LDS banks.  (Transforms
// that exchange through the line's own storage use the unpadded position.)
__host__ __device__ constexpr int xpad(int x) { return x + (x >> 4); }
template <bool PAD>
__host__ __device__ constexpr int xpos(int x) {
  return PAD ? xpad(x) : x;
}
template <int L>
__host__ __device__ constexpr int slot_buf_len() {
  return xpad(L - 1) + 1;
}

// Radix sequence of an (L, T) slot: radix E first, then the largest of 8, 4,
// 2 that divides both E and the remaining length.  E is 2, 4 or 8, or 6 or
// 10 (lengths 3 * 2^a and 5 * 2^a, whose later passes are radix 2); L / E is
// a power of two.
template <int L, int T>
struct SlotPlan {
  static_assert(L % T == 0, "slot width must divide the line length");
  static constexpr int E = L / T;
  static_assert(E == 2 || E == 4 || E == 8 || E == 6 || E == 10, "2, 4, 6, 8 or 10 elements per lane");
  static_assert((T & (T - 1)) == 0, "power-of-two slot width");
  static constexpr int rad(int p) {
    int rem = L;
    for (int i = 0;; ++i) {
      int r = E;
      if (i > 0) {
        r = 8;
        while (r > 1 && ((E % r) != 0 || (rem % r) != 0)) r >>= 1;
      }
      if (i == p) return r;
      rem /= r;
    }
  }
  static constexpr int npass() {
    int rem = L, n = 0;
    while (rem > 1) {
      rem /= rad(n);
      ++n;
    }
    return n;
  }
};

// The length-L table inside the concatenated tables.
template <int L, typename C>
__device__ __forceinline__ const cx<C>* tw_for(const cx<C>* all) {
  static_assert(L >= 4 && L <= kTwMax && (L & (L - 1)) == 0, "no twiddle table for this length");
  return all + (L - 4);
}

// v * exp(SG 2 pi i m / L), 0 <= m < L; tw is the length-L table.
template <typename C, int L, int SG>
__device__ __forceinline__ cx<C> twiddle(cx<C> v, int m, const cx<C>* __restrict__ tw) {
  cx<C> w = tw[m];
  if (SG > 0) w.im = -w.im;
  return cmul(v, w);
}

// Exchange element of a slot buffer: complex values, or with SPLIT one real
// part at a time (an FP64 exchange then needs 8 bytes per element, not 16).
template <typename C, bool SPLIT>
struct SlotXchg {
  static constexpr int bytes = SPLIT ? static_cast<int>(sizeof(C)) : static_cast<int>(sizeof(cx<C>));
};
template <typename C, int L, bool SPLIT>
__host__ __device__ constexpr int slot_buf_bytes() {
  return slot_buf_len<L>() * SlotXchg<C, SPLIT>::bytes;
}

template <typename C, int L, int T, int SG, bool SPLIT, int P, int NS, bool PAD = true>
struct SlotPass {
  using Plan = SlotPlan<L, T>;
  static constexpr int E = Plan::E;
  static constexpr int R = Plan::rad(P);
  static constexpr int NB = E / R;  // butterflies per lane
  static constexpr int NP = Plan::npass();
  static_assert(E % R == 0, "radix must divide the elements per lane");

  // u holds the inputs of this pass: butterfly b = lane + T i reads x[b + (L/R) q]
  __device__ __forceinline__ static void run(cx<C> (&v)[E], cx<C> (&u)[NB][R], void* xb, int lane,
                                             const cx<C>* __restrict__ tw) {
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      if constexpr (NS > 1) {
        const int k = static_cast<int>(static_cast<unsigned>(lane + T * i) % NS);  // unsigned: a mask for 2^a
#pragma unroll
        for (int q = 1; q < R; ++q) u[i][q] = twiddle<C, L, SG>(u[i][q], q * k * (L / (NS * R)), tw);
      }
      dft_r<C, R, false>(u[i], SG);
    }
    if constexpr (P == NP - 1) {
      // outputs of butterfly b at b + NS q = lane + T (i + NB q)
#pragma unroll
      for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int q = 0; q < R; ++q) v[i + NB * q] = u[i][q];
    } else {
      using Next = SlotPass<C, L, T, SG, SPLIT, P + 1, NS * R, PAD>;
      cx<C> w[Next::NB][Next::R];
      // outputs of butterfly b at (b/NS) NS R + b%NS + NS q; inputs of the next pass at b' + (L/R') q'
      auto put = [&](auto get, auto* buf) {
        wave_sync();  // every lane is done with the buffer's previous contents
#pragma unroll
        for (int i = 0; i < NB; ++i) {
          const int b = lane + T * i;
          const unsigned ub = static_cast<unsigned>(b);
          const int base = static_cast<int>((ub / NS) * NS * R + ub % NS);
#pragma unroll
          for (int q = 0; q < R; ++q) buf[xpos<PAD>(base + NS * q)] = get(u[i][q]);
        }
        wave_sync();
      };
      if constexpr (SPLIT) {
        C* xr = static_cast<C*>(xb);
        put([](cx<C> a) { return a.re; }, xr);
#pragma unroll
        for (int i = 0; i < Next::NB; ++i)
#pragma unroll
          for (int q = 0; q < Next::R; ++q) w[i][q].re = xr[xpos<PAD>(lane + T * i + (L / Next::R) * q)];
        put([](cx<C> a) { return a.im; }, xr);
#pragma unroll
        for (int i = 0; i < Next::NB; ++i)
#pragma unroll
          for (int q = 0; q < Next::R; ++q) w[i][q].im = xr[xpos<PAD>(lane + T * i + (L / Next::R) * q)];
      } else {
        cx<C>* xc = static_cast<cx<C>*>(xb);
        put([](cx<C> a) { return a; }, xc);
#pragma unroll
        for (int i = 0; i < Next::NB; ++i)
#pragma unroll
          for (int q = 0; q < Next::R; ++q) w[i][q] = xc[xpos<PAD>(lane + T * i + (L / Next::R) * q)];
      }
      Next::run(v, w, xb, lane, tw);
    }
  }
};

// One line transform by one slot; `xb` is the slot's buffer of
// slot_buf_bytes<C, L, SPLIT>() bytes, `tw` the length-L twiddle table
// (tw_for<L> of the concatenated tables).  All lanes of the slot call it together.
template <typename C, int L, int T, int SG, bool SPLIT = false, bool PAD = true>
__device__ __forceinline__ void slot_fft(cx<C> (&v)[L / T], void* xb, int lane, const cx<C>* __restrict__ tw) {
  constexpr int E = L / T;
  // The exchange addresses depend only on the lane; left visible, the
  // compiler hoists those of every pass of every transform in a kernel to its
  // entry and keeps them live (or spilled) throughout.  An opaque copy of the
  // lane index keeps them local to the transform.
  asm volatile("" : "+v"(lane));
  cx<C> u[1][E];  // first pass: radix E, butterfly b = lane reads x[lane + T q] = v[q]
#pragma unroll
  for (int q = 0; q < E; ++q) u[0][q] = v[q];
  SlotPass<C, L, T, SG, SPLIT, 0, 1, PAD>::run(v, u, xb, lane, tw);
}

}  // namespace cdft

// Pointwise Fundamental Measure Theory free-energy density and its partial
// derivatives, evaluated in FP64 on the device.
//
// Follows include/dft/functionals/fmt/models.hpp of the reference:
//   Phi  = -n0 f1(eta) + (n1 n2 - v0.v1) f2(eta) + Phi3 f3(eta)   (models.hpp:49-54)
//   dPhi by the chain rule (models.hpp:56-81), with the closed-form
//   derivatives of f1, f2, f3 in place of the reference's forward-mode
//   autodiff, and the same Taylor branches below eta = 1e-6
//   (models.hpp:200-202, 250-253, 282-292).
// Only the 11 independent weighted densities of one species are inputs:
// eta, n2, n_v2[3], T[6]; n1, n0 and n_v1 follow by Rosenfeld scaling
// (hard_sphere.hpp:73-97).  The outputs are already combined into the 11
// back-convolution channels of hard_sphere.hpp:115-134:
//   g0 = dPhi/deta
//   g1 = dPhi/dn2 + dPhi/dn1 / (4 pi R) + dPhi/dn0 / (4 pi R^2)
//   g2..4 = dPhi/dv1_a + dPhi/dv0_a / (4 pi R)
//   g5..10 = sym_ij dPhi/dT_ij  (xx, xy, xz, yy, yz, zz; sym = 2 off-diagonal)
//
// Both functions are templates on the scalar type: double for the evaluation,
// Dual (value + directional derivative) for the Hessian-vector product of the
// adjoint, where the derivative parts of g are sum_b d2Phi/dn_a dn_b * dn_b.
#pragma once

#include <hip/hip_runtime.h>

#include "fast_math.hpp"

namespace cdft {

// Forward-mode dual number: v + d*eps with eps^2 = 0.
struct Dual {
  double v, d;
  Dual() = default;
  __host__ __device__ constexpr Dual(double x) : v(x), d(0.0) {}
  __host__ __device__ constexpr Dual(double x, double y) : v(x), d(y) {}
};
__host__ __device__ inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__host__ __device__ inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__host__ __device__ inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
__host__ __device__ inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.v * b.d + a.d * b.v}; }
__host__ __device__ inline Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__host__ __device__ inline Dual& operator+=(Dual& a, Dual b) { return a = a + b; }

// value part (the scalar itself for double): branches test values only
__host__ __device__ inline double vof(double x) { return x; }
__host__ __device__ inline double vof(Dual x) { return x.v; }
__host__ __device__ inline double dlog(double x) { return flog(x); }
__host__ __device__ inline Dual dlog(Dual x) { return {flog(x.v), x.d / x.v}; }
__host__ __device__ inline double drcp(double x) { return frcp(x); }
__host__ __device__ inline Dual drcp(Dual x) { return Dual(1.0) / x; }
__host__ __device__ inline double dmin(double x, double c) { return fmin(x, c); }
__host__ __device__ inline Dual dmin(Dual x, double c) { return x.v < c ? x : Dual(c); }

enum FmtModelId : int {
  FMT_NONE = -1,
  FMT_ROSENFELD = 0,
  FMT_RSLT = 1,
  FMT_WHITEBEAR_I = 2,
  FMT_WHITEBEAR_II = 3,
  FMT_ESFMT = 4,
};

struct FmtParams {
  int model;
  double R;      // hard-sphere radius
  double A, B;   // esFMT coefficients (WhiteBear I/II use A=1, B=-1)
};

// f1, f2, f3 and their derivatives at packing fraction e.  The quotients
// are formed from the two reciprocals 1/(1-e) and 1/e, each taken from one
// FP64 division: a division costs about ten FP64 instructions on the device,
// a product one.
template <typename T>
__host__ __device__ inline void f_funcs(int model, T e, T* f, T* df) {
  const T om = 1.0 - e;
  const T L = dlog(om);
  const T r = drcp(e * om);
  const T io = e * r, ie = om * r;  // 1/(1-e), 1/e
  const T io2 = io * io, ie2 = ie * ie;
  // f1 (models.hpp:157-160): identical for every model
  f[0] = L;
  df[0] = -io;
  // f2
  if (model == FMT_WHITEBEAR_II) {
    if (vof(e) < 1e-6) {
      f[1] = 1.0 + e * (1.0 + e * (10.0 / 9.0 + 7.0 / 6.0 * e));
      df[1] = 1.0 + e * (20.0 / 9.0 + 3.5 * e);
    } else {
      f[1] = 1.0 / 3.0 + (4.0 / 3.0) * io + (2.0 / 3.0) * ie * L;
      df[1] = (4.0 / 3.0) * io2 - (2.0 / 3.0) * ie2 * L - (2.0 / 3.0) * ie * io;
    }
  } else {
    f[1] = io;
    df[1] = io2;
  }
  // f3
  if (model == FMT_RSLT || model == FMT_WHITEBEAR_I) {
    if (vof(e) < 1e-6) {
      if (model == FMT_RSLT) {
        f[2] = 1.0 + (2.0 / 3.0) * e + 0.5 * e * e;
        df[2] = 2.0 / 3.0 + e;
      } else {
        f[2] = 1.5 + e * (8.0 / 3.0 + e * (15.0 / 4.0 + 24.0 / 5.0 * e));
        df[2] = 8.0 / 3.0 + e * (7.5 + 14.4 * e);
      }
    } else {
      f[2] = ie * io2 + L * ie2;
      df[2] = -ie2 * io2 + 2.0 * ie * io2 * io - io * ie2 - 2.0 * L * ie2 * ie;
    }
  } else if (model == FMT_WHITEBEAR_II) {
    if (vof(e) < 1e-6) {
      f[2] = 1.5 + 7.0 / 3.0 * e + 3.25 * e * e + 4.2 * e * e * e;
      df[2] = 7.0 / 3.0 + 6.5 * e + 12.6 * e * e;
    } else {
      // -g/h - L/e^2 with g = 1 - 3e + e^2, h = e (1-e)^2
      const T g = 1.0 - 3.0 * e + e * e, gp = -3.0 + 2.0 * e;
      const T h = e * om * om, hp = om * (1.0 - 3.0 * e);
      const T ih = ie * io2;
      f[2] = -g * ih - L * ie2;
      df[2] = -(gp * h - g * hp) * ih * ih + io * ie2 + 2.0 * L * ie2 * ie;
    }
  } else {
    f[2] = io2;
    df[2] = 2.0 * io2 * io;
  }
}

// The model-specific third term Phi3(n2, n_v2, T) and its partial
// derivatives (models.hpp:18-45,200-357); identical for one species and for
// mixtures, whose n2, n_v2 and T are sums over species.
template <typename T>
struct Phi3 {
  T p3, dn2, dv[3], dT[6];
  bool tensor;
};
// REDUCED evaluates the White-Bear I/II tensor term in its reduced algebraic
// form (same values, different rounding; the mixed-precision kernels use it).
template <typename T, bool REDUCED = false>
__host__ __device__ __forceinline__ Phi3<T> fmt_phi3_t(const FmtParams& P, T n2, const T* v, T vv, const T* Tm) {
  const double PI = 3.14159265358979323846;
  Phi3<T> r;
  r.p3 = 0.0;
  r.dn2 = 0.0;
  r.tensor = false;
  for (int k = 0; k < 3; ++k) r.dv[k] = 0.0;
  for (int q = 0; q < 6; ++q) r.dT[q] = 0.0;
  if (P.model == FMT_ROSENFELD) {
    const double k24 = 1.0 / (24.0 * PI);
    r.p3 = k24 * (n2 * n2 * n2 - 3.0 * n2 * vv);
    r.dn2 = (1.0 / (8.0 * PI)) * (n2 * n2 - vv);
    for (int k = 0; k < 3; ++k) r.dv[k] = -(1.0 / (4.0 * PI)) * n2 * v[k];
  } else if (P.model == FMT_RSLT) {
    if (vof(n2) >= 1e-30) {
      const T xi = dmin(vv / (n2 * n2), 1.0);
      const T q = 1.0 - xi;
      const double k36 = 1.0 / (36.0 * PI);
      r.p3 = k36 * n2 * n2 * n2 * q * q * q;
      r.dn2 = k36 * 3.0 * n2 * n2 * q * q * (1.0 + xi);
      for (int k = 0; k < 3; ++k) r.dv[k] = k36 * (-6.0) * n2 * q * q * v[k];
    }
  } else if (P.model != FMT_NONE) {
    // tensor models: esFMT(A,B); WhiteBear I/II are esFMT(1,-1) (models.hpp:18-45,326-357)
    r.tensor = true;
    const double A = P.A, B = P.B;
    const T a = Tm[0], b = Tm[1], c = Tm[2], d = Tm[3], e = Tm[4], f = Tm[5];
    const T sxx = a * a + b * b + c * c, sxy = a * b + b * d + c * e, sxz = a * c + b * e + c * f;
    const T syy = b * b + d * d + e * e, syz = b * c + d * e + e * f, szz = c * c + e * e + f * f;
    const T trT2 = sxx + syy + szz;
    const T trT3 = a * sxx + d * syy + f * szz + 2.0 * (b * sxy + c * sxz + e * syz);
    const T Tv[3] = {a * v[0] + b * v[1] + c * v[2], b * v[0] + d * v[1] + e * v[2],
                     c * v[0] + e * v[1] + f * v[2]};
    const T qf = v[0] * Tv[0] + v[1] * Tv[1] + v[2] * Tv[2];
    const double k24 = 1.0 / (24.0 * PI);
    const double k8 = 1.0 / (8.0 * PI), k4 = 1.0 / (4.0 * PI);
    const T T2m[6] = {sxx, sxy, sxz, syy, syz, szz};
    const int ii[6] = {0, 0, 0, 1, 1, 2};
    const int jj[6] = {0, 1, 2, 1, 2, 2};
    if (REDUCED && A == 1.0 && B == -1.0) {
      // White-Bear I/II: the n2^3 terms cancel, leaving
      //   Phi3 = (n2 (tr T^2 - v.v) + v.T.v - tr T^3) / (8 pi)
      // and its derivatives in the same reduced form
      r.p3 = k8 * (n2 * (trT2 - vv) + qf - trT3);
      r.dn2 = k8 * (trT2 - vv);
      for (int k = 0; k < 3; ++k) r.dv[k] = k4 * (Tv[k] - n2 * v[k]);
#pragma unroll
      for (int q = 0; q < 6; ++q) r.dT[q] = k8 * (v[ii[q]] * v[jj[q]] - 3.0 * T2m[q] + 2.0 * n2 * Tm[q]);
    } else {
      const T n23 = n2 * n2 * n2;
      r.p3 = k24 * (A * (n23 - 3.0 * n2 * vv + 3.0 * qf - trT3) + B * (n23 - 3.0 * n2 * trT2 + 2.0 * trT3));
      r.dn2 = k24 * (A * (3.0 * n2 * n2 - 3.0 * vv) + B * (3.0 * n2 * n2 - 3.0 * trT2));
      for (int k = 0; k < 3; ++k) r.dv[k] = k24 * A * (-6.0 * n2 * v[k] + 6.0 * Tv[k]);
#pragma unroll
      for (int q = 0; q < 6; ++q)
        r.dT[q] = k8 * A * (v[ii[q]] * v[jj[q]] - T2m[q]) + k4 * B * (-n2 * Tm[q] + T2m[q]);
    }
  }
  return r;
}

// One species.  w[0..10] = eta, n2, v[3], T[xx,xy,xz,yy,yz,zz]; n1, n0 and
// n_v1 follow by Rosenfeld scaling with the species radius
// (hard_sphere.hpp:80-86).  Returns Phi and fills g[0..10] with the
// derivatives already combined for the 11 weights (hard_sphere.hpp:118-134).
template <typename T, bool REDUCED = false>
__host__ __device__ inline T fmt_phi_grad_t(const FmtParams& P, const T* w, T* g) {
  const double PI = 3.14159265358979323846;
  const T eta = w[0];
  if (!(vof(eta) >= 1e-30)) {  // hard_sphere.hpp:199-200 (skip, derivatives stay zero)
#pragma unroll
    for (int i = 0; i < 11; ++i) g[i] = T(0.0);
    return T(0.0);
  }
  const double R = P.R;
  // 1/(4 pi R) and 1/(4 pi R^2): the same for every point (hoisted out of loops)
  const double ifr = 1.0 / (4.0 * PI * R), ifr2 = ifr / R;
  const T n2 = w[1];
  const T n1 = n2 * ifr;
  const T n0 = n2 * ifr2;
  const T v[3] = {w[2], w[3], w[4]};
  const T vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  const T v0v1 = vv * ifr;

  T F[3], dF[3];
  f_funcs<T>(P.model, eta, F, dF);
  const Phi3<T> t = fmt_phi3_t<T, REDUCED>(P, n2, v, vv, w + 5);

  const T s = n1 * n2 - v0v1;
  const T phi = -n0 * F[0] + s * F[1] + t.p3 * F[2];
  const T d_eta = -n0 * dF[0] + s * dF[1] + t.p3 * dF[2];
  const T d_n0 = -F[0];
  const T d_n1 = n2 * F[1];
  const T d_n2 = n1 * F[1] + t.dn2 * F[2];
  g[0] = d_eta;
  g[1] = d_n2 + d_n1 * ifr + d_n0 * ifr2;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const T dv0 = -v[k] * F[1];                           // dPhi/dv0 = -v1 f2
    const T dv1 = -(v[k] * ifr) * F[1] + t.dv[k] * F[2];  // dPhi/dv1 = -v0 f2 + dPhi3/dv1 f3
    g[2 + k] = dv1 + dv0 * ifr;
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const double sym = (q == 0 || q == 3 || q == 5) ? 1.0 : 2.0;
    g[5 + q] = t.tensor ? sym * t.dT[q] * F[2] : T(0.0);
  }
  return phi;
}

// Mixtures (hard_sphere.hpp:174-224): m[0..15] = eta, n2, n1, n0, n_v2[3],
// n_v1[3], T[6], each already summed over species with its own radius.
// Returns Phi and fills g[0..15] with the partial derivatives in the same
// order (tensor off-diagonals doubled); the per-species weight combination
// happens in k-space.
template <typename T>
__host__ __device__ inline T fmt_phi_grad_mix_t(const FmtParams& P, const T* m, T* g) {
  const T eta = m[0];
  if (!(vof(eta) >= 1e-30)) {  // hard_sphere.hpp:199-200
#pragma unroll
    for (int i = 0; i < 16; ++i) g[i] = T(0.0);
    return T(0.0);
  }
  const T n2 = m[1], n1 = m[2], n0 = m[3];
  const T* v = m + 4;   // n_v2 (the reference's Measures::v1)
  const T* v0 = m + 7;  // n_v1 (Measures::v0)
  const T vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  const T v0v1 = v0[0] * v[0] + v0[1] * v[1] + v0[2] * v[2];

  T F[3], dF[3];
  f_funcs<T>(P.model, eta, F, dF);
  const Phi3<T> t = fmt_phi3_t<T>(P, n2, v, vv, m + 10);

  const T s = n1 * n2 - v0v1;
  const T phi = -n0 * F[0] + s * F[1] + t.p3 * F[2];
  g[0] = -n0 * dF[0] + s * dF[1] + t.p3 * dF[2];
  g[1] = n1 * F[1] + t.dn2 * F[2];
  g[2] = n2 * F[1];
  g[3] = -F[0];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    g[4 + k] = -v0[k] * F[1] + t.dv[k] * F[2];  // d/d n_v2
    g[7 + k] = -v[k] * F[1];                    // d/d n_v1
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const double sym = (q == 0 || q == 3 || q == 5) ? 1.0 : 2.0;
    g[10 + q] = t.tensor ? sym * t.dT[q] * F[2] : T(0.0);
  }
  return phi;
}

__host__ __device__ inline double fmt_phi_grad(const FmtParams& P, const double* w, double* g) {
  return fmt_phi_grad_t<double>(P, w, g);
}

}  // namespace cdft

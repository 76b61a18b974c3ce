// FP64 CPU oracle for the classicalDFT free-energy + gradient path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: the pytest suite, the
// smoke() check and the CPU-baseline leg of bench.py load it (through ctypes)
// to compare the HIP library against.  Nothing in the product links it.
//
// It restates, in plain C++ with its own mixed-radix FFT, the algorithm of the
// reference library (migduroli/classicalDFT, file:line cited per function):
//
//   grid / half-spectrum indexing ...... include/dft/grid.hpp:35-63
//   r2c / c2r (unnormalised) ........... src/math/fourier.cpp:41-57,97-103
//   convolve / back_convolve ........... include/dft/math/convolution.hpp:13-49
//   FMT weights (11 channels, 1/N) ..... include/dft/functionals/fmt/weights.hpp:91-184
//   FMT models Phi, dPhi ............... include/dft/functionals/fmt/models.hpp:22-81,151-358
//   hard_sphere ........................ include/dft/functionals/hard_sphere.hpp:50-298
//   mean-field weights + evaluation .... include/dft/functionals/mean_field.hpp:41-281
//   potentials (LJ, tWF, WRDF) ......... include/dft/physics/potentials.hpp:53-368
//   ideal gas / external field ......... include/dft/functionals/{ideal_gas,external_field}.hpp
//   total, Omega ....................... include/dft/functionals/functionals.hpp:64-95
//   bulk chemical potential ............ include/dft/functionals/bulk/thermodynamics.hpp:222-237
//   FIRE2 .............................. include/dft/algorithms/fire.hpp:60-178
//   Minimizer::grand_potential ......... include/dft/algorithms/minimization.hpp:45-110,280-353
//   DDFT integrating-factor step ....... include/dft/algorithms/dynamics.hpp:79-401,582-693
//   fixed-dt DDFT map + Gaussian wells . (no counterpart: the map the adjoint differentiates)
//   smallest Hessian eigenpair ......... include/dft/algorithms/saddle_point.hpp:28-38,159-270
//   saddle search ...................... include/dft/algorithms/saddle_point.hpp:367-589
//   split-operator / Crank-Nicolson .... include/dft/algorithms/dynamics.hpp:405-538
//   Picard ............................. include/dft/algorithms/picard.hpp:63-124
//
// Where the reference's output depends on FFTW behaviour it does not document
// (c2r of a half spectrum that is not Hermitian, which happens for the odd
// vector-weight channel on self-conjugate Nyquist entries), this restatement
// defines c2r as Re(inverse DFT of the Hermitian extension): complex inverse
// along x and y, then a z-line c2r that uses only the real part of the kz=0
// and kz=Nz/2 bins.  The HIP library implements the same convention.
//
// Parity pinning of every function is done by tests/test_oracle_*.py (numpy
// FFT, brute-force DFT, closed forms, sum rules, finite differences).

#include <algorithm>
#include <cmath>
#include <limits>
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

using cd = std::complex<double>;
constexpr double PI = 3.14159265358979323846264338327950288;

// ---------------------------------------------------------------------------
// 1D mixed-radix Stockham FFT, FP64.  sign = -1 forward (e^{-i}), +1 backward.
// ---------------------------------------------------------------------------
struct Fft1d {
  int n = 0;
  std::vector<int> radices;
  std::vector<std::vector<cd>> tw[2];  // [sign index][pass] -> r*Ns twiddles
  std::vector<std::vector<cd>> roots[2];  // [sign index][pass] -> r*r DFT matrix

  explicit Fft1d(int n_) : n(n_) {
    int m = n;
    for (int r : {4, 2, 3, 5, 7}) {
      while (m % r == 0) { radices.push_back(r); m /= r; }
    }
    for (int p = 11; m > 1; p += 2) {  // any remaining primes (generic DFT)
      while (m % p == 0) { radices.push_back(p); m /= p; }
      if (p * p > m && m > 1) { radices.push_back(m); m = 1; }
    }
    for (int s = 0; s < 2; ++s) {
      double sg = s == 0 ? -1.0 : 1.0;
      int ns = 1;
      for (int r : radices) {
        std::vector<cd> t(static_cast<size_t>(r) * ns);
        long period = static_cast<long>(ns) * r;
        for (int k = 0; k < ns; ++k)
          for (int i = 0; i < r; ++i) {
            long e = (static_cast<long>(i) * k) % period;
            double a = sg * 2.0 * PI * static_cast<double>(e) / static_cast<double>(period);
            t[static_cast<size_t>(k) * r + i] = cd(std::cos(a), std::sin(a));
          }
        tw[s].push_back(std::move(t));
        std::vector<cd> w(static_cast<size_t>(r) * r);
        for (int m = 0; m < r; ++m)
          for (int i = 0; i < r; ++i) {
            long e = (static_cast<long>(m) * i) % r;
            double a = sg * 2.0 * PI * static_cast<double>(e) / static_cast<double>(r);
            w[static_cast<size_t>(m) * r + i] = cd(std::cos(a), std::sin(a));
          }
        roots[s].push_back(std::move(w));
        ns *= r;
      }
    }
  }

  // In-place transform of `x` (length n, contiguous) using scratch `y`.
  void exec(cd* x, cd* y, int sign) const {
    const int si = sign < 0 ? 0 : 1;
    const double sg = sign < 0 ? -1.0 : 1.0;
    cd* a = x;
    cd* b = y;
    int ns = 1;
    cd v[64];
    cd out[64];
    for (size_t p = 0; p < radices.size(); ++p) {
      const int r = radices[p];
      const int nb = n / r;
      const cd* t = tw[si][p].data();
      for (int j = 0; j < nb; ++j) {
        int k = j % ns;
        for (int i = 0; i < r; ++i) v[i] = a[j + i * nb] * t[static_cast<size_t>(k) * r + i];
        // radix-r DFT
        if (r == 2) {
          out[0] = v[0] + v[1];
          out[1] = v[0] - v[1];
        } else if (r == 4) {
          cd a0 = v[0] + v[2], a1 = v[0] - v[2], b0 = v[1] + v[3], b1 = v[1] - v[3];
          cd jb1 = cd(-sg * b1.imag(), sg * b1.real());  // (sg*i) * b1
          out[0] = a0 + b0;
          out[1] = a1 + jb1;
          out[2] = a0 - b0;
          out[3] = a1 - jb1;
        } else {
          const cd* w = roots[si][p].data();
          for (int m = 0; m < r; ++m) {
            cd acc = 0.0;
            for (int i = 0; i < r; ++i) acc += v[i] * w[static_cast<size_t>(m) * r + i];
            out[m] = acc;
          }
        }
        int base = (j / ns) * ns * r + k;
        for (int m = 0; m < r; ++m) b[base + m * ns] = out[m];
      }
      std::swap(a, b);
      ns *= r;
    }
    if (a != x) std::memcpy(x, a, sizeof(cd) * static_cast<size_t>(n));
  }
};

// ---------------------------------------------------------------------------
// 3D real <-> half-complex transforms in the reference layout
//   real:  idx = iz + Nz*(iy + Ny*ix)
//   half:  idx = kz + (Nz/2+1)*(iy + Ny*ix)
// ---------------------------------------------------------------------------
struct Fft3d {
  long nx, ny, nz, nzh;
  Fft1d fx, fy, fz;
  Fft3d(long nx_, long ny_, long nz_)
      : nx(nx_), ny(ny_), nz(nz_), nzh(nz_ / 2 + 1), fx(static_cast<int>(nx_)), fy(static_cast<int>(ny_)),
        fz(static_cast<int>(nz_)) {}

  long n_real() const { return nx * ny * nz; }
  long n_half() const { return nx * ny * nzh; }

  // forward r2c, e^{-ikr}, unnormalised (fourier.cpp:41-49,97-99)
  void r2c(const double* in, cd* out) const {
#pragma omp parallel
    {
      std::vector<cd> line(static_cast<size_t>(std::max({nx, ny, nz}))), scr(line.size());
#pragma omp for schedule(static)
      for (long l = 0; l < nx * ny; ++l) {
        for (long z = 0; z < nz; ++z) line[z] = cd(in[l * nz + z], 0.0);
        fz.exec(line.data(), scr.data(), -1);
        for (long k = 0; k < nzh; ++k) out[l * nzh + k] = line[k];
      }
#pragma omp for schedule(static)
      for (long l = 0; l < nx * nzh; ++l) {  // y lines: fixed (ix, kz)
        long ix = l / nzh, kz = l % nzh;
        cd* base = out + ix * ny * nzh + kz;
        for (long y = 0; y < ny; ++y) line[y] = base[y * nzh];
        fy.exec(line.data(), scr.data(), -1);
        for (long y = 0; y < ny; ++y) base[y * nzh] = line[y];
      }
#pragma omp for schedule(static)
      for (long l = 0; l < ny * nzh; ++l) {  // x lines: fixed (iy, kz)
        cd* base = out + l;
        for (long x = 0; x < nx; ++x) line[x] = base[x * ny * nzh];
        fx.exec(line.data(), scr.data(), -1);
        for (long x = 0; x < nx; ++x) base[x * ny * nzh] = line[x];
      }
    }
  }

  // backward c2r, e^{+ikr}, unnormalised, Hermitian-projection semantics.
  // `in` is not modified (a private copy is transformed).
  void c2r(const cd* in, double* out) const {
    std::vector<cd> work(in, in + n_half());
#pragma omp parallel
    {
      std::vector<cd> line(static_cast<size_t>(std::max({nx, ny, nz}))), scr(line.size());
#pragma omp for schedule(static)
      for (long l = 0; l < ny * nzh; ++l) {
        cd* base = work.data() + l;
        for (long x = 0; x < nx; ++x) line[x] = base[x * ny * nzh];
        fx.exec(line.data(), scr.data(), +1);
        for (long x = 0; x < nx; ++x) base[x * ny * nzh] = line[x];
      }
#pragma omp for schedule(static)
      for (long l = 0; l < nx * nzh; ++l) {
        long ix = l / nzh, kz = l % nzh;
        cd* base = work.data() + ix * ny * nzh + kz;
        for (long y = 0; y < ny; ++y) line[y] = base[y * nzh];
        fy.exec(line.data(), scr.data(), +1);
        for (long y = 0; y < ny; ++y) base[y * nzh] = line[y];
      }
#pragma omp for schedule(static)
      for (long l = 0; l < nx * ny; ++l) {
        const cd* h = work.data() + l * nzh;
        for (long k = 0; k < nzh; ++k) line[k] = h[k];
        line[0] = cd(h[0].real(), 0.0);
        if (nz % 2 == 0) line[nz / 2] = cd(h[nz / 2].real(), 0.0);
        for (long k = nzh; k < nz; ++k) line[k] = std::conj(h[nz - k]);
        fz.exec(line.data(), scr.data(), +1);
        for (long z = 0; z < nz; ++z) out[l * nz + z] = line[z].real();
      }
    }
  }
};

// ---------------------------------------------------------------------------
// FMT models (models.hpp).  f1, f2, f3 with closed-form first derivatives.
// Model ids: 0 Rosenfeld, 1 RSLT, 2 WhiteBearI, 3 WhiteBearII, 4 esFMT(A,B).
// ---------------------------------------------------------------------------
struct FVal { double f, df; };

// models.hpp:157-160 (and identical f1 in every model)
inline FVal f1(int /*model*/, double e) { return {std::log(1.0 - e), -1.0 / (1.0 - e)}; }

// models.hpp:162,196,246,280-286
inline FVal f2(int model, double e) {
  if (model == 3) {
    if (e < 1e-6) return {1.0 + e * (1.0 + e * (10.0 / 9.0 + 7.0 / 6.0 * e)), 1.0 + e * (20.0 / 9.0 + 3.5 * e)};
    double om = 1.0 - e, L = std::log(om);
    double f = 1.0 / 3.0 + (4.0 / 3.0) / om + 2.0 / (3.0 * e) * L;
    double df = (4.0 / 3.0) / (om * om) - 2.0 / (3.0 * e * e) * L - 2.0 / (3.0 * e) / om;
    return {f, df};
  }
  double om = 1.0 - e;
  return {1.0 / om, 1.0 / (om * om)};
}

// models.hpp:164-167,198-205,248-256,288-295
inline FVal f3(int model, double e) {
  double om = 1.0 - e;
  switch (model) {
    case 1:  // RSLT
    case 2:  // WhiteBearI
      if (e < 1e-6) {
        if (model == 1) return {1.0 + (2.0 / 3.0) * e + 0.5 * e * e, 2.0 / 3.0 + e};
        return {1.5 + e * (8.0 / 3.0 + e * (15.0 / 4.0 + 24.0 / 5.0 * e)), 8.0 / 3.0 + e * (7.5 + 14.4 * e)};
      } else {
        double L = std::log(om);
        double f = 1.0 / (e * om * om) + L / (e * e);
        double df = -1.0 / (e * e * om * om) + 2.0 / (e * om * om * om) - 1.0 / (om * e * e) - 2.0 * L / (e * e * e);
        return {f, df};
      }
    case 3: {  // WhiteBearII
      if (e < 1e-6) return {1.5 + 7.0 / 3.0 * e + 3.25 * e * e + 4.2 * e * e * e, 7.0 / 3.0 + 6.5 * e + 12.6 * e * e};
      double L = std::log(om);
      double g = 1.0 - 3.0 * e + e * e, gp = -3.0 + 2.0 * e;
      double h = e * om * om, hp = om * (1.0 - 3.0 * e);
      double f = -g / h - L / (e * e);
      double df = -(gp * h - g * hp) / (h * h) + 1.0 / (om * e * e) + 2.0 * L / (e * e * e);
      return {f, df};
    }
    default:  // Rosenfeld, esFMT
      return {1.0 / (om * om), 2.0 / (om * om * om)};
  }
}

// Measures layout (19 doubles): eta, n0, n1, n2, v0[3], v1[3], T[9] (row-major).
// Naming follows measures.hpp:20-29: v1 holds n_v2 and v0 holds n_v1.
struct Meas {
  double eta, n0, n1, n2, v0[3], v1[3], T[9];
};

struct Inner {
  double v0v1, v1v1, trT2, trT3, qf;
};

// measures.hpp:34-43
inline Inner inner(const Meas& m) {
  Inner p{};
  for (int i = 0; i < 3; ++i) {
    p.v0v1 += m.v0[i] * m.v1[i];
    p.v1v1 += m.v1[i] * m.v1[i];
  }
  double T2[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += m.T[i * 3 + k] * m.T[k * 3 + j];
      T2[i * 3 + j] = s;
    }
  for (int i = 0; i < 3; ++i) p.trT2 += T2[i * 3 + i];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) p.trT3 += T2[i * 3 + k] * m.T[k * 3 + i];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) p.qf += m.v1[i] * m.T[i * 3 + j] * m.v1[j];
  return p;
}

struct Phi3 {
  double p3, d_n2, d_v1[3], d_T[9];
  bool tensor;
};

// models.hpp:18-45 (tensor WB), 169-182 (Rosenfeld), 207-232 (RSLT), 326-357 (esFMT)
inline Phi3 phi3(int model, double A, double B, const Meas& m, const Inner& p) {
  Phi3 r{};
  const double inv24pi = 1.0 / (24.0 * PI);
  if (model == 0) {
    r.p3 = inv24pi * (m.n2 * m.n2 * m.n2 - 3.0 * m.n2 * p.v1v1);
    r.d_n2 = (1.0 / (8.0 * PI)) * (m.n2 * m.n2 - p.v1v1);
    for (int k = 0; k < 3; ++k) r.d_v1[k] = -(1.0 / (4.0 * PI)) * m.n2 * m.v1[k];
    r.tensor = false;
    return r;
  }
  if (model == 1) {
    r.tensor = false;
    if (m.n2 < 1e-30) return r;
    double xi = std::min(p.v1v1 / (m.n2 * m.n2), 1.0);
    double q = 1.0 - xi;
    const double inv36pi = 1.0 / (36.0 * PI);
    r.p3 = inv36pi * m.n2 * m.n2 * m.n2 * q * q * q;
    r.d_n2 = inv36pi * 3.0 * m.n2 * m.n2 * q * q * (1.0 + xi);
    for (int k = 0; k < 3; ++k) r.d_v1[k] = inv36pi * (-6.0) * m.n2 * q * q * m.v1[k];
    return r;
  }
  double a = A, b = B;
  if (model == 2 || model == 3) { a = 1.0; b = -1.0; }
  r.tensor = true;
  double phia = m.n2 * m.n2 * m.n2 - 3.0 * m.n2 * p.v1v1 + 3.0 * p.qf - p.trT3;
  double phib = m.n2 * m.n2 * m.n2 - 3.0 * m.n2 * p.trT2 + 2.0 * p.trT3;
  r.p3 = inv24pi * (a * phia + b * phib);
  r.d_n2 = inv24pi * (a * (3.0 * m.n2 * m.n2 - 3.0 * p.v1v1) + b * (3.0 * m.n2 * m.n2 - 3.0 * p.trT2));
  for (int k = 0; k < 3; ++k) {
    double tv = 0.0;
    for (int j = 0; j < 3; ++j) tv += m.v1[j] * m.T[j * 3 + k] + m.v1[j] * m.T[k * 3 + j];
    r.d_v1[k] = inv24pi * a * (-6.0 * m.n2 * m.v1[k] + 3.0 * tv);
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double tt = 0.0;
      for (int k = 0; k < 3; ++k) tt += m.T[j * 3 + k] * m.T[k * 3 + i];
      r.d_T[i * 3 + j] = (1.0 / (8.0 * PI)) * a * (m.v1[i] * m.v1[j] - tt) +
                         (1.0 / (4.0 * PI)) * b * (-m.n2 * m.T[j * 3 + i] + tt);
    }
  return r;
}

// compute_phi / compute_d_phi (models.hpp:49-81)
inline double phi_dphi(int model, double A, double B, const Meas& m, Meas* dm) {
  Inner p = inner(m);
  FVal F1 = f1(model, m.eta), F2 = f2(model, m.eta), F3 = f3(model, m.eta);
  Phi3 P = phi3(model, A, B, m, p);
  double s = m.n1 * m.n2 - p.v0v1;
  double phi = -m.n0 * F1.f + s * F2.f + P.p3 * F3.f;
  if (dm) {
    dm->eta = -m.n0 * F1.df + s * F2.df + P.p3 * F3.df;
    dm->n0 = -F1.f;
    dm->n1 = m.n2 * F2.f;
    dm->n2 = m.n1 * F2.f + P.d_n2 * F3.f;
    for (int k = 0; k < 3; ++k) {
      dm->v0[k] = -m.v1[k] * F2.f;
      dm->v1[k] = -m.v0[k] * F2.f + P.d_v1[k] * F3.f;
    }
    for (int k = 0; k < 9; ++k) dm->T[k] = P.tensor ? P.d_T[k] * F3.f : 0.0;
  }
  return phi;
}

// measures.hpp:49-65
inline Meas uniform_measures(double rho, double d) {
  Meas m{};
  double r = 0.5 * d;
  m.eta = PI / 6.0 * rho * d * d * d;
  m.n2 = PI * rho * d * d;
  m.n1 = r * rho;
  m.n0 = rho;
  for (int i = 0; i < 3; ++i) m.T[i * 4] = m.n2 / 3.0;
  return m;
}

// ---------------------------------------------------------------------------
// Potentials (potentials.hpp:53-368) - attractive tails for the mean field.
// ---------------------------------------------------------------------------
struct Pot {
  int kind = 0;  // 0 LJ, 1 ten Wolde-Frenkel, 2 WRDF
  double sigma = 1, eps = 1, rc = -1, alpha = 50;
  double shift = 0, r_min = 0, v_min = 0, r_zero = 0, eps_eff = 0;

  double raw(double r) const {
    if (kind == 0) {
      double y = sigma / r, y6 = y * y * y * y * y * y;
      return 4.0 * eps * (y6 * y6 - y6);
    }
    if (kind == 1) {
      if (r < sigma) return 1e50;
      double s = r / sigma, y = 1.0 / (s * s - 1.0), y3 = y * y * y;
      return (4.0 * eps / (alpha * alpha)) * (y3 * y3 - alpha * y3);
    }
    if (r >= rc) return 0.0;
    double y = sigma / r, z = rc / r;
    return eps_eff * (y * y - 1.0) * (z * z - 1.0) * (z * z - 1.0);
  }

  void init() {
    if (kind == 0) {
      shift = rc > 0 ? raw(rc) : 0.0;
      r_min = std::pow(2.0, 1.0 / 6.0) * sigma;
      v_min = raw(r_min) - shift;
      r_zero = sigma * std::pow(0.5 * std::sqrt(1.0 + shift / eps) + 0.5, -1.0 / 6.0);
    } else if (kind == 1) {
      shift = rc > 0 ? raw(rc) : 0.0;
      r_min = sigma * std::sqrt(1.0 + std::pow(2.0 / alpha, 1.0 / 3.0));
      v_min = raw(r_min) - shift;
      r_zero = sigma * std::sqrt(1.0 + std::pow(25.0 * std::sqrt(1.0 + shift) + 25.0, -1.0 / 3.0));
    } else {
      double rcs = rc / sigma;
      eps_eff = eps * 2.0 * rcs * rcs * std::pow(2.0 * (rcs * rcs - 1.0) / 3.0, -3.0);
      r_min = rc * std::pow((1.0 + 2.0 * rcs * rcs) / 3.0, -0.5);
      v_min = raw(r_min);
      shift = 0.0;
    }
  }

  // split: 0 WCA, 1 BH
  double attractive(double r, int split) const {
    if (kind == 2) {
      if (r >= rc) return 0.0;
      if (split == 1) return r >= r_min ? raw(r) : 0.0;
      return r >= r_min ? raw(r) : v_min;
    }
    double rcut = rc > 0 ? rc : 1e30;
    if (r >= rcut) return 0.0;
    double vs = raw(r) - shift;
    if (split == 1) return r >= r_zero ? vs : 0.0;
    return r >= r_min ? vs : v_min;
  }
};

struct Interaction {
  int si = 0, sj = 0;
  Pot pot;
  int split = 0;
  int scheme = 3;  // 0 zero, 1 linearE, 2 linearF, 3 quadraticF, 4 gaussE, 5 gaussF
};

// mean_field.hpp:41-134
inline double cell_weight(const Interaction& in, double kT, double dx, long sx, long sy, long sz) {
  if (in.scheme == 0) {
    double r = std::sqrt(static_cast<double>(sx * sx + sy * sy + sz * sz) * dx * dx);
    return in.pot.attractive(r, in.split) / kT;
  }
  if (in.scheme == 3) {
    const double vv = 1.0 / 3.0;
    const double pt[3] = {-0.5, 0.0, 0.5};
    double sum = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          double rx = (static_cast<double>(sx) + pt[i]) * dx;
          double ry = (static_cast<double>(sy) + pt[j]) * dx;
          double rz = (static_cast<double>(sz) + pt[k]) * dx;
          double r = std::sqrt(rx * rx + ry * ry + rz * rz);
          sum += vv * vv * vv * in.pot.attractive(r, in.split) / kT;
        }
    return sum;
  }
  double sum = 0.0;  // linear (and the Gauss schemes, which fall back to it)
  for (int di = 0; di <= 1; ++di)
    for (int dj = 0; dj <= 1; ++dj)
      for (int dk = 0; dk <= 1; ++dk) {
        double x = (static_cast<double>(sx) - 0.5 + di) * dx;
        double y = (static_cast<double>(sy) - 0.5 + dj) * dx;
        double z = (static_cast<double>(sz) - 0.5 + dk) * dx;
        sum += in.pot.attractive(std::sqrt(x * x + y * y + z * z), in.split) / kT;
      }
  return sum / 8.0;
}

// ---------------------------------------------------------------------------
// Context: grid, species, precomputed weights.
// ---------------------------------------------------------------------------
struct Ctx {
  long nx, ny, nz, nzh, N, M;
  double dx, dv;
  int model;
  double A, B;
  double kT;
  std::vector<double> diam;
  std::vector<Interaction> inters;
  std::unique_ptr<Fft3d> fft;
  std::vector<std::vector<cd>> fmtw;  // per species: 11*M, channel-major
  std::vector<std::vector<cd>> mfw;   // per interaction: M
  std::vector<double> a_vdw;
  bool serial_pointwise = false;  // reference-faithful timing: per-point loop on one thread

  Ctx(long nx_, long ny_, long nz_, double dx_) : nx(nx_), ny(ny_), nz(nz_), nzh(nz_ / 2 + 1), dx(dx_) {
    N = nx * ny * nz;
    M = nx * ny * nzh;
    dv = dx * dx * dx;
    fft = std::make_unique<Fft3d>(nx, ny, nz);
  }

  // weights.hpp:91-184 and 229-269 (channel order w3, w2, wv2[x,y,z], wT xx,xy,xz,yy,yz,zz)
  void make_fmt_weights() {
    fmtw.clear();
    for (double d : diam) {
      std::vector<cd> w(11 * static_cast<size_t>(M));
      double R = 0.5 * d;
      double inv_n = 1.0 / static_cast<double>(N);
      double dkx = 2.0 * PI / (nx * dx), dky = 2.0 * PI / (ny * dx), dkz = 2.0 * PI / (nz * dx);
      for (long ix = 0; ix < nx; ++ix) {
        double kx = dkx * static_cast<double>(ix <= nx / 2 ? ix : ix - nx);
        for (long iy = 0; iy < ny; ++iy) {
          double ky = dky * static_cast<double>(iy <= ny / 2 ? iy : iy - ny);
          for (long iz = 0; iz < nzh; ++iz) {
            double kz = dkz * static_cast<double>(iz);
            long idx = iz + nzh * (iy + ny * ix);
            double kv[3] = {kx, ky, kz};
            double k2 = kx * kx + ky * ky + kz * kz;
            double k = std::sqrt(k2);
            double w3, w2, fv, Ac, Bc;
            if (k < 1e-12) {
              w3 = (4.0 * PI / 3.0) * R * R * R;
              w2 = 4.0 * PI * R * R;
              fv = 0.0;
              Ac = 0.0;
              Bc = (4.0 * PI / 3.0) * R * R;
            } else {
              double kR = k * R, s = std::sin(kR), c = std::cos(kR);
              w3 = 4.0 * PI * (s - kR * c) / (k * k * k);
              w2 = 4.0 * PI * R * s / k;
              fv = -4.0 * PI * (R * c - s / k) / (k * k);
              double k3 = k2 * k, R2 = R * R;
              double gp_over_k = 4.0 * PI * R * (R * c / k2 - s / k3);
              double gpp = 4.0 * PI * R * (-R2 * s / k - 2.0 * R * c / k2 + 2.0 * s / k3);
              Ac = -(gpp - gp_over_k) / R2;
              Bc = -gp_over_k / R2;
            }
            double inv_k2 = (k > 1e-12) ? 1.0 / k2 : 0.0;
            w[0 * M + idx] = w3 * inv_n;
            w[1 * M + idx] = w2 * inv_n;
            for (int a = 0; a < 3; ++a) w[(2 + a) * M + idx] = cd(0.0, 1.0) * kv[a] * fv * inv_n;
            int c6 = 0;
            for (int i = 0; i < 3; ++i)
              for (int j = i; j < 3; ++j, ++c6)
                w[(5 + c6) * M + idx] = (Ac * kv[i] * kv[j] * inv_k2 + Bc * (i == j ? 1.0 : 0.0)) * inv_n;
          }
        }
      }
      fmtw.push_back(std::move(w));
    }
  }

  // mean_field.hpp:145-196
  void make_mf_weights(std::vector<std::vector<double>>* real_out = nullptr) {
    mfw.clear();
    a_vdw.clear();
    for (const auto& in : inters) {
      std::vector<double> real(static_cast<size_t>(N), 0.0);
      double rcut = in.pot.kind == 2 ? in.pot.rc : in.pot.rc;
      double margin = std::sqrt(3.0) * 0.5 * dx;
      double rskip = rcut + margin, rskip2 = rskip * rskip;
      double inv_n = 1.0 / static_cast<double>(N);
      double wsum = 0.0;
      for (long ix = 0; ix < nx; ++ix) {
        long sx = ix <= nx / 2 ? ix : ix - nx;
        for (long iy = 0; iy < ny; ++iy) {
          long sy = iy <= ny / 2 ? iy : iy - ny;
          for (long iz = 0; iz < nz; ++iz) {
            long sz = iz <= nz / 2 ? iz : iz - nz;
            double d2 = static_cast<double>(sx * sx + sy * sy + sz * sz) * dx * dx;
            if (rcut > 0.0 && d2 > rskip2) continue;
            double w = cell_weight(in, kT, dx, sx, sy, sz);
            real[iz + nz * (iy + ny * ix)] = w * dv * inv_n;
            wsum += w * dv;
          }
        }
      }
      std::vector<cd> wk(static_cast<size_t>(M));
      fft->r2c(real.data(), wk.data());
      mfw.push_back(std::move(wk));
      a_vdw.push_back(wsum);
      if (real_out) real_out->push_back(std::move(real));
    }
  }

  // ---- pieces ----------------------------------------------------------
  // convolve (convolution.hpp:13-26): n = c2r(w .* rho_k)
  void convolve(const cd* w, const cd* rk, double* out) const {
    std::vector<cd> t(static_cast<size_t>(M));
    for (long i = 0; i < M; ++i) t[i] = w[i] * rk[i];
    fft->c2r(t.data(), out);
  }

  // hard_sphere (hard_sphere.hpp:263-298); returns F_hs, adds forces (times dV)
  // `wd_out`, when given, receives the 11 weighted-density fields of species 0.
  double hard_sphere(const double* rho, double* forces, double* wd_out) const {
    const long S = static_cast<long>(diam.size());
    std::vector<std::vector<double>> wd(S, std::vector<double>(11 * static_cast<size_t>(N)));
    std::vector<cd> rk(static_cast<size_t>(M));
    for (long s = 0; s < S; ++s) {
      fft->r2c(rho + s * N, rk.data());
      for (int c = 0; c < 11; ++c) convolve(fmtw[s].data() + c * M, rk.data(), wd[s].data() + c * N);
    }
    if (wd_out) std::memcpy(wd_out, wd[0].data(), sizeof(double) * 11 * N);

    // per-point loop (hard_sphere.hpp:174-224): derivative arrays in the
    // already-combined back-convolution form (hard_sphere.hpp:115-134):
    // 0 d_eta, 1 d_n2 + d_n1/(4piR) + d_n0/(4piR^2), 2..4 d_v1 + d_v0/(4piR),
    // 5..10 sym_ij d_T_ij.  The combination needs R per species.
    std::vector<std::vector<double>> g(S, std::vector<double>(11 * static_cast<size_t>(N), 0.0));
    double F = 0.0;
    const int nthreads_loop = serial_pointwise ? 1 : 0;
    auto body = [&](long lo, long hi, double& Floc) {
      double comp = 0.0;
      for (long idx = lo; idx < hi; ++idx) {
        Meas tot{};
        for (long s = 0; s < S; ++s) {
          double R = 0.5 * diam[s];
          double fpr = 4.0 * PI * R;
          const double* w = wd[s].data();
          tot.eta += w[0 * N + idx];
          tot.n2 += w[1 * N + idx];
          tot.n1 += w[1 * N + idx] / fpr;
          tot.n0 += w[1 * N + idx] / (fpr * R);
          for (int a = 0; a < 3; ++a) {
            tot.v1[a] += w[(2 + a) * N + idx];
            tot.v0[a] += w[(2 + a) * N + idx] / fpr;
          }
          int c6 = 0;
          for (int i = 0; i < 3; ++i)
            for (int j = i; j < 3; ++j, ++c6) {
              double t = w[(5 + c6) * N + idx];
              tot.T[i * 3 + j] += t;
              if (i != j) tot.T[j * 3 + i] += t;
            }
        }
        if (tot.eta < 1e-30) continue;
        Meas dm{};
        double phi = phi_dphi(model, A, B, tot, &dm);
        // Neumaier-compensated sum of phi*dv
        double term = phi * dv, t = Floc + term;
        comp += std::fabs(Floc) >= std::fabs(term) ? (Floc - t) + term : (term - t) + Floc;
        Floc = t;
        for (long s = 0; s < S; ++s) {
          double R = 0.5 * diam[s];
          double fpr = 4.0 * PI * R;
          double* gs = g[s].data();
          gs[0 * N + idx] = dm.eta;
          gs[1 * N + idx] = dm.n2 + dm.n1 / fpr + dm.n0 / (fpr * R);
          for (int a = 0; a < 3; ++a) gs[(2 + a) * N + idx] = dm.v1[a] + dm.v0[a] / fpr;
          int c6 = 0;
          for (int i = 0; i < 3; ++i)
            for (int j = i; j < 3; ++j, ++c6) gs[(5 + c6) * N + idx] = (i == j ? 1.0 : 2.0) * dm.T[i * 3 + j];
        }
      }
      Floc += comp;
    };
    if (nthreads_loop == 1) {
      body(0, N, F);
    } else {
#pragma omp parallel
      {
        int nt = 1, tid = 0;
#ifdef _OPENMP
        nt = omp_get_num_threads();
        tid = omp_get_thread_num();
#endif
        long lo = N * tid / nt, hi = N * (tid + 1) / nt;
        double Floc = 0.0;
        body(lo, hi, Floc);
#pragma omp critical
        F += Floc;
      }
    }

    // back-convolution (hard_sphere.hpp:103-137,226-257)
    std::vector<cd> acc(static_cast<size_t>(M)), dk(static_cast<size_t>(M));
    std::vector<double> out(static_cast<size_t>(N));
    for (long s = 0; s < S; ++s) {
      std::fill(acc.begin(), acc.end(), cd(0.0, 0.0));
      for (int c = 0; c < 11; ++c) {
        fft->r2c(g[s].data() + c * N, dk.data());
        const cd* w = fmtw[s].data() + c * M;
        bool odd = (c >= 2 && c <= 4);
        for (long i = 0; i < M; ++i) acc[i] += (odd ? std::conj(w[i]) : w[i]) * dk[i];
      }
      fft->c2r(acc.data(), out.data());
      for (long i = 0; i < N; ++i) forces[s * N + i] += out[i] * dv;
    }
    return F;
  }

  // mean_field (mean_field.hpp:233-281); returns F_mf, adds forces; conv_out gets
  // the first interaction's convolution [w * rho_j] (without dV)
  double mean_field(const double* rho, double* forces, double* conv_out) const {
    const long S = static_cast<long>(diam.size());
    std::vector<std::vector<cd>> rk(S, std::vector<cd>(static_cast<size_t>(M)));
    for (long s = 0; s < S; ++s) fft->r2c(rho + s * N, rk[s].data());
    double F = 0.0;
    std::vector<double> conv(static_cast<size_t>(N));
    for (size_t k = 0; k < inters.size(); ++k) {
      const auto& in = inters[k];
      long i = in.si, j = in.sj;
      convolve(mfw[k].data(), rk[j].data(), conv.data());
      if (k == 0 && conv_out) std::memcpy(conv_out, conv.data(), sizeof(double) * N);
      double dot = 0.0;
      for (long p = 0; p < N; ++p) dot += rho[i * N + p] * conv[p];
      F += 0.5 * dot * dv;
      if (i == j) {
        for (long p = 0; p < N; ++p) forces[i * N + p] += conv[p] * dv;
      } else {
        for (long p = 0; p < N; ++p) forces[i * N + p] += 0.5 * conv[p] * dv;
        std::vector<double> convi(static_cast<size_t>(N));
        convolve(mfw[k].data(), rk[i].data(), convi.data());
        for (long p = 0; p < N; ++p) forces[j * N + p] += 0.5 * convi[p] * dv;
      }
    }
    return F;
  }

  // total (functionals.hpp:64-95).  parts = {F_id, F_ext, F_hs, F_mf, F, Omega}
  void total(const double* rho, const double* vext, const double* mu, double* parts, double* forces) const {
    const long S = static_cast<long>(diam.size());
    std::vector<double> f(static_cast<size_t>(S * N), 0.0);
    double Fid = 0.0, Fext = 0.0;
    // ideal_gas.hpp:18-38
    for (long s = 0; s < S; ++s) {
      double acc = 0.0;
      for (long p = 0; p < N; ++p) {
        double r = rho[s * N + p];
        double lr = std::log(std::max(r, 1e-300));
        acc += r * (lr - 1.0);
        f[s * N + p] += (lr - mu[s]) * dv;
      }
      Fid += acc * dv;
    }
    // external_field.hpp:16-31
    if (vext) {
      for (long s = 0; s < S; ++s) {
        double acc = 0.0;
        for (long p = 0; p < N; ++p) {
          acc += rho[s * N + p] * vext[s * N + p];
          f[s * N + p] += vext[s * N + p] * dv;
        }
        Fext += acc * dv;
      }
    }
    double Fhs = model >= 0 ? hard_sphere(rho, f.data(), nullptr) : 0.0;  // model -1: ideal gas (+ mean field) only
    double Fmf = inters.empty() ? 0.0 : mean_field(rho, f.data(), nullptr);
    double F = Fid + Fext + Fhs + Fmf;
    double muN = 0.0;
    for (long s = 0; s < S; ++s) {
      double acc = 0.0;
      for (long p = 0; p < N; ++p) acc += rho[s * N + p];
      muN += mu[s] * acc * dv;
    }
    if (parts) {
      parts[0] = Fid;
      parts[1] = Fext;
      parts[2] = Fhs;
      parts[3] = Fmf;
      parts[4] = F;
      parts[5] = F - muN;
    }
    if (forces) std::memcpy(forces, f.data(), sizeof(double) * S * N);
  }

  // bulk chemical potential of species 0 for a one-component fluid
  // (thermodynamics.hpp:222-237, a_vdw synced from the grid as in functional.hpp:96-107)
  double bulk_mu(double rho) const {
    double d = diam[0];
    Meas m = uniform_measures(rho, d);
    Meas dm{};
    if (model >= 0) phi_dphi(model, A, B, m, &dm);
    double r = 0.5 * d;
    double mu = std::log(rho);
    mu += dm.eta * (PI / 6.0) * d * d * d + dm.n2 * PI * d * d + dm.n1 * r + dm.n0;
    if (model >= 2)
      for (int j = 0; j < 3; ++j) mu += dm.T[j * 4] * PI * d * d / 3.0;
    for (size_t k = 0; k < inters.size(); ++k)
      if (inters[k].si == 0 && inters[k].sj == 0) mu += a_vdw[k] * rho;
    return mu;
  }
};

// ---------------------------------------------------------------------------
// FIRE2 (fire.hpp:33-165) and Minimizer::grand_potential (minimization.hpp)
// ---------------------------------------------------------------------------
struct FireParams {
  double dt, dt_max, dt_min, alpha_start, f_inc, f_dec, f_alpha;
  int n_delay, max_uphill;
  double force_tolerance;
  int max_steps;
  // > 0: the library's max-force stop (no counterpart in the reference)
  double force_max_tolerance = 0.0;
};

struct FireState {
  double dt, alpha;
  int n_positive, n_negative;
  double energy, rms_force;
  int iteration;
  bool converged;
};

// Fire::step (fire.hpp:86-165): one FIRE2 step of (x, v) from the forces f
// at x; compute(x, f) returns the energy at the new x and fills f there.
// Returns false when the uphill limit is exceeded (the reference throws).
template <typename Compute>
bool fire_step(const FireParams& fp, FireState& st, std::vector<double>& x, std::vector<double>& v,
               std::vector<double>& f, Compute&& compute) {
  const size_t n = x.size();
  double P = 0.0;
  for (size_t i = 0; i < n; ++i) P += v[i] * f[i];
  if (P > 0.0 || st.iteration == 0) {
    st.n_positive++;
    st.n_negative = 0;
    if (st.n_positive > fp.n_delay) {
      st.dt = std::min(st.dt * fp.f_inc, fp.dt_max);
      st.alpha *= fp.f_alpha;
    }
  } else {
    st.n_positive = 0;
    st.n_negative++;
    if (st.n_negative > fp.max_uphill) return false;
    for (size_t i = 0; i < n; ++i) {  // back up half a step with the current dt, stop
      x[i] -= 0.5 * st.dt * v[i];
      v[i] = 0.0;
    }
    if (st.iteration > fp.n_delay) {
      st.dt = std::max(st.dt * fp.f_dec, fp.dt_min);
      st.alpha = fp.alpha_start;
    }
  }
  for (size_t i = 0; i < n; ++i) v[i] += st.dt * f[i];
  double vn = 0.0, fn = 0.0;
  for (size_t i = 0; i < n; ++i) {
    vn += v[i] * v[i];
    fn += f[i] * f[i];
  }
  vn = std::sqrt(vn);
  fn = std::sqrt(fn);
  if (fn > 0.0) {
    const double ratio = vn / fn;
    for (size_t i = 0; i < n; ++i) v[i] = (1.0 - st.alpha) * v[i] + st.alpha * ratio * f[i];
  }
  for (size_t i = 0; i < n; ++i) x[i] += st.dt * v[i];
  st.energy = compute(x, f);
  double s2 = 0.0;
  for (size_t i = 0; i < n; ++i) s2 += f[i] * f[i];
  st.rms_force = std::sqrt(s2 / static_cast<double>(n));
  st.iteration++;
  st.converged = st.rms_force < fp.force_tolerance;
  return true;
}

}  // namespace orc

using namespace orc;

namespace {
int minimize(Ctx* c, const double* fire_p, int param_kind, double rho_min, double rho_max,
                    int homogeneous_boundary, const double* rho0, const double* vext, double mu, double target,
                    double* rho_out, double* out, double* energies) {
  const bool fixed = target > 0.0;  // Minimizer::fixed_mass; otherwise grand_potential
  const long N = c->N;
  // fire_p holds 12 values: the 11 FIRE parameters and the max-force stop
  FireParams fp{fire_p[0], fire_p[1], fire_p[2], fire_p[3], fire_p[4], fire_p[5], fire_p[6],
                static_cast<int>(fire_p[7]), static_cast<int>(fire_p[8]), fire_p[9], static_cast<int>(fire_p[10]),
                fire_p[11]};
  double range = rho_max - rho_min;
  double gmax = 0.0;  // max |force| / dV of the last compute (density space, what FIRE relaxes)
  auto to_density = [&](const std::vector<double>& x, std::vector<double>& rho) {
    for (long i = 0; i < N; ++i) {
      double x2 = x[i] * x[i];
      rho[i] = param_kind == 0 ? rho_min + x2 : rho_min + range * x2 / (1.0 + x2);
    }
  };
  // boundary mask (grid.hpp:67-82)
  std::vector<char> bd(static_cast<size_t>(N), 0);
  for (long ix = 0; ix < c->nx; ++ix)
    for (long iy = 0; iy < c->ny; ++iy)
      for (long iz = 0; iz < c->nz; ++iz)
        bd[iz + c->nz * (iy + c->ny * ix)] =
            (ix == 0 || ix == c->nx - 1 || iy == 0 || iy == c->ny - 1 || iz == 0 || iz == c->nz - 1);
  std::vector<double> rho(static_cast<size_t>(N)), grad(static_cast<size_t>(N));
  double parts[6];
  const double zero = 0.0;
  // fixed mass: rho rescaled to the target, evaluated at mu = 0 (minimization.hpp:193-206)
  auto rescale = [&]() {
    double m = 0.0;
    for (long i = 0; i < N; ++i) m += rho[i];
    m *= c->dv;
    if (m > 0.0)
      for (long i = 0; i < N; ++i) rho[i] *= target / m;
  };
  auto compute = [&](const std::vector<double>& x, std::vector<double>& f) -> double {
    to_density(x, rho);
    if (fixed) rescale();
    c->total(rho.data(), vext, fixed ? &zero : &mu, parts, grad.data());
    if (homogeneous_boundary) {  // grid.hpp:113-131
      double sum = 0.0;
      long cnt = 0;
      for (long i = 0; i < N; ++i)
        if (bd[i]) { sum += grad[i]; ++cnt; }
      if (cnt > 0) {
        double avg = sum / static_cast<double>(cnt);
        for (long i = 0; i < N; ++i)
          if (bd[i]) grad[i] = avg;
      }
    }
    if (fixed) {  // Lagrange multiplier of the mass (minimization.hpp:208-210)
      double d = 0.0;
      for (long i = 0; i < N; ++i) d += grad[i] * rho[i];
      const double lambda = d / target;
      for (long i = 0; i < N; ++i) grad[i] -= lambda * c->dv;
    }
    gmax = 0.0;
    for (long i = 0; i < N; ++i)
      gmax = std::max(gmax, std::isfinite(grad[i]) ? std::fabs(grad[i]) : std::numeric_limits<double>::infinity());
    gmax /= c->dv;
    for (long i = 0; i < N; ++i) {  // transform_force and negate (minimization.hpp:82-98,300-301)
      double dr;
      if (param_kind == 0) dr = 2.0 * x[i];
      else {
        double x2 = x[i] * x[i];
        dr = 2.0 * range * x[i] / ((1.0 + x2) * (1.0 + x2));
      }
      f[i] = -(grad[i] * dr);
    }
    return parts[5];
  };
  // from_density (minimization.hpp:63-77)
  std::vector<double> x(static_cast<size_t>(N)), v(static_cast<size_t>(N), 0.0), f(static_cast<size_t>(N));
  for (long i = 0; i < N; ++i) {
    if (param_kind == 0) x[i] = std::sqrt(std::max(rho0[i] - rho_min, 0.0));
    else {
      double d = std::min(std::max(rho0[i] - rho_min, 0.0), range * (1.0 - 1e-14));
      x[i] = std::sqrt(d / (range - d));
    }
  }
  // Fire::initialize (fire.hpp:60-84), then the extra compute at minimization.hpp:313
  FireState st{};
  st.energy = compute(x, f);
  double sf2 = 0.0;
  for (long i = 0; i < N; ++i) sf2 += f[i] * f[i];
  st.dt = fp.dt;
  st.alpha = fp.alpha_start;
  st.rms_force = std::sqrt(sf2 / static_cast<double>(N));
  st.converged = st.rms_force < fp.force_tolerance;
  compute(x, f);
  if (energies) energies[0] = st.energy;
  double volume = static_cast<double>(N) * c->dv;
  double vv = 1.0;
  bool converged = false;
  int rc = 0;
  for (int it = 0; it < fp.max_steps && !converged; ++it) {
    double old_e = st.energy;
    if (!fire_step(fp, st, x, v, f, compute)) {
      rc = -1;
      break;
    }
    // convergence_monitor (minimization.hpp:105-110)
    vv = vv * 0.9 + 0.1 * (std::fabs(st.energy - old_e) / st.dt);
    double monitor = std::fabs(vv / volume);
    converged = monitor < fp.force_tolerance;
    if (fp.force_max_tolerance > 0.0 && gmax < fp.force_max_tolerance) converged = true;
    st.converged = converged;
    if (energies) energies[it + 1] = st.energy;
  }
  to_density(x, rho);
  if (fixed) rescale();  // minimization.hpp:256-260
  std::memcpy(rho_out, rho.data(), sizeof(double) * N);
  double fparts[6];
  c->total(rho.data(), vext, &mu, fparts, nullptr);
  out[0] = fparts[4];
  out[1] = fparts[5];
  out[2] = st.rms_force;
  out[3] = st.iteration;
  out[4] = st.converged ? 1.0 : 0.0;
  out[5] = st.dt;
  out[6] = st.alpha;
  return rc;
}

}  // namespace

namespace {
// EigenvalueSolver::solve (saddle_point.hpp:159-270): LOBPCG for the
// smallest eigenpair of the grand-potential Hessian with the forward
// difference products of saddle_point.hpp:28-38 (forces of total() at mu;
// boundary forces zeroed when `fixed`, grid.hpp:135-142).  The projected
// 2x2 / 3x3 problem is solved by Jacobi rotations, eigenvector sign c0 >= 0
// (arma::eig_sym's sign is unspecified).  v0: start (normalised here).
// out: eigenvalue, iterations, converged.
void lobpcg(Ctx* c, double tol, int max_it, double eps, int fixed, const double* rho_in, const double* vext,
                   double mu, const double* v0, double* v_out, double* out) {
  const long N = c->N;
  std::vector<char> bd(static_cast<size_t>(N), 0);
  for (long ix = 0; ix < c->nx; ++ix)
    for (long iy = 0; iy < c->ny; ++iy)
      for (long iz = 0; iz < c->nz; ++iz)
        bd[iz + c->nz * (iy + c->ny * ix)] =
            (ix == 0 || ix == c->nx - 1 || iy == 0 || iy == c->ny - 1 || iz == 0 || iz == c->nz - 1);
  using V = std::vector<double>;
  const V rho(rho_in, rho_in + N);
  double parts[6];
  auto force = [&](const V& r, V& f) {
    c->total(r.data(), vext, &mu, parts, f.data());
    if (fixed)
      for (long i = 0; i < N; ++i)
        if (bd[i]) f[i] = 0.0;
  };
  V f0(N), f1(N), tmp(N);
  force(rho, f0);
  auto hvp = [&](const V& x, V& out) {
    for (long i = 0; i < N; ++i) tmp[i] = rho[i] + eps * x[i];
    force(tmp, f1);
    for (long i = 0; i < N; ++i) out[i] = (f1[i] - f0[i]) / eps;
  };
  auto dot = [&](const V& a, const V& b) {
    double s = 0.0;
    for (long i = 0; i < N; ++i) s += a[i] * b[i];
    return s;
  };
  auto eig = [&](int n, double A[3][3], double* cv) {  // Jacobi rotations
    double W[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 64; ++sweep) {
      double off = 0.0, dg = 0.0;
      for (int i = 0; i < n; ++i) {
        dg += A[i][i] * A[i][i];
        for (int j = i + 1; j < n; ++j) off += A[i][j] * A[i][j];
      }
      if (off <= 1e-32 * dg || off == 0.0) break;
      for (int p = 0; p < n; ++p)
        for (int q = p + 1; q < n; ++q) {
          if (A[p][q] == 0.0) continue;
          double th = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
          double t = (th >= 0.0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
          double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
          for (int k = 0; k < n; ++k) {
            double a = A[k][p], b = A[k][q];
            A[k][p] = cs * a - sn * b;
            A[k][q] = sn * a + cs * b;
          }
          for (int k = 0; k < n; ++k) {
            double a = A[p][k], b = A[q][k];
            A[p][k] = cs * a - sn * b;
            A[q][k] = sn * a + cs * b;
          }
          for (int k = 0; k < n; ++k) {
            double a = W[k][p], b = W[k][q];
            W[k][p] = cs * a - sn * b;
            W[k][q] = sn * a + cs * b;
          }
        }
    }
    int m = 0;
    for (int i = 1; i < n; ++i)
      if (A[i][i] < A[m][m]) m = i;
    double sg = W[0][m] < 0.0 ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) cv[i] = sg * W[i][m];
  };
  V v(v0, v0 + N), hv(N), r(N), q1(N), hq1(N), q2(N), hq2(N), p, vnew(N);
  double nv = std::sqrt(dot(v, v));
  for (double& x : v) x /= nv;
  hvp(v, hv);
  double lambda = dot(v, hv);
  bool conv = false;
  int it = 0;
  for (; it < max_it; ++it) {
    for (long i = 0; i < N; ++i) r[i] = hv[i] - lambda * v[i];
    if (std::sqrt(dot(r, r)) < tol) { conv = true; break; }
    double rq = dot(r, v);
    for (long i = 0; i < N; ++i) q1[i] = r[i] - rq * v[i];
    double q1n = std::sqrt(dot(q1, q1));
    if (q1n < 1e-30) break;
    for (double& x : q1) x /= q1n;
    hvp(q1, hq1);
    int dim = 2;
    if (!p.empty()) {
      double a = dot(p, v), b = dot(p, q1);
      for (long i = 0; i < N; ++i) q2[i] = p[i] - a * v[i] - b * q1[i];
      double q2n = std::sqrt(dot(q2, q2));
      if (q2n > 1e-10) {
        for (double& x : q2) x /= q2n;
        hvp(q2, hq2);
        dim = 3;
      }
    }
    double H[3][3] = {{dot(v, hv), dot(v, hq1), 0}, {0, dot(q1, hq1), 0}, {0, 0, 0}};
    H[1][0] = H[0][1];
    if (dim == 3) {
      H[0][2] = H[2][0] = dot(v, hq2);
      H[1][2] = H[2][1] = dot(q1, hq2);
      H[2][2] = dot(q2, hq2);
    }
    double cv[3] = {0, 0, 0};
    eig(dim, H, cv);
    for (long i = 0; i < N; ++i) vnew[i] = cv[0] * v[i] + cv[1] * q1[i] + (dim == 3 ? cv[2] * q2[i] : 0.0);
    p.resize(N);
    for (long i = 0; i < N; ++i) p[i] = vnew[i] - v[i];
    double vn = std::sqrt(dot(vnew, vnew));
    for (long i = 0; i < N; ++i) v[i] = vnew[i] / vn;
    hvp(v, hv);
    lambda = dot(v, hv);
  }
  std::memcpy(v_out, v.data(), sizeof(double) * N);
  out[0] = lambda;
  out[1] = it;
  out[2] = conv ? 1.0 : 0.0;
}

}  // namespace

extern "C" {

// potential params per interaction: kind, sigma, eps, rc, alpha, split, scheme
void* orc_create(long nx, long ny, long nz, double dx, int n_species, const double* diameters, int n_inter,
                 const int* inter_ij, const double* inter_params, double kT, int fmt_model, double esA, double esB) {
  auto* c = new Ctx(nx, ny, nz, dx);
  c->model = fmt_model;
  c->A = esA;
  c->B = esB;
  c->kT = kT;
  c->diam.assign(diameters, diameters + n_species);
  for (int k = 0; k < n_inter; ++k) {
    Interaction in;
    in.si = inter_ij[2 * k];
    in.sj = inter_ij[2 * k + 1];
    const double* p = inter_params + 7 * k;
    in.pot.kind = static_cast<int>(p[0]);
    in.pot.sigma = p[1];
    in.pot.eps = p[2];
    in.pot.rc = p[3];
    in.pot.alpha = p[4];
    in.split = static_cast<int>(p[5]);
    in.scheme = static_cast<int>(p[6]);
    in.pot.init();
    c->inters.push_back(in);
  }
  c->make_fmt_weights();
  c->make_mf_weights();
  return c;
}

void orc_destroy(void* h) { delete static_cast<Ctx*>(h); }

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#endif
}

int orc_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_serial_pointwise(void* h, int flag) { static_cast<Ctx*>(h)->serial_pointwise = flag != 0; }

double orc_a_vdw(void* h, int k) { return static_cast<Ctx*>(h)->a_vdw[k]; }

void orc_r2c(void* h, const double* in, double* out) {
  auto* c = static_cast<Ctx*>(h);
  c->fft->r2c(in, reinterpret_cast<cd*>(out));
}

void orc_c2r(void* h, const double* in, double* out) {
  auto* c = static_cast<Ctx*>(h);
  c->fft->c2r(reinterpret_cast<const cd*>(in), out);
}

void orc_fmt_weights(void* h, int s, double* out) {
  auto* c = static_cast<Ctx*>(h);
  std::memcpy(out, c->fmtw[s].data(), sizeof(cd) * 11 * c->M);
}

void orc_mf_weight_k(void* h, int k, double* out) {
  auto* c = static_cast<Ctx*>(h);
  std::memcpy(out, c->mfw[k].data(), sizeof(cd) * c->M);
}

void orc_mf_weight_real(void* h, int k, double* out) {
  auto* c = static_cast<Ctx*>(h);
  std::vector<std::vector<double>> re;
  auto saved_w = c->mfw;
  auto saved_a = c->a_vdw;
  c->make_mf_weights(&re);
  std::memcpy(out, re[k].data(), sizeof(double) * c->N);
  c->mfw = saved_w;
  c->a_vdw = saved_a;
}

// 12 fields: the 11 FMT weighted densities of species 0 and the first mean-field convolution
void orc_weighted_densities(void* h, const double* rho, double* out) {
  auto* c = static_cast<Ctx*>(h);
  std::vector<double> f(static_cast<size_t>(c->diam.size() * c->N), 0.0);
  c->hard_sphere(rho, f.data(), out);
  if (!c->inters.empty()) c->mean_field(rho, f.data(), out + 11 * c->N);
}

void orc_total(void* h, const double* rho, const double* vext, const double* mu, double* parts, double* forces) {
  static_cast<Ctx*>(h)->total(rho, vext, mu, parts, forces);
}

double orc_bulk_mu(void* h, double rho) { return static_cast<Ctx*>(h)->bulk_mu(rho); }

// F_hs = sum Phi dV from given weighted-density fields of species 0 (11*N, the
// channel order of orc_weighted_densities); the per-point loop of hard_sphere.hpp:174-224.
double orc_fhs_from_fields(void* h, const double* wd) {
  auto* c = static_cast<Ctx*>(h);
  const long N = c->N;
  const double R = 0.5 * c->diam[0], fpr = 4.0 * PI * R;
  double F = 0.0;
#pragma omp parallel for reduction(+ : F)
  for (long idx = 0; idx < N; ++idx) {
    Meas m{};
    m.eta = wd[idx];
    m.n2 = wd[N + idx];
    m.n1 = m.n2 / fpr;
    m.n0 = m.n2 / (fpr * R);
    for (int a = 0; a < 3; ++a) {
      m.v1[a] = wd[(2 + a) * N + idx];
      m.v0[a] = m.v1[a] / fpr;
    }
    int c6 = 0;
    for (int i = 0; i < 3; ++i)
      for (int j = i; j < 3; ++j, ++c6) {
        m.T[i * 3 + j] = wd[(5 + c6) * N + idx];
        m.T[j * 3 + i] = m.T[i * 3 + j];
      }
    if (m.eta < 1e-30) continue;
    F += phi_dphi(c->model, c->A, c->B, m, nullptr) * c->dv;
  }
  return F;
}

double orc_phi_dphi(int model, double A, double B, const double* m, double* dm) {
  Meas mm;
  std::memcpy(&mm, m, sizeof(Meas));
  Meas d{};
  double p = phi_dphi(model, A, B, mm, &d);
  if (dm) std::memcpy(dm, &d, sizeof(Meas));
  return p;
}

// which: 1,2,3 -> f1,f2,f3; returns f, writes df
double orc_f(int model, int which, double eta, double* df) {
  FVal v = which == 1 ? f1(model, eta) : which == 2 ? f2(model, eta) : f3(model, eta);
  if (df) *df = v.df;
  return v.f;
}

// ---------------------------------------------------------------------------
// Minimizer::grand_potential (minimization.hpp:280-353) with FIRE2 (fire.hpp).
// Single species.  param_kind 0: Unbounded{rho_min}; 1: Bounded{rho_min, rho_max}.
// vext may be null (the reference's from_profile zeroes the field).
// fire_p: dt, dt_max, dt_min, alpha_start, f_inc, f_dec, f_alpha, n_delay, max_uphill, force_tol, max_steps
// out: F, Omega, rms_force, iterations, converged, final dt, final alpha
// energies (optional, length max_steps+1): state.energy after init and after each step
// Returns 0, or -1 when FIRE exceeds max_uphill (the reference throws).
// ---------------------------------------------------------------------------
int orc_minimize_gp(void* h, const double* fire_p, int param_kind, double rho_min, double rho_max,
                    int homogeneous_boundary, const double* rho0, const double* vext, double mu, double* rho_out,
                    double* out, double* energies) {
  return minimize(static_cast<Ctx*>(h), fire_p, param_kind, rho_min, rho_max, homogeneous_boundary, rho0, vext, mu,
                  -1.0, rho_out, out, energies);
}

// Minimizer::fixed_mass (minimization.hpp:182-278): same arguments plus the
// target mass; out[] as for orc_minimize_gp (F and Omega of the rescaled
// final density at mu); energies receive F (mu = 0 in the loop).
int orc_minimize_fixed_mass(void* h, const double* fire_p, int param_kind, double rho_min, double rho_max,
                            int homogeneous_boundary, const double* rho0, const double* vext, double mu,
                            double target_mass, double* rho_out, double* out, double* energies) {
  return minimize(static_cast<Ctx*>(h), fire_p, param_kind, rho_min, rho_max, homogeneous_boundary, rho0, vext, mu,
                  target_mass, rho_out, out, energies);
}

// ---------------------------------------------------------------------------
// DDFT integrating-factor scheme (dynamics.hpp:79-401) and Simulation::run
// (dynamics.hpp:582-693).  Single species, force callback = total() at
// chemical potential mu with optional external field (the reference's
// grand_potential_callback zeroes it; vext = null reproduces that).
// cfg: dt, D (ignored, as in the reference), min_density, dt_max, fp_tolerance, fp_max_iterations
// run: n_steps.  out: mass_initial, mass_final, final dt, time, final energy
// energies (optional, n_steps+1): energy after each step
// ---------------------------------------------------------------------------
namespace {
// Restart limit of one integrating-factor step, as in the library.
constexpr int kMaxRestarts = 30;

struct DdftWork {
  Ctx* c;
  const double* vext;
  double mu;
  std::vector<double> lx, ly, lz, fx, fy, fz;
  double st_dt = 0.0;
  long evals = 0;

  DdftWork(Ctx* c_, const double* v, double m) : c(c_), vext(v), mu(m) {
    double Dx = 1.0 / (c->dx * c->dx);
    lx.resize(c->nx);
    ly.resize(c->ny);
    lz.resize(c->nzh);
    for (long i = 0; i < c->nx; ++i) lx[i] = 2.0 * Dx * (std::cos(2.0 * PI * i / c->nx) - 1.0);
    for (long i = 0; i < c->ny; ++i) ly[i] = 2.0 * Dx * (std::cos(2.0 * PI * i / c->ny) - 1.0);
    for (long i = 0; i < c->nzh; ++i) lz[i] = 2.0 * Dx * (std::cos(2.0 * PI * i / c->nz) - 1.0);
  }
  void update_timestep(double dt) {
    if (std::fabs(st_dt - dt) < 1e-30) return;
    st_dt = dt;
    fx.resize(lx.size());
    fy.resize(ly.size());
    fz.resize(lz.size());
    for (size_t i = 0; i < lx.size(); ++i) fx[i] = std::exp(lx[i] * dt);
    for (size_t i = 0; i < ly.size(); ++i) fy[i] = std::exp(ly[i] * dt);
    for (size_t i = 0; i < lz.size(); ++i) fz[i] = std::exp(lz[i] * dt);
  }
  double compute(const std::vector<double>& rho, std::vector<double>& forces) {
    double parts[6];
    c->total(rho.data(), vext, &mu, parts, forces.data());
    ++evals;
    return parts[5];
  }
  // divergence_flux + subtract_laplacian + r2c (dynamics.hpp:135-229)
  void excess_rhs_fourier(const std::vector<double>& rho, const std::vector<double>& forces, cd* out) {
    std::vector<double> rhs(c->N);
    excess_rhs_real(rho, forces, rhs);
    c->fft->r2c(rhs.data(), out);
  }
  // excess_rhs_real (dynamics.hpp:233-241)
  void excess_rhs_real(const std::vector<double>& rho, const std::vector<double>& forces, std::vector<double>& rhs) {
    const long nx = c->nx, ny = c->ny, nz = c->nz, N = c->N;
    std::vector<double> x(N);
    for (long i = 0; i < N; ++i) x[i] = forces[i] / c->dv;
    double D = 0.5 / (c->dx * c->dx), D2 = 1.0 / (c->dx * c->dx);
    long nyz = ny * nz;
    for (long ix = 0; ix < nx; ++ix) {
      long ixp = ((ix + 1) % nx) * nyz, ixm = ((ix - 1 + nx) % nx) * nyz, ix0 = ix * nyz;
      for (long iy = 0; iy < ny; ++iy) {
        long iyp = ((iy + 1) % ny) * nz, iym = ((iy - 1 + ny) % ny) * nz, iy0 = iy * nz;
        for (long iz = 0; iz < nz; ++iz) {
          long izp = (iz + 1) % nz, izm = (iz - 1 + nz) % nz;
          long p = ix0 + iy0 + iz;
          long px = ixp + iy0 + iz, mx = ixm + iy0 + iz, py = ix0 + iyp + iz, my = ix0 + iym + iz;
          long pz = ix0 + iy0 + izp, mz = ix0 + iy0 + izm;
          double d0 = rho[p], x0 = x[p];
          double g = D * ((rho[px] + d0) * (x[px] - x0) - (d0 + rho[mx]) * (x0 - x[mx])) +
                     D * ((rho[py] + d0) * (x[py] - x0) - (d0 + rho[my]) * (x0 - x[my])) +
                     D * ((rho[pz] + d0) * (x[pz] - x0) - (d0 + rho[mz]) * (x0 - x[mz]));
          g -= D2 * (rho[px] + rho[mx] - 2.0 * d0) + D2 * (rho[py] + rho[my] - 2.0 * d0) +
               D2 * (rho[pz] + rho[mz] - 2.0 * d0);
          rhs[p] = g;
        }
      }
    }
  }
  // compute_k_squared (dynamics.hpp:310-318): continuous |k|^2, Nyquist positive
  std::vector<double> k_squared() const {
    std::vector<double> k2(static_cast<size_t>(c->M));
    const double dkx = 2.0 * PI / (c->nx * c->dx), dky = 2.0 * PI / (c->ny * c->dx), dkz = 2.0 * PI / (c->nz * c->dx);
    long pos = 0;
    for (long ix = 0; ix < c->nx; ++ix)
      for (long iy = 0; iy < c->ny; ++iy)
        for (long iz = 0; iz < c->nzh; ++iz, ++pos) {
          const double kx = dkx * static_cast<double>(ix <= c->nx / 2 ? ix : ix - c->nx);
          const double ky = dky * static_cast<double>(iy <= c->ny / 2 ? iy : iy - c->ny);
          const double kz = dkz * static_cast<double>(iz);
          k2[pos] = kx * kx + ky * ky + kz * kz;
        }
    return k2;
  }
  // split_operator_step (dynamics.hpp:410-462)
  double split_step(std::vector<double>& rho, double D, double dt, double lo) {
    const long N = c->N, M = c->M;
    const std::vector<double> k2 = k_squared();
    std::vector<cd> fk(M);
    std::vector<double> f(N), rhs(N), tmp(N);
    const double inv_n = 1.0 / static_cast<double>(N);
    auto half = [&](const std::vector<double>& in, std::vector<double>& out) {
      c->fft->r2c(in.data(), fk.data());
      for (long i = 0; i < M; ++i) fk[i] *= std::sqrt(std::exp(-D * k2[i] * dt));
      c->fft->c2r(fk.data(), out.data());
      for (long i = 0; i < N; ++i) out[i] = std::max(out[i] * inv_n, lo);
    };
    std::vector<double> rs(N);
    half(rho, rs);
    compute(rs, f);
    excess_rhs_real(rs, f, rhs);
    for (long i = 0; i < N; ++i) tmp[i] = std::max(rs[i] + dt * rhs[i], lo);
    half(tmp, rho);
    return compute(rho, f);
  }
  // crank_nicholson_step (dynamics.hpp:472-538)
  double cn_step(std::vector<double>& rho, double D, double dt, double lo, double tol, int max_it, int* iters) {
    const long N = c->N, M = c->M;
    const std::vector<double> k2 = k_squared();
    std::vector<cd> p0(M), r0(M), r1(M), w(M);
    std::vector<double> f(N), rhs(N), d1 = rho, nw(N);
    c->fft->r2c(rho.data(), p0.data());
    compute(rho, f);
    excess_rhs_fourier(rho, f, r0.data());
    const double inv_n = 1.0 / static_cast<double>(N);
    int it = 0;
    for (; it < max_it; ++it) {
      compute(d1, f);
      excess_rhs_fourier(d1, f, r1.data());
      for (long i = 0; i < M; ++i) {
        const double a = D * k2[i] * dt * 0.5;
        w[i] = ((1.0 - a) * p0[i] + (dt * 0.5) * (r0[i] + r1[i])) * (1.0 / (1.0 + a));
      }
      c->fft->c2r(w.data(), nw.data());
      double md = 0.0;
      for (long i = 0; i < N; ++i) {
        const double v = std::max(nw[i] * inv_n, lo);
        md = std::max(md, std::fabs(v - d1[i]));
        d1[i] = v;
      }
      if (md < tol) {
        ++it;
        break;
      }
    }
    *iters = it;
    rho = d1;
    return compute(rho, f);
  }
  // apply_propagator (dynamics.hpp:247-304)
  double apply_propagator(const cd* rho0, const cd* r0, const cd* r1, std::vector<double>& d1) {
    const long nx = c->nx, ny = c->ny, nzh = c->nzh, M = c->M, N = c->N;
    double dt = st_dt;
    std::vector<cd> w(M);
    long pos = 0;
    for (long ix = 0; ix < nx; ++ix)
      for (long iy = 0; iy < ny; ++iy)
        for (long iz = 0; iz < nzh; ++iz, ++pos) {
          double Lam = lx[ix] + ly[iy] + lz[iz];
          double e = fx[ix] * fy[iy] * fz[iz];
          double U0 = e;
          double U1 = pos == 0 ? dt : (e - 1.0) / Lam;
          double U2 = pos == 0 ? dt / 2.0 : (e - 1.0 - dt * Lam) / (Lam * Lam * dt);
          w[pos] = U0 * rho0[pos] + U1 * r0[pos] + U2 * (r1[pos] - r0[pos]);
        }
    std::vector<double> re(N);
    c->fft->c2r(w.data(), re.data());
    double inv_n = 1.0 / static_cast<double>(N);
    double md = 0.0;
    for (long i = 0; i < N; ++i) {
      double nv = re[i] * inv_n;
      double dev = std::isfinite(nv) ? std::fabs(d1[i] - nv) : std::numeric_limits<double>::infinity();
      if (dev > md) md = dev;
      d1[i] = nv;
    }
    return md;
  }
  // integrating_factor_step (dynamics.hpp:333-401); returns the final energy
  double step(std::vector<double>& rho, double* cfg, double* dt_used) {
    const long N = c->N, M = c->M;
    std::vector<double> f(N);
    compute(rho, f);
    std::vector<cd> rhs0(M), rho0k(M), rhs1(M);
    excess_rhs_fourier(rho, f, rhs0.data());
    c->fft->r2c(rho.data(), rho0k.data());
    std::vector<double> d1 = rho;
    bool restart;
    int restarts = 0;
    do {
      restart = false;
      update_timestep(cfg[0]);
      d1 = rho;
      double dev = 1e30;
      for (int it = 0; it < static_cast<int>(cfg[5]); ++it) {
        compute(d1, f);
        excess_rhs_fourier(d1, f, rhs1.data());
        double old = dev;
        dev = apply_propagator(rho0k.data(), rhs0.data(), rhs1.data(), d1);
        double mn = *std::min_element(d1.begin(), d1.end());
        // the restatement adds the library's guard: a non-finite iterate ends the run
        if (!std::isfinite(dev) || !std::isfinite(mn)) throw std::runtime_error("non-finite iterate");
        if (mn < 0.0 || (it > 0 && dev > old)) { restart = true; break; }
        if (dev < cfg[4]) break;
      }
      if (!restart && dev > cfg[4]) restart = true;
      if (restart) {
        // the library's restart limit (the reference has none, dynamics.hpp:374-392)
        if (++restarts > kMaxRestarts) throw std::runtime_error("no convergence within the restart limit");
        cfg[0] /= 10.0;
        d1 = rho;
      }
    } while (restart);
    double e = compute(d1, f);
    rho = d1;
    *dt_used = cfg[0];
    return e;
  }
  // The map the adjoint differentiates: the step above with a fixed dt,
  // exactly P fixed-point iterations, no tolerance exit and no restart.
  void fixed_step(std::vector<double>& rho, double dt, int P, double* mn) {
    const long N = c->N, M = c->M;
    std::vector<double> f(N);
    compute(rho, f);
    std::vector<cd> rhs0(M), rho0k(M), rhs1(M);
    excess_rhs_fourier(rho, f, rhs0.data());
    c->fft->r2c(rho.data(), rho0k.data());
    update_timestep(dt);
    std::vector<double> d1 = rho;
    for (int it = 0; it < P; ++it) {
      compute(d1, f);
      excess_rhs_fourier(d1, f, rhs1.data());
      apply_propagator(rho0k.data(), rhs0.data(), rhs1.data(), d1);
      *mn = std::min(*mn, *std::min_element(d1.begin(), d1.end()));
    }
    rho = d1;
  }
};
}  // namespace

// Simulation::run (dynamics.hpp:582-693).  mask (nullable) is a
// reservoir_boundary (dynamics.hpp:545-553): the marked cells are reset to
// `reservoir` on the initial state and after every step.
int orc_ddft_run(void* h, const double* cfg_in, int n_steps, const double* vext, double mu, double* rho_inout,
                 double* out, double* energies, const unsigned char* mask, double reservoir) {
  auto* c = static_cast<Ctx*>(h);
  DdftWork w(c, vext, mu);
  double cfg[6];
  std::memcpy(cfg, cfg_in, sizeof(cfg));
  std::vector<double> rho(rho_inout, rho_inout + c->N);
  auto boundary = [&]() {
    if (mask)
      for (long i = 0; i < c->N; ++i)
        if (mask[i]) rho[i] = reservoir;
  };
  boundary();
  double m0 = 0.0;
  for (double r : rho) m0 += r;
  m0 *= c->dv;
  std::vector<double> f(c->N);
  double e0 = w.compute(rho, f);
  if (energies) energies[0] = e0;
  int succ = 0;
  double time = 0.0, e = e0;
  for (int s = 1; s <= n_steps; ++s) {
    double before = cfg[0], used = 0.0;
    try {
      e = w.step(rho, cfg, &used);
    } catch (const std::runtime_error&) {
      return -2;  // no convergence within the restart limit, or a non-finite iterate
    }
    time += used;
    boundary();
    if (used < before) succ = 0;
    else ++succ;
    if (succ >= 5 && cfg[0] < cfg[3]) {
      cfg[0] = std::min(2.0 * cfg[0], cfg[3]);
      succ = 0;
    }
    if (energies) energies[s] = e;
  }
  double m1 = 0.0;
  for (double r : rho) m1 += r;
  m1 *= c->dv;
  std::memcpy(rho_inout, rho.data(), sizeof(double) * c->N);
  out[0] = m0;
  out[1] = m1;
  out[2] = cfg[0];
  out[3] = time;
  out[4] = e;
  out[5] = static_cast<double>(w.evals);
  return 0;
}

void orc_smallest_eigenpair(void* h, double tol, int max_it, double eps, int fixed, const double* rho_in,
                            const double* vext, double mu, const double* v0, double* v_out, double* out) {
  lobpcg(static_cast<Ctx*>(h), tol, max_it, eps, fixed, rho_in, vext, mu, v0, v_out, out);
}

// SplitMix64 uniforms, Box-Muller normals (cosine branch first, then the sine branch).
static std::vector<double> seeded_normals(uint64_t seed, long n) {
  const long m = (n + 1) / 2;
  std::vector<double> u(static_cast<size_t>(2 * m)), z(static_cast<size_t>(n));
  for (long i = 0; i < 2 * m; ++i) {
    uint64_t x = seed + 0x9E3779B97F4A7C15ull * static_cast<uint64_t>(i + 1);
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    u[i] = static_cast<double>(x >> 11) * (1.0 / 9007199254740992.0);
  }
  for (long k = 0; k < m; ++k) {
    const double r = std::sqrt(-2.0 * std::log(std::max(u[2 * k], 1e-300))), a = 2.0 * PI * u[2 * k + 1];
    if (k < n) z[k] = r * std::cos(a);
    if (m + k < n) z[m + k] = r * std::sin(a);
  }
  return z;
}

// Search::find (saddle_point.hpp:367-589).  cfg: fire[11] (as orc_minimize_gp),
// eigen tolerance, max iterations, hessian eps, param kind, rho_min, rho_max,
// omega tolerance, tolerance start, tolerance factor, max outer iterations,
// eigenvalue clear count.  The start of outer iteration k's eigen solve is
// the seeded normal sequence of seed + k until the clear count.
// out: Omega, eigenvalue, outer iterations (as the reference counts them), converged.
int orc_saddle_search(void* h, const double* cfg, const double* rho0, const double* vext, double mu,
                      double target_mass, unsigned long long seed, double* rho_out, double* eig_out, double* out) {
  auto* c = static_cast<Ctx*>(h);
  const long N = c->N;
  FireParams fp{cfg[0], cfg[1], cfg[2], cfg[3], cfg[4], cfg[5], cfg[6], static_cast<int>(cfg[7]),
                static_cast<int>(cfg[8]), cfg[9], static_cast<int>(cfg[10])};
  const double etol = cfg[11], eps = cfg[13];
  const int emax = static_cast<int>(cfg[12]);
  const int kind = static_cast<int>(cfg[14]);
  const double rho_min = cfg[15], range = cfg[16] - cfg[15];
  const double omega_tol = cfg[17], tol_factor = cfg[19];
  double ftol = cfg[18];
  const int max_outer = static_cast<int>(cfg[20]), clear = static_cast<int>(cfg[21]);
  std::vector<char> bd(static_cast<size_t>(N), 0);
  long nbd = 0;
  for (long ix = 0; ix < c->nx; ++ix)
    for (long iy = 0; iy < c->ny; ++iy)
      for (long iz = 0; iz < c->nz; ++iz) {
        const bool b = ix == 0 || ix == c->nx - 1 || iy == 0 || iy == c->ny - 1 || iz == 0 || iz == c->nz - 1;
        bd[iz + c->nz * (iy + c->ny * ix)] = b;
        nbd += b;
      }
  auto to_rho = [&](double x) { const double x2 = x * x; return kind == 0 ? rho_min + x2 : rho_min + range * x2 / (1.0 + x2); };
  auto drho = [&](double x) {
    if (kind == 0) return 2.0 * x;
    const double x2 = x * x;
    return 2.0 * range * x / ((1.0 + x2) * (1.0 + x2));
  };
  auto from_rho = [&](double r) {
    if (kind == 0) return std::sqrt(std::max(r - rho_min, 0.0));
    const double d = std::min(std::max(r - rho_min, 0.0), range * (1.0 - 1e-14));
    return std::sqrt(d / (range - d));
  };
  std::vector<double> rho(rho0, rho0 + N), eig(N), dir(N), grad(N), rl(N);
  double omega = 0.0, lambda = 0.0, parts[6];
  bool converged = false, have = false;
  int outer = 0;
  for (; outer < max_outer; ++outer) {
    const double omega_old = omega;
    std::vector<double> v0 = (outer >= clear && have) ? eig : seeded_normals(seed + static_cast<uint64_t>(outer), N);
    double eo[3];
    lobpcg(c, etol, emax, eps, 1, rho.data(), vext, mu, v0.data(), eig.data(), eo);
    lambda = eo[0];
    have = true;
    std::vector<double> x(N);
    for (long i = 0; i < N; ++i) x[i] = from_rho(rho[i]);
    double dn = 0.0;
    for (long i = 0; i < N; ++i) {
      dir[i] = eig[i] * drho(x[i]);
      dn += dir[i] * dir[i];
    }
    dn = std::sqrt(dn);
    if (dn > 1e-30)
      for (double& d : dir) d /= dn;
    double sb = 0.0;
    for (long i = 0; i < N; ++i)
      if (bd[i]) sb += rho[i];
    const double m_boundary = (nbd > 0 ? sb / static_cast<double>(nbd) : 0.0) * c->dv * static_cast<double>(nbd);
    const double mt = target_mass - m_boundary;
    auto compute = [&](const std::vector<double>& xx, std::vector<double>& f) {
      double mi = 0.0;
      for (long i = 0; i < N; ++i) {
        rl[i] = to_rho(xx[i]);
        if (!bd[i]) mi += rl[i];
      }
      mi *= c->dv;
      if (mi > 0.0 && mt > 0.0)
        for (long i = 0; i < N; ++i)
          if (!bd[i]) rl[i] *= mt / mi;
      c->total(rl.data(), vext, &mu, parts, grad.data());
      double muL = 0.0;
      for (long i = 0; i < N; ++i)
        if (!bd[i]) muL += grad[i] * rl[i];
      if (mt > 0.0) muL /= mt;
      double pj = 0.0;
      for (long i = 0; i < N; ++i) {
        const double gi = bd[i] ? 0.0 : grad[i] - muL * c->dv;
        f[i] = gi * drho(xx[i]);
        pj += dir[i] * f[i];
      }
      for (long i = 0; i < N; ++i) f[i] = -(f[i] - pj * dir[i]);
      return parts[5];
    };
    FireParams fc = fp;
    fc.force_tolerance = std::max(ftol, fp.force_tolerance);
    ftol /= tol_factor;
    // Fire::initialize + the extra compute, then Fire::step until the rms force test holds (fire.hpp:60-165)
    std::vector<double> v(N, 0.0), f(N);
    double energy = compute(x, f), dt = fc.dt, alpha = fc.alpha_start;
    auto rms = [&]() {
      double s2 = 0.0;
      for (double q : f) s2 += q * q;
      return std::sqrt(s2 / static_cast<double>(N));
    };
    bool fconv = rms() < fc.force_tolerance;
    int npos = 0, nneg = 0, iter = 0;
    for (int it = 0; it < fc.max_steps && !fconv; ++it) {
      double P = 0.0;
      for (long i = 0; i < N; ++i) P += v[i] * f[i];
      if (P > 0.0 || iter == 0) {
        ++npos;
        nneg = 0;
        if (npos > fc.n_delay) {
          dt = std::min(dt * fc.f_inc, fc.dt_max);
          alpha *= fc.f_alpha;
        }
      } else {
        npos = 0;
        if (++nneg > fc.max_uphill) return -1;
        for (long i = 0; i < N; ++i) {
          x[i] -= 0.5 * dt * v[i];
          v[i] = 0.0;
        }
        if (iter > fc.n_delay) {
          dt = std::max(dt * fc.f_dec, fc.dt_min);
          alpha = fc.alpha_start;
        }
      }
      for (long i = 0; i < N; ++i) v[i] += dt * f[i];
      double vn = 0.0, fn = 0.0;
      for (long i = 0; i < N; ++i) {
        vn += v[i] * v[i];
        fn += f[i] * f[i];
      }
      if (fn > 0.0) {
        const double ratio = std::sqrt(vn) / std::sqrt(fn);
        for (long i = 0; i < N; ++i) v[i] = (1.0 - alpha) * v[i] + alpha * ratio * f[i];
      }
      for (long i = 0; i < N; ++i) x[i] += dt * v[i];
      energy = compute(x, f);
      ++iter;
      fconv = rms() < fc.force_tolerance;
    }
    (void)energy;
    rho = rl;  // the last compute's density: to_density(x) with the interior rescaled
    c->total(rho.data(), vext, &mu, parts, nullptr);
    omega = parts[5];
    if (outer > 0 && std::fabs(omega - omega_old) < omega_tol) {
      converged = true;
      break;
    }
  }
  std::memcpy(rho_out, rho.data(), sizeof(double) * N);
  if (eig_out) std::memcpy(eig_out, eig.data(), sizeof(double) * N);
  out[0] = omega;
  out[1] = lambda;
  out[2] = outer + 1;
  out[3] = converged ? 1.0 : 0.0;
  return 0;
}

// split_operator_step (dynamics.hpp:405-462) n_steps times; cfg: dt, D, min_density.  Returns the final energy.
double orc_ddft_split_operator(void* h, const double* cfg, int n_steps, const double* vext, double mu,
                               double* rho_inout) {
  auto* c = static_cast<Ctx*>(h);
  DdftWork w(c, vext, mu);
  std::vector<double> rho(rho_inout, rho_inout + c->N);
  double e = 0.0;
  for (int s = 0; s < n_steps; ++s) e = w.split_step(rho, cfg[1], cfg[0], cfg[2]);
  std::memcpy(rho_inout, rho.data(), sizeof(double) * c->N);
  return e;
}

// crank_nicholson_step (dynamics.hpp:464-538) n_steps times; cfg: dt, D, min_density, tolerance, max iterations.
// out: final energy, iterations of the last step.
void orc_ddft_crank_nicolson(void* h, const double* cfg, int n_steps, const double* vext, double mu,
                             double* rho_inout, double* out) {
  auto* c = static_cast<Ctx*>(h);
  DdftWork w(c, vext, mu);
  std::vector<double> rho(rho_inout, rho_inout + c->N);
  double e = 0.0;
  int it = 0;
  for (int s = 0; s < n_steps; ++s) e = w.cn_step(rho, cfg[1], cfg[0], cfg[2], cfg[3], static_cast<int>(cfg[4]), &it);
  std::memcpy(rho_inout, rho.data(), sizeof(double) * c->N);
  out[0] = e;
  out[1] = it;
}

// Picard::solve (picard.hpp:63-124); cfg: mixing, min_density, tolerance, max iterations.
// target_mass > 0 applies fixed_mass_constraint (minimization.hpp:157-180) with background (nullable: 0).
// out: Omega, residual, iterations, converged.
// target (nullable: no constraint) is fixed_mass_constraint's target mass of
// the excess over bg, applied for any sign (minimization.hpp:159-180).
void orc_picard(void* h, const double* cfg, const double* vext, double mu, const double* target, const double* bg,
                double* rho, double* out) {
  auto* c = static_cast<Ctx*>(h);
  const long N = c->N;
  const double mixing = cfg[0], lo = cfg[1], tol = cfg[2];
  const int max_it = static_cast<int>(cfg[3]);
  std::vector<double> f(N);
  double parts[6], omega = 0.0, omega_prev = std::numeric_limits<double>::max(), residual = 0.0;
  int it = 0;
  for (; it < max_it; ++it) {
    c->total(rho, vext, &mu, parts, f.data());
    omega = parts[5];
    double s2 = 0.0;
    for (long i = 0; i < N; ++i) s2 += f[i] * f[i];
    residual = std::sqrt(s2 / static_cast<double>(N));
    if (residual < tol || (it > 0 && std::fabs(omega - omega_prev) < tol)) break;
    omega_prev = omega;
    for (long i = 0; i < N; ++i) rho[i] = std::max(rho[i] * std::exp(-mixing * f[i] / c->dv), lo);
    if (target) {
      const double target_mass = *target;
      double cur = 0.0;
      for (long i = 0; i < N; ++i) cur += rho[i] - (bg ? bg[i] : 0.0);
      cur *= c->dv;
      if (std::fabs(cur) > 1e-30)
        for (long i = 0; i < N; ++i) {
          const double b = bg ? bg[i] : 0.0;
          rho[i] = std::max(b + (target_mass / cur) * (rho[i] - b), lo);
        }
    }
  }
  out[0] = omega;
  out[1] = residual;
  out[2] = it;
  out[3] = (residual < tol || (it > 0 && std::fabs(omega - omega_prev) < tol)) ? 1.0 : 0.0;
}

// Fire::initialize then `steps` Fire::step calls (fire.hpp:60-165) over any
// force function: force(user, x, f) returns the energy at x and fills the
// forces f.  trace (steps + 1 rows of 2n + 6 values: x, v, dt, alpha,
// n_positive, n_negative, energy, rms force) records the state after
// initialisation and after every step.  Returns the steps taken, or
// -(steps taken + 1) when the uphill limit was exceeded.
int orc_fire_trace(const double* fire_p, int n, const double* x0, int steps,
                   double (*force)(void* user, const double* x, double* f), void* user, double* trace) {
  FireParams fp{fire_p[0], fire_p[1], fire_p[2], fire_p[3], fire_p[4], fire_p[5], fire_p[6],
                static_cast<int>(fire_p[7]), static_cast<int>(fire_p[8]), fire_p[9], static_cast<int>(fire_p[10])};
  std::vector<double> x(x0, x0 + n), v(static_cast<size_t>(n), 0.0), f(static_cast<size_t>(n));
  auto compute = [&](const std::vector<double>& xs, std::vector<double>& fs) { return force(user, xs.data(), fs.data()); };
  FireState st{};
  st.energy = compute(x, f);
  double s2 = 0.0;
  for (double g : f) s2 += g * g;
  st.dt = fp.dt;
  st.alpha = fp.alpha_start;
  st.rms_force = std::sqrt(s2 / static_cast<double>(n));
  const size_t row = 2 * static_cast<size_t>(n) + 6;
  auto record = [&](int k) {
    double* r = trace + row * static_cast<size_t>(k);
    for (int i = 0; i < n; ++i) {
      r[i] = x[static_cast<size_t>(i)];
      r[n + i] = v[static_cast<size_t>(i)];
    }
    r[2 * n] = st.dt;
    r[2 * n + 1] = st.alpha;
    r[2 * n + 2] = st.n_positive;
    r[2 * n + 3] = st.n_negative;
    r[2 * n + 4] = st.energy;
    r[2 * n + 5] = st.rms_force;
  };
  record(0);
  for (int k = 0; k < steps; ++k) {
    if (!fire_step(fp, st, x, v, f, compute)) return -(k + 1);
    record(k + 1);
  }
  return steps;
}

// Fixed-dt forward map of the DDFT adjoint: n_steps steps of fixed_step.
// Returns the smallest density of any iterate.
double orc_ddft_fixed_run(void* h, double dt, int fp_iterations, int n_steps, const double* vext, double mu,
                          double* rho_inout) {
  auto* c = static_cast<Ctx*>(h);
  DdftWork w(c, vext, mu);
  std::vector<double> rho(rho_inout, rho_inout + c->N);
  double mn = 1e300;
  for (int s = 0; s < n_steps; ++s) w.fixed_step(rho, dt, fp_iterations, &mn);
  std::memcpy(rho_inout, rho.data(), sizeof(double) * c->N);
  return mn;
}

// Sum of Gaussian wells V(r) = sum_j A_j exp(-|r - c_j|^2 / (2 w_j^2)) at
// r = (ix, iy, iz) dx with minimum-image distances; params: (A, w, cx, cy, cz) per well.
void orc_wells_potential(void* h, int n_wells, const double* params, double* out) {
  auto* c = static_cast<Ctx*>(h);
  const double L[3] = {c->nx * c->dx, c->ny * c->dx, c->nz * c->dx};
  for (long ix = 0; ix < c->nx; ++ix)
    for (long iy = 0; iy < c->ny; ++iy)
      for (long iz = 0; iz < c->nz; ++iz) {
        const double r[3] = {ix * c->dx, iy * c->dx, iz * c->dx};
        double v = 0.0;
        for (int j = 0; j < n_wells; ++j) {
          const double* q = params + 5 * j;
          double d2 = 0.0;
          for (int a = 0; a < 3; ++a) {
            double e = r[a] - q[2 + a];
            e -= L[a] * std::nearbyint(e / L[a]);
            d2 += e * e;
          }
          v += q[0] * std::exp(-d2 / (2.0 * q[1] * q[1]));
        }
        out[iz + c->nz * (iy + c->ny * ix)] = v;
      }
}

}  // extern "C"

// Smallest eigenpair of the Hessian of the grand potential by the LOBPCG
// iteration of EigenvalueSolver::solve
// (include/dft/algorithms/saddle_point.hpp:136-270), with the density, the
// search vectors and every Hessian-vector product kept in HBM.  Per
// iteration: the residual r = Hv - lambda v, the basis {v, r orthonormalised,
// the previous step p orthonormalised}, the 2x2 or 3x3 projected Hessian, its
// lowest eigenvector c (host), v <- normalised sum c_i q_i.
//
// Hessian-vector products:
//   hessian_eps > 0: the reference's forward difference
//                    (F(rho + eps v) - F(rho)) / eps of the forces F
//                    (saddle_point.hpp:28-38), through the fused evaluation;
//   hessian_eps = 0: the exact product (dual-number FMT, FP64 contexts).
// With fixed_boundary the boundary components of F are zeroed, as the force
// function of Search::find does (saddle_point.hpp:390-397, grid.hpp:135-142).
//
// Two choices the reference leaves to its libraries are fixed here: the
// projected eigenproblem is solved by cyclic Jacobi rotations with the
// eigenvector's sign chosen so that c_0 >= 0 (LAPACK's eig_sym leaves the sign
// unspecified; c_0 >= 0 keeps v from flipping between iterations), and the
// random start (arma::randn) is the project's SplitMix64 / Box-Muller normal
// sequence of a given seed.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "context.hpp"
#include "ddft.hpp"
#include "hessian.hpp"

namespace cdft {

// Kernels and helpers with internal linkage: each translation unit that
// includes this header gets its own copy.
namespace {

constexpr int kMaxDots = 8;

struct DotPairs {
  const double* a[kMaxDots];
  const double* b[kMaxDots];
  int n;
};

// Compensated per-block sums of a_k . b_k for up to 8 pairs into partials[blk * 8 + k].
__global__ void __launch_bounds__(256) k_dots(DotPairs p, long n, double* partials) {
  double s[kMaxDots], c[kMaxDots];
#pragma unroll
  for (int k = 0; k < kMaxDots; ++k) s[k] = c[k] = 0.0;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
#pragma unroll
    for (int k = 0; k < kMaxDots; ++k)
      if (k < p.n) acc2(s[k], c[k], p.a[k][i] * p.b[k][i]);
  }
  __shared__ double sh[kMaxDots][256];
#pragma unroll
  for (int k = 0; k < kMaxDots; ++k) sh[k][threadIdx.x] = s[k] + c[k];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < kMaxDots; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < kMaxDots) partials[static_cast<long>(blockIdx.x) * 8 + threadIdx.x] = sh[threadIdx.x][0];
}

// out = a0 x0 + a1 x1 + a2 x2 (x1, x2 nullable; out may be one of the inputs)
__global__ void __launch_bounds__(256) k_lincomb(double* out, double a0, const double* x0, double a1,
                                                 const double* x1, double a2, const double* x2, long n) {
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    double v = a0 * x0[i];
    if (x1) v += a1 * x1[i];
    if (x2) v += a2 * x2[i];
    out[i] = v;
  }
}

// out = (f1 - f0) / eps, boundary components 0 when fixed
__global__ void __launch_bounds__(256) k_fd_hvp(double* __restrict__ out, const double* __restrict__ f1,
                                                const double* __restrict__ f0, double eps, Geom g, int fixed) {
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < g.N;
       i += static_cast<long>(gridDim.x) * blockDim.x)
    out[i] = (fixed && on_face(i, g)) ? 0.0 : (f1[i] - f0[i]) / eps;
}

__global__ void __launch_bounds__(256) k_zero_boundary(double* __restrict__ f, Geom g) {
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < g.N;
       i += static_cast<long>(gridDim.x) * blockDim.x)
    if (on_face(i, g)) f[i] = 0.0;
}

// Lowest eigenpair of a symmetric n x n matrix (n <= 3) by cyclic Jacobi
// rotations; the eigenvector is normalised with c[0] >= 0.
void lowest_eigenpair(int n, const double (&Ain)[3][3], double* c, double* lambda) {
  double A[3][3], V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A[i][j] = Ain[i][j];
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < n; ++i) {
      diag += A[i][i] * A[i][i];
      for (int j = i + 1; j < n; ++j) off += A[i][j] * A[i][j];
    }
    if (off <= 1e-32 * diag || off == 0.0) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        if (A[p][q] == 0.0) continue;
        const double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < n; ++k) {  // A <- A J
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = cs * akp - sn * akq;
          A[k][q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < n; ++k) {  // A <- J^T A
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = cs * apk - sn * aqk;
          A[q][k] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < n; ++k) {  // V <- V J
          const double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = cs * vkp - sn * vkq;
          V[k][q] = sn * vkp + cs * vkq;
        }
      }
  }
  int m = 0;
  for (int i = 1; i < n; ++i)
    if (A[i][i] < A[m][m]) m = i;
  const double sg = V[0][m] < 0.0 ? -1.0 : 1.0;
  for (int i = 0; i < n; ++i) c[i] = sg * V[i][m];
  *lambda = A[m][m];
}

// SplitMix64 uniforms and Box-Muller normals: n values, the cosine branch
// for the first ceil(n/2), then the sine branch (the project's seeded
// standard-normal sequence).
std::vector<double> seeded_normals(uint64_t seed, long n) {
  const long m = (n + 1) / 2;
  std::vector<double> u(static_cast<size_t>(2 * m));
  for (long i = 0; i < 2 * m; ++i) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * static_cast<uint64_t>(i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    u[static_cast<size_t>(i)] = static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
  }
  std::vector<double> out(static_cast<size_t>(n));
  const double PI = 3.14159265358979323846;
  for (long k = 0; k < m; ++k) {
    const double r = std::sqrt(-2.0 * std::log(std::max(u[static_cast<size_t>(2 * k)], 1e-300)));
    const double a = 2.0 * PI * u[static_cast<size_t>(2 * k + 1)];
    if (k < n) out[static_cast<size_t>(k)] = r * std::cos(a);
    if (m + k < n) out[static_cast<size_t>(m + k)] = r * std::sin(a);
  }
  return out;
}

struct EigenRun {
  Context* c;
  const double* d_vext;
  double mu, eps;
  int fixed;
  long N;
  DevBuf rho, f0, f1, tmp, v, hv, r, q1, hq1, q2, hq2, p, part, red;
  Hessian* exact = nullptr;
  int products = 0;

  EigenRun(Context* c_, const double* vx, double m, double e, int fb)
      : c(c_), d_vext(vx), mu(m), eps(e), fixed(fb), N(c_->g.N), rho(8 * N), f0(8 * N), f1(8 * N), tmp(8 * N),
        v(8 * N), hv(8 * N), r(8 * N), q1(8 * N), hq1(8 * N), q2(8 * N), hq2(8 * N), p(8 * N),
        part(8 * 8 * kVecBlocks), red(8 * 8) {}

  dim3 blocks() const { return dim3(vec_blocks(N)); }

  std::vector<double> dots(std::initializer_list<std::pair<const double*, const double*>> ps) {
    DotPairs dp{};
    dp.n = 0;
    for (const auto& q : ps) {
      dp.a[dp.n] = q.first;
      dp.b[dp.n] = q.second;
      ++dp.n;
    }
    hipLaunchKernelGGL(k_dots, blocks(), dim3(256), 0, c->stream, dp, N, part.as<double>());
    launch_finalize(part.as<double>(), static_cast<int>(vec_blocks(N)), red.as<double>(), c->stream);
    CDFT_HIP(hipGetLastError());
    double h[8];
    CDFT_HIP(hipMemcpyAsync(h, red.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    CDFT_HIP(hipStreamSynchronize(c->stream));
    return std::vector<double>(h, h + dp.n);
  }
  void lincomb(double* out, double a0, const double* x0, double a1 = 0.0, const double* x1 = nullptr,
               double a2 = 0.0, const double* x2 = nullptr) {
    hipLaunchKernelGGL(k_lincomb, blocks(), dim3(256), 0, c->stream, out, a0, x0, a1, x1, a2, x2, N);
  }

  // forces at rho (saddle_point.hpp:390-397)
  void base_forces() {
    c->eval(rho.as<double>(), d_vext, mu, f0.as<double>(), AliasMap{0, 0.0, 0.0});
    if (fixed) hipLaunchKernelGGL(k_zero_boundary, blocks(), dim3(256), 0, c->stream, f0.as<double>(), c->g);
  }

  void hvp(const double* x, double* out) {
    ++products;
    if (eps > 0.0) {  // hessian_times_vector (saddle_point.hpp:28-38)
      lincomb(tmp.as<double>(), 1.0, rho.as<double>(), eps, x);
      c->eval(tmp.as<double>(), d_vext, mu, f1.as<double>(), AliasMap{0, 0.0, 0.0});
      hipLaunchKernelGGL(k_fd_hvp, blocks(), dim3(256), 0, c->stream, out, f1.as<double>(), f0.as<double>(), eps,
                         c->g, fixed);
    } else {
      exact->apply(rho.as<double>(), x, out);
      if (fixed) hipLaunchKernelGGL(k_zero_boundary, blocks(), dim3(256), 0, c->stream, out, c->g);
    }
    CDFT_HIP(hipGetLastError());
  }

  // EigenvalueSolver::solve (saddle_point.hpp:159-270); v holds the normalised start on entry
  void solve(double tol, int max_it, double* lambda_out, int* iters, int* converged) {
    double* V = v.as<double>();
    double* HV = hv.as<double>();
    hvp(V, HV);
    double lambda = dots({{V, HV}})[0];
    bool have_p = false, conv = false;
    int it = 0;
    for (; it < max_it; ++it) {
      lincomb(r.as<double>(), 1.0, HV, -lambda, V);  // residual
      const auto d1 = dots({{r.as<double>(), r.as<double>()}, {r.as<double>(), V}});
      if (std::sqrt(d1[0]) < tol) {
        conv = true;
        break;
      }
      lincomb(q1.as<double>(), 1.0, r.as<double>(), -d1[1], V);
      const double q1n = std::sqrt(dots({{q1.as<double>(), q1.as<double>()}})[0]);
      if (q1n < 1e-30) break;
      lincomb(q1.as<double>(), 1.0 / q1n, q1.as<double>());
      hvp(q1.as<double>(), hq1.as<double>());
      int dim = 2;
      if (have_p) {
        const auto d2 = dots({{p.as<double>(), V}, {p.as<double>(), q1.as<double>()}});
        lincomb(q2.as<double>(), 1.0, p.as<double>(), -d2[0], V, -d2[1], q1.as<double>());
        const double q2n = std::sqrt(dots({{q2.as<double>(), q2.as<double>()}})[0]);
        if (q2n > 1e-10) {
          lincomb(q2.as<double>(), 1.0 / q2n, q2.as<double>());
          hvp(q2.as<double>(), hq2.as<double>());
          dim = 3;
        }
      }
      double H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      if (dim == 2) {
        const auto h = dots({{V, HV}, {V, hq1.as<double>()}, {q1.as<double>(), hq1.as<double>()}});
        H[0][0] = h[0];
        H[0][1] = H[1][0] = h[1];
        H[1][1] = h[2];
      } else {
        const auto h = dots({{V, HV}, {V, hq1.as<double>()}, {q1.as<double>(), hq1.as<double>()},
                             {V, hq2.as<double>()}, {q1.as<double>(), hq2.as<double>()},
                             {q2.as<double>(), hq2.as<double>()}});
        H[0][0] = h[0];
        H[0][1] = H[1][0] = h[1];
        H[1][1] = h[2];
        H[0][2] = H[2][0] = h[3];
        H[1][2] = H[2][1] = h[4];
        H[2][2] = h[5];
      }
      double cv[3] = {0, 0, 0}, ev = 0.0;
      lowest_eigenpair(dim, H, cv, &ev);
      // v_new = sum c_i q_i (into tmp); p = v_new - v; v = v_new / |v_new|
      lincomb(tmp.as<double>(), cv[0], V, cv[1], q1.as<double>(), dim == 3 ? cv[2] : 0.0,
              dim == 3 ? q2.as<double>() : nullptr);
      lincomb(p.as<double>(), 1.0, tmp.as<double>(), -1.0, V);
      have_p = true;
      const double vn = std::sqrt(dots({{tmp.as<double>(), tmp.as<double>()}})[0]);
      lincomb(V, 1.0 / vn, tmp.as<double>());
      hvp(V, HV);
      lambda = dots({{V, HV}})[0];
    }
    *lambda_out = lambda;
    *iters = it;
    *converged = conv ? 1 : 0;
  }
};

}  // namespace

}  // namespace cdft

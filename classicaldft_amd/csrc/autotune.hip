// Autotuner of the general evaluation schedule (grids the fused kernels do
// not cover, e.g. 48^3, 80^3, 96^3, 160^3): the radix order of each axis'
// FFT and the launch shapes of its kernels (lines per block of the z and y
// line transforms, y columns per block of the x kernels, z lines per block of
// the pointwise kernel).
//
// One coordinate-descent pass: each parameter in turn takes each of its
// candidate values with the others fixed, and keeps the fastest.  A
// candidate is timed with HIP events over `reps` evaluations (after one
// untimed) on a smooth synthetic density, and rejected unless its Omega
// agrees with the default configuration's (1e-10 relative in FP64, 1e-5 in
// mixed precision: a different radix order only reorders the rounding).
// Results are kept per (grid, precision, channel set) for the process, so
// further contexts of the same shape reuse them without timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "api_guard.hpp"
#include "context.hpp"
#include "ddft.hpp"

namespace cdft {

namespace {

constexpr int kNT = 256;          // threads per block of the general kernels
constexpr long kLdsLimit = 65536;  // dynamic LDS without the per-kernel attribute

struct TuneKey {
  long nx, ny, nz;
  int precision, nf, nb, species;
  bool operator==(const TuneKey& o) const {
    return nx == o.nx && ny == o.ny && nz == o.nz && precision == o.precision && nf == o.nf && nb == o.nb &&
           species == o.species;
  }
};
struct TuneEntry {
  TuneKey key;
  FftPlan plan[3];
};
std::mutex g_tune_mu;
std::vector<TuneEntry> g_tuned;

TuneKey key_of(const Context* c) {
  return {c->shape[0], c->shape[1], c->shape[2], c->precision, c->cs.nf, c->cs.nb, c->n_species};
}

__global__ void k_tune_density(double* rho, Geom g) {
  for (long p = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; p < g.N;
       p += static_cast<long>(gridDim.x) * blockDim.x) {
    const long iz = p % g.nz, iy = (p / g.nz) % g.ny, ix = p / (static_cast<long>(g.ny) * g.nz);
    const double a = 6.283185307179586 * (static_cast<double>(ix) / g.nx + static_cast<double>(iy) / g.ny);
    rho[p] = 0.3 + 0.2 * sin(a) * cos(6.283185307179586 * static_cast<double>(iz) / g.nz);
  }
}

// Distinct orders of L's radices: the permutations of the default multiset
// and of the one with every 8 split into 4 x 2.
std::vector<std::vector<int>> radix_orders(const FftPlan& p) {
  std::vector<std::vector<int>> out;
  std::vector<int> a(p.rad, p.rad + p.np);
  std::vector<int> b;
  for (int r : a) {
    if (r == 8) {
      b.push_back(4);
      b.push_back(2);
    } else {
      b.push_back(r);
    }
  }
  for (auto* m : {&a, &b}) {
    if (static_cast<int>(m->size()) > 12) continue;
    std::vector<int> v = *m;
    std::sort(v.begin(), v.end());
    do {
      if (std::find(out.begin(), out.end(), v) == out.end()) out.push_back(v);
    } while (std::next_permutation(v.begin(), v.end()) && out.size() < 24);
  }
  return out;
}

void set_radices(FftPlan& p, const std::vector<int>& r) {
  p.np = static_cast<int>(r.size());
  for (int i = 0; i < p.np; ++i) p.rad[i] = r[static_cast<size_t>(i)];
}

// Launch shapes the kernels accept (element counts and LDS of the general kernels).
bool shapes_valid(const Context* c, const FftPlan pl[3]) {
  const long nx = c->shape[0], ny = c->shape[1], nz = c->shape[2], nzh = nz / 2 + 1;
  const long cx16 = 16, sS = static_cast<long>(c->csize());  // FP64-arithmetic and storage complex sizes
  const int zl = pl[2].lines, yl = pl[1].lines, xt = pl[0].tile, zp = pl[2].tile;
  if (zl > 0 && (zl < 2 || (zl & 1) || zl > 64 || (zl / 2) * nz > 8 * kNT ||
                 cx16 * (zl / 2) * nz + sS * zl * nzh > kLdsLimit))
    return false;
  if (yl > 0 && ((!slot_lanes(static_cast<int>(ny)) && yl * ny > 8 * kNT) || cx16 * yl * ny > kLdsLimit))
    return false;
  if (xt > 0 && (xt * nx > 16 * kNT || cx16 * xt * (nx + 1) > kLdsLimit)) return false;
  const long ppt = c->precision == CDFT_PRECISION_MIXED ? 4 : 2;
  if (zp > 0 && (zp < 2 || (zp & 1) || zp > 64 || zp * nz > ppt * kNT ||
                 cx16 * 4 * (zp / 2) * nz + sS * 4 * zp * nzh > kLdsLimit))
    return false;
  return true;
}

// The default value of each launch shape (as the launchers compute it).
int default_zlines(const Context* c) {
  int n = static_cast<int>((c->precision == CDFT_PRECISION_MIXED ? 4096 : 2048) / c->shape[2]);
  n = std::min(std::max(n, 2), 64);
  return n & ~1;
}
int default_ylines(const Context* c) {
  const int T = slot_lanes(static_cast<int>(c->shape[1]));
  return T ? kNT / T : static_cast<int>(2048 / c->shape[1]);
}
int default_zpoint(const Context* c) {
  int n = static_cast<int>((c->precision == CDFT_PRECISION_MIXED ? 4 : 2) * kNT / c->shape[2]);
  n = std::min(std::max(n, 2), 64);
  return n & ~1;
}

void apply(Context* c, const FftPlan pl[3]) {
  for (int d = 0; d < 3; ++d) c->plan[d] = pl[d];
  c->clear_graphs();
  if (c->n_species > 1) return;  // mixtures size their launches from the plans alone
  c->nblk_point = c->precision == CDFT_PRECISION_MIXED ? zpoint_blocks<float>(c->g, c->plan[2])
                                                      : zpoint_blocks<double>(c->g, c->plan[2]);
  if (c->nblk_point > c->max_point_rows) throw std::logic_error("autotune: pointwise blocks exceed the partials");
}

struct Timer {
  Context* c;
  hipEvent_t a = nullptr, b = nullptr;
  explicit Timer(Context* c_) : c(c_) {
    CDFT_HIP(hipEventCreate(&a));
    CDFT_HIP(hipEventCreate(&b));
  }
  ~Timer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  // microseconds per evaluation, and Omega of the last one
  double time(int reps, double* omega) {
    if (c->n_species > 1) return mixture_tune_time(c, reps, omega);
    const AliasMap id{0, 0.0, 0.0};
    c->eval_direct(c->d_rho, nullptr, 0.0, c->d_force, id);
    CDFT_HIP(hipEventRecord(a, c->stream));
    for (int i = 0; i < reps; ++i) c->eval_direct(c->d_rho, nullptr, 0.0, c->d_force, id);
    CDFT_HIP(hipEventRecord(b, c->stream));
    CDFT_HIP(hipGetLastError());
    cdft_result r{};
    c->fetch_result(0.0, &r);
    float ms = 0.0f;
    CDFT_HIP(hipEventElapsedTime(&ms, a, b));
    *omega = r.grand_potential;
    return 1000.0 * ms / reps;
  }
};

}  // namespace

}  // namespace cdft

using namespace cdft;

extern "C" {

int cdft_autotune(cdft_ctx* ctx, int reps, double* us_default, double* us_tuned) {
  return guarded([&] {
    check_ctx_any(ctx);
    if (reps < 1) throw std::invalid_argument("reps must be >= 1");
    Context* c = ctx->impl;
    CDFT_HIP(hipSetDevice(c->device));
    if (c->n_species == 1) {  // (a mixture context fills its own staging densities)
      hipLaunchKernelGGL(k_tune_density, dim3(vec_blocks(c->g.N)), dim3(256), 0, c->stream, c->d_rho, c->g);
      CDFT_HIP(hipGetLastError());
    }
    Timer tm(c);
    FftPlan best[3] = {c->plan[0], c->plan[1], c->plan[2]};
    double omega0 = 0.0;
    const double t0 = tm.time(reps, &omega0);
    double t_best = t0;
    if (!c->fast) {
      const TuneKey key = key_of(c);
      bool cached = false;
      {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        for (const auto& e : g_tuned)
          if (e.key == key) {
            for (int d = 0; d < 3; ++d) best[d] = e.plan[d];
            cached = true;
          }
      }
      const double tol = c->precision == CDFT_PRECISION_MIXED ? 1e-5 : 1e-10;
      auto trial = [&](const FftPlan cand[3]) {
        if (!shapes_valid(c, cand)) return;
        apply(c, cand);
        double om = 0.0;
        const double t = tm.time(reps, &om);
        if (std::fabs(om - omega0) > tol * std::fabs(omega0)) return;
        if (t < t_best) {
          t_best = t;
          for (int d = 0; d < 3; ++d) best[d] = cand[d];
        }
      };
      if (cached) {
        apply(c, best);
        double om = 0.0;
        t_best = tm.time(reps, &om);
      } else {
        for (int d = 0; d < 3; ++d) {  // radix orders
          for (const auto& r : radix_orders(c->plan[d])) {
            FftPlan cand[3] = {best[0], best[1], best[2]};
            set_radices(cand[d], r);
            trial(cand);
          }
        }
        const int zl = default_zlines(c), yl = default_ylines(c), zp = default_zpoint(c);
        for (int v : {zl / 2, zl / 4}) {
          FftPlan cand[3] = {best[0], best[1], best[2]};
          cand[2].lines = v;
          if (v > 0) trial(cand);
        }
        for (int v : {yl / 2, slot_lanes(static_cast<int>(c->shape[1])) ? 2 * yl : yl / 4}) {
          FftPlan cand[3] = {best[0], best[1], best[2]};
          cand[1].lines = v;
          if (v > 0) trial(cand);
        }
        for (int v : {8, 32}) {
          FftPlan cand[3] = {best[0], best[1], best[2]};
          cand[0].tile = v;
          trial(cand);
        }
        for (int v : {zp / 2}) {
          FftPlan cand[3] = {best[0], best[1], best[2]};
          cand[2].tile = v;
          if (v > 0) trial(cand);
        }
        apply(c, best);
        std::lock_guard<std::mutex> lock(g_tune_mu);
        TuneEntry e{key, {best[0], best[1], best[2]}};
        g_tuned.push_back(e);
      }
    }
    if (us_default) *us_default = t0;
    if (us_tuned) *us_tuned = t_best;
  });
}

}  // extern "C"

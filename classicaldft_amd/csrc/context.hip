// Context construction and the evaluation pipeline.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "context.hpp"
#include "device_common.hpp"

namespace cdft {

namespace {

// One thread per lattice cell: the mean-field weight table w(s) dV / N in the
// real layout with minimum-image offsets s, zero beyond the cut-off reach
// (mean_field.hpp:160-196), and per-block sums of w dV for a_vdw.
constexpr int kMfThreads = 256;
__global__ __launch_bounds__(kMfThreads) void k_mf_table(PairPotential pot, int split, int scheme, double kT,
                                                         double dx, int nx, int ny, int nz, double reach2,
                                                         double* __restrict__ table, double* __restrict__ wsum) {
  const long n = static_cast<long>(nx) * ny * nz;
  const long i = static_cast<long>(blockIdx.x) * kMfThreads + threadIdx.x;
  double v[1] = {0.0};
  if (i < n) {
    const long iz = i % nz, iy = (i / nz) % ny, ix = i / (static_cast<long>(nz) * ny);
    const long sx = ix <= nx / 2 ? ix : ix - nx;
    const long sy = iy <= ny / 2 ? iy : iy - ny;
    const long sz = iz <= nz / 2 ? iz : iz - nz;
    const double d2 = static_cast<double>(sx * sx + sy * sy + sz * sz) * dx * dx;
    double w = 0.0;
    if (!(pot.rc > 0.0 && d2 > reach2)) w = lattice_cell_weight(pot, split, scheme, kT, dx, sx, sy, sz);
    const double dv = dx * dx * dx;
    table[i] = w * dv / static_cast<double>(n);
    v[0] = w * dv;
  }
  block_reduce<1>(v, wsum + blockIdx.x);
}

// Real part of the weight spectrum in the working precision: FP64 values, or
// FP32 (value, residual) planes.
__global__ __launch_bounds__(256) void k_mf_store(const cx<double>* __restrict__ spec, long m, int mixed,
                                                  void* __restrict__ mf) {
  const long i = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= m) return;
  const double re = spec[i].re;
  if (mixed) {
    float* f = static_cast<float*>(mf);
    const float hi = static_cast<float>(re);
    f[i] = hi;
    f[m + i] = static_cast<float>(re - static_cast<double>(hi));
  } else {
    static_cast<double*>(mf)[i] = re;
  }
}

FftPlan make_plan(int L) {
  FftPlan p{};
  p.L = L;
  p.np = 0;
  int m = L;
  while (m % 8 == 0 && p.np < 12) { p.rad[p.np++] = 8; m /= 8; }
  while (m % 4 == 0 && p.np < 12) { p.rad[p.np++] = 4; m /= 4; }
  while (m % 2 == 0 && p.np < 12) { p.rad[p.np++] = 2; m /= 2; }
  while (m % 3 == 0 && p.np < 12) { p.rad[p.np++] = 3; m /= 3; }
  while (m % 5 == 0 && p.np < 12) { p.rad[p.np++] = 5; m /= 5; }
  if (m != 1)
    throw Unsupported("grid length " + std::to_string(L) +
                      " has a prime factor other than 2, 3, 5 (the supported FFT radices)");
  return p;
}

template <typename T>
T* dalloc(size_t count) {
  void* p = nullptr;
  if (count == 0) return nullptr;
  CDFT_HIP(hipMalloc(&p, count * sizeof(T)));
  return static_cast<T*>(p);
}

void* dalloc_bytes(size_t bytes) {
  void* p = nullptr;
  if (bytes == 0) return nullptr;
  CDFT_HIP(hipMalloc(&p, bytes));
  return p;
}

}  // namespace

void Context::clear_graphs() {
  for (auto& e : graphs) {
    (void)hipGraphExecDestroy(e.exec);
    (void)hipGraphDestroy(e.graph);
  }
  graphs.clear();
}

Context::~Context() {
  (void)hipSetDevice(device);  // a destructor has no error path; frees below are best effort
  clear_graphs();
  void* bufs[] = {d_rho,   d_vext,  d_force, d_local,  d_S0,     d_SF,     d_SO,  d_mf,  d_partials, d_res,
                  d_sum_part, d_shift, tw32[0], tw32[1], tw32[2], tw64[0], tw64[1], tw64[2], d_P, d_mfB,
                  ftw64,   ftw32,   d_XR,     d_WT,     d_D0,     d_WTg, d_mix, d_mix_mf};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (stream) (void)hipStreamDestroy(stream);
}

double mf_weight_spectrum(Context* c, const InteractionSpec& s, cx<double>* spec) {
  const Geom& g = c->g;
  double* d_tab = dalloc<double>(g.N);
  const long nblk = (g.N + kMfThreads - 1) / kMfThreads;
  double* d_wsum = nullptr;
  double a = 0.0, comp = 0.0;
  try {
    d_wsum = dalloc<double>(static_cast<size_t>(nblk));
    const double reach = s.pot.rc + 0.5 * std::sqrt(3.0) * g.dx;
    k_mf_table<<<static_cast<unsigned>(nblk), kMfThreads, 0, c->stream>>>(
        s.pot, s.split, s.scheme, c->kT, g.dx, g.nx, g.ny, g.nz, reach * reach, d_tab, d_wsum);
    const AliasMap id{0, 0.0, 0.0};
    launch_zr2c<double, double>(d_tab, spec, g, c->plan[2], c->tw<double>(2), id, nullptr, c->stream);
    launch_lines<double, double>(spec, static_cast<long>(g.nzh) * g.nx, c->plan[1], c->tw<double>(1), -1,
                                 c->stream);
    launch_xlines<double, double>(spec, 1, g, c->plan[0], c->tw<double>(0), -1, c->stream);
    CDFT_HIP(hipGetLastError());
    // a_vdw = sum of w dV: per-block sums reduced on the host in a fixed order
    std::vector<double> wsum(static_cast<size_t>(nblk));
    CDFT_HIP(hipMemcpyAsync(wsum.data(), d_wsum, sizeof(double) * nblk, hipMemcpyDeviceToHost, c->stream));
    CDFT_HIP(hipStreamSynchronize(c->stream));
    for (double x : wsum) {  // Neumaier summation
      const double t = a + x;
      comp += std::fabs(a) >= std::fabs(x) ? (a - t) + x : (x - t) + a;
      a = t;
    }
  } catch (...) {
    (void)hipFree(d_tab);
    if (d_wsum) (void)hipFree(d_wsum);
    throw;
  }
  CDFT_HIP(hipFree(d_tab));
  CDFT_HIP(hipFree(d_wsum));
  return a + comp;
}

void store_real_spectrum(const cx<double>* spec, double* out, long m, hipStream_t stream) {
  k_mf_store<<<static_cast<unsigned>((m + 255) / 256), 256, 0, stream>>>(spec, m, 0, out);
  CDFT_HIP(hipGetLastError());
}

Context* create_context(const cdft_grid* grid, const double* diameters, int n_species,
                        const cdft_interaction* interactions, int n_interactions, double kT,
                        const cdft_options* options) {
  if (!grid || !options) throw std::invalid_argument("cdft_ctx_create: grid and options are required");
  if (grid->dx <= 0.0)
    throw std::invalid_argument("make_grid: dx must be positive, got " + std::to_string(grid->dx));
  long shape[3];
  for (int d = 0; d < 3; ++d) {
    if (grid->box[d] <= 0.0)
      throw std::invalid_argument("make_grid: box dimension must be positive, got " + std::to_string(grid->box[d]));
    const double n = grid->box[d] / grid->dx;
    const long ni = std::lround(n);
    if (std::fabs(n - static_cast<double>(ni)) > 1e-10)
      throw std::invalid_argument("make_grid: box dimension " + std::to_string(grid->box[d]) +
                                  " is not commensurate with dx=" + std::to_string(grid->dx));
    shape[d] = ni;
  }
  if (n_species < 1 || n_species > kMaxSpecies)
    throw Unsupported("species count must be between 1 and " + std::to_string(kMaxSpecies) + "; got " +
                      std::to_string(n_species));
  if (!diameters) throw std::invalid_argument("hard-sphere diameters are required");
  for (int s = 0; s < n_species; ++s)
    if (!(diameters[s] > 0.0)) throw std::invalid_argument("hard-sphere diameter must be positive");
  if (n_interactions < 0) throw std::invalid_argument("negative interaction count");
  if (n_species == 1 && n_interactions > 1)
    throw Unsupported("a one-species context evaluates at most one mean-field interaction");
  if (n_interactions > kMaxMixInteractions)
    throw Unsupported("at most " + std::to_string(kMaxMixInteractions) + " mean-field interactions");
  if (options->precision != CDFT_PRECISION_FP64 && options->precision != CDFT_PRECISION_MIXED)
    throw std::invalid_argument("unknown precision mode");
  if (n_species > 1 && options->precision == CDFT_PRECISION_MIXED &&
      (options->schedule != CDFT_SCHEDULE_AUTO ||
       !fast_supported(static_cast<int>(shape[0]), static_cast<int>(shape[1]), static_cast<int>(shape[2]),
                       CDFT_PRECISION_FP64)))
    throw Unsupported("mixed precision for a mixture needs a grid the fused transforms cover (y and z extents 64 "
                      "or 128, x extent 64, 128 or 256)");
  if (options->schedule != CDFT_SCHEDULE_AUTO && options->schedule != CDFT_SCHEDULE_GENERAL)
    throw std::invalid_argument("unknown schedule");
  const int model = options->fmt_model;
  if (model < -1 || model > 4) throw std::invalid_argument("unknown FMT model id " + std::to_string(model));

  auto* c = new Context();
  try {
    c->device = options->device;
    CDFT_HIP(hipSetDevice(c->device));
    CDFT_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->precision = options->precision;
    c->dx = grid->dx;
    c->kT = kT;
    for (int d = 0; d < 3; ++d) c->shape[d] = shape[d];
    c->diam.assign(diameters, diameters + n_species);
    c->n_species = n_species;

    Geom& g = c->g;
    g.nx = static_cast<int>(shape[0]);
    g.ny = static_cast<int>(shape[1]);
    g.nz = static_cast<int>(shape[2]);
    g.nzh = g.nz / 2 + 1;
    g.N = shape[0] * shape[1] * shape[2];
    g.M = shape[0] * shape[1] * static_cast<long>(g.nzh);
    g.nxy = shape[0] * shape[1];
    g.dx = grid->dx;
    g.dv = grid->dx * grid->dx * grid->dx;
    const double PI = 3.14159265358979323846;
    g.dkx = 2.0 * PI / (g.nx * grid->dx);
    g.dky = 2.0 * PI / (g.ny * grid->dx);
    g.dkz = 2.0 * PI / (g.nz * grid->dx);
    g.inv_n = 1.0 / static_cast<double>(g.N);
    g.R = 0.5 * diameters[0];

    c->fp.model = model;
    c->fp.R = g.R;
    c->fp.A = (model == FMT_ESFMT) ? options->esfmt_A : 1.0;
    c->fp.B = (model == FMT_ESFMT) ? options->esfmt_B : (model >= FMT_WHITEBEAR_I ? -1.0 : 0.0);

    for (int k = 0; k < n_interactions; ++k) {
      const cdft_interaction& in = interactions[k];
      if (in.species_i < 0 || in.species_i >= n_species || in.species_j < 0 || in.species_j >= n_species)
        throw std::invalid_argument("interaction species out of range");
      if (in.potential < 0 || in.potential > 2) throw std::invalid_argument("unknown potential id");
      InteractionSpec s{};
      s.si = in.species_i;
      s.sj = in.species_j;
      s.split = in.split;
      s.scheme = in.weight_scheme;
      s.pot.kind = in.potential;
      s.pot.sigma = in.sigma;
      s.pot.eps = in.epsilon;
      s.pot.rc = in.r_cutoff;
      s.pot.alpha = in.alpha;
      s.pot.finish();
      c->inters.push_back(s);
    }

    c->cs.hs = model != FMT_NONE ? 1 : 0;
    c->cs.mf = n_interactions > 0 ? 1 : 0;
    c->cs.nf = 11 * c->cs.hs + c->cs.mf;
    c->cs.nb = 11 * c->cs.hs;

    for (int d = 0; d < 3; ++d) c->plan[d] = make_plan(static_cast<int>(shape[d]));
    for (int d = 0; d < 3; ++d) {
      const int L = static_cast<int>(shape[d]);
      std::vector<double> t64(2 * static_cast<size_t>(L));
      std::vector<float> t32(4 * static_cast<size_t>(L));  // [0, L) rounded values, [L, 2L) residuals
      for (int m = 0; m < L; ++m) {
        const double a = -2.0 * PI * static_cast<double>(m) / static_cast<double>(L);
        t64[2 * m] = std::cos(a);
        t64[2 * m + 1] = std::sin(a);
        for (int q = 0; q < 2; ++q) {
          const float hi = static_cast<float>(t64[2 * m + q]);
          t32[2 * m + q] = hi;
          t32[2 * L + 2 * m + q] = static_cast<float>(t64[2 * m + q] - static_cast<double>(hi));
        }
      }
      c->tw64[d] = dalloc_bytes(t64.size() * sizeof(double));
      c->tw32[d] = dalloc_bytes(t32.size() * sizeof(float));
      CDFT_HIP(hipMemcpy(c->tw64[d], t64.data(), t64.size() * sizeof(double), hipMemcpyHostToDevice));
      CDFT_HIP(hipMemcpy(c->tw32[d], t32.data(), t32.size() * sizeof(float), hipMemcpyHostToDevice));
    }

    if (n_species > 1) {  // mixtures: FP64, their own buffers; fused transforms where the grid has them
      c->fast = options->schedule == CDFT_SCHEDULE_AUTO && fast_supported(g.nx, g.ny, g.nz, CDFT_PRECISION_FP64);
      if (c->fast) {
        c->fg = make_fast_geom(g);
        fast_twiddles(&c->ftw64, &c->ftw32);
      }
      init_mixture(c);
      return c;
    }

    const size_t cs_bytes = c->csize();
    const int nf_alloc = std::max(c->cs.nf, 1);
    c->d_rho = dalloc<double>(g.N);
    c->d_vext = dalloc<double>(g.N);
    c->d_force = dalloc<double>(g.N);
    c->d_local = dalloc<double>(g.N);
    c->d_S0 = dalloc_bytes(cs_bytes * g.M);
    c->d_SF = dalloc_bytes(cs_bytes * g.M * nf_alloc);
    c->d_SO = dalloc_bytes(cs_bytes * g.M);
    c->d_mf = dalloc_bytes(8 * g.M);  // FP64 values, or FP32 (value, residual) planes
    if (c->cs.hs) {
      c->d_WTg = dalloc_bytes(16 * g.M);
      launch_wtable_general(static_cast<cx<double>*>(c->d_WTg), g, c->stream);
      CDFT_HIP(hipGetLastError());
    }
    c->fast = options->schedule == CDFT_SCHEDULE_AUTO && fast_supported(g.nx, g.ny, g.nz, c->precision);
    c->nblk_point = c->precision == CDFT_PRECISION_MIXED ? zpoint_blocks<float>(g, c->plan[2])
                                                         : zpoint_blocks<double>(g, c->plan[2]);
    if (c->fast) {
      c->fg = make_fast_geom(g);
      hipDeviceProp_t prop{};
      CDFT_HIP(hipGetDeviceProperties(&prop, c->device));
      c->fg.ncu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
      c->nblk_point = fast_point_blocks(c->fg, c->precision != CDFT_PRECISION_MIXED);
      fast_twiddles(&c->ftw64, &c->ftw32);
      c->d_P = dalloc_bytes(cs_bytes * g.M * 2);  // partial sets 0 and 1 of the backward plane kernel
      c->d_XR = dalloc_bytes(16 * g.M);
      c->d_D0 = dalloc<double>(static_cast<size_t>(g.nx) * g.ny);
      c->d_WT = dalloc_bytes(16 * g.M);
      launch_fast_wtable(c->d_WT, c->fg, c->stream);
      CDFT_HIP(hipGetLastError());
      CDFT_HIP(hipStreamSynchronize(c->stream));
    }
    // room for the most blocks a pointwise launch shape can have (2 z lines per block)
    c->max_point_rows = c->fast ? c->nblk_point : std::max<long>(c->nblk_point, (g.nxy + 1) / 2);
    c->d_partials = dalloc<double>(8 * static_cast<size_t>(c->max_point_rows));
    c->d_res = dalloc<double>(8);
    c->nblk_sum = sum_blocks(g);
    c->d_sum_part = dalloc<double>(8 * static_cast<size_t>(c->nblk_sum));
    c->d_shift = dalloc<double>(8);

    // Mean-field weight: FP64 spectrum, real part kept in the working precision
    // (and in the fused layout when that schedule runs).
    if (c->cs.mf) {
      void* d_spec = dalloc_bytes(16 * g.M);
      try {
        auto* S = static_cast<cx<double>*>(d_spec);
        c->a_vdw.push_back(mf_weight_spectrum(c, c->inters[0], S));
        k_mf_store<<<static_cast<unsigned>((g.M + 255) / 256), 256, 0, c->stream>>>(
            S, g.M, c->precision == CDFT_PRECISION_MIXED ? 1 : 0, c->d_mf);
        if (c->fast) {
          c->d_mfB = dalloc<double>(g.M);
          launch_fast_mf_layout(S, c->d_mfB, g, c->stream);
        }
        CDFT_HIP(hipGetLastError());
        CDFT_HIP(hipStreamSynchronize(c->stream));
      } catch (...) {
        (void)hipFree(d_spec);
        throw;
      }
      CDFT_HIP(hipFree(d_spec));
    }
    // k = 0 weights of the forward channels (weights.hpp:91-184, mean_field.hpp:189-191)
    {
      const double R = g.R;
      int ch = 0;
      if (c->cs.hs) {
        c->dc.v[0] = 4.0 * PI / 3.0 * R * R * R;
        c->dc.v[1] = 4.0 * PI * R * R;
        c->dc.v[5] = c->dc.v[8] = c->dc.v[10] = 4.0 * PI / 3.0 * R * R;
        ch = 11;
      }
      if (c->cs.mf) c->dc.v[ch] = c->a_vdw[0];
    }
  } catch (...) {
    delete c;
    throw;
  }
  return c;
}

const char* const kStageNames[kNumStages] = {"k_sum",    "k_zr2c",            "k_lines(y fwd rho)", "k_xfwd",
                                              "k_lines(y inv nf)", "k_zpoint", "k_lines(y fwd nb)", "k_xbwd",
                                              "k_lines(y inv 1)",  "k_zc2r",   "k_finalize"};

// the energy sums are finalised inside kf_zinv (by k_finalize in the
// stage of kf_zinv when no forces are asked for)
const char* const kFastStageNames[kNumFastStages] = {"kf_zfwd",   "kf_rho_xy",    "kf_plane_fwd",
                                                     "kf_zpoint", "kf_plane_bwd", "kf_zinv"};

template <typename S, typename CV>
void Context::eval_fast_t(const double* rho, const double* vext, double mu, double* force, AliasMap amap) {
  auto* A0 = static_cast<cx<S>*>(d_S0);
  auto* SF = static_cast<cx<S>*>(d_SF);
  auto* P = static_cast<cx<S>*>(d_P);
  const bool backward = force != nullptr && cs.nb > 0;
  auto mark = [&](int i) {
    if (stage_events) CDFT_HIP(hipEventRecord(stage_events[i], stream));
  };
  mark(0);
  if (cs.nf > 0) {
    launch_fast_zfwd<S>(rho, A0, d_D0, fg, ftw64, amap, stream);
    mark(1);
    launch_fast_rho_xy<S>(A0, d_D0, d_XR, d_shift, fg, ftw64, stream);
    mark(2);
    launch_fast_plane_fwd<S, CV>(d_XR, d_WT, SF, d_mfB, fg, ftw64, sizeof(CV) == 4 ? ftw32 : ftw64, cs, stream);
  }
  ChannelSet cz = cs;
  if (!backward) cz.nb = 0;
  double* lf = backward ? d_local : (force ? force : d_local);
  mark(3);
  launch_fast_zpoint<S, CV>(SF, rho, vext, lf, d_partials, fg, ftw64, sizeof(CV) == 4 ? ftw32 : ftw64, cz, fp, mu,
                            amap, cs.nf > 0 ? d_shift : nullptr, dc, stream);
  if (backward) {
    mark(4);
    launch_fast_plane_bwd<S, CV>(SF, d_WT, P, fg, sizeof(CV) == 4 ? ftw32 : ftw64, stream);
    mark(5);
    // the energy sums are reduced by one extra block of the last kernel
    launch_fast_zinv<S, CV>(P, force, d_local, g.dv, fg, sizeof(CV) == 4 ? ftw32 : ftw64, d_partials, nblk_point,
                            d_res, stream);
    mark(6);
  } else {
    mark(4);
    mark(5);
    launch_finalize(d_partials, nblk_point, d_res, stream);
    mark(6);
  }
}

template <typename S, typename CE, typename CL>
void Context::eval_t(const double* rho, const double* vext, double mu, double* force, AliasMap amap) {
  auto* S0 = static_cast<cx<S>*>(d_S0);
  auto* SF = static_cast<cx<S>*>(d_SF);
  auto* SO = static_cast<cx<S>*>(d_SO);
  const long ylines = static_cast<long>(g.nzh) * g.nx;
  const bool backward = force != nullptr && cs.nb > 0;
  // channel arrays in SF: the reduced sets of the kg_zpoint point kernel
  // (forward without v_z and T_zz, 9 backward) or all 12 / 11
  const bool red = zpoint_reduced(g.nz, cs);
  const int nfs = red ? cs.nf - 2 : cs.nf, nbs = red ? 9 : cs.nb;
  auto mark = [&](int i) {
    if (stage_events) CDFT_HIP(hipEventRecord(stage_events[i], stream));
  };
  mark(0);
  if (cs.nf > 0) {
    launch_sum(rho, amap, d_sum_part, g, stream);
    mark(1);
    // the z-R2C reduces the sum itself and stores it in d_shift
    launch_zr2c<S, CL>(rho, S0, g, plan[2], tw<CL>(2), amap, nullptr, stream, d_sum_part, nblk_sum, d_shift);
    mark(2);
    launch_lines<S, CL>(S0, ylines, plan[1], tw<CL>(1), -1, stream);
    mark(3);
    launch_xfwd<S, CE, CL>(S0, SF, static_cast<const S*>(d_mf), static_cast<const cx<double>*>(d_WTg), g, plan[0],
                           tw<CE>(0), tw<CL>(0), cs, red, stream);
    mark(4);
    if constexpr (std::is_same_v<S, float> && std::is_same_v<CE, double> && std::is_same_v<CL, float>) {
      // mixed: the y-inverse of the scalar channels (eta, n2, mean field) in
      // FP64, of the vector and tensor channels (arrays 2 .. 8 or 10) in FP32,
      // as in the fused schedule
      if (cs.hs)
        launch_lines_split(SF, ylines * nfs, plan[1], tw<CE>(1), tw<CL>(1), 2 * ylines, (red ? 9 : 11) * ylines, +1,
                           stream);
      else
        launch_lines<S, CE>(SF, ylines * nfs, plan[1], tw<CE>(1), +1, stream);
    } else {
      launch_lines<S, CE>(SF, ylines * nfs, plan[1], tw<CE>(1), +1, stream);
    }
  }
  ChannelSet cz = cs;
  if (!backward) cz.nb = 0;
  double* lf = backward ? d_local : (force ? force : d_local);
  mark(5);
  launch_zpoint<S, CE, CL>(SF, rho, vext, lf, d_partials, g, plan[2], tw<CE>(2), tw<CL>(2), cz, fp, mu, amap,
                           cs.nf > 0 ? d_shift : nullptr, dc, stream);
  if (backward) {
    mark(6);
    launch_lines<S, CL>(SF, ylines * nbs, plan[1], tw<CL>(1), -1, stream);
    mark(7);
    launch_xbwd<S, CL>(SF, SO, static_cast<const cx<double>*>(d_WTg), g, plan[0], tw<CL>(0), nbs, stream);
    mark(8);
    launch_lines<S, CL>(SO, ylines, plan[1], tw<CL>(1), +1, stream);
    mark(9);
    // its block 0 also reduces the point kernel's partial sums
    launch_zc2r<S, CL>(SO, 1, force, d_local, g.dv, g, plan[2], tw<CL>(2), nullptr, DcTable{}, stream, d_partials,
                       nblk_point, d_res);
  }
  mark(10);
  if (!backward) launch_finalize(d_partials, nblk_point, d_res, stream);
  mark(11);
}

// Algorithmic HBM bytes of each stage for one evaluation (stage-boundary
// model: each kernel reads its inputs and writes its outputs once).
void Context::stage_bytes(double* b, bool with_vext) const {
  const double sf = static_cast<double>(csize());  // complex element
  const double N = static_cast<double>(g.N), M = static_cast<double>(g.M);
  for (int i = 0; i < kNumStages; ++i) b[i] = 0.0;
  const double nf = cs.nf, nb = cs.nb;
  if (fast) {
    // plane channels: 10 forward (n_v,z and T_zz are formed in the z-line kernel), 9 backward
    const double nfp = cs.nf == 12 ? 10.0 : (cs.nf == 11 ? 9.0 : nf), nbp = cs.nb > 0 ? 9.0 : 0.0;
    if (cs.nf > 0) {
      b[0] = 8.0 * N + sf * M + 8.0 * g.nx * g.ny;     // rho -> rho_hat parts (+ FP64 kz = 0 plane)
      b[1] = sf * M + 16.0 * M;                        // parts -> FP64 spectrum
      b[2] = 16.0 * M + (cs.hs ? 16.0 * M : 0.0) + nfp * sf * M + (cs.mf ? 8.0 * M : 0.0);  // spectrum, weights -> channels
    }
    b[3] = nfp * sf * M + 8.0 * N * (with_vext ? 2.0 : 1.0) + nbp * sf * M + 8.0 * N;
    if (cs.nb > 0) {
      b[4] = nbp * sf * M + 16.0 * M + sf * M;         // channels, weights -> accumulated sub-planes
      b[5] = sf * M + 8.0 * N + 8.0 * N;               // sub-planes + local force -> force
    }
    return;
  }
  // the kg_zpoint point kernel's reduced channel sets
  const bool red = zpoint_reduced(g.nz, cs);
  const double nfs = red ? nf - 2.0 : nf, nbs = red ? 9.0 : nb;
  if (cs.nf > 0) {
    b[0] = 8.0 * N;                                     // FP64 sum of rho
    b[1] = 8.0 * N + sf * M;                            // rho (FP64) -> rho_hat
    b[2] = 2.0 * sf * M;                                // in place
    b[3] = sf * M + nfs * sf * M + (cs.mf ? sf * M : 0);  // rho_hat + table (value, residual) -> channels
    b[4] = 2.0 * nfs * sf * M;
  }
  b[5] = nfs * sf * M + 8.0 * N * (with_vext ? 2.0 : 1.0) + nbs * sf * M + 8.0 * N;  // channels, rho, V -> derivs, local
  if (cs.nb > 0) {
    b[6] = 2.0 * nbs * sf * M;
    b[7] = nbs * sf * M + sf * M;
    b[8] = 2.0 * sf * M;
    b[9] = sf * M + 8.0 * N + 8.0 * N;  // spectrum + local -> force
  }
  b[10] = 0.0;
}

void Context::eval_direct(const double* rho, const double* vext, double mu, double* force, AliasMap amap) {
  if (fast) {
    if (precision == CDFT_PRECISION_MIXED) eval_fast_t<float, float>(rho, vext, mu, force, amap);
    else eval_fast_t<double, double>(rho, vext, mu, force, amap);
  } else {
    if (precision == CDFT_PRECISION_MIXED) eval_t<float, double, float>(rho, vext, mu, force, amap);
    else eval_t<double, double, double>(rho, vext, mu, force, amap);
  }
}

void Context::eval(const double* rho, const double* vext, double mu, double* force, AliasMap amap) {
  auto direct = [&] { eval_direct(rho, vext, mu, force, amap); };
  if (!use_graphs || stage_events) {
    direct();
    CDFT_HIP(hipGetLastError());
    return;
  }
  // The ~12 launches of one evaluation are replayed from a captured graph:
  // one submission instead of a dozen kernel launches.
  const GraphKey key{rho, vext, force, mu, amap};
  for (auto& e : graphs) {
    if (e.key == key) {
      CDFT_HIP(hipGraphLaunch(e.exec, stream));
      return;
    }
  }
  if (graphs.size() >= 16) clear_graphs();
  GraphEntry ent{key, nullptr, nullptr};
  CDFT_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  try {
    direct();
  } catch (...) {
    hipGraph_t g0 = nullptr;
    (void)hipStreamEndCapture(stream, &g0);
    if (g0) (void)hipGraphDestroy(g0);
    throw;
  }
  CDFT_HIP(hipStreamEndCapture(stream, &ent.graph));
  CDFT_HIP(hipGraphInstantiate(&ent.exec, ent.graph, nullptr, nullptr, 0));
  graphs.push_back(ent);
  CDFT_HIP(hipGraphLaunch(ent.exec, stream));
}

void Context::result_from(const double* res, double mu, cdft_result* out) const {
  out->ideal = res[RS_IDEAL];
  out->external = res[RS_EXT];
  out->hard_sphere = res[RS_HS];
  out->mean_field = res[RS_MF];
  out->mass = res[RS_MASS];
  // functionals.hpp:79-92: F = F_id + F_ext + F_hs + F_mf, Omega = F - mu N
  out->free_energy = out->ideal + out->external + out->hard_sphere + out->mean_field;
  out->grand_potential = out->free_energy - mu * out->mass;
}

void Context::fetch_result(double mu, cdft_result* out) {
  double res[8];
  CDFT_HIP(hipMemcpyAsync(res, d_res, sizeof(res), hipMemcpyDeviceToHost, stream));
  CDFT_HIP(hipStreamSynchronize(stream));
  result_from(res, mu, out);
}

namespace {
template <typename S, typename C>
void r2c_t(Context* c, const double* d_in) {
  const AliasMap id{0, 0.0, 0.0};
  auto* Sp = static_cast<cx<S>*>(c->d_S0);
  const Geom& g = c->g;
  launch_zr2c<S, C>(d_in, Sp, g, c->plan[2], c->tw<C>(2), id, nullptr, c->stream);
  launch_lines<S, C>(Sp, static_cast<long>(g.nzh) * g.nx, c->plan[1], c->tw<C>(1), -1, c->stream);
  launch_xlines<S, C>(Sp, 1, g, c->plan[0], c->tw<C>(0), -1, c->stream);
}

template <typename S, typename C>
void c2r_t(Context* c, double* d_out) {
  auto* Sp = static_cast<cx<S>*>(c->d_S0);
  const Geom& g = c->g;
  launch_xlines<S, C>(Sp, 1, g, c->plan[0], c->tw<C>(0), +1, c->stream);
  launch_lines<S, C>(Sp, static_cast<long>(g.nzh) * g.nx, c->plan[1], c->tw<C>(1), +1, c->stream);
  launch_zc2r<S, C>(Sp, 1, d_out, nullptr, 1.0, g, c->plan[2], c->tw<C>(2), nullptr, DcTable{}, c->stream);
}

template <typename S, typename CE, typename CL>
void wd_t(Context* c, const double* d_rho_in, double* d_out) {
  const AliasMap id{0, 0.0, 0.0};
  const Geom& g = c->g;
  const long ylines = static_cast<long>(g.nzh) * g.nx;
  auto* S0 = static_cast<cx<S>*>(c->d_S0);
  auto* SF = static_cast<cx<S>*>(c->d_SF);
  launch_sum(d_rho_in, id, c->d_sum_part, g, c->stream);
  launch_finalize(c->d_sum_part, c->nblk_sum, c->d_shift, c->stream);
  launch_zr2c<S, CL>(d_rho_in, S0, g, c->plan[2], c->tw<CL>(2), id, c->d_shift, c->stream);
  launch_lines<S, CL>(S0, ylines, c->plan[1], c->tw<CL>(1), -1, c->stream);
  launch_xfwd<S, CE, CE>(S0, SF, static_cast<const S*>(c->d_mf), static_cast<const cx<double>*>(c->d_WTg), g,
                         c->plan[0], c->tw<CE>(0), c->tw<CE>(0), c->cs, false, c->stream);
  launch_lines<S, CE>(SF, ylines * c->cs.nf, c->plan[1], c->tw<CE>(1), +1, c->stream);
  launch_zc2r<S, CE>(SF, c->cs.nf, d_out, nullptr, 1.0, g, c->plan[2], c->tw<CE>(2), c->d_shift, c->dc, c->stream);
}
}  // namespace

void Context::r2c(const double* d_in) {
  if (precision == CDFT_PRECISION_MIXED) r2c_t<float, float>(this, d_in);
  else r2c_t<double, double>(this, d_in);
  CDFT_HIP(hipGetLastError());
}

void Context::c2r(double* d_out) {
  if (precision == CDFT_PRECISION_MIXED) c2r_t<float, float>(this, d_out);
  else c2r_t<double, double>(this, d_out);
  CDFT_HIP(hipGetLastError());
}

void Context::weighted_densities(const double* d_rho_in, double* d_out) {
  if (cs.nf == 0) return;
  if (precision == CDFT_PRECISION_MIXED) wd_t<float, double, float>(this, d_rho_in, d_out);
  else wd_t<double, double, double>(this, d_rho_in, d_out);
  CDFT_HIP(hipGetLastError());
}

// BulkThermodynamics::chemical_potential (bulk/thermodynamics.hpp:222-237) for
// one species: ln rho + mu_ex^FMT (uniform measures, measures.hpp:49-65) + a_vdw rho.
double Context::bulk_mu(double rho) const {
  const double PI = 3.14159265358979323846;
  const double d = diam[0];
  double mu = std::log(rho);
  if (fp.model != FMT_NONE) {
    const double n2 = PI * rho * d * d;
    double w[11] = {PI / 6.0 * rho * d * d * d, n2, 0, 0, 0, n2 / 3.0, 0, 0, n2 / 3.0, 0, n2 / 3.0};
    double gg[11];
    fmt_phi_grad(fp, w, gg);
    mu += gg[0] * PI / 6.0 * d * d * d + gg[1] * PI * d * d;
    if (fp.model >= FMT_WHITEBEAR_I) mu += (gg[5] + gg[8] + gg[10]) * PI * d * d / 3.0;
  }
  for (size_t k = 0; k < a_vdw.size(); ++k) mu += a_vdw[k] * rho;
  return mu;
}

}  // namespace cdft

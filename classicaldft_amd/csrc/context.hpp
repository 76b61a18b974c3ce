// Library context: geometry, FFT plans, device buffers, and the evaluation
// pipeline shared by the C ABI entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "cdft.h"
#include "eval_fast.hpp"
#include "pair_potential.hpp"
#include "kernels.hpp"

namespace cdft {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct Unsupported : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define CDFT_HIP(call)                                                                                \
  do {                                                                                                \
    hipError_t e_ = (call);                                                                           \
    if (e_ != hipSuccess)                                                                             \
      throw ::cdft::HipError(std::string(#call) + " failed: " + hipGetErrorString(e_) + " at " +      \
                             __FILE__ + ":" + std::to_string(__LINE__));                              \
  } while (0)

struct InteractionSpec {
  int si, sj, split, scheme;
  PairPotential pot;
};

constexpr int kNumStages = 11;  // stages of the general schedule (the most of any path)
extern const char* const kStageNames[kNumStages];
constexpr int kNumFastStages = 6;
extern const char* const kFastStageNames[kNumFastStages];

// Results of one evaluation, in the order the fused kernel accumulates them.
enum ResultSlot { RS_IDEAL = 0, RS_EXT = 1, RS_HS = 2, RS_MF = 3, RS_MASS = 4 };

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  int precision = CDFT_PRECISION_FP64;
  long shape[3] = {0, 0, 0};
  double dx = 0, kT = 1;
  Geom g{};
  FmtParams fp{};
  ChannelSet cs{};
  std::vector<double> diam;
  std::vector<InteractionSpec> inters;
  std::vector<double> a_vdw;
  FftPlan plan[3]{};
  void* tw32[3] = {nullptr, nullptr, nullptr};
  void* tw64[3] = {nullptr, nullptr, nullptr};

  // device buffers
  double* d_rho = nullptr;    // staging of host inputs / alias-mapped density
  double* d_vext = nullptr;
  double* d_force = nullptr;
  double* d_local = nullptr;  // local (ideal + external + mean-field) force
  void* d_S0 = nullptr;       // M complex
  void* d_SF = nullptr;       // nf * M complex (derivative channels reuse it)
  void* d_SO = nullptr;       // M complex
  void* d_mf = nullptr;       // M real: mean-field k-space weight
  void* d_WTg = nullptr;      // M complex FP64: radial FMT weights (w3, w2) in the general schedule's order
  double* d_partials = nullptr;
  double* d_res = nullptr;
  double* d_sum_part = nullptr;  // block partials of the FP64 density sum
  double* d_shift = nullptr;     // [0] = sum of rho (mean-density split)
  int nblk_point = 0;  // rows of 8 partial sums written by the point kernel
  long max_point_rows = 0;  // rows d_partials has room for
  int nblk_sum = 0;
  DcTable dc{};                  // analytic k = 0 weight of every forward channel
  hipEvent_t* stage_events = nullptr;  // when set, eval records nstages()+1 events

  // Fused five-kernel schedule (power-of-two grids it has kernels for).
  bool fast = false;
  FastGeom fg{};
  void* ftw64 = nullptr;      // shared twiddle tables of the fused kernels
  void* ftw32 = nullptr;
  void* d_P = nullptr;        // 2 * M complex: partial sets 0 and 1 of the backward plane kernel
  void* d_XR = nullptr;       // M complex FP64: rho_hat for the forward plane kernel
  double* d_D0 = nullptr;     // nx * ny FP64: kz = 0 plane (line sums) of the z-transformed rho
  void* d_WT = nullptr;       // M complex FP64: radial FMT weights (w3, w2) in the plane kernels' order
  double* d_mfB = nullptr;    // M: mean-field weight in the plane kernel's order

  // Mixtures (more than one species): FP64 general schedule.  d_mix is one
  // block holding the species spectra, channel spectra and fields, backward
  // spectra, local forces, input staging and partial sums (mixture.hip);
  // d_mix_mf holds the real mean-field weight spectrum of every interaction.
  int n_species = 1;
  void* d_mix = nullptr;
  double* d_mix_mf = nullptr;
  int nstages() const { return fast ? kNumFastStages : kNumStages; }
  const char* stage_name(int i) const {
    return (i < 0 || i >= nstages()) ? "" : (fast ? kFastStageNames[i] : kStageNames[i]);
  }

  // Captured evaluation graphs, keyed by the arguments baked into them.
  struct GraphKey {
    const double* rho;
    const double* vext;
    double* force;
    double mu;
    AliasMap amap;
    bool operator==(const GraphKey& o) const {
      return rho == o.rho && vext == o.vext && force == o.force && mu == o.mu && amap.kind == o.amap.kind &&
             amap.rho_min == o.amap.rho_min && amap.range == o.amap.range;
    }
  };
  struct GraphEntry {
    GraphKey key;
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs;
  bool use_graphs = true;
  void clear_graphs();

  ~Context();
  size_t csize() const { return precision == CDFT_PRECISION_MIXED ? 8 : 16; }

  // Twiddle table of one axis for arithmetic type C (FP32: value + residual pairs).
  template <typename C>
  const cx<C>* tw(int axis) const {
    return static_cast<const cx<C>*>(sizeof(C) == 4 ? tw32[axis] : tw64[axis]);
  }

  // One F + dF/drho evaluation on device arrays (rho may be an alias variable
  // mapped through `amap`).  Results land in d_res (device).  When `force`
  // is null only the energy is computed.
  void eval(const double* rho, const double* vext, double mu, double* force, AliasMap amap);
  // The same launches issued directly on `stream` (no cached graph), for
  // callers that capture a larger graph around the evaluation.
  void eval_direct(const double* rho, const double* vext, double mu, double* force, AliasMap amap);
  // S: spectral storage type; CE: arithmetic of the energy path's inverse
  // transforms; CL: arithmetic of the other transforms.
  template <typename S, typename CE, typename CL>
  void eval_t(const double* rho, const double* vext, double mu, double* force, AliasMap amap);
  // The fused schedule: S storage, CV arithmetic of the non-scalar channels
  // and of the backward path.
  template <typename S, typename CV>
  void eval_fast_t(const double* rho, const double* vext, double mu, double* force, AliasMap amap);

  // Read back d_res and assemble a cdft_result.
  void fetch_result(double mu, cdft_result* out);
  void result_from(const double* res, double mu, cdft_result* out) const;

  // Plain 3D transforms between a device real field and S0 (plane-major).
  void r2c(const double* d_in);
  void c2r(double* d_out);
  void weighted_densities(const double* d_rho_in, double* d_out /* nf * N */);

  double bulk_mu(double rho) const;
  void stage_bytes(double* b, bool with_vext) const;
};

Context* create_context(const cdft_grid* grid, const double* diameters, int n_species,
                        const cdft_interaction* interactions, int n_interactions, double kT,
                        const cdft_options* options);

// Species and interaction limits of a context (kernel argument arrays).
constexpr int kMaxSpecies = 8;
constexpr int kMaxMixInteractions = 36;  // every pair of 8 species, self pairs included

// Mean-field weight spectrum of one interaction (M complex FP64, plane-major,
// real-valued), built on the device; returns a_vdw = sum of w dV.
double mf_weight_spectrum(Context* c, const InteractionSpec& s, cx<double>* spec);
// out[i] = spec[i].re for i < m (FP64), on `stream`.
void store_real_spectrum(const cx<double>* spec, double* out, long m, hipStream_t stream);
// Buffers and weights of a mixture context (defined with the mixture kernels).
void init_mixture(Context* c);
// Autotuning of a mixture context: `reps` evaluations (with forces) of a
// fixed smooth density in the context's staging buffers; microseconds per
// evaluation, and Omega of the last.
double mixture_tune_time(Context* c, int reps, double* omega);

void set_error(const std::string& msg);

}  // namespace cdft

struct cdft_ctx {
  cdft::Context* impl;
};

// C++ facade of the free-energy + gradient path over the C ABI of cdft.h.
//
// It restates the reference's C++ interfaces for this path, with
// std::vector<double> (dft::vec) in place of arma::vec, so that code written
// against the reference's
//   dft::make_grid, dft::physics::Model, dft::functionals::{make_weights,
//   total, make_functional, Functional}, dft::algorithms::fire::Fire,
//   dft::algorithms::minimization::Minimizer and
//   dft::algorithms::dynamics::Simulation
// compiles against it with the container type changed.  Every evaluation
// runs in libcdft.so on the GPU; errors surface as the exception types the
// reference throws (std::invalid_argument for bad arguments, std::runtime_error
// otherwise).  Header-only, C++17; link with -lcdft.
//
// Limits of the GPU path (std::invalid_argument when exceeded): up to 8
// species for functionals::total (the algorithms act on one species, as the
// reference's do), grid lengths with prime factors 2, 3, 5.  The algorithms
// take any force callable: the functional's own grand_potential_callback
// evaluates on the device with no host round trip; another callable runs on
// the host with the transforms, stencils and propagators on the device.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <exception>
#include <functional>
#include <limits>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <variant>
#include <vector>

#include "cdft.h"

namespace dft {

using vec = std::vector<double>;

namespace detail {
// Maps a cdft status to the reference's exception types.
inline void check(int rc) {
  if (rc == CDFT_OK) return;
  const std::string msg = cdft_last_error();
  if (rc == CDFT_EINVAL || rc == CDFT_EUNSUPPORTED) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}
}  // namespace detail

// ------------------------------------------------------------------ grid.hpp:26-103
struct Grid {
  double dx;
  std::array<double, 3> box_size;
  std::array<long, 3> shape;

  double cell_volume() const noexcept { return dx * dx * dx; }
  long total_points() const noexcept { return shape[0] * shape[1] * shape[2]; }
  long flat_index(long ix, long iy, long iz) const noexcept { return iz + shape[2] * (iy + shape[1] * ix); }
};

// Same validation and messages as the reference's make_grid (grid.hpp:87-103).
inline Grid make_grid(double dx, std::array<double, 3> box) {
  if (dx <= 0.0) throw std::invalid_argument("make_grid: dx must be positive, got " + std::to_string(dx));
  std::array<long, 3> shape{};
  for (int d = 0; d < 3; ++d) {
    if (box[d] <= 0.0)
      throw std::invalid_argument("make_grid: box dimension must be positive, got " + std::to_string(box[d]));
    const double n = box[d] / dx;
    const long ni = static_cast<long>(std::round(n));
    if (std::abs(n - static_cast<double>(ni)) > 1e-10)
      throw std::invalid_argument("make_grid: box dimension " + std::to_string(box[d]) +
                                  " is not commensurate with dx=" + std::to_string(dx));
    shape[d] = ni;
  }
  return Grid{dx, box, shape};
}

// ------------------------------------------------------------------ types.hpp:17-48
struct Density {
  vec values;
  vec external_field;
};
struct Species {
  std::string name;
  double hard_sphere_diameter;
};
struct SpeciesState {
  Density density;
  vec force;
  double chemical_potential{0.0};
  std::optional<double> fixed_mass;
};
struct State {
  std::vector<SpeciesState> species;
  double temperature;
};

// ------------------------------------------------------------------ physics
namespace physics {
namespace potentials {
// potentials.hpp:18-21
enum class SplitScheme { WeeksChandlerAndersen, BarkerHenderson };
// potentials.hpp:57-63, 122-128, 188-194 (parameters only: the attractive
// tails and lattice weights are built inside the library)
struct LennardJones {
  double sigma{1.0}, epsilon{1.0}, r_cutoff{-1.0};
};
struct TenWoldeFrenkel {
  double sigma{1.0}, epsilon{1.0}, r_cutoff{-1.0}, alpha{50.0};
};
struct WangRamirezDobnikarFrenkel {
  double sigma{1.0}, epsilon{1.0}, r_cutoff{3.0};
};
using Potential = std::variant<LennardJones, TenWoldeFrenkel, WangRamirezDobnikarFrenkel>;

inline LennardJones make_lennard_jones(double sigma, double epsilon, double r_cutoff = -1.0) {
  return {sigma, epsilon, r_cutoff};
}
inline TenWoldeFrenkel make_ten_wolde_frenkel(double sigma, double epsilon, double r_cutoff = -1.0,
                                              double alpha = 50.0) {
  return {sigma, epsilon, r_cutoff, alpha};
}
inline WangRamirezDobnikarFrenkel make_wang_ramirez_dobnikar_frenkel(double sigma, double epsilon,
                                                                     double r_cutoff = 3.0) {
  return {sigma, epsilon, r_cutoff};
}
}  // namespace potentials

// interactions.hpp:11-31
enum class WeightScheme {
  InterpolationZero,
  InterpolationLinearE,
  InterpolationLinearF,
  InterpolationQuadraticF,
  GaussE,
  GaussF,
};
struct Interaction {
  int species_i{0};
  int species_j{0};
  potentials::Potential potential;
  potentials::SplitScheme split{potentials::SplitScheme::WeeksChandlerAndersen};
  WeightScheme weight_scheme{WeightScheme::InterpolationQuadraticF};
  int gauss_order{5};
};

// model.hpp:15-20
struct Model {
  Grid grid;
  std::vector<Species> species;
  std::vector<Interaction> interactions;
  double temperature{1.0};
};
}  // namespace physics

// ------------------------------------------------------------------ functionals
namespace functionals {
namespace fmt {
// models.hpp:153-419: the model tags and FMTModel::from_name
struct Rosenfeld {};
struct RSLT {};
struct WhiteBearI {};
struct WhiteBearII {};
struct EsFMT {
  double A{1.0};
  double B{0.0};
};

class FMTModel {
 public:
  using VariantType = std::variant<Rosenfeld, RSLT, WhiteBearI, WhiteBearII, EsFMT>;
  FMTModel(Rosenfeld m) : data_(m) {}
  FMTModel(RSLT m) : data_(m) {}
  FMTModel(WhiteBearI m) : data_(m) {}
  FMTModel(WhiteBearII m) : data_(m) {}
  FMTModel(EsFMT m) : data_(m) {}

  static FMTModel from_name(const std::string& name) {
    if (name == "Rosenfeld") return Rosenfeld{};
    if (name == "RSLT") return RSLT{};
    if (name == "WhiteBearI") return WhiteBearI{};
    if (name == "WhiteBearII") return WhiteBearII{};
    if (name == "esFMT") return EsFMT{};
    throw std::invalid_argument("Unknown FMT model: " + name);
  }
  const VariantType& variant() const { return data_; }

  // cdft_options.fmt_model and the esFMT coefficients
  int id() const { return static_cast<int>(data_.index()); }
  double esfmt_a() const { return std::holds_alternative<EsFMT>(data_) ? std::get<EsFMT>(data_).A : 1.0; }
  double esfmt_b() const { return std::holds_alternative<EsFMT>(data_) ? std::get<EsFMT>(data_).B : -1.0; }

 private:
  VariantType data_;
};

// The cdft_options.fmt_model values follow the variant order.
static_assert(CDFT_FMT_ROSENFELD == 0 && CDFT_FMT_RSLT == 1 && CDFT_FMT_WHITEBEAR_I == 2 &&
                  CDFT_FMT_WHITEBEAR_II == 3 && CDFT_FMT_ESFMT == 4,
              "FMT model numbering");
}  // namespace fmt

// types.hpp:11-23
struct Contribution {
  double free_energy{0.0};
  std::vector<vec> forces;
};
struct Result {
  double free_energy{0.0};
  double grand_potential{0.0};
  std::vector<vec> forces;
};

// Precision of the device evaluation (no counterpart in the reference, which
// is FP64 throughout): FP64, or FP32 fields and transforms with FP64
// pointwise terms, reductions and state.
enum class Precision { FP64 = CDFT_PRECISION_FP64, Mixed = CDFT_PRECISION_MIXED };

// FMTWeights / MeanFieldWeights (weights.hpp, mean_field.hpp:23-35): handles
// on the same device context, for hard_sphere() and mean_field().
struct FMTWeights {
  std::shared_ptr<cdft_ctx> ctx;
  int fmt_model_id{CDFT_FMT_WHITEBEAR_II};  // the model the context evaluates
};
struct MeanFieldWeights {
  std::shared_ptr<cdft_ctx> ctx;
};

// functionals.hpp:19-33: the weights of a model.  Here they live on the GPU,
// in a cdft context, and Weights shares ownership of it.
struct Weights {
  fmt::FMTModel fmt_model = fmt::WhiteBearII{};
  std::shared_ptr<cdft_ctx> ctx;
  long n_points{0};
  FMTWeights fmt;
  MeanFieldWeights mean_field;
};

namespace detail {
inline cdft_interaction to_c(const physics::Interaction& in) {
  cdft_interaction c{};
  c.species_i = in.species_i;
  c.species_j = in.species_j;
  c.split = in.split == physics::potentials::SplitScheme::WeeksChandlerAndersen ? 0 : 1;
  c.weight_scheme = static_cast<int>(in.weight_scheme);
  c.alpha = 50.0;
  if (auto* lj = std::get_if<physics::potentials::LennardJones>(&in.potential)) {
    c.potential = 0;
    c.sigma = lj->sigma;
    c.epsilon = lj->epsilon;
    c.r_cutoff = lj->r_cutoff;
  } else if (auto* tw = std::get_if<physics::potentials::TenWoldeFrenkel>(&in.potential)) {
    c.potential = 1;
    c.sigma = tw->sigma;
    c.epsilon = tw->epsilon;
    c.r_cutoff = tw->r_cutoff;
    c.alpha = tw->alpha;
  } else {
    const auto& w = std::get<physics::potentials::WangRamirezDobnikarFrenkel>(in.potential);
    c.potential = 2;
    c.sigma = w.sigma;
    c.epsilon = w.epsilon;
    c.r_cutoff = w.r_cutoff;
  }
  return c;
}

inline bool all_zero(const vec& v) {
  for (double x : v)
    if (x != 0.0) return false;
  return true;
}
}  // namespace detail

// functionals.hpp:27-33 (+ the grid-synchronised a_vdw of functional.hpp:93-107).
inline Weights make_weights(const fmt::FMTModel& fmt_model, const physics::Model& model,
                            Precision precision = Precision::FP64, int device = 0) {
  if (model.species.empty()) throw std::invalid_argument("the model has no species");
  cdft_grid g{model.grid.dx, {model.grid.box_size[0], model.grid.box_size[1], model.grid.box_size[2]}};
  std::vector<double> diam;
  for (const auto& s : model.species) diam.push_back(s.hard_sphere_diameter);
  std::vector<cdft_interaction> inter;
  for (const auto& i : model.interactions) inter.push_back(detail::to_c(i));
  cdft_options o{};
  o.fmt_model = fmt_model.id();
  o.esfmt_A = fmt_model.esfmt_a();
  o.esfmt_B = fmt_model.esfmt_b();
  o.precision = static_cast<int>(precision);
  o.device = device;
  o.schedule = CDFT_SCHEDULE_AUTO;
  cdft_ctx* raw = nullptr;
  ::dft::detail::check(cdft_ctx_create(&g, diam.data(), static_cast<int>(diam.size()), inter.data(),
                                       static_cast<int>(inter.size()), model.temperature, &o, &raw));
  std::shared_ptr<cdft_ctx> ctx(raw, cdft_ctx_destroy);
  return Weights{fmt_model, ctx, model.grid.total_points(), FMTWeights{ctx, fmt_model.id()}, MeanFieldWeights{ctx}};
}

namespace detail {
// functionals.hpp:64-95 for more than one species: the densities, fields and
// forces travel species-major through cdft_total_species.
inline Result total_species(const physics::Model& model, const State& state, const Weights& weights) {
  const size_t S = state.species.size();
  if (S != model.species.size())
    throw std::invalid_argument("the state has " + std::to_string(S) + " species, the model " +
                                std::to_string(model.species.size()));
  const size_t n = static_cast<size_t>(model.grid.total_points());
  vec rho(S * n), v(S * n, 0.0), f(S * n), mu(S);
  bool has_v = false;
  for (size_t k = 0; k < S; ++k) {
    const Density& d = state.species[k].density;
    if (d.values.size() != n) throw std::invalid_argument("density size does not match the grid");
    std::copy(d.values.begin(), d.values.end(), rho.begin() + static_cast<long>(k * n));
    if (!d.external_field.empty() && !all_zero(d.external_field)) {
      if (d.external_field.size() != n) throw std::invalid_argument("external field size does not match the grid");
      std::copy(d.external_field.begin(), d.external_field.end(), v.begin() + static_cast<long>(k * n));
      has_v = true;
    }
    mu[k] = state.species[k].chemical_potential;
  }
  cdft_result cr{};
  ::dft::detail::check(cdft_total_species(weights.ctx.get(), rho.data(), has_v ? v.data() : nullptr, mu.data(), &cr,
                                          f.data()));
  Result r;
  for (size_t k = 0; k < S; ++k)
    r.forces.emplace_back(f.begin() + static_cast<long>(k * n), f.begin() + static_cast<long>((k + 1) * n));
  r.free_energy = cr.free_energy;
  r.grand_potential = cr.grand_potential;
  return r;
}
}  // namespace detail

// functionals.hpp:64-95: Omega, F and delta(beta Omega)/delta rho * dV of a
// state (external fields included when non-zero).
inline Result total(const physics::Model& model, const State& state, const Weights& weights) {
  if (state.species.size() != 1) return detail::total_species(model, state, weights);
  const SpeciesState& s = state.species[0];
  const long n = model.grid.total_points();
  if (static_cast<long>(s.density.values.size()) != n)
    throw std::invalid_argument("density has " + std::to_string(s.density.values.size()) + " values, grid has " +
                                std::to_string(n));
  const bool has_v = !s.density.external_field.empty() && !detail::all_zero(s.density.external_field);
  if (has_v && static_cast<long>(s.density.external_field.size()) != n)
    throw std::invalid_argument("external field size does not match the grid");
  Result r;
  r.forces.assign(1, vec(static_cast<size_t>(n)));
  cdft_result cr{};
  ::dft::detail::check(cdft_total(weights.ctx.get(), s.density.values.data(),
                                  has_v ? s.density.external_field.data() : nullptr, s.chemical_potential, &cr,
                                  r.forces[0].data()));
  r.free_energy = cr.free_energy;
  r.grand_potential = cr.grand_potential;
  return r;
}

namespace detail {
// One term of functionals::total through cdft_contributions: term 2 hard
// spheres, 3 mean field; the state's external fields and chemical potentials
// are read as total() reads them.
inline Contribution device_term(const std::shared_ptr<cdft_ctx>& ctx, const Grid& grid, const State& state,
                                int term) {
  if (!ctx) throw std::invalid_argument("the weights hold no device context");
  const size_t S = state.species.size(), n = static_cast<size_t>(grid.total_points());
  vec rho(S * n), v(S * n, 0.0), f(S * n), mu(S);
  bool has_v = false;
  for (size_t k = 0; k < S; ++k) {
    const Density& d = state.species[k].density;
    if (d.values.size() != n) throw std::invalid_argument("density size does not match the grid");
    std::copy(d.values.begin(), d.values.end(), rho.begin() + static_cast<long>(k * n));
    if (!d.external_field.empty() && !all_zero(d.external_field)) {
      if (d.external_field.size() != n) throw std::invalid_argument("external field size does not match the grid");
      std::copy(d.external_field.begin(), d.external_field.end(), v.begin() + static_cast<long>(k * n));
      has_v = true;
    }
    mu[k] = state.species[k].chemical_potential;
  }
  int n_species = 0;
  ::dft::detail::check(cdft_species_count(ctx.get(), &n_species));
  if (static_cast<size_t>(n_species) != S)
    throw std::invalid_argument("the state has " + std::to_string(S) + " species, the weights " +
                                std::to_string(n_species));
  cdft_result cr{};
  ::dft::detail::check(cdft_contributions(ctx.get(), rho.data(), has_v ? v.data() : nullptr, mu.data(), &cr, nullptr,
                                          nullptr, term == 2 ? f.data() : nullptr, term == 3 ? f.data() : nullptr));
  Contribution c;
  c.free_energy = term == 2 ? cr.hard_sphere : cr.mean_field;
  for (size_t k = 0; k < S; ++k)
    c.forces.emplace_back(f.begin() + static_cast<long>(k * n), f.begin() + static_cast<long>((k + 1) * n));
  return c;
}
}  // namespace detail

// hard_sphere.hpp:263-298: F_hs and delta F_hs / delta rho dV of every species
// (the model and species are those the weights were built with).
inline Contribution hard_sphere(const fmt::FMTModel& model, const Grid& grid, const State& state,
                                const std::vector<Species>& species, const FMTWeights& weights) {
  if (model.id() != weights.fmt_model_id)
    throw std::invalid_argument("hard_sphere: the FMT model differs from the one the weights were built for");
  if (species.size() != state.species.size()) throw std::invalid_argument("species and state sizes differ");
  return detail::device_term(weights.ctx, grid, state, 2);
}

// mean_field.hpp:233-281: F_mf and its forces for every species.
inline Contribution mean_field(const Grid& grid, const State& state, const std::vector<Species>& species,
                               const MeanFieldWeights& weights) {
  if (species.size() != state.species.size()) throw std::invalid_argument("species and state sizes differ");
  return detail::device_term(weights.ctx, grid, state, 3);
}
}  // namespace functionals

// init.hpp:11-60 (namespace dft::init)
namespace init {
// a uniform density for every species of the model (init.hpp:13-25)
inline State homogeneous(const physics::Model& model, double density) {
  const auto n = static_cast<size_t>(model.grid.total_points());
  State st{{}, model.temperature};
  for (size_t k = 0; k < model.species.size(); ++k)
    st.species.push_back(SpeciesState{Density{vec(n, density), vec(n, 0.0)}, vec(n, 0.0), 0.0, std::nullopt});
  return st;
}
inline State from_profile(const physics::Model& model, const vec& rho) {
  const auto n = static_cast<size_t>(model.grid.total_points());
  State st{{}, model.temperature};
  st.species.push_back(SpeciesState{Density{rho, vec(n, 0.0)}, vec(n, 0.0), 0.0, std::nullopt});
  return st;
}
inline State from_profiles(const physics::Model& model, const std::vector<vec>& profiles) {
  const auto n = static_cast<size_t>(model.grid.total_points());
  State st{{}, model.temperature};
  for (const auto& rho : profiles)
    st.species.push_back(SpeciesState{Density{rho, vec(n, 0.0)}, vec(n, 0.0), 0.0, std::nullopt});
  return st;
}
}  // namespace init

namespace algorithms {
namespace dynamics {
// dynamics.hpp:21
using ForceCallback = std::function<std::pair<double, std::vector<vec>>(const std::vector<vec>&)>;
}  // namespace dynamics
}  // namespace algorithms

namespace functionals {
namespace bulk {
// bulk/thermodynamics.hpp:200-255: the homogeneous-fluid equation of state of
// the model's species with the grid-synchronised a_vdw (one scalar evaluation
// on the host per call).  As in the reference, the free-energy density counts
// a cross pair as a rho_i rho_j and the chemical potentials give it half.
struct BulkThermodynamics {
  std::vector<Species> species;
  fmt::FMTModel fmt_model;
  std::shared_ptr<cdft_ctx> ctx;

  double free_energy_density(const vec& rho) const {
    check_size(rho);
    double f = 0.0;
    ::dft::detail::check(cdft_bulk_free_energy_density(ctx.get(), rho.data(), &f));
    return f;
  }
  vec chemical_potentials(const vec& rho) const {
    check_size(rho);
    vec mu(rho.size());
    ::dft::detail::check(cdft_bulk_chemical_potentials(ctx.get(), rho.data(), mu.data()));
    return mu;
  }
  double chemical_potential(const vec& rho, size_t species_idx) const {
    if (species_idx >= species.size()) throw std::invalid_argument("species index out of range");
    return chemical_potentials(rho)[species_idx];
  }
  double pressure(const vec& rho) const {
    check_size(rho);
    double p = 0.0;
    ::dft::detail::check(cdft_bulk_pressure(ctx.get(), rho.data(), &p));
    return p;
  }
  double grand_potential_density(const vec& rho) const { return -pressure(rho); }

 private:
  void check_size(const vec& rho) const {
    if (rho.size() != species.size())
      throw std::invalid_argument("expected " + std::to_string(species.size()) + " bulk densities, got " +
                                  std::to_string(rho.size()));
  }
};

// make_bulk_thermodynamics(species, weights) (thermodynamics.hpp:261-275)
inline BulkThermodynamics make_bulk_thermodynamics(const std::vector<Species>& species, const Weights& weights) {
  return BulkThermodynamics{species, weights.fmt_model, weights.ctx};
}
}  // namespace bulk

struct Functional;

// The callable that Functional::grand_potential_callback returns; Simulation
// recognises it and runs the whole DDFT loop on the device.
struct GrandPotentialCallback {
  const Functional* functional;
  double chemical_potential;
  std::pair<double, std::vector<vec>> operator()(const std::vector<vec>& densities) const;
};

// functional.hpp:22-107
struct Functional {
  physics::Model model;
  Weights weights;

  Result evaluate(const State& state) const { return total(model, state, weights); }

  Result evaluate(const vec& rho, double chemical_potential) const {
    auto state = init::from_profile(model, rho);
    state.species[0].chemical_potential = chemical_potential;
    return total(model, state, weights);
  }

  double grand_potential(const vec& rho, double chemical_potential) const {
    return evaluate(rho, chemical_potential).grand_potential;
  }

  // BulkThermodynamics::chemical_potential of the one-component fluid with the
  // grid-synchronised a_vdw (bulk/thermodynamics.hpp:141-151).
  double bulk_chemical_potential(double rho) const {
    double mu = 0.0;
    ::dft::detail::check(cdft_bulk_chemical_potential(weights.ctx.get(), rho, &mu));
    return mu;
  }

  algorithms::dynamics::ForceCallback grand_potential_callback(double chemical_potential) const {
    return GrandPotentialCallback{this, chemical_potential};
  }

  // functional.hpp:48-50
  bulk::BulkThermodynamics bulk() const { return bulk::make_bulk_thermodynamics(model.species, weights); }
};

inline std::pair<double, std::vector<vec>> GrandPotentialCallback::operator()(const std::vector<vec>& densities) const {
  auto state = init::from_profiles(functional->model, densities);
  for (auto& sp : state.species) sp.chemical_potential = chemical_potential;
  auto result = total(functional->model, state, functional->weights);
  return {result.grand_potential, std::move(result.forces)};
}

// functional.hpp:93-107
inline Functional make_functional(const fmt::FMTModel& fmt_model, physics::Model model,
                                  Precision precision = Precision::FP64, int device = 0) {
  auto w = make_weights(fmt_model, model, precision, device);
  return Functional{std::move(model), std::move(w)};
}
}  // namespace functionals

namespace algorithms {
namespace fire {
// fire.hpp:15-25: the optimiser's state between steps.
struct FireState {
  std::vector<vec> x;
  std::vector<vec> v;
  double dt{0.0};
  double alpha{0.0};
  int n_positive{0};
  int n_negative{0};
  double energy{0.0};
  double rms_force{0.0};
  int iteration{0};
  bool converged{false};
};

// fire.hpp:30: (energy, forces = -gradient) at positions x.
using ForceFunction = std::function<std::pair<double, std::vector<vec>>(const std::vector<vec>&)>;

// fire.hpp:33-178: FIRE2 over any force function (defaults as there).  The
// update runs on the host over the caller's vectors; `compute` does the heavy
// work (the functional's callbacks evaluate on the device).
// Minimizer::grand_potential runs its whole FIRE loop on the device instead.
struct Fire {
  double dt{1e-3};
  double dt_max{1e-2};
  double dt_min{1e-8};
  double alpha_start{0.1};
  double f_inc{1.1};
  double f_dec{0.5};
  double f_alpha{0.99};
  int n_delay{5};
  int max_uphill{20};
  double force_tolerance{0.1};
  int max_steps{10000};

  FireState initialize(std::vector<vec> x0, const ForceFunction& compute) const {
    auto ef = compute(x0);
    FireState st;
    st.v.reserve(x0.size());
    for (const auto& xs : x0) st.v.emplace_back(xs.size(), 0.0);
    st.x = std::move(x0);
    st.dt = dt;
    st.alpha = alpha_start;
    st.energy = ef.first;
    st.rms_force = rms(ef.second);
    st.converged = st.rms_force < force_tolerance;
    return st;
  }

  // One FIRE2 step from `forces` (the forces at state.x); returns the new
  // state and the forces at its positions.
  std::pair<FireState, std::vector<vec>> step(FireState st, const std::vector<vec>& forces,
                                              const ForceFunction& compute) const {
    if (forces.size() != st.x.size()) throw std::invalid_argument("fire::step: force and position counts differ");
    double power = 0.0;  // P = v . f
    for (size_t s = 0; s < st.x.size(); ++s) power += dot(st.v[s], forces[s]);
    if (power > 0.0 || st.iteration == 0) {
      ++st.n_positive;
      st.n_negative = 0;
      if (st.n_positive > n_delay) {
        st.dt = std::min(st.dt * f_inc, dt_max);
        st.alpha *= f_alpha;
      }
    } else {
      st.n_positive = 0;
      if (++st.n_negative > max_uphill) throw std::runtime_error("fire::step: exceeded max uphill steps");
      // back up half a step with the current dt and stop, then shrink dt (FIRE2)
      for (size_t s = 0; s < st.x.size(); ++s)
        for (size_t i = 0; i < st.x[s].size(); ++i) {
          st.x[s][i] -= 0.5 * st.dt * st.v[s][i];
          st.v[s][i] = 0.0;
        }
      if (st.iteration > n_delay) {
        st.dt = std::max(st.dt * f_dec, dt_min);
        st.alpha = alpha_start;
      }
    }
    // semi-implicit Euler: velocity, mixing, then position
    double vv = 0.0, ff = 0.0;
    for (size_t s = 0; s < st.x.size(); ++s) {
      for (size_t i = 0; i < st.v[s].size(); ++i) st.v[s][i] += st.dt * forces[s][i];
      vv += dot(st.v[s], st.v[s]);
      ff += dot(forces[s], forces[s]);
    }
    const double vn = std::sqrt(vv), fn = std::sqrt(ff);
    if (fn > 0.0) {
      const double c = st.alpha * (vn / fn);
      for (size_t s = 0; s < st.x.size(); ++s)
        for (size_t i = 0; i < st.v[s].size(); ++i) st.v[s][i] = (1.0 - st.alpha) * st.v[s][i] + c * forces[s][i];
    }
    for (size_t s = 0; s < st.x.size(); ++s)
      for (size_t i = 0; i < st.x[s].size(); ++i) st.x[s][i] += st.dt * st.v[s][i];
    auto ef = compute(st.x);
    st.energy = ef.first;
    st.rms_force = rms(ef.second);
    ++st.iteration;
    st.converged = st.rms_force < force_tolerance;
    return {std::move(st), std::move(ef.second)};
  }

  // Steps until converged or max_steps (fire.hpp:166-176; the forces at x0
  // are computed twice, as there).
  FireState minimize(std::vector<vec> x0, const ForceFunction& compute) const {
    FireState st = initialize(std::move(x0), compute);
    std::vector<vec> forces = compute(st.x).second;
    for (int i = 0; i < max_steps && !st.converged; ++i) {
      auto next = step(std::move(st), forces, compute);
      st = std::move(next.first);
      forces = std::move(next.second);
    }
    return st;
  }

 private:
  static double dot(const vec& a, const vec& b) {
    if (a.size() != b.size()) throw std::invalid_argument("fire: vector sizes differ");
    double t = 0.0;
    for (size_t i = 0; i < a.size(); ++i) t += a[i] * b[i];
    return t;
  }
  static double rms(const std::vector<vec>& f) {
    double t = 0.0, n = 0.0;
    for (const auto& fs : f) {
      t += dot(fs, fs);
      n += static_cast<double>(fs.size());
    }
    return std::sqrt(t / n);
  }
};
}  // namespace fire

namespace minimization {
// minimization.hpp:27-39
struct Unbounded {
  double rho_min{1e-18};
};
struct Bounded {
  double rho_min{1e-18};
  double rho_max{1.0};
};
using Parametrization = std::variant<Unbounded, Bounded>;

// minimization.hpp:114-121
struct Result {
  std::vector<vec> densities;
  double free_energy{0.0};
  double grand_potential{0.0};
  double rms_force{0.0};
  int iterations{0};
  bool converged{false};
};

// minimization.hpp:123-134: grand_potential runs FIRE2 in the alias variable
// on the device (the uphill limit throws std::runtime_error, as fire.hpp:106-108).
struct Minimizer {
  fire::Fire fire{};
  Parametrization param = Unbounded{};
  bool use_homogeneous_boundary{true};
  int log_interval{100};

  // minimization.hpp:280-353
  Result grand_potential(const physics::Model& model, const functionals::Weights& weights, const vec& initial_density,
                         double chemical_potential) const {
    return run(model, weights, initial_density, chemical_potential, -1.0);
  }

  // minimization.hpp:182-278
  Result fixed_mass(const physics::Model& model, const functionals::Weights& weights, const vec& initial_density,
                    double chemical_potential, double target_mass) const {
    return run(model, weights, initial_density, chemical_potential, target_mass);
  }

 private:
  Result run(const physics::Model& model, const functionals::Weights& weights, const vec& initial_density,
             double chemical_potential, double target_mass) const {
    if (static_cast<long>(initial_density.size()) != model.grid.total_points())
      throw std::invalid_argument("initial density size does not match the grid");
    cdft_fire f{fire.dt,      fire.dt_max,  fire.dt_min,     fire.alpha_start,    fire.f_inc,   fire.f_dec,
                fire.f_alpha, fire.n_delay, fire.max_uphill, fire.force_tolerance, fire.max_steps, 0.0};
    cdft_parametrization p{};
    if (auto* u = std::get_if<Unbounded>(&param)) {
      p.kind = 0;
      p.rho_min = u->rho_min;
      p.rho_max = 1.0;
    } else {
      const auto& b = std::get<Bounded>(param);
      p.kind = 1;
      p.rho_min = b.rho_min;
      p.rho_max = b.rho_max;
    }
    Result r;
    r.densities.assign(1, vec(initial_density.size()));
    cdft_min_result cr{};
    const int homog = use_homogeneous_boundary ? 1 : 0;
    if (target_mass > 0.0)
      ::dft::detail::check(cdft_minimize_fixed_mass(weights.ctx.get(), &f, &p, homog, initial_density.data(), nullptr,
                                                    chemical_potential, target_mass, r.densities[0].data(), &cr,
                                                    nullptr));
    else
      ::dft::detail::check(cdft_minimize_grand_potential(weights.ctx.get(), &f, &p, homog, initial_density.data(),
                                                         nullptr, chemical_potential, r.densities[0].data(), &cr,
                                                         nullptr));
    r.free_energy = cr.free_energy;
    r.grand_potential = cr.grand_potential;
    r.rms_force = cr.rms_force;
    r.iterations = cr.iterations;
    r.converged = cr.converged != 0;
    return r;
  }
};
}  // namespace minimization

namespace saddle_point {
// EigenvalueResult (saddle_point.hpp:147-152)
struct EigenvalueResult {
  vec eigenvector;
  double eigenvalue{0.0};
  int iterations{0};
  bool converged{false};
};

// EigenvalueSolver (saddle_point.hpp:154-270).  The reference's solve() takes
// any ForceFunction; on the device the force function is the functional's own
// grand-potential gradient at mu (with the boundary forces zeroed when
// fixed_boundary, as Search::find does), so solve() takes the weights and mu.
// hessian_eps = 0 selects exact Hessian-vector products (FP64 functionals).
struct EigenvalueSolver {
  double tolerance{1e-4};
  int max_iterations{300};
  double hessian_eps{1e-6};
  int log_interval{0};
  bool fixed_boundary{false};
  unsigned long long seed{0};  // start vector when no initial guess is given (stands in for arma::randn)

  EigenvalueResult solve(const functionals::Weights& weights, double chemical_potential, const vec& rho,
                         const vec& initial_guess = {}) const {
    if (static_cast<long>(rho.size()) != weights.n_points)
      throw std::invalid_argument("density size does not match the grid");
    cdft_eigen_config cfg{tolerance, max_iterations, hessian_eps, fixed_boundary ? 1 : 0};
    cdft_eigen_result cr{};
    EigenvalueResult r;
    r.eigenvector.assign(rho.size(), 0.0);
    const double* guess = initial_guess.size() == rho.size() ? initial_guess.data() : nullptr;
    ::dft::detail::check(cdft_smallest_eigenpair(weights.ctx.get(), &cfg, rho.data(), nullptr, chemical_potential,
                                                 guess, seed, r.eigenvector.data(), &cr));
    r.eigenvalue = cr.eigenvalue;
    r.iterations = cr.iterations;
    r.converged = cr.converged != 0;
    return r;
  }
};
// Result (saddle_point.hpp:313-320)
struct Result {
  vec density;
  vec eigenvector;
  double grand_potential{0.0};
  double eigenvalue{0.0};
  int outer_iterations{0};
  bool converged{false};
};

// Search (saddle_point.hpp:324-347): the critical cluster by eigenvector
// following at fixed mass.  The eigen solves use the start vectors of
// seed + outer iteration until eigenvalue_clear_count.
struct Search {
  fire::Fire fire{1e-3, 1.0, 1e-8, 0.01, 1.1, 0.5, 1.0, 5, 20, 1e-4, 100000};
  EigenvalueSolver eigenvalue{1e-4, 300, 1e-6};
  minimization::Parametrization param = minimization::Unbounded{1e-18};
  double omega_tolerance{1e-5};
  double tolerance_start{1e-3};
  double tolerance_factor{10.0};
  int max_outer_iterations{20};
  int eigenvalue_clear_count{2};
  unsigned long long seed{0};

  Result find(const physics::Model& model, const functionals::Weights& weights, const vec& initial_density,
              double chemical_potential, double target_mass) const {
    if (static_cast<long>(initial_density.size()) != model.grid.total_points())
      throw std::invalid_argument("initial density size does not match the grid");
    cdft_saddle_config cfg{};
    cfg.fire = {fire.dt,      fire.dt_max,  fire.dt_min,     fire.alpha_start,    fire.f_inc,   fire.f_dec,
                fire.f_alpha, fire.n_delay, fire.max_uphill, fire.force_tolerance, fire.max_steps, 0.0};
    cfg.eigen = {eigenvalue.tolerance, eigenvalue.max_iterations, eigenvalue.hessian_eps, 1};
    if (auto* u = std::get_if<minimization::Unbounded>(&param)) {
      cfg.param = {0, u->rho_min, 1.0};
    } else {
      const auto& b = std::get<minimization::Bounded>(param);
      cfg.param = {1, b.rho_min, b.rho_max};
    }
    cfg.omega_tolerance = omega_tolerance;
    cfg.tolerance_start = tolerance_start;
    cfg.tolerance_factor = tolerance_factor;
    cfg.max_outer_iterations = max_outer_iterations;
    cfg.eigenvalue_clear_count = eigenvalue_clear_count;
    Result r;
    r.density.assign(initial_density.size(), 0.0);
    r.eigenvector.assign(initial_density.size(), 0.0);
    cdft_saddle_result cr{};
    ::dft::detail::check(cdft_saddle_search(weights.ctx.get(), &cfg, initial_density.data(), nullptr,
                                            chemical_potential, target_mass, seed, r.density.data(),
                                            r.eigenvector.data(), &cr));
    r.grand_potential = cr.grand_potential;
    r.eigenvalue = cr.eigenvalue;
    r.outer_iterations = cr.outer_iterations;
    r.converged = cr.converged != 0;
    return r;
  }
};
}  // namespace saddle_point

namespace dynamics {
// dynamics.hpp:26-33
struct StepConfig {
  double dt{1e-4};
  double diffusion_coefficient{1.0};
  double min_density{1e-18};
  double dt_max{1.0};
  double fp_tolerance{1e-4};
  int fp_max_iterations{100};
};
// dynamics.hpp:43-49
struct Snapshot {
  int step{0};
  double time{0.0};
  double energy{0.0};
  std::vector<vec> densities;
};
// dynamics.hpp:52-58
struct SimulationResult {
  std::vector<vec> densities;
  std::vector<double> times;
  std::vector<double> energies;
  std::vector<Snapshot> snapshots;
  double mass_initial{0.0};
  double mass_final{0.0};
};
using BoundaryCondition = std::function<void(std::vector<vec>&)>;
using StopCondition = std::function<bool(int step, double time, double energy)>;

namespace detail {
// The host side of a cdft_hooks: the reference's std::function callbacks
// behind C function pointers.  An exception thrown by a callback is kept and
// rethrown once the library call has returned its error status.
struct Callbacks {
  size_t n = 0;
  const ForceCallback* force = nullptr;
  const BoundaryCondition* boundary = nullptr;
  const std::function<std::vector<vec>(const std::vector<vec>&)>* constraint = nullptr;
  const StopCondition* stop = nullptr;
  std::vector<Snapshot>* snapshots = nullptr;
  std::exception_ptr error;

  cdft_hooks hooks(int snapshot_interval = 0) {
    cdft_hooks h{};
    h.user = this;
    if (force) h.force = &force_fn;
    if (boundary || constraint) h.project = &project_fn;
    if (stop) h.stop = &stop_fn;
    if (snapshots) h.snapshot = &snapshot_fn;
    h.snapshot_interval = snapshot_interval;
    return h;
  }
  // the status of a library call; a callback's own exception comes first
  void check(int rc) const {
    if (rc != CDFT_OK && error) std::rethrow_exception(error);
    ::dft::detail::check(rc);
  }

 private:
  template <typename F>
  static int guard(void* user, F&& fn) {
    auto* self = static_cast<Callbacks*>(user);
    try {
      fn(*self);
      return 0;
    } catch (...) {
      self->error = std::current_exception();
      return 1;
    }
  }
  static int force_fn(void* user, const double* rho, double* energy, double* forces) {
    return guard(user, [&](Callbacks& c) {
      std::vector<vec> d{vec(rho, rho + c.n)};
      auto r = (*c.force)(d);
      if (r.second.empty() || r.second[0].size() != c.n)
        throw std::invalid_argument("the force callback returned forces of the wrong size");
      *energy = r.first;
      std::copy(r.second[0].begin(), r.second[0].end(), forces);
    });
  }
  static int project_fn(void* user, double* rho) {
    return guard(user, [&](Callbacks& c) {
      std::vector<vec> d{vec(rho, rho + c.n)};
      if (c.boundary) (*c.boundary)(d);
      else d = (*c.constraint)(d);
      if (d.size() != 1 || d[0].size() != c.n)
        throw std::invalid_argument("the boundary condition or constraint changed the densities' shape");
      std::copy(d[0].begin(), d[0].end(), rho);
    });
  }
  static int stop_fn(void* user, int step, double time, double energy, int* flag) {
    return guard(user, [&](Callbacks& c) { *flag = (*c.stop)(step, time, energy) ? 1 : 0; });
  }
  static int snapshot_fn(void* user, int step, double time, double energy, const double* rho) {
    return guard(user, [&](Callbacks& c) {
      c.snapshots->push_back(Snapshot{step, time, energy, {vec(rho, rho + c.n)}});
    });
  }
};

// A device context that only transforms (no FMT model, no interactions): the
// grid's FFTs, stencils and propagators for steps driven by a host force
// callback.
inline std::shared_ptr<cdft_ctx> transform_context(const Grid& grid, int device = 0) {
  cdft_grid g{grid.dx, {grid.box_size[0], grid.box_size[1], grid.box_size[2]}};
  const double diameter = 1.0;
  cdft_options o{};
  o.fmt_model = CDFT_FMT_NONE;
  o.esfmt_A = 1.0;
  o.esfmt_B = 0.0;
  o.precision = CDFT_PRECISION_FP64;
  o.device = device;
  o.schedule = CDFT_SCHEDULE_AUTO;
  cdft_ctx* raw = nullptr;
  ::dft::detail::check(cdft_ctx_create(&g, &diameter, 1, nullptr, 0, 1.0, &o, &raw));
  return std::shared_ptr<cdft_ctx>(raw, cdft_ctx_destroy);
}

// The device context and chemical potential a force callback evaluates with:
// the functional's own for its grand_potential_callback (no host callback),
// a transform-only context of the grid otherwise (`cb` is then set to
// receive the callable).
inline std::shared_ptr<cdft_ctx> context_for(const ForceCallback& compute, const Grid& grid,
                                             const std::vector<vec>& densities, Callbacks& cb, double* mu) {
  if (!compute) throw std::invalid_argument("the force callback is empty");
  if (densities.size() != 1 || static_cast<long>(densities[0].size()) != grid.total_points())
    throw std::invalid_argument("the device steps take one density of the grid's size");
  cb.n = densities[0].size();
  if (const auto* g = compute.target<functionals::GrandPotentialCallback>()) {
    if (g->functional->weights.n_points != grid.total_points())
      throw std::invalid_argument("the functional's grid differs from the step's grid");
    *mu = g->chemical_potential;
    return g->functional->weights.ctx;
  }
  cb.force = &compute;
  *mu = 0.0;
  return transform_context(grid);
}
}  // namespace detail

// dynamics.hpp:564-693: Lutsko's integrating-factor scheme with fixed-point
// iteration and adaptive dt, with the stencils, transforms and propagator on
// the device.  A Functional::grand_potential_callback is evaluated on the
// device; any other force callback, the boundary condition and the stop
// condition run on the host between device stages.  Snapshots are recorded
// as there: the initial state, then every snapshot_interval steps.
struct Simulation {
  StepConfig step;
  int n_steps{1000};
  int snapshot_interval{100};
  int log_interval{50};
  double energy_offset{0.0};
  BoundaryCondition boundary{};
  StopCondition stop_condition{};

  SimulationResult run(std::vector<vec> densities, const Grid& grid, const ForceCallback& force_fn) const {
    if (n_steps < 0) throw std::invalid_argument("Simulation::run: negative n_steps");
    SimulationResult r;
    detail::Callbacks cb;
    double mu = 0.0;
    const auto ctx = detail::context_for(force_fn, grid, densities, cb, &mu);
    if (boundary) cb.boundary = &boundary;
    if (stop_condition) cb.stop = &stop_condition;
    cb.snapshots = &r.snapshots;
    const cdft_hooks hooks = cb.hooks(snapshot_interval);
    cdft_step_config c{step.dt, step.diffusion_coefficient, step.min_density, step.dt_max, step.fp_tolerance,
                       step.fp_max_iterations};
    std::vector<double> e(static_cast<size_t>(n_steps) + 1), t(static_cast<size_t>(n_steps) + 1);
    cdft_sim_result cr{};
    cb.check(cdft_ddft_run(ctx.get(), &c, n_steps, nullptr, mu, &hooks, densities[0].data(), &cr, e.data(),
                           t.data()));
    // recorded like the reference: the initial state, then every log_interval
    // steps and the last step (dynamics.hpp:603-604,650-653)
    for (int s = 0; s <= cr.steps; ++s) {
      if (s == 0 || (log_interval > 0 && (s % log_interval == 0 || s == n_steps))) {
        r.times.push_back(t[static_cast<size_t>(s)]);
        r.energies.push_back(e[static_cast<size_t>(s)]);
      }
    }
    r.densities = std::move(densities);
    r.mass_initial = cr.mass_initial;
    r.mass_final = cr.mass_final;
    return r;
  }
};

// dynamics.hpp:395-401
struct StepResult {
  std::vector<vec> densities;
  double energy{0.0};
  double dt_used{0.0};
};

// compute_k_squared (dynamics.hpp:310-318): |k|^2 of the half spectrum, Nyquist positive
inline vec compute_k_squared(const Grid& grid) {
  const long nx = grid.shape[0], ny = grid.shape[1], nzh = grid.shape[2] / 2 + 1;
  const double PI = 3.14159265358979323846;
  const double dk[3] = {2.0 * PI / (nx * grid.dx), 2.0 * PI / (ny * grid.dx), 2.0 * PI / (grid.shape[2] * grid.dx)};
  vec k2(static_cast<size_t>(nx * ny * nzh));
  size_t i = 0;
  for (long ix = 0; ix < nx; ++ix)
    for (long iy = 0; iy < ny; ++iy)
      for (long iz = 0; iz < nzh; ++iz, ++i) {
        const double kx = dk[0] * static_cast<double>(ix <= nx / 2 ? ix : ix - nx);
        const double ky = dk[1] * static_cast<double>(iy <= ny / 2 ? iy : iy - ny);
        const double kz = dk[2] * static_cast<double>(iz);
        k2[i] = kx * kx + ky * ky + kz * kz;
      }
  return k2;
}

// diffusion_propagator (dynamics.hpp:320-322)
inline vec diffusion_propagator(const vec& k_squared, double D, double dt) {
  vec p(k_squared.size());
  for (size_t i = 0; i < p.size(); ++i) p[i] = std::exp(-D * k_squared[i] * dt);
  return p;
}

// split_operator_step (dynamics.hpp:410-462) on the device.  The device forms
// sqrt(exp(-D k^2 dt)) itself from config, so `propagator` must be
// diffusion_propagator(k_squared, config.diffusion_coefficient, config.dt),
// as every reference caller passes it.  A callback other than the
// functional's own runs on the host, with a transform context of the grid
// made for the call.
inline StepResult split_operator_step(const std::vector<vec>& densities, const Grid& grid, const vec& k_squared,
                                      const vec& propagator, const ForceCallback& compute, const StepConfig& config) {
  if (k_squared.size() != propagator.size()) throw std::invalid_argument("k_squared and propagator sizes differ");
  detail::Callbacks cb;
  double mu = 0.0;
  const auto ctx = detail::context_for(compute, grid, densities, cb, &mu);
  const cdft_hooks hooks = cb.hooks();
  cdft_step_config c{config.dt, config.diffusion_coefficient, config.min_density, config.dt_max,
                     config.fp_tolerance, config.fp_max_iterations};
  StepResult r{densities, 0.0, config.dt};
  cb.check(cdft_ddft_split_operator(ctx.get(), &c, 1, nullptr, mu, &hooks, r.densities[0].data(), &r.energy));
  return r;
}

namespace _internal {
// dynamics.hpp:70-77.  The device forms the propagator factors from the grid
// itself, so the state keeps the grid and the last dt, plus the transform
// context that steps driven by a host force callback run on (made at the
// first such step and reused by the later ones).
struct IntegratingFactorState {
  std::vector<long> shape;
  double dx{0.0};
  long fourier_total{0};
  double dt{0.0};
  std::shared_ptr<cdft_ctx> transforms;
};
// dynamics.hpp:79-129
inline IntegratingFactorState make_if_state(const Grid& grid) {
  return IntegratingFactorState{{grid.shape[0], grid.shape[1], grid.shape[2]},
                                grid.dx,
                                grid.shape[0] * grid.shape[1] * (grid.shape[2] / 2 + 1),
                                0.0,
                                nullptr};
}
}  // namespace _internal

// integrating_factor_step (dynamics.hpp:333-400) on the device: one step of
// the integrating-factor scheme with its fixed-point iteration; a restart
// divides config.dt by 10, as there, and the step never lengthens dt.  The
// functional's grand_potential_callback evaluates on the device; any other
// callback is called on the host (once per force evaluation: the reference's
// repeated evaluation of the unchanged initial density is taken once).
inline StepResult integrating_factor_step(const std::vector<vec>& densities, const Grid& grid,
                                          _internal::IntegratingFactorState& st, const ForceCallback& compute,
                                          StepConfig& config) {
  if (st.shape.size() != 3 || st.shape[0] != grid.shape[0] || st.shape[1] != grid.shape[1] ||
      st.shape[2] != grid.shape[2])
    throw std::invalid_argument("integrating_factor_step: the state was made for another grid");
  detail::Callbacks cb;
  double mu = 0.0;
  std::shared_ptr<cdft_ctx> ctx;
  if (const auto* g = compute.target<functionals::GrandPotentialCallback>()) {
    ctx = detail::context_for(compute, grid, densities, cb, &mu);
    (void)g;
  } else {
    if (!compute) throw std::invalid_argument("the force callback is empty");
    if (densities.size() != 1 || static_cast<long>(densities[0].size()) != grid.total_points())
      throw std::invalid_argument("the device steps take one density of the grid's size");
    if (!st.transforms) st.transforms = detail::transform_context(grid);
    ctx = st.transforms;
    cb.n = densities[0].size();
    cb.force = &compute;
  }
  const cdft_hooks hooks = cb.hooks();
  cdft_step_config c{config.dt, config.diffusion_coefficient, config.min_density, config.dt_max,
                     config.fp_tolerance, config.fp_max_iterations};
  StepResult r{densities, 0.0, config.dt};
  cdft_sim_result cr{};
  cb.check(cdft_ddft_run(ctx.get(), &c, 1, nullptr, mu, &hooks, r.densities[0].data(), &cr, nullptr, nullptr));
  config.dt = cr.dt;
  st.dt = cr.dt;
  r.energy = cr.energy;
  r.dt_used = cr.dt;
  return r;
}

// crank_nicholson_step (dynamics.hpp:472-538) on the device (callbacks as for
// split_operator_step)
inline StepResult crank_nicholson_step(const std::vector<vec>& densities, const Grid& grid, const vec& k_squared,
                                       const ForceCallback& compute, const StepConfig& config) {
  (void)k_squared;  // formed on the device from the grid
  detail::Callbacks cb;
  double mu = 0.0;
  const auto ctx = detail::context_for(compute, grid, densities, cb, &mu);
  const cdft_hooks hooks = cb.hooks();
  cdft_step_config c{config.dt, config.diffusion_coefficient, config.min_density, config.dt_max,
                     config.fp_tolerance, config.fp_max_iterations};
  StepResult r{densities, 0.0, config.dt};
  cb.check(cdft_ddft_crank_nicolson(ctx.get(), &c, 1, nullptr, mu, &hooks, r.densities[0].data(), &r.energy,
                                    nullptr));
  return r;
}
}  // namespace dynamics

namespace picard {
using ForceFunction = dynamics::ForceCallback;
using Constraint = std::function<std::vector<vec>(const std::vector<vec>&)>;

// fixed_mass_constraint (minimization.hpp:157-180) as a callable the device
// solver recognises (one species)
struct FixedMassConstraint {
  double target_mass;
  vec background;
  double cell_volume;
  double min_density;
  std::vector<vec> operator()(const std::vector<vec>& densities) const {
    auto result = densities;
    if (result.empty()) return result;
    if (result[0].size() != background.size())
      throw std::invalid_argument("fixed_mass_constraint: density and background sizes differ");
    double cur = 0.0;
    for (size_t i = 0; i < result[0].size(); ++i) cur += result[0][i] - background[i];
    cur *= cell_volume;
    if (std::abs(cur) > 1e-30)
      for (size_t i = 0; i < result[0].size(); ++i)
        result[0][i] = std::max(background[i] + target_mass / cur * (result[0][i] - background[i]), min_density);
    return result;
  }
};
inline Constraint fixed_mass_constraint(std::vector<double> target_masses, std::vector<vec> backgrounds,
                                        double cell_volume, double min_density = 1e-30) {
  if (target_masses.size() != 1 || backgrounds.size() != 1)
    throw std::invalid_argument("fixed_mass_constraint: one species on the device");
  return FixedMassConstraint{target_masses[0], std::move(backgrounds[0]), cell_volume, min_density};
}

// PicardResult (picard.hpp:17-23)
struct PicardResult {
  std::vector<vec> densities;
  double grand_potential{0.0};
  double residual{0.0};
  int iterations{0};
  bool converged{false};
};

// Picard (picard.hpp:25-33).  With the functional's grand_potential_callback
// solve() runs on the device, the fixed_mass_constraint too; another
// constraint is called on the host after every device update.  Picard has no
// transforms of its own, so with any other force callable the pointwise
// update runs on the host over the caller's vectors, as fire::Fire does.
struct Picard {
  double mixing{0.01};
  double min_density{1e-30};
  double tolerance{1e-6};
  int max_iterations{5000};
  int log_interval{500};

  PicardResult solve(std::vector<vec> densities, const ForceFunction& compute, double cell_volume,
                     const Constraint& constraint = {}) const {
    if (!compute) throw std::invalid_argument("Picard::solve: the force function is empty");
    const auto* cb = compute.target<functionals::GrandPotentialCallback>();
    if (cb == nullptr) return solve_host(std::move(densities), compute, cell_volume, constraint);
    if (densities.size() != 1) throw std::invalid_argument("Picard::solve on the GPU takes one density");
    const long n = cb->functional->weights.n_points;
    if (static_cast<long>(densities[0].size()) != n)
      throw std::invalid_argument("Picard::solve: the density has " + std::to_string(densities[0].size()) +
                                  " values, the functional's grid " + std::to_string(n));
    (void)cell_volume;  // the functional's grid fixes it
    const double* target = nullptr;
    const double* bg = nullptr;
    dynamics::detail::Callbacks hooks_cb;
    hooks_cb.n = densities[0].size();
    if (constraint) {
      if (const auto* fm = constraint.target<FixedMassConstraint>()) {
        if (static_cast<long>(fm->background.size()) != n)
          throw std::invalid_argument("fixed_mass_constraint: the background has " +
                                      std::to_string(fm->background.size()) + " values, the grid " +
                                      std::to_string(n));
        target = &fm->target_mass;
        bg = fm->background.data();
      } else {
        hooks_cb.constraint = &constraint;
      }
    }
    const cdft_hooks hooks = hooks_cb.hooks();
    cdft_picard c{mixing, min_density, tolerance, max_iterations};
    cdft_picard_result cr{};
    hooks_cb.check(cdft_picard_solve(cb->functional->weights.ctx.get(), &c, nullptr, cb->chemical_potential, target,
                                     bg, &hooks, densities[0].data(), &cr));
    return PicardResult{std::move(densities), cr.grand_potential, cr.residual, cr.iterations, cr.converged != 0};
  }

 private:
  // picard.hpp:63-124 over any force function, on the host
  PicardResult solve_host(std::vector<vec> densities, const ForceFunction& compute, double cell_volume,
                          const Constraint& constraint) const {
    double residual = 0.0, omega = 0.0, omega_prev = std::numeric_limits<double>::max();
    int iter = 0;
    for (iter = 0; iter < max_iterations; ++iter) {
      auto ef = compute(densities);
      omega = ef.first;
      if (ef.second.size() != densities.size())
        throw std::invalid_argument("Picard::solve: the force function returned another species count");
      double sum_f2 = 0.0, dof = 0.0;
      for (const auto& f : ef.second) {
        for (double x : f) sum_f2 += x * x;
        dof += static_cast<double>(f.size());
      }
      residual = std::sqrt(sum_f2 / dof);
      if (residual < tolerance || (iter > 0 && std::abs(omega - omega_prev) < tolerance)) break;
      omega_prev = omega;
      for (size_t s = 0; s < densities.size(); ++s) {
        if (ef.second[s].size() != densities[s].size())
          throw std::invalid_argument("Picard::solve: force and density sizes differ");
        for (size_t i = 0; i < densities[s].size(); ++i)
          densities[s][i] = std::max(densities[s][i] * std::exp(-mixing * ef.second[s][i] / cell_volume), min_density);
      }
      if (constraint) densities = constraint(densities);
    }
    const bool conv = residual < tolerance || (iter > 0 && std::abs(omega - omega_prev) < tolerance);
    return PicardResult{std::move(densities), omega, residual, iter, conv};
  }
};
}  // namespace picard
}  // namespace algorithms

}  // namespace dft

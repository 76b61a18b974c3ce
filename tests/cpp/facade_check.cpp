// Exercises the C++ facade (include/cdft/dft.hpp) the way the reference's own
// callers use its interfaces (make_functional -> evaluate / callback,
// Minimizer::grand_potential, Simulation::run).
//
//   facade_check cpu                       argument validation and error mapping (no GPU calls)
//   facade_check gpu RHO_IN MU_RHO OUT_DIR 16^3 LJ droplet: writes forces.bin, fire_rho.bin,
//                                          ddft_rho.bin, ... to OUT_DIR and one JSON line to stdout
//   facade_check callbacks RHO_IN OUT_DIR  the algorithms with callables other than the functional's own:
//                                          the ideal-gas callback of the reference's dynamics tests on
//                                          8^3, a total()-wrapping lambda, a reservoir boundary, a stop
//                                          condition, snapshots, Picard constraints; one JSON line
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>

#include "cdft/dft.hpp"

namespace {

int failures = 0;

template <typename E, typename F>
void expect_throw(const char* what, F&& f, const char* needle) {
  try {
    f();
  } catch (const E& e) {
    if (needle && std::string(e.what()).find(needle) == std::string::npos) {
      std::printf("FAIL %s: message '%s' lacks '%s'\n", what, e.what(), needle);
      ++failures;
    }
    return;
  } catch (const std::exception& e) {
    std::printf("FAIL %s: wrong exception type: %s\n", what, e.what());
    ++failures;
    return;
  }
  std::printf("FAIL %s: no exception\n", what);
  ++failures;
}

dft::physics::Model lj_model(double dx, double box, int n_species = 1) {
  using namespace dft;
  physics::Model m;
  m.grid = make_grid(dx, {box, box, box});
  for (int i = 0; i < n_species; ++i) m.species.push_back(Species{"argon", 1.0});
  physics::Interaction in;
  in.potential = physics::potentials::make_lennard_jones(1.0, 1.0, 2.5);
  m.interactions.push_back(in);
  m.temperature = 0.8;
  return m;
}

// fire.cpp:12-19 of the reference tests: E = sum x^2 / 2, f = -x
std::pair<double, std::vector<dft::vec>> quadratic_force(const std::vector<dft::vec>& x) {
  double e = 0.0;
  std::vector<dft::vec> f(x.size());
  for (size_t s = 0; s < x.size(); ++s) {
    f[s].resize(x[s].size());
    for (size_t i = 0; i < x[s].size(); ++i) {
      e += 0.5 * x[s][i] * x[s][i];
      f[s][i] = -x[s][i];
    }
  }
  return {e, f};
}

double norm(const dft::vec& v) {
  double t = 0.0;
  for (double a : v) t += a * a;
  return std::sqrt(t);
}

void check(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAIL %s\n", what);
    ++failures;
  }
}

// The known-answer tests of the reference's fire.cpp:21-137 on the facade's host FIRE2.
void fire_checks() {
  using dft::algorithms::fire::Fire;
  {
    Fire cfg;
    const auto st = cfg.initialize({{1.0, 2.0, 3.0}}, quadratic_force);
    check(std::abs(st.energy - 7.0) < 1e-15 && st.rms_force > 0.0 && st.iteration == 0 && !st.converged &&
              st.v.size() == 1 && norm(st.v[0]) == 0.0,
          "fire initialize");
  }
  {
    Fire cfg;
    cfg.dt = 0.01;
    cfg.force_tolerance = 1e-6;
    const auto st = cfg.initialize({{1.0, 2.0, 3.0}}, quadratic_force);
    const auto next = cfg.step(st, quadratic_force(st.x).second, quadratic_force);
    check(next.first.energy < st.energy && next.first.iteration == 1, "fire step reduces energy");
  }
  for (const auto& x0 : std::vector<std::vector<dft::vec>>{{{5.0}}, {{3.0, -2.0, 1.0}}, {{2.0, 1.0}, {-1.0, 3.0}}}) {
    Fire cfg;
    cfg.dt = 0.01;
    cfg.force_tolerance = 1e-6;
    cfg.max_steps = 5000;
    const auto r = cfg.minimize(x0, quadratic_force);
    bool small = true;
    for (const auto& xs : r.x) small = small && norm(xs) < 1e-3;
    check(r.converged && small && r.energy < 1e-6, "fire minimize quadratic");
  }
  {
    auto rosenbrock = [](const std::vector<dft::vec>& x) -> std::pair<double, std::vector<dft::vec>> {
      const double a = x[0][0], b = x[0][1];
      const double e = (1.0 - a) * (1.0 - a) + 100.0 * (b - a * a) * (b - a * a);
      const double gx = -2.0 * (1.0 - a) - 400.0 * a * (b - a * a), gy = 200.0 * (b - a * a);
      return {e, {{-gx, -gy}}};
    };
    Fire cfg;
    cfg.dt = 1e-4;
    cfg.dt_max = 1e-3;
    cfg.force_tolerance = 0.1;
    cfg.max_steps = 50000;
    const auto r = cfg.minimize({{-1.0, 1.0}}, rosenbrock);
    check(r.energy < 0.1 * rosenbrock({{-1.0, 1.0}}).first, "fire Rosenbrock progress");
  }
  {
    Fire cfg;
    cfg.force_tolerance = 1e-6;
    check(cfg.initialize({{0.0, 0.0, 0.0}}, quadratic_force).converged, "fire at minimum");
  }
  {
    int count = 0;
    auto alternating = [&count](const std::vector<dft::vec>& x) -> std::pair<double, std::vector<dft::vec>> {
      ++count;
      const double sg = (count % 2 == 0) ? -1.0 : 1.0;
      const double e = sg * (x[0][0] * x[0][0] + x[0][1] * x[0][1]);
      return {e, {{sg * -2.0 * x[0][0], sg * -2.0 * x[0][1]}}};
    };
    Fire cfg;
    cfg.dt = 0.5;
    cfg.dt_max = 0.5;
    cfg.n_delay = 0;
    cfg.max_uphill = 3;
    cfg.force_tolerance = 1e-10;
    cfg.max_steps = 5000;
    expect_throw<std::runtime_error>("fire uphill", [&] { (void)cfg.minimize({{1.0, 1.0}}, alternating); },
                                     "exceeded max uphill steps");
  }
}

int run_cpu() {
  using namespace dft;
  fire_checks();
  const Grid g = make_grid(0.1, {3.2, 1.6, 0.8});
  if (g.shape[0] != 32 || g.shape[1] != 16 || g.shape[2] != 8 || g.total_points() != 4096 ||
      g.flat_index(1, 2, 3) != 3 + 8 * (2 + 16 * 1)) {
    std::printf("FAIL make_grid shape/index\n");
    ++failures;
  }
  expect_throw<std::invalid_argument>("dx", [] { (void)make_grid(0.0, {1.0, 1.0, 1.0}); }, "dx must be positive");
  expect_throw<std::invalid_argument>("box", [] { (void)make_grid(0.1, {1.0, -1.0, 1.0}); },
                                      "box dimension must be positive");
  expect_throw<std::invalid_argument>("commensurate", [] { (void)make_grid(0.3, {1.0, 1.0, 1.0}); },
                                      "not commensurate");
  expect_throw<std::invalid_argument>("fmt name", [] { (void)functionals::fmt::FMTModel::from_name("WhiteBearIII"); },
                                      "Unknown FMT model: WhiteBearIII");
  if (functionals::fmt::FMTModel::from_name("WhiteBearII").id() != CDFT_FMT_WHITEBEAR_II) {
    std::printf("FAIL fmt id\n");
    ++failures;
  }
  // rejected before any device work: a state whose species count differs from the model's
  expect_throw<std::invalid_argument>("species", [] {
    const auto model = lj_model(0.1, 1.6, 1);
    const auto n = static_cast<size_t>(model.grid.total_points());
    const auto state = init::from_profiles(model, {vec(n, 0.1), vec(n, 0.1)});
    (void)functionals::total(model, state, functionals::Weights{});
  }, "the state has 2 species, the model 1");
  {
    const auto st = init::homogeneous(lj_model(0.1, 1.6, 3), 0.25);
    if (st.species.size() != 3 || st.species[2].density.values.size() != 4096 ||
        st.species[1].density.values[17] != 0.25 || st.temperature != 0.8) {
      std::printf("FAIL init::homogeneous\n");
      ++failures;
    }
  }
  std::printf(failures ? "facade cpu checks: %d failures\n" : "facade cpu checks: ok\n", failures);
  return failures ? 1 : 0;
}

dft::vec read_doubles(const std::string& path, long n) {
  dft::vec v(static_cast<size_t>(n));
  std::ifstream f(path, std::ios::binary);
  f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(8 * n));
  if (!f) throw std::runtime_error("cannot read " + path);
  return v;
}

void write_doubles(const std::string& path, const dft::vec& v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(8 * v.size()));
}

int run_gpu(const std::string& rho_path, double mu_rho, const std::string& out) {
  using namespace dft;
  const auto model = lj_model(0.25, 4.0);  // 16^3
  auto func = functionals::make_functional(functionals::fmt::WhiteBearII{}, model);
  const long n = model.grid.total_points();
  const vec rho = read_doubles(rho_path, n);
  const double mu = func.bulk_chemical_potential(mu_rho);

  const auto res = func.evaluate(rho, mu);
  write_doubles(out + "/forces.bin", res.forces[0]);
  // the callback the dynamics use gives the same numbers
  const auto cb = func.grand_potential_callback(mu);
  const auto [omega_cb, forces_cb] = cb({rho});
  double dmax = 0.0;
  for (long i = 0; i < n; ++i) dmax = std::max(dmax, std::abs(forces_cb[0][static_cast<size_t>(i)] - res.forces[0][static_cast<size_t>(i)]));

  algorithms::minimization::Minimizer mz;
  mz.fire.force_tolerance = 1e-9;
  mz.fire.max_steps = 40;
  const auto mres = mz.grand_potential(model, func.weights, rho, mu);
  write_doubles(out + "/fire_rho.bin", mres.densities[0]);
  double mass0 = 0.0;
  for (double r : rho) mass0 += r;
  mass0 *= model.grid.cell_volume();
  const auto fres = mz.fixed_mass(model, func.weights, rho, mu, 0.9 * mass0);
  write_doubles(out + "/fixed_mass_rho.bin", fres.densities[0]);
  algorithms::saddle_point::EigenvalueSolver es;
  es.tolerance = 1e-6;
  es.seed = 17;
  const auto eres = es.solve(func.weights, mu, rho);
  algorithms::saddle_point::Search search;
  search.fire.max_steps = 300;
  search.eigenvalue.max_iterations = 100;
  search.max_outer_iterations = 3;
  search.seed = 1;
  const auto sres_saddle = search.find(model, func.weights, rho, mu, mass0);
  algorithms::dynamics::StepConfig sc;
  sc.dt = 1e-3;
  const auto k2 = algorithms::dynamics::compute_k_squared(model.grid);
  const auto prop = algorithms::dynamics::diffusion_propagator(k2, sc.diffusion_coefficient, sc.dt);
  const auto split = algorithms::dynamics::split_operator_step({rho}, model.grid, k2, prop, cb, sc);
  sc.fp_tolerance = 1e-9;
  const auto cn = algorithms::dynamics::crank_nicholson_step({rho}, model.grid, k2, cb, sc);
  algorithms::picard::Picard pc;
  pc.mixing = 0.02;
  pc.tolerance = 1e-8;
  pc.max_iterations = 200;
  const auto pres = pc.solve({rho}, cb, model.grid.cell_volume(),
                             algorithms::picard::fixed_mass_constraint({mass0}, {vec(rho.size(), 0.0)},
                                                                       model.grid.cell_volume()));
  write_doubles(out + "/split_rho.bin", split.densities[0]);
  write_doubles(out + "/cn_rho.bin", cn.densities[0]);
  write_doubles(out + "/picard_rho.bin", pres.densities[0]);

  // a binary mixture (diameters 1.0 and 0.8, a self and a cross pair) through the same facade calls
  physics::Model mix = model;
  mix.species.push_back(Species{"krypton", 0.8});
  physics::Interaction cross;
  cross.species_i = 0;
  cross.species_j = 1;
  cross.potential = physics::potentials::make_lennard_jones(0.9, 0.7, 2.5);
  mix.interactions.push_back(cross);
  auto mix_func = functionals::make_functional(functionals::fmt::WhiteBearII{}, mix);
  vec rho_b(rho.size());
  for (size_t i = 0; i < rho.size(); ++i) rho_b[i] = 0.5 * rho[i] + 0.02;
  State mix_state = init::from_profiles(mix, {rho, rho_b});
  mix_state.species[0].chemical_potential = mu;
  mix_state.species[1].chemical_potential = -0.3;
  const auto mres_mix = functionals::total(mix, mix_state, mix_func.weights);
  vec mix_forces(mres_mix.forces[0]);
  mix_forces.insert(mix_forces.end(), mres_mix.forces[1].begin(), mres_mix.forces[1].end());
  write_doubles(out + "/mix_forces.bin", mix_forces);
  // the hard-sphere and mean-field terms of the mixture through the reference's signatures
  const auto hs_term = functionals::hard_sphere(functionals::fmt::WhiteBearII{}, mix.grid, mix_state, mix.species,
                                                mix_func.weights.fmt);
  const auto mf_term = functionals::mean_field(mix.grid, mix_state, mix.species, mix_func.weights.mean_field);
  vec terms(hs_term.forces[0]);
  for (const auto* f : {&hs_term.forces[1], &mf_term.forces[0], &mf_term.forces[1]})
    terms.insert(terms.end(), f->begin(), f->end());
  write_doubles(out + "/mix_terms.bin", terms);
  expect_throw<std::invalid_argument>("hard_sphere model", [&] {
    (void)functionals::hard_sphere(functionals::fmt::Rosenfeld{}, mix.grid, mix_state, mix.species,
                                   mix_func.weights.fmt);
  }, "FMT model differs");
  const auto eos = mix_func.bulk();
  const vec mix_mu = eos.chemical_potentials({0.3, 0.2});
  const double mix_p = eos.pressure({0.3, 0.2});

  // three single integrating-factor steps at fixed dt
  auto if_state = algorithms::dynamics::_internal::make_if_state(model.grid);
  algorithms::dynamics::StepConfig if_cfg;
  if_cfg.dt = 1e-3;
  if_cfg.dt_max = 1e-3;
  std::vector<vec> if_rho{rho};
  double if_energy = 0.0;
  for (int k = 0; k < 3; ++k) {
    const auto sr = algorithms::dynamics::integrating_factor_step(if_rho, model.grid, if_state, cb, if_cfg);
    if_rho = sr.densities;
    if_energy = sr.energy;
  }
  write_doubles(out + "/ifstep_rho.bin", if_rho[0]);

  algorithms::dynamics::Simulation sim;
  sim.step.dt = 1e-3;
  sim.step.dt_max = 1e-3;
  sim.n_steps = 6;
  sim.log_interval = 2;
  const auto sres = sim.run({rho}, model.grid, cb);
  write_doubles(out + "/ddft_rho.bin", sres.densities[0]);
  // a density of the wrong size is refused before any device work
  bool rejected = false;
  try {
    (void)sim.run({vec(10, 0.1)}, model.grid, cb);
  } catch (const std::invalid_argument&) {
    rejected = true;
  }

  std::printf("{\"mu\": %.17g, \"free_energy\": %.17g, \"grand_potential\": %.17g, \"callback_omega\": %.17g, "
              "\"callback_force_maxdiff\": %.3g, \"fire_iterations\": %d, \"fire_grand_potential\": %.17g, "
              "\"fixed_mass_target\": %.17g, \"fixed_mass_iterations\": %d, \"fixed_mass_free_energy\": %.17g, "
              "\"eigenvalue\": %.17g, \"eigen_iterations\": %d, \"saddle_omega\": %.17g, "
              "\"saddle_eigenvalue\": %.17g, \"saddle_outer\": %d, \"split_energy\": %.17g, "
              "\"cn_energy\": %.17g, \"picard_iterations\": %d, \"mixture_grand_potential\": %.17g, "
              "\"mixture_bulk_mu\": [%.17g, %.17g], \"mixture_bulk_pressure\": %.17g, "
              "\"ifstep_energy\": %.17g, \"ifstep_dt\": %.17g, \"mixture_hs_energy\": %.17g, "
              "\"mixture_mf_energy\": %.17g, "
              "\"ddft_times\": [",
              mu, res.free_energy, res.grand_potential, omega_cb, dmax, mres.iterations, mres.grand_potential,
              0.9 * mass0, fres.iterations, fres.free_energy, eres.eigenvalue, eres.iterations,
              sres_saddle.grand_potential, sres_saddle.eigenvalue, sres_saddle.outer_iterations, split.energy,
              cn.energy, pres.iterations, mres_mix.grand_potential, mix_mu[0], mix_mu[1], mix_p, if_energy,
              if_cfg.dt, hs_term.free_energy, mf_term.free_energy);
  for (size_t i = 0; i < sres.times.size(); ++i) std::printf("%s%.17g", i ? ", " : "", sres.times[i]);
  std::printf("], \"ddft_energies\": [");
  for (size_t i = 0; i < sres.energies.size(); ++i) std::printf("%s%.17g", i ? ", " : "", sres.energies[i]);
  std::printf("], \"ddft_mass\": [%.17g, %.17g], \"wrong_size_rejected\": %s}\n", sres.mass_initial,
              sres.mass_final, rejected ? "true" : "false");
  return 0;
}

// The reference's dynamics tests' force callback (tests/unit/algorithms/dynamics.cpp:17-27):
// ideal gas only, F = sum rho (ln rho - 1) dV, force ln(rho) dV, densities clamped at 1e-18.
std::pair<double, std::vector<dft::vec>> ideal_gas_forces(const std::vector<dft::vec>& densities, double dv) {
  double energy = 0.0;
  std::vector<dft::vec> forces(densities.size());
  for (size_t s = 0; s < densities.size(); ++s) {
    forces[s].resize(densities[s].size());
    for (size_t i = 0; i < densities[s].size(); ++i) {
      const double r = std::max(densities[s][i], 1e-18);
      energy += r * (std::log(r) - 1.0) * dv;
      forces[s][i] = std::log(r) * dv;
    }
  }
  return {energy, forces};
}

void write_json_array(const char* key, const std::vector<double>& v, bool last = false) {
  std::printf("\"%s\": [", key);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? ", " : "", v[i]);
  std::printf(last ? "]" : "], ");
}

int run_callbacks(const std::string& rho_path, const std::string& out) {
  using namespace dft;
  namespace dyn = algorithms::dynamics;
  // 1. the ideal-gas callback on the reference test's grid (dx = 0.5, box 4: 8^3) with a Gaussian bump
  const Grid grid{0.5, {4.0, 4.0, 4.0}, {8, 8, 8}};
  const double dv = grid.cell_volume();
  const long n8 = grid.total_points();
  vec rho8(static_cast<size_t>(n8), 0.5);
  for (long ix = 0; ix < 8; ++ix)
    for (long iy = 0; iy < 8; ++iy)
      for (long iz = 0; iz < 8; ++iz) {
        const double x = (ix - 4.0) * 0.5, y = (iy - 4.0) * 0.5, z = (iz - 4.0) * 0.5;
        rho8[static_cast<size_t>(grid.flat_index(ix, iy, iz))] += 0.1 * std::exp(-(x * x + y * y + z * z));
      }
  write_doubles(out + "/ideal_rho0.bin", rho8);
  dyn::ForceCallback ideal = [dv](const std::vector<vec>& d) { return ideal_gas_forces(d, dv); };
  dyn::Simulation sim;
  sim.step.dt = 1e-2;
  sim.step.dt_max = 1e-2;
  sim.n_steps = 5;
  sim.log_interval = 1;
  sim.snapshot_interval = 2;
  const auto ig = sim.run({rho8}, grid, ideal);
  write_doubles(out + "/ideal_ddft_rho.bin", ig.densities[0]);
  const auto k2 = dyn::compute_k_squared(grid);
  dyn::StepConfig sc;
  sc.dt = 1e-3;
  const auto prop = dyn::diffusion_propagator(k2, sc.diffusion_coefficient, sc.dt);
  const auto igs = dyn::split_operator_step({rho8}, grid, k2, prop, ideal, sc);
  sc.fp_tolerance = 1e-10;
  const auto igc = dyn::crank_nicholson_step({rho8}, grid, k2, ideal, sc);
  write_doubles(out + "/ideal_split_rho.bin", igs.densities[0]);
  write_doubles(out + "/ideal_cn_rho.bin", igc.densities[0]);
  auto st8 = dyn::_internal::make_if_state(grid);
  dyn::StepConfig ic;
  ic.dt = 1e-2;
  ic.dt_max = 1e-2;
  std::vector<vec> r8{rho8};
  for (int k = 0; k < 3; ++k) r8 = dyn::integrating_factor_step(r8, grid, st8, ideal, ic).densities;
  write_doubles(out + "/ideal_ifstep_rho.bin", r8[0]);

  // 2. the LJ droplet functional behind a total()-wrapping lambda (nucleation.cpp:279-291) and behind
  //    its own callback: the same trajectory
  const auto model = lj_model(0.25, 4.0);  // 16^3
  auto func = functionals::make_functional(functionals::fmt::WhiteBearII{}, model);
  const long n = model.grid.total_points();
  const vec rho = read_doubles(rho_path, n);
  const double mu = func.bulk_chemical_potential(0.05);
  int lambda_calls = 0;
  dyn::ForceCallback wrapped = [&](const std::vector<vec>& d) {
    ++lambda_calls;
    auto state = init::from_profile(model, d[0]);
    state.species[0].chemical_potential = mu;
    auto result = functionals::total(model, state, func.weights);
    return std::make_pair(result.grand_potential, result.forces);
  };
  dyn::Simulation sim16;
  sim16.step.dt = 1e-3;
  sim16.step.dt_max = 1e-3;
  sim16.n_steps = 6;
  sim16.log_interval = 2;
  const auto via_lambda = sim16.run({rho}, model.grid, wrapped);
  const auto via_device = sim16.run({rho}, model.grid, func.grand_potential_callback(mu));
  double traj_diff = 0.0;
  for (long i = 0; i < n; ++i)
    traj_diff = std::max(traj_diff, std::abs(via_lambda.densities[0][static_cast<size_t>(i)] -
                                             via_device.densities[0][static_cast<size_t>(i)]));
  double energy_diff = 0.0;
  for (size_t i = 0; i < via_lambda.energies.size(); ++i)
    energy_diff = std::max(energy_diff, std::abs(via_lambda.energies[i] - via_device.energies[i]));

  // 3. a reservoir boundary (dynamics.hpp:545-553: marked cells reset to a reservoir density after each
  //    step) on the x = 0 face, a stop condition after step 4 and snapshots every 2 steps
  std::vector<unsigned char> mask(static_cast<size_t>(n), 0);
  for (long iy = 0; iy < model.grid.shape[1]; ++iy)
    for (long iz = 0; iz < model.grid.shape[2]; ++iz) mask[static_cast<size_t>(model.grid.flat_index(0, iy, iz))] = 1;
  dyn::Simulation simb = sim16;
  simb.n_steps = 10;
  simb.snapshot_interval = 2;
  simb.boundary = [&](std::vector<vec>& d) {
    for (auto& r : d)
      for (size_t i = 0; i < r.size(); ++i)
        if (mask[i]) r[i] = 0.05;
  };
  std::vector<int> stop_steps;
  simb.stop_condition = [&](int step, double, double) {
    stop_steps.push_back(step);
    return step >= 4;
  };
  const auto sb = simb.run({rho}, model.grid, func.grand_potential_callback(mu));
  write_doubles(out + "/boundary_rho.bin", sb.densities[0]);
  for (const auto& snap : sb.snapshots)
    write_doubles(out + "/snapshot_" + std::to_string(snap.step) + ".bin", snap.densities[0]);
  std::vector<double> snap_steps, snap_times, snap_energies;
  for (const auto& snap : sb.snapshots) {
    snap_steps.push_back(snap.step);
    snap_times.push_back(snap.time);
    snap_energies.push_back(snap.energy);
  }

  // 4. an exception thrown by a callback reaches the caller unchanged
  bool rethrown = false;
  try {
    dyn::ForceCallback bad = [](const std::vector<vec>&) -> std::pair<double, std::vector<vec>> {
      throw std::domain_error("callback says no");
    };
    (void)sim16.run({rho}, model.grid, bad);
  } catch (const std::domain_error& e) {
    rethrown = std::string(e.what()) == "callback says no";
  }

  // 5. Picard: the functional's callback with the fixed-mass constraint as an opaque lambda (a host
  //    projection between device updates), and the total()-wrapping lambda (the host loop)
  double mass0 = 0.0;
  for (double r : rho) mass0 += r;
  mass0 *= model.grid.cell_volume();
  algorithms::picard::Picard pc;
  pc.mixing = 0.02;
  pc.tolerance = 1e-8;
  pc.max_iterations = 200;
  const auto fm = algorithms::picard::fixed_mass_constraint({mass0}, {vec(rho.size(), 0.0)}, model.grid.cell_volume());
  algorithms::picard::Constraint opaque = [&fm](const std::vector<vec>& d) { return fm(d); };
  const auto p_device = pc.solve({rho}, func.grand_potential_callback(mu), model.grid.cell_volume(), fm);
  const auto p_opaque = pc.solve({rho}, func.grand_potential_callback(mu), model.grid.cell_volume(), opaque);
  const auto p_host = pc.solve({rho}, wrapped, model.grid.cell_volume(), fm);
  double pd_opaque = 0.0, pd_host = 0.0;
  for (long i = 0; i < n; ++i) {
    const size_t k = static_cast<size_t>(i);
    pd_opaque = std::max(pd_opaque, std::abs(p_opaque.densities[0][k] - p_device.densities[0][k]));
    pd_host = std::max(pd_host, std::abs(p_host.densities[0][k] - p_device.densities[0][k]));
  }
  // a background of the wrong size is refused
  bool bg_rejected = false;
  try {
    (void)pc.solve({rho}, func.grand_potential_callback(mu), model.grid.cell_volume(),
                   algorithms::picard::fixed_mass_constraint({mass0}, {vec(10, 0.0)}, model.grid.cell_volume()));
  } catch (const std::invalid_argument&) {
    bg_rejected = true;
  }
  bool size_rejected = false;
  try {
    (void)pc.solve({vec(10, 0.1)}, func.grand_potential_callback(mu), model.grid.cell_volume());
  } catch (const std::invalid_argument&) {
    size_rejected = true;
  }

  std::printf("{");
  write_json_array("ideal_times", ig.times);
  write_json_array("ideal_energies", ig.energies);
  std::vector<double> ig_snaps;
  for (const auto& snap : ig.snapshots) ig_snaps.push_back(snap.step);
  write_json_array("ideal_snapshot_steps", ig_snaps);
  std::printf("\"ideal_mass\": [%.17g, %.17g], \"ideal_split_energy\": %.17g, \"ideal_cn_energy\": %.17g, "
              "\"lambda_traj_maxdiff\": %.3g, \"lambda_energy_maxdiff\": %.3g, \"lambda_calls\": %d, ",
              ig.mass_initial, ig.mass_final, igs.energy, igc.energy, traj_diff, energy_diff, lambda_calls);
  write_json_array("boundary_times", sb.times);
  write_json_array("boundary_energies", sb.energies);
  write_json_array("snapshot_steps", snap_steps);
  write_json_array("snapshot_times", snap_times);
  write_json_array("snapshot_energies", snap_energies);
  std::vector<double> ss(stop_steps.begin(), stop_steps.end());
  write_json_array("stop_calls", ss);
  std::printf("\"boundary_mass\": [%.17g, %.17g], \"exception_rethrown\": %s, \"picard_iterations\": [%d, %d, %d], "
              "\"picard_opaque_maxdiff\": %.3g, \"picard_host_maxdiff\": %.3g, \"picard_bg_rejected\": %s, "
              "\"picard_size_rejected\": %s}\n",
              sb.mass_initial, sb.mass_final, rethrown ? "true" : "false", p_device.iterations, p_opaque.iterations,
              p_host.iterations, pd_opaque, pd_host, bg_rejected ? "true" : "false", size_rejected ? "true" : "false");
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    if (argc >= 2 && std::strcmp(argv[1], "cpu") == 0) return run_cpu();
    if (argc >= 5 && std::strcmp(argv[1], "gpu") == 0) return run_gpu(argv[2], std::stod(argv[3]), argv[4]);
    if (argc >= 4 && std::strcmp(argv[1], "callbacks") == 0) return run_callbacks(argv[2], argv[3]);
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    return 2;
  }
  std::printf("usage: facade_check cpu | gpu RHO_IN MU_RHO OUT_DIR | callbacks RHO_IN OUT_DIR\n");
  return 2;
}

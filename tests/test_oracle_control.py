"""Pins the control logic of the FP64 CPU oracle's algorithms (CPU only):

- FIRE2 (fire.hpp:86-165) against a trajectory worked out by hand on a 1-D
  quadratic: the n_delay crossover of the dt growth, uphill steps before and
  after n_delay (half-step backtrack with the current dt, velocity reset, dt
  cut, alpha reset), the uphill limit; and one 2-D step that exercises the
  velocity mixing;
- Minimizer::fixed_mass (minimization.hpp:182-278) against an independent
  restatement of the legacy FIRE at fixed mass (legacy/algorithms.hpp:208-271),
  as nucleation.cpp:378-383 pins it (there to 1e-12);
- Simulation::run / integrating_factor_step (dynamics.hpp:333-401,582-693)
  against an independent restatement of the legacy DDFT step
  (legacy/algorithms.hpp:627-677) with the same dt adaptation, as
  nucleation.cpp:397-401 pins it (there to 1e-3);
- Search::find (saddle_point.hpp:367-589): the returned density is a saddle
  of the grand potential at fixed interior mass - its interior residual
  orthogonal to the eigenvector vanishes to the FIRE tolerance and the dense
  (central-difference) interior Hessian has exactly one negative eigenvalue,
  the one returned, with the returned eigenvector.
"""
import math

import numpy as np
import pytest

from classicaldft_amd import workloads as W
from legacy_restatement import boundary_mask, ddft_step, fire_minimize_fixed_mass
from oracle import oracle as O

LJ3 = dict(potential="LJ", sigma=1.0, epsilon=1.0, r_cutoff=3.0, split="WCA", scheme="quadraticF")


# ------------------------------------------------------------------ FIRE2 by hand
def _quadratic(k):
    return lambda x: (0.5 * k * float(x @ x), -k * x)


FIRE_1D = dict(dt=0.1, dt_max=1.0, dt_min=1e-8, alpha_start=0.1, f_inc=1.1, f_dec=0.5, f_alpha=0.99, n_delay=1,
               max_uphill=5)

# E = 75 x^2 (f = -150 x) from x = 1, v = 0.  In one dimension v stays parallel to f after every update, so
# the mixing v <- (1 - a) v + a |v|/|f| f leaves v unchanged and each step is
#   power test -> (uphill: x -= dt v / 2, v = 0, and dt, a cut if iteration > n_delay)
#   -> v += dt f(old x) -> x += dt v.
# Rows: (x, v, dt, alpha, n_positive, n_negative) after each step, worked out by hand:
HAND_1D = [
    # step 1 (iteration 0, counted positive): v = 0.1 (-150) = -15, x = 1 - 1.5
    (-0.5, -15.0, 0.1, 0.1, 1, 0),
    # step 2: P = -15 * 75 < 0, uphill 1 at iteration 1 = n_delay (no cut): x = -0.5 + 0.75 = 0.25, v = 0,
    # v = 0.1 * 75 (the force at -0.5) = 7.5, x = 0.25 + 0.75
    (1.0, 7.5, 0.1, 0.1, 0, 1),
    # step 3: P = 7.5 * -150 < 0, uphill 2 at iteration 2 > n_delay: x = 1 - 0.375 = 0.625, v = 0, dt = 0.05,
    # v = 0.05 * -150 = -7.5, x = 0.625 - 0.375
    (0.25, -7.5, 0.05, 0.1, 0, 2),
    # step 4: P = -7.5 * -37.5 > 0, n_positive = 1 (not > n_delay): v = -7.5 - 1.875, x = 0.25 - 0.46875
    (-0.21875, -9.375, 0.05, 0.1, 1, 0),
    # step 5: P < 0, uphill: x = -0.21875 + 0.234375 = 0.015625, dt = 0.025, v = 0.025 * 32.8125,
    # x = 0.015625 + 0.0205078125
    (0.0361328125, 0.8203125, 0.025, 0.1, 0, 1),
    # step 6: P = 0.8203125 * -5.419921875 < 0, uphill 2: x = 0.0361328125 - 0.01025390625, dt = 0.0125,
    # v = 0.0125 * -5.419921875, x = 0.02587890625 - 0.00084686279296875
    (0.02503204345703125, -0.0677490234375, 0.0125, 0.1, 0, 2),
    # step 7: P > 0, n_positive = 1: v = -0.0677490234375 + 0.0125 * -3.7548065185546875
    (0.023598492145538330078125, -0.11468410491943359375, 0.0125, 0.1, 1, 0),
    # step 8: P > 0, n_positive = 2 > n_delay: dt = 0.0125 * 1.1 = 0.01375, alpha = 0.099,
    # v = -0.11468410491943359375 + 0.01375 * -3.53977382183074951171875
    (0.023598492145538330078125 - 0.01375 * 0.1633559949696063995361328125, -0.1633559949696063995361328125,
     0.01375, 0.099, 2, 0),
]


def test_fire2_hand_computed_trajectory():
    trace, st = O.fire_trace(np.array([1.0]), _quadratic(150.0), len(HAND_1D), **FIRE_1D)
    assert st == len(HAND_1D)
    # after initialisation: x0, v = 0, dt, alpha_start, E = 75, rms force 150
    np.testing.assert_allclose(trace[0], [1.0, 0.0, 0.1, 0.1, 0, 0, 75.0, 150.0], rtol=0, atol=0)
    for k, (x, v, dt, a, npos, nneg) in enumerate(HAND_1D, start=1):
        row = trace[k]
        assert row[0] == pytest.approx(x, rel=1e-13, abs=1e-16), k
        assert row[1] == pytest.approx(v, rel=1e-13, abs=1e-16), k
        assert row[2] == pytest.approx(dt, rel=1e-14), k
        assert row[3] == pytest.approx(a, rel=1e-14), k
        assert (row[4], row[5]) == (npos, nneg), k
        assert row[6] == pytest.approx(75.0 * x * x, rel=1e-12), k


def test_fire2_uphill_limit():
    # with max_uphill = 1 the second consecutive uphill step (step 3 above) exceeds the limit
    trace, st = O.fire_trace(np.array([1.0]), _quadratic(150.0), 8, **dict(FIRE_1D, max_uphill=1))
    assert st == -3
    assert trace[2][0] == pytest.approx(1.0, rel=1e-14)


def test_fire2_mixing_step_by_hand():
    # E = (x^2 + 4 y^2) / 2 from (1, 1), default n_delay = 5, alpha = 0.1, dt = 0.1.
    # step 1: v = dt f = (-0.1, -0.4) (parallel to f: unchanged by the mixing), x = (0.99, 0.96),
    #         f = (-0.99, -3.84).
    # step 2: P = 0.099 + 1.536 > 0; v = (-0.1 - 0.099, -0.4 - 0.384) = (-0.199, -0.784);
    #         |v|^2 = 0.039601 + 0.614656, |f|^2 = 0.9801 + 14.7456; r = |v| / |f|;
    #         v = 0.9 (-0.199, -0.784) + 0.1 r (-0.99, -3.84); x = (0.99, 0.96) + 0.1 v.
    fn = lambda x: (0.5 * (x[0] ** 2 + 4 * x[1] ** 2), np.array([-x[0], -4.0 * x[1]]))
    trace, st = O.fire_trace(np.array([1.0, 1.0]), fn, 2, dt=0.1, alpha_start=0.1, dt_max=1.0)
    assert st == 2
    np.testing.assert_allclose(trace[1][:4], [0.99, 0.96, -0.1, -0.4], rtol=1e-14)
    r = math.sqrt(0.654257 / 15.7257)
    v = (0.9 * -0.199 + 0.1 * r * -0.99, 0.9 * -0.784 + 0.1 * r * -3.84)
    np.testing.assert_allclose(trace[2][2:4], v, rtol=1e-13)
    np.testing.assert_allclose(trace[2][:2], (0.99 + 0.1 * v[0], 0.96 + 0.1 * v[1]), rtol=1e-13)
    assert (trace[2][6], trace[2][7]) == (2, 0)  # n_positive 2, not yet above n_delay: dt unchanged
    assert trace[2][4] == pytest.approx(0.1, rel=1e-15)


# ------------------------------------------------------------------ legacy cross-checks
def _cluster_model():
    # RSLT FMT + LJ (r_c = 3, WCA, kT = 0.8) as the reference's nucleation test, on 16^3 with dx = 0.4
    shape, dx = (16, 16, 16), 0.4
    orc = O.Oracle(shape, dx, interactions=[LJ3], kT=0.8, model="RSLT")
    ax = [np.arange(n) * dx - n * dx / 2 for n in shape]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    return shape, dx, orc, np.sqrt(X * X + Y * Y + Z * Z)


def test_fixed_mass_minimizer_matches_legacy_fire():
    shape, dx, orc, r = _cluster_model()
    rho0 = np.where(r < 1.6, 0.75, 0.01425)  # a step-function cluster, as nucleation.cpp:64-68
    dv = dx ** 3
    target = float(np.sum(rho0)) * dv

    def free_energy(rho):  # the legacy force function: functionals::total at mu = 0
        parts, f = orc.total(rho.reshape(shape), 0.0)
        return parts[4], f.reshape(-1)

    legacy_e = []
    rl, Fl, _, itl, convl = fire_minimize_fixed_mass(free_energy, rho0, target, dv, shape, force_tolerance=1e-6,
                                                     max_steps=3000, energies=legacy_e)
    ro, reso, eo = orc.minimize_fixed_mass(rho0, 0.0, target, rho_min=1e-99, homogeneous_boundary=True,
                                           fire=dict(dt=0.1, dt_max=1.0, alpha_start=0.01, f_alpha=0.99,
                                                     force_tolerance=1e-6, max_steps=3000))
    print(f"fixed mass: oracle {reso['iterations']} / legacy {itl} iterations; F {reso['free_energy']:.15e} / "
          f"{Fl:.15e}", flush=True)
    assert convl and reso["converged"] and itl == reso["iterations"]
    # the trajectory includes uphill steps (energy increases), so the backtrack branch is compared too
    assert np.any(np.diff(eo) > 0)
    np.testing.assert_allclose(eo[1:], legacy_e, rtol=1e-12, atol=1e-12)
    assert np.max(np.abs(ro.reshape(-1) - rl)) < 1e-12
    assert reso["free_energy"] == pytest.approx(Fl, rel=1e-12)


def test_ddft_steps_match_legacy_ddft():
    shape, dx, orc, r = _cluster_model()
    rho_v = 0.01425
    rho0 = rho_v + 0.5 * (0.75 - rho_v) * (1.0 - np.tanh((r - 1.6) / 0.4))
    mu = orc.bulk_mu(rho_v)
    dv = dx ** 3

    def omega(rho):
        parts, f = orc.total(rho, mu)
        return parts[5], f

    steps, dt, dt_max, succ = 20, 1e-2, 1.0, 0
    rl, energies, used_dts = rho0.copy(), [], []
    for _ in range(steps):  # the dt adaptation of nucleation.cpp:345-369 around the legacy step
        before = dt
        rl, e, dt = ddft_step(rl, omega, dv, dx, 1e-4, 100, dt)
        succ = 0 if dt < before else succ + 1
        used_dts.append(dt)
        if succ >= 5 and dt < dt_max:
            dt, succ = min(2.0 * dt, dt_max), 0
        energies.append(e)
    ro, reso, eo = orc.ddft_run(rho0, mu, steps, dt=1e-2, dt_max=dt_max, fp_tolerance=1e-4, fp_max_iterations=100)
    print(f"ddft: dt used {used_dts}; max |oracle - legacy| {np.max(np.abs(ro - rl)):.2e}", flush=True)
    # the run exercises dt doubling and a restart (dt / 10)
    assert max(used_dts) >= 0.04 and used_dts[-1] < max(used_dts) / 2
    assert reso["dt"] == pytest.approx(dt, rel=1e-15)
    assert np.max(np.abs(ro - rl)) < 1e-12
    np.testing.assert_allclose(eo[1:], energies, rtol=1e-12)


# ------------------------------------------------------------------ Search::find
def test_saddle_search_returns_a_saddle():
    shape, dx = (16, 16, 16), 0.25
    orc = O.Oracle(shape, dx, interactions=[W.LJ_WCA], kT=0.8, model="WhiteBearII")
    rho0 = W.droplet(shape, dx, radius=1.1, width=0.4, rho_l=0.6, rho_v=0.05, noise=0.0, seed=3)
    target = float(np.sum(rho0)) * dx ** 3
    mu = orc.bulk_mu(0.05)
    rho, vec, res = orc.saddle_search(rho0, mu, target, seed=1, fire=dict(max_steps=300),
                                      eigen=dict(max_iterations=100), max_outer_iterations=3)
    assert res["converged"]
    rf, ef = rho.reshape(-1), vec.reshape(-1)
    bd = boundary_mask(shape)
    inner = np.nonzero(~bd)[0]
    assert float(np.sum(rf)) * dx ** 3 == pytest.approx(target, rel=1e-12)
    # stationary at fixed interior mass: the interior force less its Lagrange multiplier, orthogonal to the
    # (alias-space) eigen direction, is below the last FIRE run's rms tolerance (1e-4); the droplet it
    # started from is at 1e-2
    _, g = orc.total(rho, mu)
    g = g.reshape(-1)
    lam = float(g[inner] @ rf[inner]) / (float(np.sum(rf[inner])) * dx ** 3)
    res_int = np.zeros_like(g)
    res_int[inner] = g[inner] - lam * dx ** 3
    x = np.sqrt(rf - 1e-18)
    fa, d = 2 * x * res_int, np.zeros_like(g)
    d[inner] = 2 * x[inner] * ef[inner]
    d /= np.linalg.norm(d)
    perp = fa - float(d @ fa) * d
    assert math.sqrt(float(np.mean(perp ** 2))) < 1.5e-4
    # the dense interior Hessian (central differences of the forces; the fixed boundary's zeroed rows only
    # add zero eigenvalues): one negative eigenvalue, the one returned, with the returned vector
    n, eps = inner.size, 1e-5
    H = np.zeros((n, n))
    for k, j in enumerate(inner):
        up, dn = rf.copy(), rf.copy()
        up[j] += eps
        dn[j] -= eps
        H[:, k] = (orc.total(up.reshape(shape), mu)[1].reshape(-1)[inner] -
                   orc.total(dn.reshape(shape), mu)[1].reshape(-1)[inner]) / (2 * eps)
    assert np.max(np.abs(H - H.T)) < 1e-9 * np.max(np.abs(H))
    w, U = np.linalg.eigh(0.5 * (H + H.T))
    print(f"saddle: returned eigenvalue {res['eigenvalue']:.8e}, dense {w[:2]}", flush=True)
    assert w[0] < 0.0 < w[1]
    # the returned value comes from forward differences (eps 1e-6) and a 1e-4 residual tolerance
    assert res["eigenvalue"] == pytest.approx(w[0], rel=2e-3)
    assert abs(float(U[:, 0] @ ef[inner])) / np.linalg.norm(ef[inner]) > 0.9999


def test_max_force_stop():
    # the library's stop on max |delta(beta Omega)/delta rho| (no dV factor): the returned density meets it,
    # checked against the functional itself, and the energy monitor alone would have run on
    shape, dx = (8, 8, 8), 0.5
    orc = O.Oracle(shape, dx, interactions=[W.LJ_WCA], kT=0.9, model="WhiteBearII")
    rng = np.random.default_rng(4)
    rho0 = 0.3 * (1.0 + 0.05 * rng.standard_normal(shape))
    mu = orc.bulk_mu(0.3)
    fire = dict(dt=1e-2, dt_max=0.1, force_tolerance=0.0, force_max_tolerance=1e-7)
    r, res, e = orc.minimize_grand_potential(rho0, mu, fire=fire, max_steps=20000, homogeneous_boundary=False)
    assert res["converged"] and res["iterations"] < 20000
    _, f = orc.total(r, mu)
    assert np.max(np.abs(f)) / dx ** 3 < 1e-7
    # one step earlier the criterion was not met
    r1, res1, _ = orc.minimize_grand_potential(rho0, mu, fire=dict(fire, force_max_tolerance=0.0),
                                               max_steps=res["iterations"] - 1, homogeneous_boundary=False)
    _, f1 = orc.total(r1, mu)
    assert np.max(np.abs(f1)) / dx ** 3 >= 1e-7

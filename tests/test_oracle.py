"""Pins the FP64 CPU oracle (oracle/oracle.cpp) before the GPU tests compare against it.

Each check is chosen so that a plausible restatement mistake (dropped term,
wrong sign or index, transposed operand, wrong normalisation) fails it:
numpy's pocketfft and a brute-force DFT for the FFT, numerical quadrature for
the weight closed forms, Carnahan-Starling / Percus-Yevick bulk limits for the
FMT models, central differences for every derivative, the reference's own
known-answer tests (uniform fluid, sum rules, a_vdw), and a functional-
derivative finite-difference test of the complete grand potential.
"""
import math

import numpy as np
import pytest
from scipy import integrate

from oracle import oracle as O

PI = math.pi


def test_r2c_matches_numpy():
    rng = np.random.default_rng(1)
    for shape in [(4, 4, 4), (6, 4, 8), (5, 5, 5), (8, 6, 10), (3, 5, 4), (16, 16, 16), (12, 10, 6)]:
        orc = O.Oracle(shape, 0.1, model="none")
        x = rng.standard_normal(shape)
        got = orc.r2c(x)
        ref = np.fft.rfftn(x)
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref)), shape


def test_fft_brute_force_dft():
    rng = np.random.default_rng(2)
    shape = (3, 4, 5)
    orc = O.Oracle(shape, 0.1, model="none")
    x = rng.standard_normal(shape)
    nx, ny, nz = shape
    X = np.zeros((nx, ny, nz // 2 + 1), dtype=complex)
    r = np.indices(shape).reshape(3, -1)
    for kx in range(nx):
        for ky in range(ny):
            for kz in range(nz // 2 + 1):
                ph = np.exp(-2j * PI * (kx * r[0] / nx + ky * r[1] / ny + kz * r[2] / nz))
                X[kx, ky, kz] = np.sum(x.reshape(-1) * ph)
    assert np.max(np.abs(orc.r2c(x) - X)) < 1e-11


def test_c2r_round_trip_unnormalised():
    rng = np.random.default_rng(3)
    for shape in [(8, 8, 8), (6, 10, 4), (5, 7, 9)]:
        orc = O.Oracle(shape, 0.1, model="none")
        x = rng.standard_normal(shape)
        y = orc.c2r(orc.r2c(x))
        n = np.prod(shape)
        assert np.max(np.abs(y / n - x)) < 1e-13


def test_c2r_non_hermitian_projection_matches_numpy():
    # pocketfft's irfftn: complex inverse over x,y then c2r along z that ignores the
    # imaginary parts of the kz=0 and Nyquist bins - the convention the oracle defines.
    rng = np.random.default_rng(4)
    for shape in [(4, 6, 8), (6, 6, 6), (5, 4, 7)]:
        orc = O.Oracle(shape, 0.1, model="none")
        X = rng.standard_normal((shape[0], shape[1], shape[2] // 2 + 1)) + 1j * rng.standard_normal(
            (shape[0], shape[1], shape[2] // 2 + 1))
        got = orc.c2r(X)
        ref = np.fft.irfftn(X, s=shape, axes=(0, 1, 2)) * np.prod(shape)
        assert np.max(np.abs(got - ref)) < 1e-11, shape
        # and equals Re(IDFT) of the explicit Hermitian extension
        nz = shape[2]
        full = np.zeros(shape, dtype=complex)
        full[:, :, : nz // 2 + 1] = X
        for kz in range(nz // 2 + 1, nz):
            full[:, :, kz] = np.conj(X[(-np.arange(shape[0])) % shape[0]][:, (-np.arange(shape[1])) % shape[1], nz - kz])
        ext = np.real(np.fft.ifftn(full)) * np.prod(shape)
        assert np.max(np.abs(got - ext)) < 1e-11


# ---------------------------------------------------------------- weights


def _kgrid(shape, dx):
    nx, ny, nz = shape
    kx = 2 * PI / (nx * dx) * np.where(np.arange(nx) <= nx // 2, np.arange(nx), np.arange(nx) - nx)
    ky = 2 * PI / (ny * dx) * np.where(np.arange(ny) <= ny // 2, np.arange(ny), np.arange(ny) - ny)
    kz = 2 * PI / (nz * dx) * np.arange(nz // 2 + 1)
    return np.meshgrid(kx, ky, kz, indexing="ij")


def test_weight_dc_values():
    shape, dx, d = (16, 16, 16), 0.1, 1.0
    orc = O.Oracle(shape, dx, diameters=(d,), model="WhiteBearII")
    w = orc.fmt_weights()
    N = np.prod(shape)
    R = d / 2
    assert w[0, 0, 0, 0] == pytest.approx(4 * PI / 3 * R ** 3 / N, rel=1e-14)
    assert w[1, 0, 0, 0] == pytest.approx(4 * PI * R ** 2 / N, rel=1e-14)
    for a in range(3):
        assert w[2 + a, 0, 0, 0] == 0
    diag = [5, 8, 10]
    for c in range(6):
        expect = 4 * PI / 3 * R ** 2 / N if 5 + c in diag else 0.0
        assert w[5 + c, 0, 0, 0] == pytest.approx(expect, abs=1e-18)


def test_weight_closed_forms_against_quadrature():
    """w3, w2, the vector prefactor and the tensor A, B against direct integrals over the shell/ball."""
    shape, dx, d = (10, 12, 14), 0.13, 1.1
    orc = O.Oracle(shape, dx, diameters=(d,), model="WhiteBearII")
    w = orc.fmt_weights() * np.prod(shape)
    R = d / 2
    KX, KY, KZ = _kgrid(shape, dx)
    rng = np.random.default_rng(5)
    picks = [tuple(rng.integers(0, s) for s in w.shape[1:]) for _ in range(6)]
    for p in picks:
        kv = np.array([KX[p], KY[p], KZ[p]])
        k = np.linalg.norm(kv)
        if k < 1e-9:
            continue
        # ball: 4 pi int_0^R r^2 j0(kr) dr
        w3 = 4 * PI * integrate.quad(lambda r: r * r * np.sinc(k * r / PI), 0, R, epsabs=1e-14)[0]
        assert w[(0,) + p].real == pytest.approx(w3, rel=1e-9, abs=1e-12)
        # shell: 4 pi R^2 j0(kR)
        assert w[(1,) + p].real == pytest.approx(4 * PI * R * R * np.sinc(k * R / PI), rel=1e-10, abs=1e-12)
        # vector: the e^{-ik.r} transform of delta(R-r) r_hat is -4 pi i R^2 j1(kR) k_hat.  The
        # reference's channel (weights.hpp:110-119,164-168) carries the opposite overall sign; every
        # physical output is even in it.  Check magnitude and the stated sign.
        x = k * R
        j1 = math.sin(x) / x ** 2 - math.cos(x) / x
        std = -4j * PI * R * R * j1 * kv / k
        for a in range(3):
            assert w[(2 + a,) + p] == pytest.approx(-std[a], rel=1e-9, abs=1e-12)
        # tensor: k along e_z gives wT_zz = 2 pi R^2 int mu^2 e^{-ikR mu}, wT_xx = pi R^2 int (1-mu^2) e^{..}
        tzz = 2 * PI * R * R * integrate.quad(lambda mu: mu * mu * math.cos(x * mu), -1, 1, epsabs=1e-14)[0]
        txx = PI * R * R * integrate.quad(lambda mu: (1 - mu * mu) * math.cos(x * mu), -1, 1, epsabs=1e-14)[0]
        B, A = txx, tzz - txx
        ch = {(0, 0): 5, (0, 1): 6, (0, 2): 7, (1, 1): 8, (1, 2): 9, (2, 2): 10}
        for (i, j), c in ch.items():
            expect = A * kv[i] * kv[j] / k ** 2 + (B if i == j else 0.0)
            assert w[(c,) + p].real == pytest.approx(expect, rel=1e-8, abs=1e-11)


def test_uniform_sum_rules():
    shape, dx, d, rho0 = (16, 16, 16), 0.1, 1.0, 0.5
    orc = O.Oracle(shape, dx, diameters=(d,), model="WhiteBearII")
    wd = orc.weighted_densities(np.full(shape, rho0))
    n2 = PI * rho0 * d * d
    assert np.allclose(wd[0], PI / 6 * rho0 * d ** 3, rtol=1e-13)
    assert np.allclose(wd[1], n2, rtol=1e-13)
    assert np.max(np.abs(wd[2:5])) < 1e-14
    for c, diag in zip(range(5, 11), [1, 0, 0, 1, 0, 1]):
        assert np.allclose(wd[c], n2 / 3 * diag, atol=1e-13)


# ---------------------------------------------------------------- FMT models


def _cs(eta):
    return eta * (4 - 3 * eta) / (1 - eta) ** 2


def _pyc(eta):
    return -math.log(1 - eta) + 1.5 * eta * (2 - eta) / (1 - eta) ** 2


def _uniform(rho, d=1.0):
    m = np.zeros(19)
    m[0] = PI / 6 * rho * d ** 3
    m[3] = PI * rho * d * d
    m[2] = 0.5 * d * rho
    m[1] = rho
    m[10] = m[14] = m[18] = m[3] / 3
    return m


@pytest.mark.parametrize("model,closed", [("WhiteBearI", _cs), ("WhiteBearII", _cs), ("Rosenfeld", _pyc),
                                          ("RSLT", _cs)])
def test_bulk_free_energy_matches_eos(model, closed):
    for rho in (0.1, 0.3, 0.5, 0.7, 0.9):
        eta = PI / 6 * rho
        phi, _ = O.phi_dphi(model, _uniform(rho))
        assert phi / rho == pytest.approx(closed(eta), abs=1e-10)


def _sympy_f(model, which):
    """The reference's f1/f2/f3 (models.hpp:157-295) as sympy expressions (closed-form branch)."""
    import sympy as sp

    e = sp.symbols("eta")
    f1 = sp.log(1 - e)
    if model == "WhiteBearII":
        f2 = sp.Rational(1, 3) + sp.Rational(4, 3) / (1 - e) + 2 / (3 * e) * sp.log(1 - e)
        f3 = -(1 - 3 * e + e * e) / (e * (1 - e) ** 2) - sp.log(1 - e) / e ** 2
    elif model in ("RSLT", "WhiteBearI"):
        f2 = 1 / (1 - e)
        f3 = 1 / (e * (1 - e) ** 2) + sp.log(1 - e) / e ** 2
    else:
        f2 = 1 / (1 - e)
        f3 = 1 / (1 - e) ** 2
    f = {1: f1, 2: f2, 3: f3}[which]
    return e, f, sp.diff(f, e)


@pytest.mark.parametrize("model", ["Rosenfeld", "RSLT", "WhiteBearI", "WhiteBearII", "esFMT"])
def test_f_values_and_derivatives_symbolic(model):
    """Closed-form f and f' against sympy differentiation evaluated at 40 digits."""
    for which in (1, 2, 3):
        e, f, df = _sympy_f(model, which)
        for eta in (5e-4, 0.0105, 0.05, 0.2, 0.35, 0.5, 0.6):
            v, d = O.f_func(model, which, eta)
            ev = float(f.evalf(40, subs={e: eta}))
            ed = float(df.evalf(40, subs={e: eta}))
            # FP64 closed forms cancel near small eta (ln(1-eta)/eta^2): allow eps/eta^2
            tol = 1e-13 + 4e-16 / eta ** 2
            assert v == pytest.approx(ev, rel=1e-12, abs=tol), (model, which, eta)
            assert d == pytest.approx(ed, rel=1e-12, abs=tol / eta), (model, which, eta)


@pytest.mark.parametrize("model,which", [("RSLT", 3), ("WhiteBearI", 3), ("WhiteBearII", 2), ("WhiteBearII", 3)])
def test_f_taylor_branch_derivatives(model, which):
    """Below eta=1e-6 the reference differentiates its series (models.hpp:200-202,250-253,282-292)."""
    for _ in (0,):
        for eta in (1e-9, 3e-7, 6e-7):
            v, d = O.f_func(model, which, eta)
            h = 0.5 * eta  # central differences of the cubic series: error O(h^2), far below rel 1e-5
            fd = (O.f_func(model, which, eta + h)[0] - O.f_func(model, which, eta - h)[0]) / (2 * h)
            assert d == pytest.approx(fd, rel=1e-5), (model, which, eta)


def test_taylor_branch_values_exact():
    # WB-I/WB-II series (models.hpp:250-253,282-292) equal the exact function below the switch
    # (the FP64 closed form just above it cancels to ~1e-16/eta^2, as in the reference).
    for model in ("WhiteBearI", "WhiteBearII"):
        for which in (2, 3):
            if model == "WhiteBearI" and which == 2:
                continue
            e, f, _ = _sympy_f(model, which)
            for eta in (1e-9, 5e-7, 9.9e-7):
                exact = float(f.evalf(40, subs={e: eta}))
                assert O.f_func(model, which, eta)[0] == pytest.approx(exact, rel=1e-15, abs=1e-15)


def _test_measures(eta, n2, v, tdiag, R=0.5):
    m = np.zeros(19)
    m[0], m[3] = eta, n2
    m[2] = n2 / (4 * PI * R)
    m[1] = n2 / (4 * PI * R * R)
    v1 = np.array([v, -0.3 * v, 0.1 * v])
    m[7:10] = v1
    m[4:7] = v1 / (4 * PI * R)
    T = np.array([[tdiag, 0.1, -0.05], [0.1, tdiag * 0.8, 0.02], [-0.05, 0.02, tdiag * 0.6]])
    m[10:19] = T.reshape(-1)
    return m


@pytest.mark.parametrize("model", ["Rosenfeld", "RSLT", "WhiteBearI", "WhiteBearII", "esFMT"])
def test_dphi_central_difference(model):
    # the reference's non-uniform measure sets (tests/integration/functionals/fmt.cpp:74-89)
    for args in [(0.3, 2.5, 0.4, 0.3), (0.25, 2.0, 0.3, 0.25), (0.0105, 0.05, 0.01, 0.015), (3e-7, 1e-5, 1e-6, 3e-6)]:
        m = _test_measures(*args)
        A, B = (1.0, -1.0) if model == "esFMT" else (1.0, 0.0)
        _, dm = O.phi_dphi(model, m, A, B)
        for c in range(19):
            h = 1e-6 * max(abs(m[c]), 1e-3)
            mp, mm = m.copy(), m.copy()
            mp[c] += h
            mm[c] -= h
            fd = (O.phi_dphi(model, mp, A, B)[0] - O.phi_dphi(model, mm, A, B)[0]) / (2 * h)
            assert dm[c] == pytest.approx(fd, rel=1e-5, abs=1e-8), (model, args, c)


def test_tensor_phi3_explicit_form():
    # WB Phi3 = (1/8pi)(n2 tr T^2 - n2 v.v + v.T.v - tr T^3) (models.hpp:18-27): compare
    # Phi - (Phi with f3 -> 0) through the esFMT(1,-1) identity at eta where f3 is known.
    m = _test_measures(0.25, 2.0, 0.3, 0.25)
    T = m[10:19].reshape(3, 3)
    v = m[7:10]
    n2 = m[3]
    p3 = (n2 * np.trace(T @ T) - n2 * v @ v + v @ T @ v - np.trace(T @ T @ T)) / (8 * PI)
    eta = m[0]
    f1, _ = O.f_func("WhiteBearI", 1, eta)
    f2, _ = O.f_func("WhiteBearI", 2, eta)
    f3, _ = O.f_func("WhiteBearI", 3, eta)
    phi_expected = -m[1] * f1 + (m[2] * n2 - m[4:7] @ v) * f2 + p3 * f3
    assert O.phi_dphi("WhiteBearI", m)[0] == pytest.approx(phi_expected, rel=1e-13)
    assert O.phi_dphi("esFMT", m, 1.0, -1.0)[0] == pytest.approx(
        -m[1] * f1 + (m[2] * n2 - m[4:7] @ v) / (1 - eta) + p3 / (1 - eta) ** 2, rel=1e-13)


def test_phi_rotation_invariant():
    rng = np.random.default_rng(6)
    m = _test_measures(0.3, 2.5, 0.4, 0.3)
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    r = m.copy()
    r[4:7] = Q @ m[4:7]
    r[7:10] = Q @ m[7:10]
    r[10:19] = (Q @ m[10:19].reshape(3, 3) @ Q.T).reshape(-1)
    for model in ("WhiteBearI", "WhiteBearII", "RSLT", "Rosenfeld"):
        assert O.phi_dphi(model, r)[0] == pytest.approx(O.phi_dphi(model, m)[0], rel=1e-13)


@pytest.mark.parametrize("model", ["Rosenfeld", "WhiteBearI", "WhiteBearII"])
def test_bulk_mu_numerical_derivative(model):
    orc = O.Oracle((8, 8, 8), 0.25, model=model)
    rho = 0.5
    h = 1e-5
    f = lambda r: O.phi_dphi(model, _uniform(r))[0] + r * (math.log(r) - 1)  # noqa: E731
    assert orc.bulk_mu(rho) == pytest.approx((f(rho + h) - f(rho - h)) / (2 * h), abs=1e-7)


# ---------------------------------------------------------------- functional


@pytest.mark.parametrize("model", ["Rosenfeld", "RSLT", "WhiteBearI", "WhiteBearII"])
def test_uniform_hard_sphere_matches_bulk(model):
    shape, dx = (16, 16, 16), 0.1
    orc = O.Oracle(shape, dx, model=model)
    rho0 = 0.5
    parts, forces = orc.total(np.full(shape, rho0), mu=0.0)
    V = np.prod(shape) * dx ** 3
    phi, _ = O.phi_dphi(model, _uniform(rho0))
    assert parts[2] == pytest.approx(phi * V, rel=1e-10)
    # force = (ln rho + mu_ex) dV at mu = 0
    mu_total = orc.bulk_mu(rho0)
    assert np.allclose(forces, mu_total * dx ** 3, rtol=1e-10)


def test_zero_density_gives_zero_hard_sphere():
    orc = O.Oracle((8, 8, 8), 0.2, model="WhiteBearII")
    parts, _ = orc.total(np.zeros((8, 8, 8)), mu=0.0)
    assert parts[2] == 0.0


def _lj(rc=2.5, scheme="quadraticF"):
    return dict(potential="LJ", sigma=1.0, epsilon=1.0, r_cutoff=rc, split="WCA", scheme=scheme)


def test_mean_field_uniform():
    shape, dx, kT = (16, 16, 16), 0.4, 0.8
    orc = O.Oracle(shape, dx, interactions=[_lj()], kT=kT, model="none")
    rho0 = 0.3
    parts, forces = orc.total(np.full(shape, rho0), mu=0.0)
    a = orc.a_vdw()
    V = np.prod(shape) * dx ** 3
    assert parts[3] == pytest.approx(0.5 * a * rho0 ** 2 * V, rel=1e-12)
    assert np.allclose(forces, (math.log(rho0) + a * rho0) * dx ** 3, rtol=1e-12)


def _legacy_qf_avdw(sigma, eps, rc, kT, dx):
    """Restatement of the legacy cube sum (legacy/classicaldft.hpp:696-759) for an independent a_vdw."""
    def watt(r2):
        shift = 4 * eps * ((sigma / rc) ** 12 - (sigma / rc) ** 6)
        rmin = 2 ** (1 / 6) * sigma
        vmin = 4 * eps * ((sigma / rmin) ** 12 - (sigma / rmin) ** 6) - shift
        if r2 >= rc * rc:
            return 0.0
        if r2 < rmin * rmin:
            return vmin
        y6 = (sigma * sigma / r2) ** 3
        return 4 * eps * (y6 * y6 - y6) - shift

    lim = 1 + int(rc / dx)
    pts = (-0.5, 0.0, 0.5)
    total = 0.0
    cache = {}
    for ix in range(-(lim + 1), lim + 2):
        for iy in range(-(lim + 1), lim + 2):
            for iz in range(-(lim + 1), lim + 2):
                key = tuple(sorted((abs(ix), abs(iy), abs(iz))))
                if key not in cache:
                    s = 0.0
                    for a in pts:
                        for b in pts:
                            for c in pts:
                                s += watt(((key[0] + a) ** 2 + (key[1] + b) ** 2 + (key[2] + c) ** 2) * dx * dx) / 27
                    cache[key] = s
                total += cache[key]
    return total * dx ** 3 / kT


@pytest.mark.parametrize("dx", [0.5, 0.4])
def test_a_vdw_matches_legacy_qf_sum(dx):
    kT, rc = 0.8, 2.5
    n = int(round(8.0 / dx))
    orc = O.Oracle((n, n, n), dx, interactions=[_lj(rc)], kT=kT, model="none")
    assert orc.a_vdw() == pytest.approx(_legacy_qf_avdw(1.0, 1.0, rc, kT, dx), rel=1e-10)
    # and the continuum value 2 * (2 pi / kT) int r^2 V_att(r) dr within 5% (mean_field.cpp test)
    rmin = 2 ** (1 / 6)
    shift = 4 * (rc ** -12 - rc ** -6)
    vmin = 4 * (rmin ** -12 - rmin ** -6) - shift
    inner = integrate.quad(lambda r: r * r * vmin, 0, rmin)[0] + integrate.quad(
        lambda r: r * r * (4 * (r ** -12 - r ** -6) - shift), rmin, rc)[0]
    assert orc.a_vdw() == pytest.approx(2 * 2 * PI / kT * inner, rel=0.05)


def test_total_parts_and_omega():
    rng = np.random.default_rng(7)
    shape, dx, kT = (8, 8, 8), 0.3, 0.9
    orc = O.Oracle(shape, dx, interactions=[_lj()], kT=kT, model="WhiteBearII")
    rho = 0.3 + 0.1 * rng.random(shape)
    V = 0.5 * np.cos(2 * PI * np.arange(8) / 8)[None, None, :] * np.ones(shape)
    mu = -1.3
    parts, _ = orc.total(rho, mu, vext=V)
    dv = dx ** 3
    assert parts[0] == pytest.approx(np.sum(rho * (np.log(rho) - 1)) * dv, rel=1e-12)
    assert parts[1] == pytest.approx(np.sum(rho * V) * dv, rel=1e-12)
    assert parts[4] == pytest.approx(parts[0] + parts[1] + parts[2] + parts[3], rel=1e-14)
    assert parts[5] == pytest.approx(parts[4] - mu * np.sum(rho) * dv, rel=1e-14)


@pytest.mark.parametrize("model", ["WhiteBearII", "WhiteBearI", "RSLT"])
def test_functional_derivative_finite_difference(model):
    """force_p = d Omega / d rho_p for the full functional - pins the adjoint, the conj on the
    vector channel, the off-diagonal x2, the Rosenfeld scaling and the Nyquist convention."""
    rng = np.random.default_rng(8)
    shape, dx, kT = (8, 6, 10), 0.25, 0.8
    orc = O.Oracle(shape, dx, interactions=[_lj(1.2)], kT=kT, model=model)
    rho = 0.35 + 0.25 * rng.random(shape)
    V = 0.3 * rng.standard_normal(shape)
    mu = 0.7
    _, forces = orc.total(rho, mu, vext=V)
    for _ in range(6):
        p = tuple(rng.integers(0, s) for s in shape)
        h = 1e-6
        rp, rm = rho.copy(), rho.copy()
        rp[p] += h
        rm[p] -= h
        fd = (orc.total(rp, mu, vext=V)[0][5] - orc.total(rm, mu, vext=V)[0][5]) / (2 * h)
        assert forces[p] == pytest.approx(fd, rel=1e-6, abs=1e-10), p


def test_uniform_state_at_bulk_mu_has_zero_force():
    shape, dx, kT = (12, 12, 12), 0.2, 0.8
    orc = O.Oracle(shape, dx, interactions=[_lj()], kT=kT, model="WhiteBearII")
    rho0 = 0.45
    _, f = orc.total(np.full(shape, rho0), orc.bulk_mu(rho0))
    assert np.max(np.abs(f)) < 1e-12


# ---------------------------------------------------------------- algorithms


def test_fire_minimizer_relaxes_to_bulk():
    rng = np.random.default_rng(9)
    shape, dx = (8, 8, 8), 0.25
    orc = O.Oracle(shape, dx, model="WhiteBearII")
    rho_b = 0.4
    mu = orc.bulk_mu(rho_b)
    rho0 = rho_b * (1 + 0.02 * rng.standard_normal(shape))
    rho, res, energies = orc.minimize_grand_potential(rho0, mu, fire=dict(force_tolerance=1e-12),
                                                      homogeneous_boundary=False, max_steps=3000)
    assert np.max(np.abs(rho - rho_b)) < 1e-4
    omega_bulk = orc.total(np.full(shape, rho_b), mu)[0][5]
    assert res["grand_potential"] == pytest.approx(omega_bulk, rel=1e-8)
    assert energies[-1] <= energies[0]


def test_ddft_ideal_gas_mode_decay_and_mass():
    # docs/dynamics/check/main.cpp:99-160: ideal gas, a small cosine mode decays as exp(Lambda_1 t)
    shape, dx = (16, 4, 4), 0.5
    orc = O.Oracle(shape, dx, model="none")
    rho0 = 0.2
    eps = 1e-4
    x = np.arange(16)
    rho = (rho0 * (1 + eps * np.cos(2 * PI * x / 16)))[:, None, None] * np.ones(shape)
    dt, n = 1e-2, 20
    out, res, _ = orc.ddft_run(rho, 0.0, n, dt=dt, dt_max=dt, fp_tolerance=1e-12)
    lam = 2 / dx ** 2 * (math.cos(2 * PI / 16) - 1)
    amp = np.sum(out[:, 0, 0] * np.cos(2 * PI * x / 16)) / 8
    assert amp / (rho0 * eps) == pytest.approx(math.exp(lam * res["time"]), rel=1e-3)
    assert abs(res["mass_final"] - res["mass_initial"]) / res["mass_initial"] < 1e-12


# ------------------------------------------------------------------ fixed-dt DDFT map (adjoint forward)
def _wells_numpy(shape, dx, prm):
    L = np.array(shape) * dx
    ax = [np.arange(n) * dx for n in shape]
    X = np.meshgrid(*ax, indexing="ij")
    V = np.zeros(shape)
    for A, w, *c in prm:
        d2 = np.zeros(shape)
        for a in range(3):
            e = X[a] - c[a]
            e -= L[a] * np.round(e / L[a])
            d2 += e * e
        V += A * np.exp(-d2 / (2 * w * w))
    return V


def test_wells_potential_matches_numpy():
    from classicaldft_amd import workloads as W

    shape, dx = (12, 10, 8), 0.3
    orc = O.Oracle(shape, dx, model="none")
    prm = W.wells(shape, dx, n_wells=5, seed=3)
    np.testing.assert_allclose(orc.wells_potential(prm), _wells_numpy(shape, dx, prm), rtol=1e-14, atol=1e-15)


def _rhs_numpy(rho, f, dx):
    # div(rho grad(f/dV)) - Lap(rho) with the face fluxes of dynamics.hpp:135-229
    g = f / dx ** 3
    D, D2 = 0.5 / dx ** 2, 1.0 / dx ** 2
    r = np.zeros_like(rho)
    for a in range(3):
        rp, rm = np.roll(rho, -1, a), np.roll(rho, 1, a)
        gp, gm = np.roll(g, -1, a), np.roll(g, 1, a)
        r += D * ((rp + rho) * (gp - g) - (rho + rm) * (g - gm))
        r -= D2 * (rp + rm - 2 * rho)
    return r


def _fixed_map_numpy(orc, rho0, mu, dt, P, vext):
    # the propagator of dynamics.hpp:247-289 written with numpy's FFT:
    # d^(i+1) = irfft(U0 rho0_k + U1 R0_k + U2 (R(d^(i))_k - R0_k)), d^(0) = rho0
    shape, dx = rho0.shape, orc.dx
    lam = [2 / dx ** 2 * (np.cos(2 * PI * np.arange(n) / n) - 1) for n in shape]
    L = lam[0][:, None, None] + lam[1][None, :, None] + lam[2][None, None, : shape[2] // 2 + 1]
    E = np.exp(L * dt)
    with np.errstate(divide="ignore", invalid="ignore"):
        U1 = np.where(L == 0, dt, (E - 1) / L)
        U2 = np.where(L == 0, dt / 2, (E - 1 - dt * L) / (L * L * dt))
    R0 = np.fft.rfftn(_rhs_numpy(rho0, orc.total(rho0, mu, vext)[1], dx))
    P0 = np.fft.rfftn(rho0)
    d = rho0
    for _ in range(P):
        Ri = np.fft.rfftn(_rhs_numpy(d, orc.total(d, mu, vext)[1], dx))
        d = np.fft.irfftn(E * P0 + U1 * R0 + U2 * (Ri - R0), s=shape)
    return d


@pytest.mark.parametrize("P", [1, 2, 3])
def test_fixed_ddft_step_matches_numpy_propagator(P):
    from classicaldft_amd import workloads as W

    shape, dx = (12, 12, 12), 0.25
    orc = O.Oracle(shape, dx, interactions=[W.LJ_WCA], kT=0.9, model="WhiteBearII")
    rho0, prm, _ = W.adjoint_problem(shape, dx, n_wells=3, rho_bulk=0.3, seed=4)
    V = orc.wells_potential(prm)
    mu = orc.bulk_mu(0.3)
    dt = 2e-3
    got, mn = orc.ddft_fixed_run(rho0, mu, 1, dt, P, vext=V)
    want = _fixed_map_numpy(orc, rho0, mu, dt, P, V)
    assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-12
    # the smallest density of every iterate, the final one included
    assert mn <= got.min() and (P > 1 or mn == got.min())


def test_fixed_ddft_converges_to_reference_step():
    # with many iterations the fixed map reaches the fixed point the reference's
    # tolerance-controlled step stops at (dt_max = dt: no adaptation)
    from classicaldft_amd import workloads as W

    shape, dx = (12, 12, 12), 0.25
    orc = O.Oracle(shape, dx, interactions=[W.LJ_WCA], kT=0.9, model="WhiteBearII")
    rho0, prm, _ = W.adjoint_problem(shape, dx, n_wells=3, rho_bulk=0.3, seed=4)
    V = orc.wells_potential(prm)
    mu = orc.bulk_mu(0.3)
    dt = 2e-3
    ref, res, _ = orc.ddft_run(rho0, mu, 3, dt=dt, dt_max=dt, fp_tolerance=1e-11, fp_max_iterations=40, vext=V)
    assert res["dt"] == dt
    got, _ = orc.ddft_fixed_run(rho0, mu, 3, dt, 40, vext=V)
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-10


# ------------------------------------------------------------------ Minimizer::fixed_mass
def test_fixed_mass_minimizer_keeps_mass_and_lowers_F():
    from classicaldft_amd import workloads as W

    shape, dx = (12, 12, 12), 0.25
    orc = O.Oracle(shape, dx, interactions=[W.LJ_WCA], kT=0.8, model="WhiteBearII")
    rho0 = W.droplet(shape, dx, radius=1.0, width=0.4, rho_l=0.6, rho_v=0.05, noise=1e-3, seed=3)
    target = 0.9 * float(np.sum(rho0)) * dx ** 3
    mu = orc.bulk_mu(0.05)
    rho, res, en = orc.minimize_fixed_mass(rho0, mu, target, fire=dict(force_tolerance=1e-12), max_steps=80)
    assert np.sum(rho) * dx ** 3 == pytest.approx(target, rel=1e-13)
    # energies are F (the loop runs at mu = 0); the final result is at mu
    parts, _ = orc.total(rho, mu)
    assert res["free_energy"] == pytest.approx(parts[4], rel=1e-12)
    assert res["grand_potential"] == pytest.approx(parts[4] - mu * target, rel=1e-10)
    assert en[-1] < en[0]


def test_fixed_mass_converges_to_a_stationary_profile():
    # Euler-Lagrange condition of the constrained problem: at convergence
    # dF/drho is uniform and equals the multiplier lambda = sum(g rho) / N
    from classicaldft_amd import workloads as W

    shape, dx = (12, 12, 12), 0.25
    orc = O.Oracle(shape, dx, interactions=[W.LJ_WCA], kT=1.2, model="WhiteBearII")
    rho0 = 0.05 * (1.0 + 0.05 * (2.0 * W.uniform(9, 12 ** 3).reshape(shape) - 1.0))
    target = 0.05 * 12 ** 3 * dx ** 3
    rho, res, _ = orc.minimize_fixed_mass(rho0, 0.0, target, fire=dict(force_tolerance=1e-14, dt=0.02, dt_max=0.2),
                                          max_steps=3000, homogeneous_boundary=False)
    _, g = orc.total(rho, 0.0)
    lam = float(np.sum(g * rho)) / target
    dev = float(np.max(np.abs(g / dx ** 3 - lam)))
    print(f"fixed mass: {res['iterations']} steps, max |dF/drho - lambda| = {dev:.2e}, lambda = {lam:.6f}")
    assert dev < 1e-6 * abs(lam)
    # lambda is the bulk chemical potential of the uniform fluid at the mean density
    assert lam == pytest.approx(orc.bulk_mu(0.05), rel=1e-6)


# ------------------------------------------------------------------ EigenvalueSolver (saddle_point.hpp)
def test_eigen_solver_matches_dense_hessian():
    # LOBPCG with forward-difference products against the lowest eigenvalue of
    # the dense (symmetrised) finite-difference Hessian of the same forces
    from classicaldft_amd import workloads as W

    shape, dx = (8, 8, 8), 0.3
    orc = O.Oracle(shape, dx, interactions=[W.LJ_WCA], kT=0.9, model="WhiteBearII")
    rho = W.droplet(shape, dx, radius=0.7, width=0.3, rho_l=0.5, rho_v=0.05, noise=1e-3, seed=2)
    mu = orc.bulk_mu(0.05)
    n = rho.size
    eps = 1e-6
    _, f0 = orc.total(rho, mu)
    H = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        H[:, j] = ((orc.total(rho + eps * e.reshape(shape), mu)[1] - f0) / eps).ravel()
    lam_dense = np.linalg.eigvalsh(0.5 * (H + H.T))[0]
    v0 = W.normal(5, n)
    v, lam, it, conv = orc.smallest_eigenpair(rho, mu, v0, tolerance=1e-7, max_iterations=2000, hessian_eps=eps)
    print(f"eigen: LOBPCG {lam:.10e} in {it} iterations, dense {lam_dense:.10e}")
    assert conv
    assert lam == pytest.approx(lam_dense, rel=1e-6, abs=1e-9)
    hv = 0.5 * (H + H.T) @ v.ravel()
    assert np.linalg.norm(hv - lam * v.ravel()) < 1e-5 * abs(lam_dense) + 1e-8


# ------------------------------------------------------------------ split-operator, Crank-Nicolson, Picard
def _ideal_mode(shape, dx, rho0=0.2, eps=1e-4):
    x = np.arange(shape[0])
    return (rho0 * (1 + eps * np.cos(2 * PI * x / shape[0])))[:, None, None] * np.ones(shape)


def _mode_amplitude(rho, n):
    x = np.arange(n)
    return np.sum(rho[:, 0, 0] * np.cos(2 * PI * x / n)) / (n / 2)


def test_split_operator_ideal_gas_mode_decay():
    # ideal gas: the excess term vanishes to O(eps^2), so a mode decays by exp(-D k^2 dt) per step
    shape, dx = (16, 4, 4), 0.5
    orc = O.Oracle(shape, dx, model="none")
    rho = _ideal_mode(shape, dx)
    k2 = (2 * PI / (16 * dx)) ** 2
    dt, n = 1e-2, 5
    out, _ = orc.ddft_split_operator(rho, math.log(0.2), n, dt=dt, diffusion_coefficient=1.0)
    ratio = _mode_amplitude(out, 16) / _mode_amplitude(rho, 16)
    assert ratio == pytest.approx(math.exp(-k2 * dt * n), rel=1e-6)
    assert float(np.sum(out)) == pytest.approx(float(np.sum(rho)), rel=1e-13)


def test_crank_nicolson_ideal_gas_mode_decay():
    shape, dx = (16, 4, 4), 0.5
    orc = O.Oracle(shape, dx, model="none")
    rho = _ideal_mode(shape, dx)
    k2 = (2 * PI / (16 * dx)) ** 2
    dt, n = 1e-2, 5
    out, _, it = orc.ddft_crank_nicolson(rho, math.log(0.2), n, dt=dt, fp_tolerance=1e-14)
    a = k2 * dt / 2
    ratio = _mode_amplitude(out, 16) / _mode_amplitude(rho, 16)
    assert ratio == pytest.approx(((1 - a) / (1 + a)) ** n, rel=1e-6)
    assert it >= 2


def test_picard_ideal_gas_reaches_exp_mu():
    # ln rho - mu = 0: rho = exp(mu) everywhere; with a fixed mass the uniform density target / V
    shape, dx = (8, 8, 8), 0.5
    orc = O.Oracle(shape, dx, model="none")
    rho0 = 0.1 * (1.0 + 0.2 * np.random.default_rng(3).random(shape))
    # one update: rho exp(-mixing (ln rho - mu) dV / dV), exactly
    one, _ = orc.picard(rho0, math.log(0.3), mixing=0.5, tolerance=1e-30, max_iterations=1)
    np.testing.assert_allclose(one, rho0 * np.exp(-0.5 * (np.log(rho0) - math.log(0.3))), rtol=1e-13)
    # converged (the change of Omega stops it once it is below the tolerance)
    out, res = orc.picard(rho0, math.log(0.3), mixing=0.5, tolerance=1e-12, max_iterations=2000)
    assert res["converged"]
    np.testing.assert_allclose(out, 0.3, rtol=1e-6)
    target = 0.25 * rho0.size * dx ** 3
    out, res = orc.picard(rho0, math.log(0.3), mixing=0.5, tolerance=1e-12, max_iterations=2000, target_mass=target)
    np.testing.assert_allclose(out, 0.25, rtol=1e-6)
    assert float(np.sum(out)) * dx ** 3 == pytest.approx(target, rel=1e-13)
    # a negative excess over a uniform background 0.5 (fixed_mass_constraint takes any sign): 0.5 - 0.1
    V = rho0.size * dx ** 3
    out, res = orc.picard(rho0, math.log(0.3), mixing=0.5, tolerance=1e-12, max_iterations=2000,
                          target_mass=-0.1 * V, background=np.full(shape, 0.5))
    np.testing.assert_allclose(out, 0.4, rtol=1e-6)
    assert float(np.sum(out - 0.5)) * dx ** 3 == pytest.approx(-0.1 * V, rel=1e-12)


# ---------------------------------------------------------------------------
# Mixtures (hard_sphere.hpp:174-254, mean_field.hpp:233-281, functionals.hpp:64-95)
# ---------------------------------------------------------------------------

LJ_MIX = dict(potential="LJ", sigma=1.0, epsilon=1.0, r_cutoff=2.5, split="WCA", scheme="quadraticF")

from closed_forms import rosenfeld_mixture  # noqa: E402


def test_oracle_mixture_uniform_matches_scaled_particle_theory():
    # binary fluid of the reference's hard_sphere.cpp:208-248 (d = 1.0, 0.8; rho = 0.3, 0.2)
    shape, dx = (16, 16, 16), 0.2
    diams, rhos, mus = (1.0, 0.8), (0.3, 0.2), (0.4, -0.7)
    orc = O.Oracle(shape, dx, diameters=diams, model="Rosenfeld")
    rho = np.stack([np.full(shape, r) for r in rhos])
    parts, f = orc.total(rho, mus)
    phi, mu_ex = rosenfeld_mixture(rhos, diams)
    vol, dv = np.prod(shape) * dx ** 3, dx ** 3
    assert parts[2] == pytest.approx(phi * vol, rel=1e-12)
    assert parts[0] == pytest.approx(sum(r * (np.log(r) - 1) for r in rhos) * vol, rel=1e-13)
    assert parts[5] == pytest.approx(parts[4] - sum(m * r for m, r in zip(mus, rhos)) * vol, rel=1e-13)
    for s in range(2):
        expect = (np.log(rhos[s]) - mus[s] + mu_ex[s]) * dv
        assert np.max(np.abs(f[s] - expect)) < 1e-12 * abs(expect) + 1e-15
    assert abs(mu_ex[0] - mu_ex[1]) > 0.1  # the species differ


def test_oracle_mixture_equal_diameters_reduce_to_one_species():
    # two species of one diameter: the summed measures are those of rho_a + rho_b
    from classicaldft_amd import workloads as W

    shape, dx = (20, 18, 16), 0.2
    ra = W.smooth_noise_fluid(shape, dx, rho0=0.35, amp=0.05, seed=3)
    rb = W.smooth_noise_fluid(shape, dx, rho0=0.25, amp=0.05, seed=4)
    for model in ("WhiteBearII", "RSLT"):
        mix = O.Oracle(shape, dx, diameters=(1.0, 1.0), model=model)
        one = O.Oracle(shape, dx, diameters=(1.0,), model=model)
        pm, fm = mix.total(np.stack([ra, rb]), (0.0, 0.0))
        p1, f1 = one.total(ra + rb, 0.0)
        assert pm[2] == pytest.approx(p1[2], rel=1e-12)
        dv = dx ** 3
        hs1 = f1 - np.log(ra + rb) * dv
        for s, r in enumerate((ra, rb)):
            hs = fm[s] - np.log(r) * dv
            assert np.max(np.abs(hs - hs1)) < 1e-11 * np.max(np.abs(hs1))


@pytest.mark.parametrize("model", ["WhiteBearII", "Rosenfeld", "esFMT"])
def test_oracle_mixture_forces_match_finite_differences(model):
    # d Omega / d rho_s(p) = force_s(p) with unequal radii and a cross pair: pins the
    # per-species radius in the weight combination and the halves of the cross term
    from classicaldft_amd import workloads as W

    shape, dx = (12, 12, 12), 0.25
    diams = (1.0, 0.6)
    inter = [dict(LJ_MIX, si=0, sj=0), dict(LJ_MIX, si=0, sj=1, epsilon=0.7)]
    orc = O.Oracle(shape, dx, diameters=diams, interactions=inter, kT=1.2, model=model, A=1.2, B=-0.6)
    rho = np.stack([W.smooth_noise_fluid(shape, dx, rho0=0.3, amp=0.06, seed=7),
                    W.smooth_noise_fluid(shape, dx, rho0=0.2, amp=0.06, seed=8)])
    mus = (0.3, -0.2)
    _, f = orc.total(rho, mus)
    h = 1e-6
    rng = np.random.default_rng(1)
    for s in range(2):
        for _ in range(4):
            p = tuple(int(i) for i in rng.integers(0, 12, size=3))
            up, dn = rho.copy(), rho.copy()
            up[(s,) + p] += h
            dn[(s,) + p] -= h
            fd = (orc.total(up, mus)[0][5] - orc.total(dn, mus)[0][5]) / (2 * h)
            assert fd == pytest.approx(f[(s,) + p], rel=2e-6, abs=1e-9)


def test_oracle_mixture_cross_mean_field_closed_form():
    # mean_field.cpp:183-229 of the reference: F = 1/2 a rho_a rho_b V, forces 1/2 a rho_other dV
    shape, dx = (16, 16, 16), 0.2
    orc = O.Oracle(shape, dx, diameters=(1.0, 0.8), interactions=[dict(LJ_MIX, si=0, sj=1)], kT=1.0,
                   model="none")
    ra, rb = 0.3, 0.2
    parts, f = orc.total(np.stack([np.full(shape, ra), np.full(shape, rb)]), (0.0, 0.0))
    a = orc.a_vdw(0)
    vol, dv = np.prod(shape) * dx ** 3, dx ** 3
    assert parts[3] == pytest.approx(0.5 * a * ra * rb * vol, rel=1e-12)
    assert np.max(np.abs(f[0] - (np.log(ra) + 0.5 * a * rb) * dv)) < 1e-14
    assert np.max(np.abs(f[1] - (np.log(rb) + 0.5 * a * ra) * dv)) < 1e-14


def test_ddft_restart_limit_and_non_finite_iterate():
    # the restatement's guards (the library's): a step that restarts more than 30 times, or whose iterate
    # is not finite, ends the run with an error
    shape, dx = (8, 8, 8), 0.5
    orc = O.Oracle(shape, dx, model="none")
    x = np.arange(8)
    rho = (0.3 * (1 + 0.1 * np.cos(2 * PI * x / 8)))[:, None, None] * np.ones(shape)
    with pytest.raises(RuntimeError):
        orc.ddft_run(rho, 0.0, 1, dt=1e-2, fp_tolerance=1e-300, fp_max_iterations=1)
    bad = rho.copy()
    bad[1, 2, 3] = np.nan
    with pytest.raises(RuntimeError):
        orc.ddft_run(bad, 0.0, 1, dt=1e-2)
    # a reachable tolerance runs normally
    out, res, _ = orc.ddft_run(rho, 0.0, 2, dt=1e-2, dt_max=1e-2)
    assert res["time"] == pytest.approx(2e-2, rel=1e-14)


def test_ddft_reservoir_boundary():
    # reservoir_boundary (dynamics.hpp:545-553): marked cells are reset after every step; no mark leaves
    # the run unchanged
    shape, dx = (8, 8, 8), 0.5
    orc = O.Oracle(shape, dx, model="none")
    rng = np.random.default_rng(8)
    rho = 0.3 + 0.05 * rng.random(shape)
    mask = np.zeros(shape, dtype=bool)
    mask[0] = True
    out, res, _ = orc.ddft_run(rho, 0.0, 3, dt=1e-2, dt_max=1e-2, reservoir_mask=mask, reservoir_density=0.2)
    assert np.all(out[0] == 0.2)
    start = rho.copy()
    start[0] = 0.2
    assert res["mass_initial"] == pytest.approx(float(np.sum(start)) * dx ** 3, rel=1e-14)
    plain, _, _ = orc.ddft_run(rho, 0.0, 3, dt=1e-2, dt_max=1e-2)
    none, _, _ = orc.ddft_run(rho, 0.0, 3, dt=1e-2, dt_max=1e-2, reservoir_mask=np.zeros(shape, dtype=bool))
    assert np.array_equal(plain, none)
    # mass flows: the unmarked cells differ from the closed run
    assert np.max(np.abs(out[1:] - plain[1:])) > 1e-6

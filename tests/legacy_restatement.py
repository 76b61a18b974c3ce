"""Independent numpy restatements of the reference's legacy algorithms, used by
the CPU tests to pin the oracle's FIRE / Minimizer::fixed_mass and DDFT
integrating-factor control logic the way the reference pins its own
(tests/integration/algorithms/nucleation.cpp:378-401 compares
Minimizer::fixed_mass with legacy FIRE to 1e-12 and integrating_factor_step
with the legacy DDFT step to 1e-3).

Test infrastructure only.  The two functions follow the legacy formulations,
which differ in arrangement from the library's: the legacy FIRE updates the
velocity with the alias-space gradient (v -= dt dF) and scales before mixing,
and counts iterations from 1; the legacy DDFT step divides the whole flux
term by dV after the stencil and iterates under a combined loop condition.
"""
from __future__ import annotations

import numpy as np

DMIN = 1e-99  # legacy alias floor (legacy/algorithms.hpp:116)


def boundary_mask(shape):
    """Face points of the box (legacy/algorithms.hpp:131-142)."""
    m = np.zeros(shape, dtype=bool)
    m[0], m[-1] = True, True
    m[:, 0], m[:, -1] = True, True
    m[:, :, 0], m[:, :, -1] = True, True
    return m.reshape(-1)


def _evaluate_fixed_mass(force_fn, x, target_mass, dv, mask):
    """legacy/algorithms.hpp:161-181: density from the alias, rescaled to the target mass; forces averaged over
    the boundary; Lagrange multiplier removed; gradient in alias space."""
    rho = DMIN + x * x
    rho = rho * (target_mass / (np.sum(rho) * dv))
    F, dF = force_fn(rho)
    dF = dF.copy()
    if np.any(mask):
        dF[mask] = np.sum(dF[mask]) / np.count_nonzero(mask)
    lam = float(np.dot(dF, rho)) / target_mass
    dF -= lam * dv
    return F, 2.0 * x * dF, lam, rho


def fire_minimize_fixed_mass(force_fn, rho_init, target_mass, dv, shape, dt=0.1, dt_max=1.0, alpha_start=0.01,
                             alpha_fac=0.99, force_tolerance=1e-4, max_steps=1000000, n_delay=5, f_dec=0.5,
                             f_inc=1.1, energies=None):
    """legacy/algorithms.hpp:208-271 (fire_minimize_fixed_mass).  force_fn(rho) -> (F, dF/drho dV) of a flat
    density.  Returns (density, F, lambda, iterations, converged); `energies` (a list) receives F after every
    iteration."""
    n = rho_init.size
    volume = n * dv
    mask = boundary_mask(shape)
    rho0 = np.asarray(rho_init, dtype=np.float64).reshape(-1)
    rho0 = rho0 * (target_mass / (np.sum(rho0) * dv))
    x = np.sqrt(np.maximum(0.0, rho0 - DMIN))
    v = np.zeros(n)
    F, dF, lam, rho = _evaluate_fixed_mass(force_fn, x, target_mass, dv, mask)
    alpha, n_pos, vv, converged = alpha_start, 0, 1.0, False
    it = 0
    for it in range(1, max_steps + 1):
        P = -float(np.dot(v, dF))
        if P > 0 or it == 1:
            n_pos += 1
            if n_pos > n_delay:
                dt = min(dt * f_inc, dt_max)
                alpha *= alpha_fac
        else:
            n_pos = 0
            x = x - 0.5 * dt * v
            v = np.zeros(n)
            if it >= n_delay:
                dt *= f_dec
                alpha = alpha_start
        v = v - dt * dF
        vnorm = float(np.linalg.norm(v))
        fnorm = float(np.linalg.norm(dF))
        v = v * (1.0 - alpha)
        if fnorm > 0:
            v = v - (alpha * vnorm / fnorm) * dF
        x = x + dt * v
        F_old = F
        F, dF, lam, rho = _evaluate_fixed_mass(force_fn, x, target_mass, dv, mask)
        if energies is not None:
            energies.append(F)
        vv = vv * 0.9 + 0.1 * (abs(F - F_old) / dt)
        if abs(vv / volume) < force_tolerance:
            converged = True
            break
    return rho, F, lam, it, converged


def _lambdas(shape, dx):
    """Eigenvalues of the periodic 3-point Laplacian per axis (legacy/algorithms.hpp:442-461)."""
    nx, ny, nz = shape
    D = 1.0 / (dx * dx)
    lx = 2.0 * D * (np.cos(2 * np.pi * np.arange(nx) / nx) - 1.0)
    ly = 2.0 * D * (np.cos(2 * np.pi * np.arange(ny) / ny) - 1.0)
    lz = 2.0 * D * (np.cos(2 * np.pi * np.arange(nz // 2 + 1) / nz) - 1.0)
    return lx[:, None, None] + ly[None, :, None] + lz[None, None, :]


def _excess_rhs_k(rho, forces, dv, dx):
    """legacy/algorithms.hpp:476-555: g_dot_x(rho, f) / dV - Laplacian(rho), forward transformed."""
    D = 0.5 / (dx * dx)
    gx = np.zeros_like(rho)
    for ax in range(3):
        rp, rm = np.roll(rho, -1, ax), np.roll(rho, 1, ax)
        fp, fm = np.roll(forces, -1, ax), np.roll(forces, 1, ax)
        gx += D * ((rp + rho) * (fp - forces) - (rho + rm) * (forces - fm))
    rhs = gx / dv
    D2 = 1.0 / (dx * dx)
    for ax in range(3):
        rhs -= D2 * (np.roll(rho, -1, ax) + np.roll(rho, 1, ax) - 2.0 * rho)
    return np.fft.rfftn(rhs)


def ddft_step(rho0, force_fn, dv, dx, fp_tolerance, fp_max_iterations, dt):
    """legacy/algorithms.hpp:627-677 (ddft_step) on a 3D density.  force_fn(rho) -> (energy, forces).  Returns
    (density, energy, dt used)."""
    shape = rho0.shape
    L = _lambdas(shape, dx)
    energy, forces = force_fn(rho0)
    r0k = _excess_rhs_k(rho0, forces, dv, dx)
    p0k = np.fft.rfftn(rho0)
    while True:
        E = np.exp(L * dt)
        with np.errstate(divide="ignore", invalid="ignore"):
            U1 = (E - 1.0) / L
            U2 = (E - 1.0 - dt * L) / (L * L * dt)
        U1[0, 0, 0] = dt
        U2[0, 0, 0] = dt / 2.0
        d1 = rho0.copy()
        deviation, restart, i = 1e30, False, 0
        while i < fp_max_iterations and deviation > fp_tolerance and not restart:
            _, f1 = force_fn(d1)
            r1k = _excess_rhs_k(d1, f1, dv, dx)
            new = np.fft.irfftn(E * p0k + U1 * r0k + U2 * (r1k - r0k), s=shape, axes=(0, 1, 2))
            old = deviation
            deviation = float(np.max(np.abs(d1 - new)))
            d1 = new
            if d1.min() < 0 or (i > 0 and old < deviation):
                restart = True
            i += 1
        if restart or deviation > fp_tolerance:
            dt /= 10.0
            continue
        break
    e_final, _ = force_fn(d1)
    return d1, e_final, dt

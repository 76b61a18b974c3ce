"""Seeded synthetic inputs for the benchmark and parity workloads.

A project-owned SplitMix64 generator keeps every input reproducible across
machines and languages (no implementation-defined std:: distributions).
"""
from __future__ import annotations

import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n consecutive SplitMix64 outputs (uint64) from `seed`."""
    with np.errstate(over="ignore"):
        state = np.uint64(seed) + _GOLDEN * (np.arange(1, n + 1, dtype=np.uint64))
        z = state
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, n: int) -> np.ndarray:
    """53-bit uniforms in [0, 1)."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int) -> np.ndarray:
    """Box-Muller standard normals."""
    m = (n + 1) // 2
    u = uniform(seed, 2 * m)
    u1 = np.maximum(u[0::2], 1e-300)
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * math.pi * u2), r * np.sin(2 * math.pi * u2)])
    return z[:n]


def droplet(shape, dx, radius=4.0, width=0.5, rho_l=0.70, rho_v=0.02, noise=1e-3, seed=2024):
    """Spherical tanh droplet at the box centre plus uniform noise in [-noise, noise]."""
    nx, ny, nz = shape
    ax = [(np.arange(n) - n / 2) * dx for n in shape]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    rho = rho_v + 0.5 * (rho_l - rho_v) * (1.0 - np.tanh((r - radius) / width))
    rho += noise * (2.0 * uniform(seed, nx * ny * nz).reshape(shape) - 1.0)
    return rho


def slab(shape, dx, fraction=0.4, width=0.5, rho_l=0.70, rho_v=0.02, noise=1e-3, seed=2025, axis=0):
    """Planar liquid slab over `fraction` of the box along `axis`, tanh interfaces."""
    n = shape[axis]
    L = n * dx
    x = (np.arange(n) + 0.5) * dx - L / 2
    half = fraction * L / 2
    prof = rho_v + 0.5 * (rho_l - rho_v) * (np.tanh((x + half) / width) - np.tanh((x - half) / width))
    sh = [1, 1, 1]
    sh[axis] = n
    rho = np.broadcast_to(prof.reshape(sh), shape).copy()
    rho += noise * (2.0 * uniform(seed, int(np.prod(shape))).reshape(shape) - 1.0)
    return rho


def smooth_noise_fluid(shape, dx, rho0=0.5, amp=0.01, sigma=0.3, seed=7):
    """rho0 (1 + amp xi), xi = N(0,1) noise smoothed by a periodic Gaussian of width sigma, unit RMS."""
    xi = normal(seed, int(np.prod(shape))).reshape(shape)
    k = [2 * math.pi * np.fft.fftfreq(n, d=dx) for n in shape]
    KX, KY, KZ = np.meshgrid(*k, indexing="ij")
    g = np.exp(-0.5 * sigma * sigma * (KX ** 2 + KY ** 2 + KZ ** 2))
    xs = np.real(np.fft.ifftn(np.fft.fftn(xi) * g))
    xs /= np.sqrt(np.mean(xs * xs))
    return rho0 * (1.0 + amp * xs)


LJ_WCA = dict(potential="LJ", sigma=1.0, epsilon=1.0, r_cutoff=2.5, split="WCA", scheme="quadraticF")


BATCH_SIZES = (48, 64, 80, 96, 128, 160)


def batch_state_points(n, seed=5, sizes=BATCH_SIZES):
    """n seeded state points of uneven grid size for the batched form: cubic grid of a size drawn from `sizes`
    (dx = 0.1), kT ~ U(0.7, 1.2), bulk density ~ U(0.05, 0.80), a tanh droplet or slab between that density and a
    liquid / vapour partner.  Returns a list of dicts with shape, dx, kT, rho_bulk and the initial profile."""
    u = uniform(seed, 5 * n).reshape(n, 5)
    pts = []
    for i in range(n):
        L = sizes[min(int(u[i, 0] * len(sizes)), len(sizes) - 1)]
        shape, dx = (L, L, L), 0.1
        kT = 0.7 + 0.5 * u[i, 1]
        rb = 0.05 + 0.75 * u[i, 2]
        other = 0.7 if rb < 0.4 else 0.05
        if u[i, 3] < 0.5:
            prof = droplet(shape, dx, radius=0.25 * L * dx, width=0.5, rho_l=other if rb < 0.4 else rb,
                           rho_v=rb if rb < 0.4 else other, noise=1e-3, seed=seed + 17 * i)
        else:
            prof = slab(shape, dx, fraction=0.4, width=0.5, rho_l=max(rb, other), rho_v=min(rb, other), noise=1e-3,
                        seed=seed + 17 * i)
        pts.append(dict(shape=shape, dx=dx, kT=kT, rho_bulk=rb, rho0=prof))
    return pts


def wells(shape, dx, n_wells=16, seed=11, amp=(-2.0, 0.0), width=(0.5, 1.5)):
    """Seeded Gaussian wells (n_wells, 5): amplitude A ~ U(amp) kT, width w ~ U(width) sigma, centre uniform in
    the periodic box.  V(r) = sum_j A_j exp(-|r - c_j|^2 / (2 w_j^2)) with minimum-image distances."""
    u = uniform(seed, 5 * n_wells).reshape(n_wells, 5)
    L = np.array(shape, dtype=np.float64) * dx
    prm = np.empty((n_wells, 5))
    prm[:, 0] = amp[0] + (amp[1] - amp[0]) * u[:, 0]
    prm[:, 1] = width[0] + (width[1] - width[0]) * u[:, 1]
    prm[:, 2:5] = u[:, 2:5] * L
    return prm


def adjoint_problem(shape=(128, 128, 128), dx=0.1, n_wells=16, rho_bulk=0.3, seed=11, perturb=0.1):
    """The DDFT adjoint workload: a fluid at rho_bulk with 1e-3 uniform noise, the seeded wells, and wells whose
    amplitudes are scaled by (1 + perturb u), u ~ U(-1, 1), to generate the target profile.  Returns
    (rho0, params, target_params)."""
    n = int(np.prod(shape))
    rho0 = rho_bulk * (1.0 + 1e-3 * (2.0 * uniform(seed + 1, n).reshape(shape) - 1.0))
    prm = wells(shape, dx, n_wells, seed)
    tgt = prm.copy()
    tgt[:, 0] *= 1.0 + perturb * (2.0 * uniform(seed + 2, n_wells) - 1.0)
    return rho0, prm, tgt

"""Closed forms shared by the oracle and GPU tests (not a test module)."""
import numpy as np


def rosenfeld_mixture(rhos, diams):
    """Bulk Rosenfeld (scaled-particle / Percus-Yevick compressibility) mixture:
    Phi and d Phi / d rho_s from n3 = sum rho pi d^3/6, n2 = sum rho pi d^2,
    n1 = sum rho d/2, n0 = sum rho."""
    rhos, diams = np.asarray(rhos, float), np.asarray(diams, float)
    dn = np.stack([np.pi * diams ** 3 / 6, np.pi * diams ** 2, diams / 2, np.ones_like(diams)])  # [alpha, s]
    n3, n2, n1, n0 = dn @ rhos
    e = 1.0 - n3
    phi = -n0 * np.log(e) + n1 * n2 / e + n2 ** 3 / (24 * np.pi * e ** 2)
    d_n3 = n0 / e + n1 * n2 / e ** 2 + n2 ** 3 / (12 * np.pi * e ** 3)
    d_n2 = n1 / e + n2 ** 2 / (8 * np.pi * e ** 2)
    d_n1 = n2 / e
    d_n0 = -np.log(e)
    mu_ex = np.array([d_n3, d_n2, d_n1, d_n0]) @ dn
    return phi, mu_ex

"""Writes the golden fixtures of this folder from the FP64 CPU oracle.

The reference library cannot be built here (C++23 with FFTW3, Armadillo,
GSL and network-fetched autodiff/Catch2; none available) and holds no stored
numeric vectors for this path, so the expected outputs come from the oracle,
which tests/test_oracle.py pins against closed forms, brute-force DFTs and
finite differences.  Each fixture is a small .npz (no pickles): the seeded
inputs, the model settings (JSON in `settings`) and the expected outputs
(`parts` = [F_id, F_ext, F_hs, F_mf, F, Omega], `forces`, `a_vdw`, and for one
species the 12 weighted-density fields).

    python3 tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from classicaldft_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402

LJ = dict(potential="LJ", sigma=1.0, epsilon=1.0, r_cutoff=2.5, split="WCA", scheme="quadraticF")

# name -> (shape, dx, diameters, interactions, kT, model, rho builder, mu, external field amplitude)
CASES = {
    "droplet_8cubed_wb2": dict(shape=(8, 8, 8), dx=0.25, diameters=(1.0,), inter=[LJ], kT=0.8,
                               model="WhiteBearII", mu=(-2.1,), vamp=0.4),
    "fluid_12x10x8_rslt": dict(shape=(12, 10, 8), dx=0.25, diameters=(1.0,), inter=[], kT=1.0,
                               model="RSLT", mu=(0.3,), vamp=0.0),
    "binary_16cubed_rosenfeld": dict(shape=(16, 16, 16), dx=0.2, diameters=(1.0, 0.8),
                                     inter=[dict(LJ, si=0, sj=0), dict(LJ, si=0, sj=1, sigma=0.9, epsilon=0.7)],
                                     kT=1.1, model="Rosenfeld", mu=(0.4, -0.3), vamp=0.2),
}


def inputs(name, c):
    shape, dx = c["shape"], c["dx"]
    S = len(c["diameters"])
    if name.startswith("droplet"):
        rho = W.droplet(shape, dx, radius=0.6, width=0.3, rho_l=0.6, rho_v=0.05, noise=1e-3, seed=11)[None]
    else:
        rho = np.stack([W.smooth_noise_fluid(shape, dx, rho0=0.3 - 0.1 * s, amp=0.03, seed=21 + s)
                        for s in range(S)])
    z = (np.arange(shape[2]) + 0.5) * dx
    v1 = np.broadcast_to(np.cos(2 * np.pi * z / (shape[2] * dx)), shape)
    vext = np.stack([c["vamp"] * (1.0 - 0.5 * s) * v1 for s in range(S)])
    return rho, vext


def main():
    for name, c in CASES.items():
        rho, vext = inputs(name, c)
        orc = O.Oracle(c["shape"], c["dx"], diameters=c["diameters"], interactions=c["inter"], kT=c["kT"],
                       model=c["model"])
        one = len(c["diameters"]) == 1
        v = None if c["vamp"] == 0.0 else (vext[0] if one else vext)
        parts, forces = orc.total(rho[0] if one else rho, c["mu"][0] if one else c["mu"], vext=v)
        settings = dict(shape=c["shape"], dx=c["dx"], diameters=c["diameters"], interactions=c["inter"],
                        kT=c["kT"], model=c["model"], has_vext=c["vamp"] != 0.0)
        out = dict(settings=np.array(json.dumps(settings)), rho=rho, vext=vext, mu=np.asarray(c["mu"], float),
                   parts=parts,
                   forces=np.asarray(forces).reshape(rho.shape),
                   a_vdw=np.array([orc.a_vdw(k) for k in range(len(c["inter"]))]))
        if one:
            out["weighted_densities"] = orc.weighted_densities(rho[0])
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print(name, "Omega", parts[5])


if __name__ == "__main__":
    main()

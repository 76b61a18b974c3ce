"""GPU parity of the per-term split of functionals::total (cdft_contributions:
ideal_gas, external_field, hard_sphere and mean_field forces) against the FP64
CPU oracle, through the C ABI.

The oracle's terms: ideal (ln rho - mu) dV and external V dV in closed form;
the mean field from the oracle's convolution field (one species) or from an
oracle without FMT (mixtures); hard spheres as the oracle's total less those.
Tolerance: 1e-11 relative to the largest force of the term.
"""
import numpy as np
import pytest

from classicaldft_amd import workloads as W

pytestmark = pytest.mark.gpu

LJ = W.LJ_WCA


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("shape,dx", [((64, 64, 64), 0.1), ((30, 24, 20), 0.2)])
def test_one_species_terms_match_oracle(shape, dx):
    import classicaldft_amd as cd
    from oracle import oracle as O

    ctx = cd.Context(shape, dx, interactions=[LJ], kT=0.8, model="WhiteBearII")
    orc = O.Oracle(shape, dx, interactions=[LJ], kT=0.8, model="WhiteBearII")
    rho = W.droplet(shape, dx, radius=1.2, width=0.4, rho_l=0.6, rho_v=0.05, noise=1e-3, seed=7)
    z = (np.arange(shape[2]) + 0.5) * dx
    vext = np.broadcast_to(0.3 * np.cos(2 * np.pi * z / (shape[2] * dx)), shape).copy()
    mu = -1.7
    print(f"terms {shape} ({ctx.schedule})", flush=True)
    res, t = ctx.contributions(rho, mu, vext=vext)
    parts, fo = orc.total(rho, mu, vext=vext)
    dv = dx ** 3
    ideal = (np.log(rho) - mu) * dv
    ext = vext * dv
    mf = orc.weighted_densities(rho)[11] * dv
    hs = fo - ideal - ext - mf
    assert _rel(t["ideal"][0], ideal) < 1e-13
    assert _rel(t["external"][0], ext) < 1e-13
    assert _rel(t["mean_field"][0], mf) < 1e-11
    assert _rel(t["hard_sphere"][0], hs) < 1e-11
    for i, k in enumerate(("ideal", "external", "hard_sphere", "mean_field")):
        assert res[k] == pytest.approx(parts[i], rel=1e-12)
    # the terms add up to the total force
    _, ftot = ctx.total(rho, mu, vext=vext)
    assert _rel(sum(t[k][0] for k in t), ftot) < 1e-14


def test_mixture_terms_match_oracle():
    import classicaldft_amd as cd
    from oracle import oracle as O

    shape, dx = (24, 24, 24), 0.2
    diams = (1.0, 0.8)
    inter = [dict(LJ, si=0, sj=0), dict(LJ, si=0, sj=1, sigma=0.9, epsilon=0.8)]
    ctx = cd.Context(shape, dx, diameters=diams, interactions=inter, kT=1.0, model="RSLT")
    orc = O.Oracle(shape, dx, diameters=diams, interactions=inter, kT=1.0, model="RSLT")
    orc_mf = O.Oracle(shape, dx, diameters=diams, interactions=inter, kT=1.0, model="none")
    rho = np.stack([W.smooth_noise_fluid(shape, dx, rho0=r, amp=0.08 * r, seed=50 + s)
                    for s, r in enumerate((0.3, 0.2))])
    mus = (0.2, -0.4)
    res, t = ctx.contributions(rho, mus)
    _, fo = orc.total(rho, mus)
    _, fn = orc_mf.total(rho, mus)  # ideal + mean field only
    dv = dx ** 3
    for s in range(2):
        ideal = (np.log(rho[s]) - mus[s]) * dv
        assert _rel(t["ideal"][s], ideal) < 1e-13
        assert np.all(t["external"][s] == 0.0)
        assert _rel(t["mean_field"][s], fn[s] - ideal) < 1e-11
        assert _rel(t["hard_sphere"][s], fo[s] - fn[s]) < 1e-11
    r2, _ = ctx.total_species(rho, mus)
    assert res == r2


def test_ideal_only_context_has_zero_hard_sphere_term():
    import classicaldft_amd as cd

    shape, dx = (16, 16, 16), 0.2
    ctx = cd.Context(shape, dx, interactions=[LJ], kT=1.0, model="none")
    rho = W.smooth_noise_fluid(shape, dx, rho0=0.3, amp=0.02, seed=3)
    res, t = ctx.contributions(rho, 0.0)
    assert res["hard_sphere"] == 0.0 and np.all(t["hard_sphere"] == 0.0)


def test_mixed_precision_terms_on_the_fused_schedule():
    # the mean-field term comes from the weighted-density pass, the total from the fused kernels:
    # the hard-sphere term keeps the mixed-precision force tolerance
    import classicaldft_amd as cd
    from oracle import oracle as O

    shape, dx = (64, 64, 64), 0.1
    ctx = cd.Context(shape, dx, interactions=[LJ], kT=0.8, model="WhiteBearII", precision="mixed")
    assert ctx.schedule == "fused"
    orc = O.Oracle(shape, dx, interactions=[LJ], kT=0.8, model="WhiteBearII")
    rho = W.droplet(shape, dx, radius=1.2, width=0.4, rho_l=0.6, rho_v=0.05, noise=1e-3, seed=8)
    mu = ctx.bulk_mu(0.05)
    res, t = ctx.contributions(rho, mu)
    parts, fo = orc.total(rho, mu)
    dv = dx ** 3
    ideal = (np.log(rho) - mu) * dv
    mf = orc.weighted_densities(rho)[11] * dv
    assert _rel(t["ideal"][0], ideal) < 1e-13
    assert _rel(t["mean_field"][0], mf) < 2e-5
    assert _rel(t["hard_sphere"][0], fo - ideal - mf) < 2e-5
    assert res["grand_potential"] == pytest.approx(parts[5], rel=1e-6)

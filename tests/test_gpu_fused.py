"""GPU parity of the fused six-kernel evaluation schedule (y and z extents 64
or 128, x extent 64, 128 or 256) against the FP64 CPU oracle and against the
general schedule, through the C ABI.

Tolerances: FP64 fields and forces 1e-11 relative to their maximum, energies
1e-12; mixed precision Omega 1e-6 relative, forces 2e-5 relative to max|force|.
"""
import numpy as np
import pytest

from classicaldft_amd import workloads as W

pytestmark = pytest.mark.gpu

LJ = W.LJ_WCA


def _ctx(shape, dx, model="WhiteBearII", inter=(LJ,), kT=0.8, precision="fp64", schedule="auto", **kw):
    import classicaldft_amd as cd

    return cd.Context(shape, dx, interactions=list(inter), kT=kT, model=model, precision=precision,
                      schedule=schedule, **kw)


def _orc(shape, dx, model="WhiteBearII", inter=(LJ,), kT=0.8, **kw):
    from oracle import oracle as O

    return O.Oracle(shape, dx, interactions=list(inter), kT=kT, model=model, **kw)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


PARTS = ("ideal", "external", "hard_sphere", "mean_field", "free_energy", "grand_potential")


def _check_parts(res, parts, rtol, atol):
    # oracle parts: [F_id, F_ext, F_hs, F_mf, F, Omega]
    for i, k in enumerate(PARTS):
        assert res[k] == pytest.approx(parts[i], rel=rtol, abs=atol), k


@pytest.mark.parametrize("shape", [(64, 64, 64), (128, 64, 64), (64, 128, 64), (64, 64, 128), (256, 64, 64),
                                   (256, 64, 128)])
def test_fused_fp64_matches_oracle(shape):
    dx = 0.1
    ctx = _ctx(shape, dx)
    assert ctx.schedule == "fused"
    orc = _orc(shape, dx)
    rho = W.droplet(shape, dx, radius=1.6, width=0.4, rho_l=0.7, rho_v=0.02, noise=1e-3, seed=11)
    mu = orc.bulk_mu(0.02)
    res, f = ctx.total(rho, mu)
    ref, fr = orc.total(rho, mu)
    e = _rel(f, fr)
    print(f"fused fp64 {shape}: force rel err {e:.2e}, Omega {res['grand_potential']:.12e} "
          f"vs {ref[5]:.12e}", flush=True)
    assert e < 1e-11
    _check_parts(res, ref, 1e-11, 1e-10)


@pytest.mark.parametrize("model", ["WhiteBearII", "WhiteBearI", "Rosenfeld", "RSLT", "esFMT"])
def test_fused_models_with_external_field(model):
    shape, dx = (64, 64, 64), 0.1
    kw = dict(model=model, **(dict(A=1.0, B=-0.5) if model == "esFMT" else {}))
    ctx = _ctx(shape, dx, **kw)
    orc = _orc(shape, dx, **kw)
    rho = W.droplet(shape, dx, radius=1.5, width=0.4, rho_l=0.6, rho_v=0.05, noise=1e-3, seed=4)
    V = 0.5 * np.cos(2 * np.pi * np.arange(shape[2]) / shape[2])[None, None, :] * np.ones(shape)
    mu = orc.bulk_mu(0.05)
    res, f = ctx.total(rho, mu, vext=V)
    ref, fr = orc.total(rho, mu, vext=V)
    e = _rel(f, fr)
    print(f"fused {model} + V: force rel err {e:.2e}", flush=True)
    assert e < 1e-11
    _check_parts(res, ref, 1e-11, 1e-10)


@pytest.mark.parametrize("model,inter", [("WhiteBearII", ()), ("none", (LJ,)), ("none", ())])
def test_fused_channel_sets(model, inter):
    shape, dx = (64, 64, 64), 0.1
    ctx = _ctx(shape, dx, model=model, inter=inter)
    orc = _orc(shape, dx, model=model, inter=inter)
    rho = W.droplet(shape, dx, radius=1.5, width=0.4, rho_l=0.6, rho_v=0.05, noise=1e-3, seed=8)
    mu = -1.3
    res, f = ctx.total(rho, mu)
    ref, fr = orc.total(rho, mu)
    assert _rel(f, fr) < 1e-11
    _check_parts(res, ref, 1e-11, 1e-10)
    # energy-only evaluation: same energies
    res2, _ = ctx.total(rho, mu, forces=False)
    _check_parts(res2, ref, 1e-11, 1e-10)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fused_matches_general_schedule(precision):
    shape, dx = (64, 64, 128), 0.1
    rho = W.droplet(shape, dx, radius=1.8, width=0.5, rho_l=0.7, rho_v=0.02, noise=1e-3, seed=21)
    a = _ctx(shape, dx, precision=precision)
    b = _ctx(shape, dx, precision=precision, schedule="general")
    assert a.schedule == "fused" and b.schedule == "general"
    mu = a.bulk_mu(0.02)
    ra, fa = a.total(rho, mu)
    rb, fb = b.total(rho, mu)
    tol = 1e-12 if precision == "fp64" else 2e-5
    e = _rel(fa, fb)
    print(f"fused vs general {precision}: force {e:.2e}, Omega {ra['grand_potential']:.12e} "
          f"vs {rb['grand_potential']:.12e}", flush=True)
    assert e < tol
    assert ra["grand_potential"] == pytest.approx(rb["grand_potential"], rel=1e-12 if precision == "fp64" else 1e-6)


@pytest.mark.parametrize("shape,precision", [((256, 128, 128), "mixed"), ((256, 64, 128), "fp64"),
                                             ((256, 128, 64), "mixed")])
def test_fused_x256_matches_general_schedule(shape, precision):
    # x extent 256 (the DDFT slab grid is 256 x 128 x 128)
    dx = 0.1
    rho = W.slab(shape, dx, fraction=0.4, width=0.5, rho_l=0.7, rho_v=0.02, noise=1e-3, axis=0)
    a = _ctx(shape, dx, precision=precision)
    b = _ctx(shape, dx, precision=precision, schedule="general")
    assert a.schedule == "fused" and b.schedule == "general"
    mu = a.bulk_mu(0.02)
    ra, fa = a.total(rho, mu)
    rb, fb = b.total(rho, mu)
    e = _rel(fa, fb)
    print(f"fused vs general {shape} {precision}: force {e:.2e}, Omega {ra['grand_potential']:.12e} "
          f"vs {rb['grand_potential']:.12e}", flush=True)
    assert e < (1e-12 if precision == "fp64" else 2e-5)
    assert ra["grand_potential"] == pytest.approx(rb["grand_potential"], rel=1e-12 if precision == "fp64" else 1e-6)


def test_slab_256x128x128_mixed_matches_oracle():
    # the DDFT slab workload's grid and profile (LJ slab, rho_l 0.70 over 0.4 L, rho_v 0.02, 1e-3 noise),
    # fused mixed precision against the FP64 oracle
    shape, dx = (256, 128, 128), 0.1
    rho = W.slab(shape, dx, fraction=0.4, width=0.5, rho_l=0.7, rho_v=0.02, noise=1e-3, axis=0)
    ctx = _ctx(shape, dx, precision="mixed")
    assert ctx.schedule == "fused"
    orc = _orc(shape, dx)
    mu = orc.bulk_mu(0.02)
    res, f = ctx.total(rho, mu)
    ref, fr = orc.total(rho, mu)
    e = _rel(f, fr)
    om = abs(res["grand_potential"] - ref[5]) / abs(ref[5])
    print(f"slab 256x128x128 mixed: Omega rel err {om:.2e}, force rel err {e:.2e}", flush=True)
    assert om < 1e-6
    assert e < 2e-5


def test_fp64_256x128_plane_uses_general_schedule():
    # its FP64 plane kernels would need more LDS than a CU has
    assert _ctx((256, 128, 64), 0.1).schedule == "general"
    assert _ctx((256, 128, 64), 0.1, precision="mixed").schedule == "fused"


def test_fused_mixed_gate_64():
    shape, dx = (64, 64, 64), 0.1
    rho = W.droplet(shape, dx, radius=2.0, width=0.5, seed=5)
    orc = _orc(shape, dx)
    mu = orc.bulk_mu(0.02)
    ref, fr = orc.total(rho, mu)
    ctx = _ctx(shape, dx, precision="mixed")
    res, f = ctx.total(rho, mu)
    eo = abs(res["grand_potential"] - ref[5]) / abs(ref[5])
    ef = _rel(f, fr)
    print(f"fused mixed 64^3: Omega rel err {eo:.3e}, force rel err {ef:.3e}", flush=True)
    assert eo < 1e-6
    assert ef < 2e-5


def test_fused_minimizer_trajectory_fp64():
    shape, dx = (64, 64, 64), 0.1
    ctx = _ctx(shape, dx)
    orc = _orc(shape, dx)
    rho0 = W.droplet(shape, dx, radius=1.6, width=0.4, rho_l=0.6, rho_v=0.05, noise=1e-3, seed=3)
    mu = orc.bulk_mu(0.05)
    kw = dict(param="unbounded", rho_max=1.2, homogeneous_boundary=True, max_steps=6)
    rg, resg, eg = ctx.minimize_grand_potential(rho0, mu, fire=dict(force_tolerance=1e-12), **kw)
    ro, reso, eo = orc.minimize_grand_potential(rho0, mu, fire=dict(force_tolerance=1e-12), **kw)
    assert resg["iterations"] == reso["iterations"]
    np.testing.assert_allclose(eg, eo, rtol=1e-10, atol=1e-10)
    assert _rel(rg, ro) < 1e-10

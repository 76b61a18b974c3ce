"""GPU checks of the DDFT adjoint (FP64) through the C ABI.

- Hessian-vector products H(rho) u against central differences of the FP64
  CPU oracle's forces (every FMT model), symmetry <u, H v> = <v, H u>, and on
  a fused-schedule grid against central differences of the GPU's own forces.
- The fixed-dt forward map against the oracle's.
- The adjoint gradient (well parameters, rho0 and V as fields) against
  central finite differences (h = 1e-5, FP64) of the oracle's forward map on
  a small grid and of the GPU's forward map on a fused-schedule grid, and its
  independence of the checkpoint interval.
"""
import numpy as np
import pytest

from classicaldft_amd import workloads as W

pytestmark = pytest.mark.gpu

KT = 0.9


def _esfmt(model):
    return dict(A=1.0, B=-0.5) if model == "esFMT" else {}


def _ctx(shape, dx, model="WhiteBearII", inter=(W.LJ_WCA,), precision="fp64"):
    import classicaldft_amd as cd

    return cd.Context(shape, dx, interactions=list(inter), kT=KT, model=model, precision=precision, **_esfmt(model))


def _orc(shape, dx, model="WhiteBearII", inter=(W.LJ_WCA,)):
    from oracle import oracle as O

    return O.Oracle(shape, dx, interactions=list(inter), kT=KT, model=model, **_esfmt(model))


def _loss(rho, target, dx):
    return float(np.sum((rho - target) ** 2) * dx ** 3)


def test_wells_potential_matches_oracle():
    shape, dx = (24, 20, 16), 0.25
    prm = W.wells(shape, dx, n_wells=6, seed=9)
    got = _ctx(shape, dx).wells_potential(prm)
    want = _orc(shape, dx).wells_potential(prm)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("model", ["Rosenfeld", "RSLT", "WhiteBearI", "WhiteBearII", "esFMT", "none"])
def test_hessian_vector_matches_oracle_differences(model):
    shape, dx = (16, 16, 16), 0.25
    ctx, orc = _ctx(shape, dx, model), _orc(shape, dx, model)
    rho = W.droplet(shape, dx, radius=1.2, width=0.4, rho_l=0.6, rho_v=0.05, noise=1e-3, seed=3)
    rng = np.random.default_rng(1)
    u = 0.05 * rho * rng.standard_normal(shape)
    v = 0.05 * rho * rng.standard_normal(shape)
    mu = orc.bulk_mu(0.05)
    hu = ctx.hessian_vector(rho, u)
    hv = ctx.hessian_vector(rho, v)
    sym = abs(np.sum(u * hv) - np.sum(v * hu)) / (np.linalg.norm(u) * np.linalg.norm(hv))
    h = 1e-5
    fd = (orc.total(rho + h * u, mu)[1] - orc.total(rho - h * u, mu)[1]) / (2 * h)
    err = np.max(np.abs(hu - fd)) / np.max(np.abs(fd))
    print(f"hvp {model}: symmetry {sym:.1e}, vs oracle differences {err:.1e}", flush=True)
    assert sym < 1e-12
    assert err < 1e-7


def test_hessian_vector_on_fused_grid_matches_gpu_differences():
    shape, dx = (64, 64, 64), 0.1
    ctx = _ctx(shape, dx)
    rho = W.droplet(shape, dx, radius=2.0, width=0.5, rho_l=0.7, rho_v=0.02, noise=1e-3, seed=5)
    u = 0.05 * rho * np.random.default_rng(2).standard_normal(shape)
    mu = ctx.bulk_mu(0.02)
    h = 1e-5
    fd = (ctx.total(rho + h * u, mu)[1] - ctx.total(rho - h * u, mu)[1]) / (2 * h)
    hu = ctx.hessian_vector(rho, u)
    err = np.max(np.abs(hu - fd)) / np.max(np.abs(fd))
    print(f"hvp 64^3 vs fused-schedule differences {err:.1e}", flush=True)
    assert err < 1e-7


def test_fixed_run_matches_oracle():
    shape, dx = (16, 16, 16), 0.25
    ctx, orc = _ctx(shape, dx), _orc(shape, dx)
    rho0, prm, _ = W.adjoint_problem(shape, dx, n_wells=4, rho_bulk=0.3, seed=4)
    V = orc.wells_potential(prm)
    mu = orc.bulk_mu(0.3)
    got, mg = ctx.ddft_fixed_run(rho0, mu, 3, 2e-3, 2, vext=V)
    want, mo = orc.ddft_fixed_run(rho0, mu, 3, 2e-3, 2, vext=V)
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    print(f"fixed DDFT map vs oracle {err:.1e}", flush=True)
    assert err < 1e-11
    assert mg == pytest.approx(mo, rel=1e-11)


def _oracle_problem(shape, dx, n_wells, K, P, dt, seed=4):
    orc = _orc(shape, dx)
    rho0, prm, tprm = W.adjoint_problem(shape, dx, n_wells=n_wells, rho_bulk=0.3, seed=seed)
    mu = orc.bulk_mu(0.3)
    target, _ = orc.ddft_fixed_run(rho0, mu, K, dt, P, vext=orc.wells_potential(tprm))
    return orc, rho0, prm, target, mu


def test_adjoint_gradient_matches_oracle_differences():
    shape, dx, K, P, dt = (12, 12, 12), 0.25, 3, 2, 2e-3
    orc, rho0, prm, target, mu = _oracle_problem(shape, dx, 3, K, P, dt)
    ctx = _ctx(shape, dx)
    grad, res = ctx.ddft_adjoint_wells(rho0, target, prm, mu, K, dt, P, checkpoint_interval=2)

    def oracle_loss(p):
        rho, _ = orc.ddft_fixed_run(rho0, mu, K, dt, P, vext=orc.wells_potential(p))
        return _loss(rho, target, dx)

    fd = np.zeros_like(prm)
    h = 1e-5
    for j in range(prm.shape[0]):
        for k in range(5):
            pp, pm = prm.copy(), prm.copy()
            pp[j, k] += h
            pm[j, k] -= h
            fd[j, k] = (oracle_loss(pp) - oracle_loss(pm)) / (2 * h)
    err = np.max(np.abs(grad - fd)) / np.max(np.abs(fd))
    print(f"adjoint vs oracle differences: {err:.1e} (loss {res['loss']:.3e}, {res['checkpoints']} checkpoints)",
          flush=True)
    assert res["loss"] == pytest.approx(oracle_loss(prm), rel=1e-10)
    assert res["checkpoints"] == 2
    assert res["hessian_products"] == K * P
    assert err < 1e-6


def test_adjoint_independent_of_checkpoint_interval():
    shape, dx, K, P, dt = (16, 16, 16), 0.25, 5, 2, 2e-3
    _, rho0, prm, target, mu = _oracle_problem(shape, dx, 4, K, P, dt)
    ctx = _ctx(shape, dx)
    grads = []
    for C in (1, 2, 5, 0):
        g, res = ctx.ddft_adjoint_wells(rho0, target, prm, mu, K, dt, P, checkpoint_interval=C)
        assert res["checkpoints"] == (K + (C or 3) - 1) // (C or 3)
        grads.append(g)
    for g in grads[1:]:
        np.testing.assert_allclose(g, grads[0], rtol=1e-12, atol=1e-18)


def test_adjoint_field_gradients_match_differences():
    shape, dx, K, P, dt = (16, 16, 16), 0.25, 3, 2, 2e-3
    _, rho0, prm, target, mu = _oracle_problem(shape, dx, 4, K, P, dt)
    ctx = _ctx(shape, dx)
    grad, res, (g_rho0, g_vext, rho_K) = ctx.ddft_adjoint_wells(rho0, target, prm, mu, K, dt, P, want_fields=True)
    V = ctx.wells_potential(prm)
    fwd, _ = ctx.ddft_fixed_run(rho0, mu, K, dt, P, vext=V)
    np.testing.assert_array_equal(rho_K, fwd)
    assert res["loss"] == pytest.approx(_loss(fwd, target, dx), rel=1e-12)
    rng = np.random.default_rng(7)
    h = 1e-6
    u = rho0 * rng.standard_normal(shape)
    lp = _loss(ctx.ddft_fixed_run(rho0 + h * u, mu, K, dt, P, vext=V)[0], target, dx)
    lm = _loss(ctx.ddft_fixed_run(rho0 - h * u, mu, K, dt, P, vext=V)[0], target, dx)
    d_rho = (lp - lm) / (2 * h)
    w = rng.standard_normal(shape)
    lp = _loss(ctx.ddft_fixed_run(rho0, mu, K, dt, P, vext=V + h * w)[0], target, dx)
    lm = _loss(ctx.ddft_fixed_run(rho0, mu, K, dt, P, vext=V - h * w)[0], target, dx)
    d_v = (lp - lm) / (2 * h)
    print(f"dL/drho0.u {np.sum(g_rho0 * u):.6e} vs {d_rho:.6e}; dL/dV.w {np.sum(g_vext * w):.6e} vs {d_v:.6e}",
          flush=True)
    assert np.sum(g_rho0 * u) == pytest.approx(d_rho, rel=1e-6)
    assert np.sum(g_vext * w) == pytest.approx(d_v, rel=1e-6)
    # the well gradient is the field gradient projected on dV/dtheta: dL/dA_j = sum_r dL/dV(r) G_j(r)
    G0 = ctx.wells_potential(np.r_[[1.0], prm[0, 1:]].reshape(1, 5))
    assert grad[0, 0] == pytest.approx(np.sum(g_vext * G0), rel=1e-12)


def test_adjoint_gradient_on_fused_grid_matches_gpu_differences():
    shape, dx, K, P, dt = (64, 64, 64), 0.1, 4, 2, 5e-4
    ctx = _ctx(shape, dx)
    rho0, prm, tprm = W.adjoint_problem(shape, dx, n_wells=4, rho_bulk=0.3, seed=12)
    mu = ctx.bulk_mu(0.3)
    target, _ = ctx.ddft_fixed_run(rho0, mu, K, dt, P, vext=ctx.wells_potential(tprm))
    grad, res = ctx.ddft_adjoint_wells(rho0, target, prm, mu, K, dt, P)

    def gpu_loss(p):
        return _loss(ctx.ddft_fixed_run(rho0, mu, K, dt, P, vext=ctx.wells_potential(p))[0], target, dx)

    fd = np.zeros_like(prm)
    h = 1e-5
    for j in range(prm.shape[0]):
        for k in range(5):
            pp, pm = prm.copy(), prm.copy()
            pp[j, k] += h
            pm[j, k] -= h
            fd[j, k] = (gpu_loss(pp) - gpu_loss(pm)) / (2 * h)
    err = np.max(np.abs(grad - fd)) / np.max(np.abs(fd))
    print(f"adjoint 64^3 vs GPU differences: {err:.1e} (loss {res['loss']:.3e})", flush=True)
    assert err < 1e-6


def test_adjoint_rejects_mixed_precision_and_bad_config():
    import classicaldft_amd as cd

    shape, dx = (16, 16, 16), 0.25
    rho0, prm, _ = W.adjoint_problem(shape, dx, n_wells=2, rho_bulk=0.3, seed=4)
    mixed = _ctx(shape, dx, precision="mixed")
    with pytest.raises(cd.CdftError) as ei:
        mixed.ddft_adjoint_wells(rho0, rho0, prm, 0.0, 2, 1e-3, 2)
    assert ei.value.code == -4
    ctx = _ctx(shape, dx)
    with pytest.raises(cd.InvalidArgument):
        ctx.ddft_adjoint_wells(rho0, rho0, prm, 0.0, 0, 1e-3, 2)
    with pytest.raises(cd.InvalidArgument):
        ctx.ddft_fixed_run(rho0, 0.0, 2, 1e-3, 0)
    bad = prm.copy()
    bad[0, 1] = 0.0
    with pytest.raises(cd.InvalidArgument):
        ctx.wells_potential(bad)

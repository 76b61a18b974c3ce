"""GPU checks of the general-schedule autotuner through the C ABI: tuning
changes no result beyond rounding (FP64 forces 1e-11 relative to the oracle
after tuning; mixed Omega 1e-6), a second context of the same grid takes the
cached choice, and a fused-schedule context is left as it is."""
import numpy as np
import pytest

from classicaldft_amd import workloads as W

pytestmark = pytest.mark.gpu


def _ctx(shape, precision):
    import classicaldft_amd as cd

    return cd.Context(shape, 0.1, interactions=[W.LJ_WCA], kT=0.8, model="WhiteBearII", precision=precision)


@pytest.mark.parametrize("shape,precision", [((48, 48, 48), "fp64"), ((40, 48, 36), "mixed"), ((80, 80, 80), "mixed")])
def test_autotune_keeps_results(shape, precision):
    rho = W.droplet(shape, 0.1, radius=1.5, width=0.5, rho_l=0.7, rho_v=0.02, noise=1e-3, seed=8)
    ctx = _ctx(shape, precision)
    assert ctx.schedule == "general"
    mu = ctx.bulk_mu(0.02)
    r0, f0 = ctx.total(rho, mu)
    t_def, t_tuned = ctx.autotune(reps=2)
    r1, f1 = ctx.total(rho, mu)
    e = float(np.max(np.abs(f1 - f0)) / np.max(np.abs(f0)))
    print(f"autotune {shape} {precision}: {t_def:.1f} -> {t_tuned:.1f} us per evaluation; force change {e:.1e}",
          flush=True)
    assert t_tuned <= t_def * 1.02
    if precision == "fp64":
        from oracle import oracle as O

        orc = O.Oracle(shape, 0.1, interactions=[W.LJ_WCA], kT=0.8, model="WhiteBearII")
        ref, fr = orc.total(rho, mu)
        assert float(np.max(np.abs(f1 - fr)) / np.max(np.abs(fr))) < 1e-11
        assert r1["grand_potential"] == pytest.approx(ref[5], rel=1e-11)
    else:
        assert e < 2e-5
        assert r1["grand_potential"] == pytest.approx(r0["grand_potential"], rel=1e-6)
    # a second context of the grid takes the cached choice
    ctx2 = _ctx(shape, precision)
    t_def2, t_tuned2 = ctx2.autotune(reps=2)
    r2, f2 = ctx2.total(rho, mu)
    np.testing.assert_array_equal(f2, f1)


def test_autotune_leaves_fused_context():
    ctx = _ctx((64, 64, 64), "mixed")
    assert ctx.schedule == "fused"
    t_def, t_tuned = ctx.autotune(reps=2)
    assert t_def == t_tuned

"""GPU parity of multi-species evaluation (functionals::total for a mixture,
functionals.hpp:64-95 with hard_sphere.hpp:174-254 and mean_field.hpp:233-281)
against the FP64 CPU oracle, through the C ABI (cdft_total_species).

Tolerances: energies 1e-12 relative (1e-14 absolute), forces 1e-11 relative
to max|force| per species; mixed precision (FP32 real-space fields, fused
grids): Omega and its parts 1e-6, forces 2e-5.
"""
import numpy as np
import pytest

from classicaldft_amd import workloads as W

pytestmark = pytest.mark.gpu

LJ = W.LJ_WCA
PARTS = ("ideal", "external", "hard_sphere", "mean_field", "free_energy", "grand_potential")


def _pair(shape, diams, inter, model, kT=1.0, **kw):
    import classicaldft_amd as cd
    from oracle import oracle as O

    dx = kw.pop("dx")
    ctx = cd.Context(shape, dx, diameters=diams, interactions=list(inter), kT=kT, model=model, **kw)
    orc = O.Oracle(shape, dx, diameters=diams, interactions=list(inter), kT=kT, model=model,
                   A=kw.get("A", 1.0), B=kw.get("B", 0.0))
    return ctx, orc


def _densities(shape, dx, rho0s, seed):
    return np.stack([W.smooth_noise_fluid(shape, dx, rho0=r, amp=0.08 * r, seed=seed + s)
                     for s, r in enumerate(rho0s)])


def _compare(ctx, orc, rho, mus, vext=None, tag=""):
    res, f = ctx.total_species(rho, mus, vext=vext)
    parts, fo = orc.total(rho, mus, vext=vext)
    S = rho.shape[0]
    print(f"{tag}: Omega gpu {res['grand_potential']:.15e} oracle {parts[5]:.15e}", flush=True)
    for i, k in enumerate(PARTS):
        assert res[k] == pytest.approx(parts[i], rel=1e-12, abs=1e-14), k
    assert res["mass"] == pytest.approx(float(np.sum(rho)) * ctx.dx ** 3, rel=1e-13)
    for s in range(S):
        e = float(np.max(np.abs(f[s] - fo[s])) / np.max(np.abs(fo[s])))
        assert e < 1e-11, (s, e)


@pytest.mark.parametrize("model", ["Rosenfeld", "RSLT", "WhiteBearI", "WhiteBearII", "esFMT"])
def test_binary_mixture_matches_oracle(model):
    shape, dx = (32, 32, 32), 0.15
    inter = [dict(LJ, si=0, sj=0), dict(LJ, si=1, sj=1, sigma=0.8, epsilon=0.6),
             dict(LJ, si=0, sj=1, sigma=0.9, epsilon=0.8)]
    ctx, orc = _pair(shape, (1.0, 0.8), inter, model, kT=1.1, dx=dx, A=1.2, B=-0.7)
    assert ctx.n_species == 2 and ctx.schedule == "general"
    for k in range(3):
        assert ctx.a_vdw(k) == pytest.approx(orc.a_vdw(k), rel=1e-12)
    rho = _densities(shape, dx, (0.3, 0.2), seed=11)
    _compare(ctx, orc, rho, (0.4, -0.3), tag=f"binary {model}")


def test_ternary_mixture_with_external_field_on_a_ragged_grid():
    # 2,3,5-smooth extents exercise the radix-3/5 transforms; one cross pair only
    shape, dx = (30, 24, 20), 0.2
    inter = [dict(LJ, si=2, sj=0, epsilon=0.5)]
    ctx, orc = _pair(shape, (1.0, 0.7, 1.3), inter, "WhiteBearII", kT=0.9, dx=dx)
    rho = _densities(shape, dx, (0.2, 0.15, 0.1), seed=21)
    z = (np.arange(shape[2]) + 0.5) * dx
    vext = np.stack([np.broadcast_to(a * np.cos(2 * np.pi * z / (shape[2] * dx)), shape)
                     for a in (0.5, -0.3, 0.2)])
    _compare(ctx, orc, rho, (0.1, 0.0, -0.5), vext=vext, tag="ternary")


def test_mixture_with_a_vacuum_region_and_ideal_only():
    # eta < 1e-30 points are skipped (hard_sphere.hpp:199); ideal + mean field without FMT
    shape, dx = (24, 24, 24), 0.2
    rho = _densities(shape, dx, (0.3, 0.2), seed=31)
    rho[:, :, :, :6] = 0.0
    ctx, orc = _pair(shape, (1.0, 0.8), [dict(LJ, si=0, sj=1)], "WhiteBearII", dx=dx)
    _compare(ctx, orc, rho, (0.0, 0.0), tag="vacuum")
    ctx2, orc2 = _pair(shape, (1.0, 0.8), [dict(LJ, si=0, sj=1)], "none", dx=dx)
    _compare(ctx2, orc2, np.maximum(rho, 1e-6), (0.0, 0.0), tag="ideal+mf")


def test_uniform_binary_matches_reference_mixture_test():
    # hard_sphere.cpp:208-248: uniform forces that differ between species
    shape, dx = (16, 16, 16), 0.2
    ctx, orc = _pair(shape, (1.0, 0.8), [], "Rosenfeld", dx=dx)
    rho = np.stack([np.full(shape, 0.3), np.full(shape, 0.2)])
    res, f = ctx.total_species(rho, (0.0, 0.0))
    assert res["hard_sphere"] > 0.0
    for s in range(2):
        assert np.max(np.abs(f[s] - f[s].mean())) < 1e-8 * abs(f[s].mean())
    assert abs(f[0].flat[0] - f[1].flat[0]) > 1e-10
    _compare(ctx, orc, rho, (0.0, 0.0), tag="uniform")


def test_one_species_context_through_species_entry_matches_total():
    import classicaldft_amd as cd

    shape, dx = (32, 32, 32), 0.15
    ctx = cd.Context(shape, dx, interactions=[LJ], kT=0.8, model="WhiteBearII")
    rho = W.smooth_noise_fluid(shape, dx, rho0=0.4, amp=0.03, seed=5)
    r1, f1 = ctx.total(rho, 0.2)
    r2, f2 = ctx.total_species(rho[None], (0.2,))
    assert r1 == r2
    assert np.array_equal(f1, f2[0])


def test_mixture_contexts_refuse_one_species_entry_points_and_bad_input():
    import classicaldft_amd as cd

    shape, dx = (16, 16, 16), 0.2
    ctx = cd.Context(shape, dx, diameters=(1.0, 0.8), model="WhiteBearII")
    rho = np.full(shape, 0.3)
    with pytest.raises(cd.CdftError) as e:
        ctx.total(rho, 0.0)
    assert e.value.code == -4  # CDFT_EUNSUPPORTED
    with pytest.raises(cd.CdftError) as e:
        cd.Context(shape, dx, diameters=(1.0, 0.8), model="WhiteBearII", precision="mixed")
    assert e.value.code == -4
    with pytest.raises(cd.InvalidArgument):
        cd.Context(shape, dx, diameters=(1.0, 0.8), interactions=[dict(LJ, si=0, sj=2)])
    with pytest.raises(cd.InvalidArgument):
        cd.Context(shape, dx, diameters=(1.0, -0.8))
    with pytest.raises(ValueError):
        ctx.total_species(rho, (0.0, 0.0))  # one species' worth of values
    with pytest.raises(cd.InvalidArgument):
        ctx.total_species(np.stack([rho, rho]), (0.0, float("nan")))


# ---------------------------------------------------------------------------
# Bulk thermodynamics (bulk/thermodynamics.hpp:17-255) for any species count
# ---------------------------------------------------------------------------

def test_bulk_one_species_matches_bulk_mu():
    import classicaldft_amd as cd

    for model in ("WhiteBearII", "Rosenfeld", "RSLT", "none"):
        ctx = cd.Context((16, 16, 16), 0.2, interactions=[LJ], kT=0.9, model=model)
        for rho in (0.05, 0.4, 0.8):
            assert ctx.bulk_chemical_potentials(rho)[0] == pytest.approx(ctx.bulk_mu(rho), rel=1e-13, abs=1e-14)


def test_bulk_binary_rosenfeld_matches_scaled_particle_theory():
    import classicaldft_amd as cd
    from closed_forms import rosenfeld_mixture

    diams, rhos = (1.0, 0.8), np.array([0.3, 0.2])
    ctx = cd.Context((16, 16, 16), 0.2, diameters=diams, model="Rosenfeld")
    phi, mu_ex = rosenfeld_mixture(rhos, diams)
    f_id = float(np.sum(rhos * (np.log(rhos) - 1.0)))
    assert ctx.bulk_free_energy_density(rhos) == pytest.approx(f_id + phi, rel=1e-13)
    assert np.allclose(ctx.bulk_chemical_potentials(rhos), np.log(rhos) + mu_ex, rtol=1e-13, atol=0)
    mu = ctx.bulk_chemical_potentials(rhos)
    assert ctx.bulk_pressure(rhos) == pytest.approx(float(rhos @ mu) - f_id - phi, rel=1e-13)


@pytest.mark.parametrize("model", ["WhiteBearII", "esFMT", "RSLT"])
def test_bulk_chemical_potentials_are_derivatives_with_self_pairs(model):
    # with self pairs only, the reference's free energy and chemical potentials are consistent
    import classicaldft_amd as cd

    inter = [dict(LJ, si=0, sj=0), dict(LJ, si=1, sj=1, sigma=0.8, epsilon=0.6)]
    ctx = cd.Context((16, 16, 16), 0.2, diameters=(1.0, 0.7), interactions=inter, kT=1.1, model=model,
                     A=1.3, B=-0.4)
    rhos = np.array([0.25, 0.3])
    mu = ctx.bulk_chemical_potentials(rhos)
    h = 1e-6
    for s in range(2):
        up, dn = rhos.copy(), rhos.copy()
        up[s] += h
        dn[s] -= h
        fd = (ctx.bulk_free_energy_density(up) - ctx.bulk_free_energy_density(dn)) / (2 * h)
        assert fd == pytest.approx(mu[s], rel=1e-8)


def test_bulk_matches_the_functional_on_a_uniform_mixture():
    # uniform state: the functional's force / dV is mu_bulk - mu; its F / V equals the bulk
    # free-energy density except for the reference's extra half of every cross pair
    import classicaldft_amd as cd
    from oracle import oracle as O

    shape, dx = (16, 16, 16), 0.2
    inter = [dict(LJ, si=0, sj=0), dict(LJ, si=0, sj=1, sigma=0.9, epsilon=0.8)]
    ctx = cd.Context(shape, dx, diameters=(1.0, 0.8), interactions=inter, kT=1.0, model="WhiteBearII")
    rhos = np.array([0.3, 0.2])
    mus = (0.1, -0.4)
    res, f = ctx.total_species(np.stack([np.full(shape, r) for r in rhos]), mus)
    mu_b = ctx.bulk_chemical_potentials(rhos)
    dv, vol = dx ** 3, np.prod(shape) * dx ** 3
    for s in range(2):
        assert np.max(np.abs(f[s] / dv - (mu_b[s] - mus[s]))) < 1e-11
    a01 = ctx.a_vdw(1)
    assert ctx.bulk_free_energy_density(rhos) == pytest.approx(
        res["free_energy"] / vol + 0.5 * a01 * rhos[0] * rhos[1], rel=1e-12)
    # and the a_vdw the bulk uses is the oracle's
    orc = O.Oracle(shape, dx, diameters=(1.0, 0.8), interactions=inter, kT=1.0, model="WhiteBearII")
    assert a01 == pytest.approx(orc.a_vdw(1), rel=1e-12)


def test_bulk_rejects_bad_densities():
    import classicaldft_amd as cd

    ctx = cd.Context((16, 16, 16), 0.2, diameters=(1.0, 0.8), model="WhiteBearII")
    with pytest.raises(ValueError):
        ctx.bulk_chemical_potentials([0.3])
    with pytest.raises(cd.InvalidArgument):
        ctx.bulk_chemical_potentials([0.3, 0.0])


def test_eight_species_with_every_pair_matches_oracle():
    # the capacity of a context: 8 species and all 36 pairs (self pairs included)
    import classicaldft_amd as cd

    shape, dx = (16, 16, 16), 0.25
    diams = tuple(0.6 + 0.1 * s for s in range(8))
    inter = [dict(LJ, si=i, sj=j, sigma=0.5 * (diams[i] + diams[j]), epsilon=0.3 + 0.05 * (i + j))
             for i in range(8) for j in range(i, 8)]
    assert len(inter) == 36
    ctx, orc = _pair(shape, diams, inter, "WhiteBearI", kT=1.5, dx=dx)
    rho = _densities(shape, dx, tuple(0.02 + 0.01 * s for s in range(8)), seed=41)
    _compare(ctx, orc, rho, tuple(0.1 * s - 0.3 for s in range(8)), tag="8 species")
    with pytest.raises(cd.CdftError) as e:
        cd.Context(shape, dx, diameters=(1.0,) * 9)
    assert e.value.code == -4
    with pytest.raises(cd.CdftError) as e:
        cd.Context(shape, dx, diameters=diams, interactions=inter + [dict(LJ, si=0, sj=1)])
    assert e.value.code == -4


@pytest.mark.parametrize("model,diams", [("WhiteBearII", (1.0,)), ("esFMT", (1.0, 0.7, 1.2)), ("none", (1.0, 0.8))])
def test_bulk_batch_on_the_device_matches_the_host_calls(model, diams):
    import classicaldft_amd as cd

    S = len(diams)
    inter = [dict(LJ, si=i, sj=j, epsilon=0.5 + 0.1 * (i + j)) for i in range(S) for j in range(i, S)]
    ctx = cd.Context((16, 16, 16), 0.2, diameters=diams, interactions=inter, kT=1.1, model=model, A=1.2, B=-0.5)
    rng = np.random.default_rng(3)
    rho = rng.uniform(0.01, 0.7 / S, size=(257, S))
    f, mu, p = ctx.bulk_batch(rho)
    for i in range(0, 257, 16):
        assert f[i] == pytest.approx(ctx.bulk_free_energy_density(rho[i]), rel=1e-13, abs=1e-15)
        assert np.allclose(mu[i], ctx.bulk_chemical_potentials(rho[i]), rtol=1e-13, atol=1e-15)
        assert p[i] == pytest.approx(ctx.bulk_pressure(rho[i]), rel=1e-12, abs=1e-14)
    if S == 1:
        assert np.allclose(mu[:, 0], [ctx.bulk_mu(r) for r in rho[:, 0]], rtol=1e-13, atol=1e-15)
    with pytest.raises(cd.InvalidArgument):
        ctx.bulk_batch(np.zeros((2, S)))


def test_binary_mixture_at_64cubed_matches_oracle():
    # a grid the fused kernels cover: the mixture's transforms run on them (spectra in the sub-plane order)
    shape, dx = (64, 64, 64), 0.1
    inter = [dict(LJ, si=0, sj=0), dict(LJ, si=1, sj=1, sigma=0.8, epsilon=0.6),
             dict(LJ, si=0, sj=1, sigma=0.9, epsilon=0.8)]
    ctx, orc = _pair(shape, (1.0, 0.8), inter, "WhiteBearII", kT=1.0, dx=dx)
    assert ctx.schedule == "fused"
    rho = np.stack([W.droplet(shape, dx, radius=1.5, width=0.4, rho_l=0.5, rho_v=0.03, noise=1e-3, seed=61),
                    W.droplet(shape, dx, radius=1.2, width=0.4, rho_l=0.2, rho_v=0.05, noise=1e-3, seed=62)])
    _compare(ctx, orc, rho, (-1.5, -1.2), tag="binary 64^3")


@pytest.mark.parametrize("model", ["Rosenfeld", "RSLT", "WhiteBearI", "esFMT"])
def test_fused_binary_mixture_models_match_oracle(model):
    shape, dx = (64, 64, 64), 0.12
    inter = [dict(LJ, si=0, sj=0), dict(LJ, si=0, sj=1, sigma=0.9, epsilon=0.8)]
    ctx, orc = _pair(shape, (1.0, 0.8), inter, model, kT=1.1, dx=dx, A=1.2, B=-0.7)
    assert ctx.schedule == "fused"
    rho = _densities(shape, dx, (0.3, 0.2), seed=81)
    _compare(ctx, orc, rho, (0.4, -0.3), tag=f"fused binary {model}")


@pytest.mark.parametrize("shape", [(64, 64, 128), (256, 64, 64), (128, 128, 64)])
def test_fused_ternary_mixture_with_external_field_matches_oracle(shape):
    # every plane shape class of the fused field transforms: x 64 / 128 / 256, y 64 / 128, z 64 / 128
    dx = 0.15
    inter = [dict(LJ, si=0, sj=0), dict(LJ, si=1, sj=2, sigma=0.9, epsilon=0.7)]
    ctx, orc = _pair(shape, (1.0, 0.7, 1.3), inter, "WhiteBearII", kT=0.9, dx=dx)
    assert ctx.schedule == "fused"
    rho = _densities(shape, dx, (0.2, 0.15, 0.1), seed=91)
    z = (np.arange(shape[2]) + 0.5) * dx
    vext = np.stack([np.broadcast_to(a * np.cos(2 * np.pi * z / (shape[2] * dx)), shape)
                     for a in (0.3, -0.2, 0.1)])
    _compare(ctx, orc, rho, (0.1, -0.2, 0.3), vext=vext, tag=f"fused ternary {shape}")


def test_fused_mixture_matches_its_general_schedule():
    # the same mixture on the fused and on the general transforms: equal to rounding
    import classicaldft_amd as cd

    shape, dx = (128, 64, 64), 0.1
    inter = [dict(LJ, si=0, sj=0), dict(LJ, si=1, sj=1, sigma=0.8, epsilon=0.6),
             dict(LJ, si=0, sj=1, sigma=0.9, epsilon=0.8)]
    kw = dict(diameters=(1.0, 0.8), interactions=inter, kT=1.0, model="WhiteBearII")
    fused = cd.Context(shape, dx, **kw)
    general = cd.Context(shape, dx, schedule="general", **kw)
    assert fused.schedule == "fused" and general.schedule == "general"
    rho = _densities(shape, dx, (0.3, 0.2), seed=101)
    rf, ff = fused.total_species(rho, (0.2, -0.1))
    rg, fg = general.total_species(rho, (0.2, -0.1))
    for k in PARTS:
        assert rf[k] == pytest.approx(rg[k], rel=1e-12, abs=1e-14), k
    for s in range(2):
        assert np.max(np.abs(ff[s] - fg[s])) / np.max(np.abs(fg[s])) < 1e-12


def test_autotuned_mixture_still_matches_oracle():
    # cdft_autotune on a mixture context times the mixture evaluation; results change only by rounding
    shape, dx = (30, 24, 20), 0.2
    inter = [dict(LJ, si=0, sj=1, epsilon=0.6)]
    ctx, orc = _pair(shape, (1.0, 0.7), inter, "WhiteBearII", kT=1.0, dx=dx)
    us0, us1 = ctx.autotune(reps=3)
    print(f"mixture autotune {shape}: {us0:.1f} -> {us1:.1f} us", flush=True)
    assert us1 <= us0 * 1.05
    rho = _densities(shape, dx, (0.25, 0.15), seed=71)
    _compare(ctx, orc, rho, (0.0, -0.2), tag="autotuned mixture")


def test_mixture_contexts_and_calls_return_their_device_memory():
    # free device memory from the HIP runtime itself (hipMemGetInfo), around repeated mixture
    # contexts and calls; the runtime's last-error slot must be clear when the test starts
    import ctypes as C
    import gc

    import classicaldft_amd as cd

    hip = C.CDLL("libamdhip64.so")
    hip.hipGetErrorName.restype = C.c_char_p
    last = hip.hipGetLastError()
    n = C.c_int()
    rc = hip.hipGetDeviceCount(C.byref(n))
    print(f"hip last error at start: {last} ({hip.hipGetErrorName(last).decode()}); device count rc {rc}, "
          f"{n.value}", flush=True)

    def free_bytes():
        assert hip.hipDeviceSynchronize() == 0
        f, t = C.c_size_t(), C.c_size_t()
        assert hip.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value

    shape, dx = (64, 64, 64), 0.1
    inter = [dict(LJ, si=0, sj=1)]
    rho = _densities(shape, dx, (0.3, 0.2), seed=81)

    def cycle():
        ctx = cd.Context(shape, dx, diameters=(1.0, 0.8), interactions=inter, kT=1.0, model="WhiteBearII")
        ctx.total_species(rho, (0.0, 0.0))
        ctx.contributions(rho, (0.0, 0.0))
        ctx.bulk_batch(np.full((1000, 2), 0.2))
        ctx.close()

    cycle()  # first use: code objects and caches
    gc.collect()
    free0 = free_bytes()
    for _ in range(5):
        cycle()
    gc.collect()
    free1 = free_bytes()
    print(f"device memory change over 5 mixture contexts: {(free0 - free1) / 2**20:.1f} MiB", flush=True)
    assert free0 - free1 < 64 * 2**20
    # the slot was read (and so cleared) at the start; what it holds now comes from the cycles above
    after = hip.hipGetLastError()
    print(f"hip last error after the cycles: {after} ({hip.hipGetErrorName(after).decode()})", flush=True)
    assert after == 0, "a mixture context or call left an error in the HIP runtime's last-error slot"
    assert last == 0, "an earlier call in this process left an error in the HIP runtime's last-error slot"


def _compare_mixed(ctx, orc, rho, mus, vext=None, tag=""):
    # mixed-precision gate: Omega and its parts 1e-6 relative, forces 2e-5 relative to max|force| per species
    res, f = ctx.total_species(rho, mus, vext=vext)
    parts, fo = orc.total(rho, mus, vext=vext)
    eo = abs(res["grand_potential"] - parts[5]) / abs(parts[5])
    ef = max(float(np.max(np.abs(f[s] - fo[s])) / np.max(np.abs(fo[s]))) for s in range(rho.shape[0]))
    print(f"{tag}: Omega rel err {eo:.2e}, force rel err {ef:.2e}", flush=True)
    assert eo <= 1e-6 and ef <= 2e-5
    for i, k in enumerate(PARTS):
        assert res[k] == pytest.approx(parts[i], rel=1e-6, abs=1e-6), k


def test_mixed_precision_binary_mixture_matches_oracle():
    # FP32 real-space measure, derivative and potential fields; transforms, spectra and sums in FP64
    shape, dx = (64, 64, 64), 0.1
    inter = [dict(LJ, si=0, sj=0), dict(LJ, si=1, sj=1, sigma=0.8, epsilon=0.6),
             dict(LJ, si=0, sj=1, sigma=0.9, epsilon=0.8)]
    ctx, orc = _pair(shape, (1.0, 0.8), inter, "WhiteBearII", kT=1.0, dx=dx, precision="mixed")
    assert ctx.schedule == "fused"
    rho = np.stack([W.droplet(shape, dx, radius=1.5, width=0.4, rho_l=0.5, rho_v=0.03, noise=1e-3, seed=61),
                    W.droplet(shape, dx, radius=1.2, width=0.4, rho_l=0.2, rho_v=0.05, noise=1e-3, seed=62)])
    _compare_mixed(ctx, orc, rho, (-1.5, -1.2), tag="mixed binary 64^3")


def test_mixed_precision_ternary_mixture_with_external_field_matches_oracle():
    shape, dx = (256, 64, 64), 0.15
    inter = [dict(LJ, si=0, sj=0), dict(LJ, si=1, sj=2, sigma=0.9, epsilon=0.7)]
    ctx, orc = _pair(shape, (1.0, 0.7, 1.3), inter, "esFMT", kT=0.9, dx=dx, A=1.2, B=-0.7, precision="mixed")
    rho = _densities(shape, dx, (0.2, 0.15, 0.1), seed=111)
    z = (np.arange(shape[2]) + 0.5) * dx
    vext = np.stack([np.broadcast_to(a * np.cos(2 * np.pi * z / (shape[2] * dx)), shape)
                     for a in (0.3, -0.2, 0.1)])
    _compare_mixed(ctx, orc, rho, (0.1, -0.2, 0.3), vext=vext, tag="mixed ternary 256x64x64")


def test_mixed_precision_mixture_terms_match_fp64():
    # cdft_contributions reads the FP32 mean-field potentials of a mixed-precision mixture
    import classicaldft_amd as cd

    shape, dx = (64, 64, 64), 0.12
    inter = [dict(LJ, si=0, sj=0), dict(LJ, si=0, sj=1, sigma=0.9, epsilon=0.8)]
    kw = dict(diameters=(1.0, 0.8), interactions=inter, kT=1.1, model="WhiteBearII")
    mixed = cd.Context(shape, dx, precision="mixed", **kw)
    fp64 = cd.Context(shape, dx, **kw)
    rho = _densities(shape, dx, (0.3, 0.2), seed=121)
    rm, tm = mixed.contributions(rho, (0.4, -0.3))
    rd, td = fp64.contributions(rho, (0.4, -0.3))
    assert rm["grand_potential"] == pytest.approx(rd["grand_potential"], rel=1e-6)
    for k in ("ideal", "external", "hard_sphere", "mean_field"):
        scale = max(float(np.max(np.abs(td[k]))), 1e-300)
        assert float(np.max(np.abs(tm[k] - td[k]))) / scale <= 2e-5, k

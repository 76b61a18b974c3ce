"""CPU checks of the Python binding's argument validation: every wrapper that
hands a field to the native library (which reads or writes exactly
n_species * N doubles) rejects an array of any other size with ValueError
before the library is called.  A context shell without a device handle
stands in for a real context, so no GPU is touched."""
import numpy as np
import pytest


def _shell(shape=(4, 4, 4), n_species=1):
    import classicaldft_amd as cd

    ctx = object.__new__(cd.Context)
    ctx._h = None
    ctx.shape = tuple(shape)
    ctx.N = int(np.prod(shape))
    ctx.M = shape[0] * shape[1] * (shape[2] // 2 + 1)
    ctx.nf = 12
    ctx.n_species = n_species
    return ctx


GOOD = np.full((4, 4, 4), 0.3)
BAD = np.full(63, 0.3)

CALLS = [
    ("total", lambda c, a: c.total(a, 0.0)),
    ("total vext", lambda c, a: c.total(GOOD, 0.0, vext=a)),
    ("contributions", lambda c, a: c.contributions(a, 0.0)),
    ("contributions vext", lambda c, a: c.contributions(GOOD, 0.0, vext=a)),
    ("weighted_densities", lambda c, a: c.weighted_densities(a)),
    ("r2c", lambda c, a: c.r2c(a)),
    ("minimize_grand_potential", lambda c, a: c.minimize_grand_potential(a, 0.0)),
    ("minimize_grand_potential vext", lambda c, a: c.minimize_grand_potential(GOOD, 0.0, vext=a)),
    ("minimize_fixed_mass", lambda c, a: c.minimize_fixed_mass(a, 0.0, 1.0)),
    ("ddft_run", lambda c, a: c.ddft_run(a, 0.0, 1)),
    ("ddft_run vext", lambda c, a: c.ddft_run(GOOD, 0.0, 1, vext=a)),
    ("smallest_eigenpair", lambda c, a: c.smallest_eigenpair(a, 0.0)),
    ("smallest_eigenpair guess", lambda c, a: c.smallest_eigenpair(GOOD, 0.0, initial_guess=a)),
    ("saddle_search", lambda c, a: c.saddle_search(a, 0.0, 1.0)),
    ("ddft_split_operator", lambda c, a: c.ddft_split_operator(a, 0.0, 1)),
    ("ddft_crank_nicolson vext", lambda c, a: c.ddft_crank_nicolson(GOOD, 0.0, 1, vext=a)),
    ("picard", lambda c, a: c.picard(a, 0.0)),
    ("picard background", lambda c, a: c.picard(GOOD, 0.0, target_mass=1.0, background=a)),
    ("hessian_vector rho", lambda c, a: c.hessian_vector(a, GOOD)),
    ("hessian_vector u", lambda c, a: c.hessian_vector(GOOD, a)),
    ("ddft_fixed_run", lambda c, a: c.ddft_fixed_run(a, 0.0, 1, 1e-4, 2)),
    ("ddft_adjoint_wells target", lambda c, a: c.ddft_adjoint_wells(GOOD, a, np.zeros((1, 5)), 0.0, 1, 1e-4, 2)),
]


@pytest.mark.parametrize("name,call", CALLS, ids=[n for n, _ in CALLS])
def test_wrong_sized_field_is_rejected(name, call):
    with pytest.raises(ValueError, match="expected 1 x 64"):
        call(_shell(), BAD)


def test_wrong_sized_spectrum_is_rejected():
    with pytest.raises(ValueError, match="spectrum"):
        _shell().c2r(np.zeros(10, dtype=complex))


def test_mixture_fields_need_every_species():
    c = _shell(n_species=2)
    with pytest.raises(ValueError, match="expected 2 x 64"):
        c.total_species(GOOD, (0.0, 0.0))
    with pytest.raises(ValueError, match="expected 2 x 64"):
        c.total_species(np.stack([GOOD, GOOD]), (0.0, 0.0), vext=GOOD)

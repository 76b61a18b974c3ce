"""cdft_weighted_densities_noalloc (no device allocation; the library's
workspace) against cdft_weighted_densities on the same context, through the C
ABI: two calls in a row, the second on a smaller grid, each bit-identical to
the allocating entry point (the same kernels in the same order), and a grid
whose fields exceed the workspace refused with an error before any kernel."""
import numpy as np
import pytest

from classicaldft_amd import workloads as W

pytestmark = pytest.mark.gpu


def _ctx(shape, precision):
    import classicaldft_amd as cd

    return cd.Context(shape, 0.1, interactions=[W.LJ_WCA], kT=0.8, model="WhiteBearII", precision=precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_allocating_entry_point(precision):
    for shape, seed in (((64, 64, 64), 3), ((16, 24, 20), 4)):  # the second call is smaller
        ctx = _ctx(shape, precision)
        rho = W.droplet(shape, 0.1, radius=1.2, width=0.4, rho_l=0.7, rho_v=0.02, noise=1e-3, seed=seed)
        print(f"noalloc {precision} {shape}", flush=True)
        a = ctx.weighted_densities(rho, noalloc=True)
        b = ctx.weighted_densities(rho)
        assert a.shape == (12,) + shape
        assert np.array_equal(a, b)
        assert np.all(np.isfinite(a))


def test_noalloc_refuses_fields_beyond_the_workspace():
    import classicaldft_amd as cd

    ctx = _ctx((96, 96, 96), "fp64")  # 12 fields of 96^3 doubles: 85 MB > 64 MiB
    rho = np.full(ctx.shape, 0.3)
    with pytest.raises(cd.InvalidArgument, match="workspace"):
        ctx.weighted_densities(rho, noalloc=True)
    # the context is unaffected, and a fitting grid still works afterwards
    ref = ctx.weighted_densities(rho)
    assert np.all(np.isfinite(ref))
    small = _ctx((16, 16, 16), "fp64")
    r = np.full(small.shape, 0.3)
    assert np.array_equal(small.weighted_densities(r, noalloc=True), small.weighted_densities(r))

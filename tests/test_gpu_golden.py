"""GPU parity against the committed golden fixtures (tests/golden) through the
C ABI, without running the oracle: energies 1e-12 relative, forces and
weighted densities 1e-11 relative to their maximum (FP64); mixed precision
Omega 1e-6, forces 2e-5."""
import numpy as np
import pytest

from golden_cases import load, names

pytestmark = pytest.mark.gpu

PARTS = ("ideal", "external", "hard_sphere", "mean_field", "free_energy", "grand_potential")


def _ctx(s, precision="fp64"):
    import classicaldft_amd as cd

    return cd.Context(tuple(s["shape"]), s["dx"], diameters=tuple(s["diameters"]), interactions=s["interactions"],
                      kT=s["kT"], model=s["model"], precision=precision)


@pytest.mark.parametrize("name", names())
def test_gpu_matches_fixture(name):
    g = load(name)
    s = g["settings"]
    print(f"golden {name}", flush=True)
    ctx = _ctx(s)
    v = g["vext"] if s["has_vext"] else None
    res, f = ctx.total_species(g["rho"], g["mu"], vext=v)
    for i, k in enumerate(PARTS):
        assert res[k] == pytest.approx(g["parts"][i], rel=1e-12, abs=1e-14), k
    assert np.max(np.abs(f - g["forces"])) <= 1e-11 * np.max(np.abs(g["forces"]))
    for k in range(len(g["a_vdw"])):
        assert ctx.a_vdw(k) == pytest.approx(g["a_vdw"][k], rel=1e-12)
    if "weighted_densities" in g:
        wd = ctx.weighted_densities(g["rho"][0])
        ref = g["weighted_densities"][: wd.shape[0]]
        assert np.max(np.abs(wd - ref)) <= 1e-11 * np.max(np.abs(ref))


@pytest.mark.parametrize("name", [n for n in names() if "binary" not in n])
def test_gpu_mixed_precision_matches_fixture(name):
    g = load(name)
    s = g["settings"]
    ctx = _ctx(s, precision="mixed")
    v = g["vext"][0] if s["has_vext"] else None
    res, f = ctx.total(g["rho"][0], g["mu"][0], vext=v)
    assert res["grand_potential"] == pytest.approx(g["parts"][5], rel=1e-6)
    assert np.max(np.abs(f - g["forces"][0])) <= 2e-5 * np.max(np.abs(g["forces"][0]))

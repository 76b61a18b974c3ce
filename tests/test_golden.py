"""The committed golden fixtures (tests/golden, written by make_golden.py from
the FP64 oracle): their internal identities, and the oracle reproducing them
(a regression pin of the oracle)."""
import numpy as np
import pytest

from golden_cases import load, names


def test_fixtures_present():
    assert len(names()) >= 3


@pytest.mark.parametrize("name", names())
def test_fixture_identities(name):
    g = load(name)
    s = g["settings"]
    dv = s["dx"] ** 3
    p = g["parts"]
    assert p[4] == pytest.approx(p[0] + p[1] + p[2] + p[3], rel=1e-14)
    mass = [float(np.sum(r)) * dv for r in g["rho"]]
    assert p[5] == pytest.approx(p[4] - sum(m * n for m, n in zip(g["mu"], mass)), rel=1e-13)
    ideal = sum(float(np.sum(r * (np.log(r) - 1.0))) * dv for r in g["rho"])
    assert p[0] == pytest.approx(ideal, rel=1e-13)
    if s["has_vext"]:
        assert p[1] == pytest.approx(float(np.sum(g["rho"] * g["vext"])) * dv, rel=1e-13)


@pytest.mark.parametrize("name", names())
def test_oracle_reproduces_fixture(name):
    from oracle import oracle as O

    g = load(name)
    s = g["settings"]
    orc = O.Oracle(tuple(s["shape"]), s["dx"], diameters=tuple(s["diameters"]), interactions=s["interactions"],
                   kT=s["kT"], model=s["model"])
    one = len(s["diameters"]) == 1
    v = (g["vext"][0] if one else g["vext"]) if s["has_vext"] else None
    parts, f = orc.total(g["rho"][0] if one else g["rho"], g["mu"][0] if one else g["mu"], vext=v)
    assert np.allclose(parts, g["parts"], rtol=1e-13, atol=0)
    f = np.asarray(f).reshape(g["forces"].shape)
    assert np.max(np.abs(f - g["forces"])) <= 1e-13 * np.max(np.abs(g["forces"]))
    for k in range(len(g["a_vdw"])):
        assert orc.a_vdw(k) == pytest.approx(g["a_vdw"][k], rel=1e-14)
    if one:
        wd = orc.weighted_densities(g["rho"][0])
        assert np.max(np.abs(wd - g["weighted_densities"])) <= 1e-13 * np.max(np.abs(g["weighted_densities"]))

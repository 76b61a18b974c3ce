"""CPU-side checks of the C ABI: the library loads, exports every symbol the
header declares, and rejects invalid arguments before touching a GPU."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "cdft.h")).read()
    return sorted(set(re.findall(r"CDFT_API\s+[\w\s\*]*?\b(cdft_\w+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = _declared()
    for n in ("cdft_ctx_create", "cdft_total", "cdft_minimize_grand_potential", "cdft_ddft_run"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import classicaldft_amd as cd

    L = cd.lib()
    for n in _declared():
        assert hasattr(L, n), n
    assert sorted(cd.EXPORTED) == _declared()


@pytest.mark.parametrize("dx,box", [(0.0, 1.6), (0.1, -1.0), (0.3, 1.0)])
def test_make_grid_validation(dx, box):
    # grid.hpp:87-108: non-positive dx/box or incommensurate box -> std::invalid_argument
    import ctypes as C

    import classicaldft_amd as cd

    g = cd._Grid(dx, (C.c_double * 3)(box, box, box))
    opt = cd._Options(3, 1.0, 0.0, 0, 0)
    d = (C.c_double * 1)(1.0)
    h = C.c_void_p()
    rc = cd.lib().cdft_ctx_create(C.byref(g), d, 1, None, 0, 1.0, C.byref(opt), C.byref(h))
    assert rc == -1
    assert "make_grid" in cd.lib().cdft_last_error().decode()


def test_unknown_model_rejected():
    import ctypes as C

    import classicaldft_amd as cd

    g = cd._Grid(0.1, (C.c_double * 3)(1.6, 1.6, 1.6))
    opt = cd._Options(9, 1.0, 0.0, 0, 0)
    d = (C.c_double * 1)(1.0)
    h = C.c_void_p()
    assert cd.lib().cdft_ctx_create(C.byref(g), d, 1, None, 0, 1.0, C.byref(opt), C.byref(h)) == -1


def _create(diams, n_inter=0, precision=0, inter=None):
    import ctypes as C

    import classicaldft_amd as cd

    g = cd._Grid(0.1, (C.c_double * 3)(1.6, 1.6, 1.6))
    opt = cd._Options(3, 1.0, 0.0, precision, 0)
    d = (C.c_double * max(1, len(diams)))(*diams)
    its = inter
    if its is None and n_inter > 0:
        its = (cd._Interaction * n_inter)()
        for k in range(n_inter):
            its[k] = cd._Interaction(0, 0, 0, 1.0, 1.0, 2.5, 50.0, 0, 3)
    h = C.c_void_p()
    rc = cd.lib().cdft_ctx_create(C.byref(g), d, len(diams), its, n_inter, 1.0, C.byref(opt), C.byref(h))
    return rc, cd.lib().cdft_last_error().decode()


@pytest.mark.parametrize("diams,n_inter,precision,code,needle", [
    ((), 0, 0, -4, "species count"),                  # no species
    ((1.0,) * 9, 0, 0, -4, "species count"),          # beyond the 8 a context holds
    ((1.0, -0.8), 0, 0, -1, "diameter"),              # hard_sphere_diameter must be positive
    ((1.0,), 2, 0, -4, "at most one"),                # one species: at most one interaction
    ((1.0, 0.8), 37, 0, -4, "at most 36"),            # every pair of 8 species is 36
    ((1.0, 0.8), 0, 1, -4, "fused transforms"),       # a mixed-precision mixture needs a fused-schedule grid
])
def test_species_and_interaction_limits(diams, n_inter, precision, code, needle):
    # rejected before any device work (cdft_ctx_create validates first)
    rc, msg = _create(diams, n_inter, precision)
    assert rc == code and needle in msg, (rc, msg)


def test_new_entry_points_reject_a_null_context():
    import ctypes as C

    import classicaldft_amd as cd

    L = cd.lib()
    x = (C.c_double * 4)(0.1, 0.2, 0.3, 0.4)
    r = cd._Result()
    n = C.c_int()
    v = C.c_double()
    assert L.cdft_species_count(None, C.byref(n)) == -1
    assert L.cdft_total_species(None, x, None, x, C.byref(r), None) == -1
    assert L.cdft_contributions(None, x, None, x, C.byref(r), None, None, None, None) == -1
    assert L.cdft_bulk_chemical_potentials(None, x, x) == -1
    assert L.cdft_bulk_free_energy_density(None, x, C.byref(v)) == -1
    assert L.cdft_bulk_pressure(None, x, C.byref(v)) == -1
    assert "null context" in L.cdft_last_error().decode()

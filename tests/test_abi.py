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

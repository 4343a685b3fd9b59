"""CPU checks of the C-ABI library: it loads and exports every symbol include/vacc.h declares.
No compute calls (there is no GPU here)."""
import os
import re

import pytest

from vlasiator_amd import vacc as V

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header="vacc.h"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(vacc_[a-z0-9_]+)\s*\(", txt)))


HEADERS = ("vacc.h", "vacc_vlsv.h", "vacc_trans.h")


def test_header_declares_bound_symbols():
    from vlasiator_amd import trans, vlsv
    assert sorted(V.EXPORTS) == declared_symbols()
    assert sorted(vlsv.EXPORTS) == declared_symbols("vacc_vlsv.h")
    assert sorted(trans.EXPORTS) == declared_symbols("vacc_trans.h")


def test_library_exports_declared_symbols():
    if not os.path.exists(V.LIB_PATH):
        pytest.skip("libvacc.so not built")
    lib = V.lib()
    for h in HEADERS:
        for s in declared_symbols(h):
            assert hasattr(lib, s), s


def test_host_intersections_match_oracle():
    """The product's FP64 transform + intersections (host build of the same code the device
    runs) against the oracle restatement."""
    if not os.path.exists(V.LIB_PATH):
        pytest.skip("libvacc.so not built")
    import numpy as np
    import oracle as O
    from vlasiator_amd import workloads as W
    w = W.config(0, ncells=16)
    mesh = V.make_mesh(w.grid_length, w.vmin, w.vmax)
    sp = V.make_species(w.mass, w.charge, w.min_value, w.add_width)
    op = V.make_options(V.PPM)
    om = O.make_mesh(w.grid_length, w.vmin, w.vmax)
    for c in w.cells:
        for mo in range(3):
            got = V.cell_intersections_host(mesh, sp, op, V.make_fields(c.B, c.V, c.dt), mo)
            T = O.transform(O.make_fields(c.B, c.V), w.mass, w.charge, c.dt, 1.60217653e-19)
            want = O.intersections(om, T, mo)
            scale = np.array([2e6, 2e4, 2e4, 2e4] * 3)
            assert np.all(np.abs(got - want) <= 1e-12 * scale)


def test_max_v_dt_matches_oracle():
    if not os.path.exists(V.LIB_PATH):
        pytest.skip("libvacc.so not built")
    import oracle as O
    sp = V.make_species(1.67262158e-27, 1.60217653e-19)
    op = V.make_options(max_rotation_deg=10.0)
    import ctypes as C
    import numpy as np
    B = np.array([3e-9, -1e-8, 2e-9])
    got = V.lib().vacc_max_v_dt(C.byref(sp), C.byref(op), B.ctypes.data_as(C.c_void_p))
    assert got == O.max_v_dt(B, 1.67262158e-27, 1.60217653e-19, 10.0)

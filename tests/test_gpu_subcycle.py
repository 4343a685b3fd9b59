"""GPU parity of acceleration subcycling (vlasovmover.cpp:398-475, :481-565) against the oracle.

max_rotation_deg = 25 and dt = 2.5 x the smallest max_v_dt of the cells, so the cell with the
strongest field takes 3 subcycles (25, 25 and 12.5 degrees) and the others fewer.  Every
subcycle step is checked: the device intersections against the oracle's transform with the
subcycle dt the reference's rule gives each cell (a wrong dt or a wrong active set moves them
far beyond the tolerance), then the remaps and adjust bit for bit against the oracle driven
by those intersections.  Cells past their last subcycle must come out unchanged.
"""
import math

import numpy as np
import pytest

import oracle as O
from tests import cases
from vlasiator_amd import vacc as V
from vlasiator_amd import workloads as W

pytestmark = pytest.mark.gpu

MAXROT = 25.0
HALL_MIN = 1.60217653e-19


def _subcycle_dt(step, maxdt, dt):
    """vlasovmover.cpp:429-456"""
    d = max(abs(dt) - step * maxdt, 0.0) if (step + 1) * maxdt > abs(dt) else maxdt
    return -d if dt < 0 else d


@pytest.mark.parametrize("method,map_order", [(O.PPM, 1), (O.PQM, 2)])
def test_subcycled_step_matches_oracle(method, map_order):
    w = W.config(0, ncells=6)
    cells = cases.build_cells(w)
    m = cases.oracle_mesh(w)
    maxdt = [O.max_v_dt(c.B, w.mass, w.charge, MAXROT) for c in w.cells]
    dt = 2.5 * min(maxdt)
    nsub = [max(math.ceil(dt / x), 1) for x in maxdt]
    print(f"subcycles method={method} order={map_order} per cell {nsub}", flush=True)
    assert max(nsub) == 3 and min(nsub) < 3
    ctx = V.Context(V.make_mesh(w.grid_length, w.vmin, w.vmax),
                    V.make_species(w.mass, w.charge, w.min_value, w.add_width),
                    V.make_options(method, max_rotation_deg=MAXROT))
    ctx.upload_cells(cells)
    ctx.set_fields([V.make_fields(c.B, c.V, 0.0) for c in w.cells])
    assert ctx.subcycle_count(dt) == 3
    ref = [(g.copy(), f.copy()) for g, f in cells]
    for s in range(3):
        loss = np.zeros(len(cells))
        st = ctx.subcycle_step(s, map_order, use_moments=True, adjust=True, rho_loss=loss)
        assert st == 0
        isx = ctx.get_intersections(len(cells))
        new, want_loss = [], []
        for i, (g, f) in enumerate(ref):
            if s < nsub[i]:
                mo = O.moments(m, g, f)
                piv = tuple(mo[1:] / mo[0])
                T = O.transform(O.make_fields(w.cells[i].B, piv, rhoq=mo[0] * w.charge), w.mass, w.charge,
                                _subcycle_dt(s, maxdt[i], dt), HALL_MIN)
                want = O.intersections(m, T, map_order)
                scale = np.abs(want).max() + 1e-300
                assert np.all(np.abs(isx[i] - want) <= 1e-9 * scale), (s, i)
                g, f, _ = O.accelerate_cell(m, g, f, isx[i], map_order, method, w.min_value)
            g3, f3, l3 = O.adjust(m, g, f, w.min_value, w.add_width)
            new.append((g3, f3))
            want_loss.append(l3)
        ref = new
        cases.assert_cells_equal(ctx.download_cells(), ref, f"subcycle {s}")
        assert np.allclose(loss, want_loss, rtol=1e-12, atol=0)
    # the one-call form from the same start gives the same state
    ctx2 = V.Context(V.make_mesh(w.grid_length, w.vmin, w.vmax),
                     V.make_species(w.mass, w.charge, w.min_value, w.add_width),
                     V.make_options(method, max_rotation_deg=MAXROT))
    ctx2.upload_cells(cells)
    ctx2.set_fields([V.make_fields(c.B, c.V, 0.0) for c in w.cells])
    tot = np.zeros(len(cells))
    st, n = ctx2.accelerate_subcycled(dt, map_order, use_moments=True, rho_loss=tot)
    assert n == 3 and st == 0
    cases.assert_cells_equal(ctx2.download_cells(), ref, "vacc_accelerate_subcycled")
    ctx.close()
    ctx2.close()


def test_subcycle_step_needs_counts_of_the_current_fields():
    w = W.config(0, ncells=2)
    cells = cases.build_cells(w)
    ctx = V.Context(V.make_mesh(w.grid_length, w.vmin, w.vmax),
                    V.make_species(w.mass, w.charge, w.min_value, w.add_width), V.make_options(O.PPM))
    ctx.upload_cells(cells)
    ctx.set_fields([V.make_fields(c.B, c.V, 0.0) for c in w.cells])
    with pytest.raises(V.VaccError, match="subcycle"):
        ctx.subcycle_step(0, 0)
    ctx.subcycle_count(1.0)
    ctx.set_fields([V.make_fields(c.B, c.V, 0.0) for c in w.cells])   # new fields: counts are stale
    with pytest.raises(V.VaccError, match="subcycle"):
        ctx.subcycle_step(0, 0)
    ctx.close()

"""The large two-species workload (100^3-block mesh over [-4e6, 4e6] m/s, protons plus He2+ at
4 % density with 4x the proton mass-to-charge ratio, PQM, adjust every step) on a few cells,
stepped with vacc_accelerate_populations and compared with the oracle.

Each step advances the populations in order; population p pivots on the plasma moments of
both populations as they stand (arch_moments.cpp:406-440: V = sum m nV / sum m n,
RHOQ = sum q n).  The oracle computes that pivot from its own FP64 moments and its own FP64
transform and intersections; the device's intersections must agree within 1e-12 of the
mesh scale (the FP64 sums run in a different order).  The remap and the adjust are then run
in the oracle with the device's intersections, so that f and the block lists compare bit for
bit: FP32 storage directly, BF16 / FP16 storage against the oracle's FP32 remap with every
written block re-quantized (initial state quantized the same way).
"""
import numpy as np
import pytest

import oracle as O
from tests import cases
from vlasiator_amd import vacc as V
from vlasiator_amd import workloads as W

pytestmark = pytest.mark.gpu

HALL_MIN = 1.60217653e-19
NCELLS, STEPS = 6, 3
STORES = {"f32": (V.STORE_F32, O.ST_F32), "bf16": (V.STORE_BF16, O.ST_BF16), "f16": (V.STORE_F16, O.ST_F16)}


def _pivot(m, ws, state, i):
    """Plasma bulk velocity and charge density of cell i from the oracle's current state."""
    rhom, nv, rhoq = 0.0, np.zeros(3), 0.0
    for w, cells in zip(ws, state):
        g, f = cells[i]
        mo = O.moments(m, g, f)
        rhom += mo[0] * w.mass
        nv += mo[1:] * w.mass
        rhoq += mo[0] * w.charge
    return nv / rhom, rhoq


@pytest.mark.parametrize("store", ["f32", "bf16", "f16"])
def test_two_species_steps_match_oracle(store):
    st, ost = STORES[store]
    ws = W.populations(4, ncells=NCELLS)
    assert ws[0].grid_length == (100, 100, 100) and ws[1].mass == 4.0 * ws[0].mass
    m = cases.oracle_mesh(ws[0])
    cells = [cases.build_cells(w) for w in ws]
    if ost != O.ST_F32:
        cells = [[(g, O.requantize(f, ost)) for g, f in cw] for cw in cells]
    ctxs = []
    for w, cw in zip(ws, cells):
        c = V.Context(V.make_mesh(w.grid_length, w.vmin, w.vmax),
                      V.make_species(w.mass, w.charge, w.min_value, w.add_width),
                      V.make_options(w.method, storage=st))
        c.upload_cells(cw)
        c.set_fields([V.make_fields(s.B, s.V, s.dt) for s in w.cells])
        ctxs.append(c)
    # the oracle starts from the device's stored state: a 16-bit load encodes the given
    # (already requantized) values again, and the block scale S_f / S_code of decoded values
    # can move by an ulp, which moves the moment pivot at the 1e-11 level
    state = [c.download_cells() for c in ctxs]
    scale = np.array([4e6, 8e4, 8e4, 8e4] * 3)
    for step in range(STEPS):
        mo = step % 3
        print(f"two species {store} step {step}", flush=True)
        V.accelerate_populations(ctxs, mo, use_moments=True, adjust=True)
        for p, w in enumerate(ws):
            isx = ctxs[p].get_intersections(NCELLS)
            nxt = []
            for i, s in enumerate(w.cells):
                Vp, rq = _pivot(m, ws, state, i)
                T = O.transform(O.make_fields(s.B, Vp, rhoq=rq), w.mass, w.charge, s.dt, HALL_MIN)
                want = O.intersections(m, T, mo)
                assert np.all(np.abs(isx[i] - want) <= 1e-12 * scale), (p, i, isx[i] - want)
                g, f = state[p][i]
                g, f, _ = O.accelerate_cell(m, g, f, isx[i], mo, w.method, w.min_value, storage=ost)
                g, f, _ = O.adjust(m, g, f, w.min_value, w.add_width)
                nxt.append((g, f))
            state[p] = nxt
        for p in range(2):
            cases.assert_cells_equal(ctxs[p].download_cells(), state[p], f"{store} pop {p} step {step}")
    assert sum(g.size for g, _ in state[0]) > 1000 and all(g.size for g, _ in state[1])
    for c in ctxs:
        c.close()

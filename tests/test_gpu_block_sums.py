"""Block sums of the tolerance form (vacc_options.sums_from_f = 0): the last remap pass of a step
stores FP32 sums of each block (sum f and the in-block index sums), and the step's removed mass
and the next moments are formed from them.  Checked against the same quantities summed from f
(a context with sums_from_f = 1 that runs the same steps) and against the oracle's FP64
moments of the stored f."""
import numpy as np
import pytest

import oracle as O
from tests import cases
from vlasiator_amd import vacc as V
from vlasiator_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _ctx(w, storage, sums_from_f, use_graph=False):
    return V.Context(V.make_mesh(w.grid_length, w.vmin, w.vmax),
                     V.make_species(w.mass, w.charge, w.min_value, w.add_width),
                     V.make_options(w.method, storage=storage, fast_arith=True, sums_from_f=sums_from_f,
                                    use_graph=use_graph))


def _vmax(w):
    return float(np.max(np.abs(np.concatenate([np.asarray(w.vmin), np.asarray(w.vmax)]))))


@pytest.mark.parametrize("storage", [V.STORE_F32, V.STORE_BF16])
def test_moments_and_removed_mass_from_block_sums(storage):
    """One step of two contexts that differ only in where the sums come from: equal f and
    block lists (the first step's pivot is taken from f in both); the removed mass within
    1e-5 relative per cell (plus 1e-7 of the cell's n); the moments of the stored state within
    1e-6 of the cell's n (n V: of n times the mesh's largest |v|) of the oracle's FP64 sums."""
    w = W.config(2, ncells=6)
    cells = cases.build_cells(w)
    fields = [V.make_fields(c.B, c.V, c.dt) for c in w.cells]
    m = cases.oracle_mesh(w)
    n = len(cells)
    a, b = _ctx(w, storage, False), _ctx(w, storage, True)
    for c in (a, b):
        c.upload_cells(cells)
        c.set_fields(fields)
    _say(f"block sums storage={storage}")
    la, lb = np.zeros(n), np.zeros(n)
    assert a.step(0, use_moments=True, adjust=True, rho_loss=la) == 0
    assert b.step(0, use_moments=True, adjust=True, rho_loss=lb) == 0
    # moments first: a download compacts the pool, after which the sums no longer apply
    ma, mb = a.moments(n), b.moments(n)
    state = a.download_cells()
    cases.assert_cells_equal(state, b.download_cells(), "block sums: state")
    assert lb.sum() > 0.0
    assert np.all(np.abs(la - lb) <= 1e-5 * np.abs(lb) + 1e-7 * mb[:, 0]), (la, lb)
    # different arithmetic: the FP32 block sums round differently from the FP64 sums of f
    assert not np.array_equal(ma, mb)
    want = np.array([O.moments(m, g, f) for g, f in state])
    tol = 1e-6 * np.abs(want[:, :1]) * np.array([[1.0, _vmax(w), _vmax(w), _vmax(w)]])
    assert np.all(np.abs(mb - want) <= 1e-3 * tol)
    assert np.all(np.abs(ma - want) <= tol), np.max(np.abs(ma - want) / tol)
    a.close()
    b.close()


def test_block_sums_follow_the_state():
    """Sums are used only for the pool they describe: after a lone 1-D remap (which takes no
    sums) and a block adjust that follows it, the moments equal the FP64 sums of f bit for
    bit; after an acceleration, and an adjust that follows it, they come from the sums."""
    w = W.config(2, ncells=6)
    cells = cases.build_cells(w)
    n = len(cells)
    isx = cases.oracle_intersections(w, 0)
    a, b = _ctx(w, V.STORE_F32, False), _ctx(w, V.STORE_F32, True)
    for c in (a, b):
        c.upload_cells(cells)
        c.set_intersections(isx)
    for c in (a, b):
        c.accelerate(0)
    assert not np.array_equal(a.moments(n), b.moments(n))
    cases.assert_cells_equal(a.download_cells(), b.download_cells(), "accelerate")
    for c in (a, b):
        c.map_1d(1)
    assert np.array_equal(a.moments(n), b.moments(n))
    for c in (a, b):
        c.set_intersections(isx)
        c.accelerate(0)
        c.adjust(n)
    assert not np.array_equal(a.moments(n), b.moments(n))   # the adjust keeps the pool
    for c in (a, b):
        c.map_1d(2)
        c.adjust(n)
    assert np.array_equal(a.moments(n), b.moments(n))
    a.close()
    b.close()


def test_block_sums_in_graph_replays():
    """Captured steps with block sums (the sums' validity is part of the captured state):
    replays equal eager steps bit for bit over 8 steps, removed mass and moments included."""
    w = W.config(2, ncells=6)
    cells = cases.build_cells(w)
    fields = [V.make_fields(c.B, c.V, c.dt) for c in w.cells]
    n = len(cells)
    eager, graph = _ctx(w, V.STORE_F32, False), _ctx(w, V.STORE_F32, False, use_graph=True)
    for c in (eager, graph):
        c.upload_cells(cells)
        c.set_fields(fields)
    for step in range(8):
        _say(f"block sums graph step {step}")
        le, lg = np.zeros(n), np.zeros(n)
        eager.step(step % 3, use_moments=True, adjust=True, rho_loss=le)
        graph.step(step % 3, use_moments=True, adjust=True, rho_loss=lg)
        assert np.array_equal(le, lg), f"step {step} loss"
        if step % 3 == 2:
            assert np.array_equal(eager.moments(n), graph.moments(n))
    cases.assert_cells_equal(graph.download_cells(), eager.download_cells(), "graph vs eager")
    eager.close()
    graph.close()


@pytest.mark.parametrize("storage", [V.STORE_F32, V.STORE_BF16])
def test_block_sums_of_wide_and_pass_through_cells(storage):
    """The sums are also stored by the exact kernel's instances that take the units of wide
    deposit spans (22 to 118 cells) and by pass-through copies (span above 118 cells): an
    acceleration in map order Y-Z-X whose last (x) pass has its cross shears scaled so that
    the four cells have spans of about 30, 70, 1 and 150 cells.  Moments and removed mass
    from the sums against those from f, as above."""
    w = W.config(0, ncells=4)
    cells = cases.build_cells(w)
    n = len(cells)
    nx, ny = w.grid_length[0], w.grid_length[1]
    dv = (w.vmax[0] - w.vmin[0]) / (4 * nx)
    isx = cases.oracle_intersections(w, 1)
    spans = []
    for i, factor in enumerate([6.0, 14.0, 1.0, 30.0]):
        g = cells[i][0]
        ic, jc = 4 * (g // (nx * ny)).mean() + 2, 4 * (g // nx % ny).mean() + 2
        I, dk, dj, di = isx[i, 0:4]
        di2, dj2 = di * factor + 2.0 * dv * factor, dj * factor
        isx[i, 0] = I + ic * di + jc * dj - (ic * di2 + jc * dj2)
        isx[i, 2], isx[i, 3] = dj2, di2
        spans.append((dv + 3 * abs(di2) + 3 * abs(dj2)) / dk)
    assert 22 < spans[0] < 118 and 22 < spans[1] < 118 and spans[2] < 22 and spans[3] > 118, spans
    _say(f"block sums, wide and pass-through cells, spans {[round(x, 1) for x in spans]}, storage={storage}")
    a, b = _ctx(w, storage, False), _ctx(w, storage, True)
    for c in (a, b):
        c.upload_cells(cells)
        c.set_intersections(isx)
        c.accelerate(1)
    assert bool(a.status(n)[3] & V.STATUS_DK)
    ma, mb = a.moments(n), b.moments(n)
    assert not np.array_equal(ma, mb)
    tol = 1e-6 * np.abs(mb[:, :1]) * np.array([[1.0, _vmax(w), _vmax(w), _vmax(w)]])
    assert np.all(np.abs(ma - mb) <= tol), np.max(np.abs(ma - mb) / tol)
    la, lb = a.adjust(n), b.adjust(n)
    assert np.all(np.abs(la - lb) <= 1e-5 * np.abs(lb) + 1e-7 * mb[:, 0]), (la, lb)
    cases.assert_cells_equal(a.download_cells(), b.download_cells(), "wide and pass-through")
    a.close()
    b.close()

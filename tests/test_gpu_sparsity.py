"""GPU parity of the sparsity options against the oracle.

Dynamic threshold (<pop>_sparse.dynamicAlgorithm 1 and 2, SpatialCell::updateSparseMinValue,
spatial_cell_cpu.cpp:699-727): before each adjust every cell's threshold is set from its
density (1) or block count (2); that threshold is the adjust's content test and the
reconstruction threshold of the next remaps (cpu_acc_map.cpp:539-547).  Three (accelerate,
adjust) steps are compared bit for bit with the oracle run with the same per-cell thresholds.

sparse_conserve_mass (block_adjust_cpu.cpp:96-113): the kept blocks are scaled by the ratio
of the FP64 sums of f before and after the adjust.  The device sums in another order, so the
ratio can differ from the oracle's in its last bits: block lists must be equal and f within
one FP32 rounding (2^-23 relative).
"""
import numpy as np
import pytest

import oracle as O
from tests import cases
from vlasiator_amd import vacc as V
from vlasiator_amd import workloads as W

pytestmark = pytest.mark.gpu


def _ctx(w, method, **kw):
    return V.Context(V.make_mesh(w.grid_length, w.vmin, w.vmax),
                     V.make_species(w.mass, w.charge, w.min_value, w.add_width, **kw), V.make_options(method))


@pytest.mark.parametrize("alg,bulk,dmin", [(1, (0.9e6, 1.1e6), (5e-16, 5e-15)), (2, (300.0, 900.0), (2e-16, 3e-15))])
def test_dynamic_threshold_bit_exact(alg, bulk, dmin):
    w = W.config(0, ncells=6)
    cells = cases.build_cells(w)
    m = cases.oracle_mesh(w)
    ctx = _ctx(w, O.PPM, dynamic=alg, bulk=bulk, dyn_min=dmin)
    ctx.upload_cells(cells)
    thr = [w.min_value] * len(cells)   # a loaded cell starts from minValue
    ref = [(g.copy(), f.copy()) for g, f in cells]
    seen = set()
    for step in range(3):
        mo = step % 3
        isx = cases.oracle_intersections(w, mo)
        print(f"dynamic threshold alg={alg} step {step} thresholds {['%.3g' % t for t in thr]}", flush=True)
        ctx.set_intersections(isx)
        ctx.accelerate(mo)
        ctx.adjust(len(cells))
        new = []
        for i, (g, f) in enumerate(ref):
            g2, f2, _ = O.accelerate_cell(m, g, f, isx[i], mo, O.PPM, thr[i])
            rho = O.moments(m, g2, f2)[0]
            thr[i] = O.sparse_min_value(alg, w.min_value, bulk[0], bulk[1], dmin[0], dmin[1], rho, len(g2))
            seen.add(thr[i])
            g3, f3, _ = O.adjust(m, g2, f2, thr[i], w.add_width)
            new.append((g3, f3))
        ref = new
        cases.assert_cells_equal(ctx.download_cells(), ref, f"dynamic threshold alg {alg} step {step}")
    assert len(seen) > 2   # the cells got different thresholds
    ctx.close()


@pytest.mark.parametrize("storage", [V.STORE_F32, V.STORE_BF16])
def test_conserve_mass(storage):
    w = W.config(0, ncells=4)
    cells = cases.build_cells(w)
    m = cases.oracle_mesh(w)
    thr = 3e-13   # a high threshold removes blocks with mass
    ctx = V.Context(V.make_mesh(w.grid_length, w.vmin, w.vmax),
                    V.make_species(w.mass, w.charge, thr, w.add_width, conserve_mass=True),
                    V.make_options(O.PPM, storage=storage))
    ctx.upload_cells(cells)
    held = ctx.download_cells()   # the stored (requantized) state
    loss = ctx.adjust(len(cells))
    got = ctx.download_cells()
    print(f"conserve mass storage={storage} removed mass {loss}", flush=True)
    assert np.all(loss > 0)
    for i, (g, f) in enumerate(held):
        g2, f2, _ = O.adjust(m, g, f, thr, w.add_width)
        f3 = O.conserve_mass(f, f2)
        g1, f1 = got[i]
        assert np.array_equal(g1, g2)
        if storage == V.STORE_F32:
            assert np.all(np.abs(f1.astype(np.float64) - f3) <= 2.0 ** -23 * np.abs(f3.astype(np.float64)))
        else:   # the block scale is scaled: within one FP32 rounding of the scale per value
            assert np.all(np.abs(f1.astype(np.float64) - f3) <= 2.0 ** -22 * np.abs(f3.astype(np.float64)) + 1e-45)
        assert abs(f1.astype(np.float64).sum() / f.astype(np.float64).sum() - 1) < 1e-6
    ctx.close()

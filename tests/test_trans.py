"""CPU checks of the spatial translation: the oracle's PPM on non-uniform widths against the
reference's own header (oracle/_ref, and the golden vectors it wrote), the pencil builder's
structure, and the oracle's pencil transport against the invariants that pin it.

What pins what:
- compute_ppm_coeff_nonuniform (cpu_1d_ppm_nonuniform.hpp:40-62 with the face estimate and
  filter of cpu_face_estimates.hpp:179-191, 327-372): bit for bit against the reference
  header compiled in place (oracle/_ref) and against tests/golden/translation_ppm_nu.npz.
- the pencil loop (cpu_trans_map_amr.cpp:36-149, 214-441): a zero time step is the
  identity; mass (sum of f times the cell width) is kept on a periodic line whose cells
  all hold the block; an isolated cell (flattened to a constant by the limiter) sends the
  fraction v dt / dx of its content downstream and keeps the rest; a uniform field stays
  uniform; cells with target ratio 0 keep their values.
- the builder (cpu_trans_pencils.cpp, refinement level 0): every propagated cell is in
  exactly one pencil's propagated run, runs are contiguous along the dimension, stencil
  positions hold the true neighbours or the nearest good cell, ratios as the reference sets.
"""
import os

import numpy as np
import pytest

import oracle as O
from vlasiator_amd import vacc as V

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "translation_ppm_nu.npz")
HAVE_LIB = os.path.exists(V.LIB_PATH)


def _stencils(n=300, seed=3):
    rng = np.random.default_rng(seed)
    parts = [10 ** rng.uniform(-16, -10, (n, 5)),
             np.where(rng.random((n, 5)) < 0.4, 0.0, 10 ** rng.uniform(-14, -11, (n, 5))),
             np.sort(10 ** rng.uniform(-15, -11, (n, 5)), axis=1),
             1e-15 * rng.uniform(0.5, 2.0, (n, 5))]
    s = np.concatenate(parts).astype(np.float32)
    s[:3] = 0.0
    return s


def test_ppm_nonuniform_matches_golden_vectors():
    g = np.load(GOLDEN)
    for i, h in enumerate(g["widths"]):
        for j, thr in enumerate(g["thresholds"]):
            got = O.ppm_nu_batch(h, g["stencils"], float(thr))
            assert np.array_equal(got.view(np.uint32), g["coeffs"][i, j].view(np.uint32)), (h, thr)


def test_ppm_nonuniform_matches_reference_header():
    if O.ref_lib() is None:
        pytest.skip("oracle/_ref not built (needs the reference checkout)")
    rng = np.random.default_rng(4)
    s = _stencils(200, seed=9)
    for _ in range(20):
        h = (10 ** rng.uniform(4, 7, 5)).astype(np.float32)
        for thr in (1e-15, 0.0):
            a = O.ppm_nu_batch(h, s, thr)
            b = O.ref_ppm_nu_batch(h, s, thr)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _line_cells(n, gids_per_cell, rng, fill=None):
    cells = []
    for _ in range(n):
        g = np.array(gids_per_cell, np.uint32)
        f = (rng.random((len(g), 64)) * 1e-12).astype(np.float32) if fill is None else np.full((len(g), 64), fill,
                                                                                               np.float32)
        cells.append((g, f))
    return cells


def _periodic_line(n, dx=1.0e6):
    """One periodic pencil over cells 0..n-1 (positions: 2 stencil, n cells, 2 stencil)."""
    cells = [(i - 2) % n for i in range(n + 4)]
    ratio = [0.0] + [1.0] * (n + 2) + [0.0]
    return (np.array([n + 4], np.uint32), np.array([0], np.uint64), np.array(cells, np.uint32),
            np.full(n + 4, dx, np.float32), np.array(ratio, np.float32))


def _mesh():
    return O.make_mesh((20, 20, 20), (-1e6,) * 3, (1e6,) * 3)


def test_zero_time_step_is_identity():
    m = _mesh()
    rng = np.random.default_rng(1)
    cells = _line_cells(7, [5, 1210, 4000, 4001], rng)
    for dim in range(3):
        out = O.trans_map_1d(m, cells, _periodic_line(7), dim, 0.0, 1e-15)
        for (g0, f0), (g1, f1) in zip(cells, out):
            assert np.array_equal(g0, g1) and np.array_equal(f0, f1)


@pytest.mark.parametrize("dim", [0, 1, 2])
def test_periodic_line_keeps_mass(dim):
    m = _mesh()
    rng = np.random.default_rng(2)
    cells = _line_cells(9, [3, 421, 4210, 7999], rng)
    out = O.trans_map_1d(m, cells, _periodic_line(9), dim, 0.4, 1e-15)
    before = sum(f.astype(np.float64).sum(0) for _, f in cells)
    after = sum(f.astype(np.float64).sum(0) for _, f in out)
    np.testing.assert_allclose(after, before, rtol=2e-6)
    assert not all(np.array_equal(a[1], b[1]) for a, b in zip(cells, out))


def test_isolated_cell_moves_its_cfl_fraction_downstream():
    """f in one cell, zero elsewhere: the limiter flattens the profile to a constant, so the
    neighbour downstream receives f |v| dt / dx and the cell keeps the rest."""
    m = _mesh()
    dx, dt = 1.0e6, 0.5
    n = 6
    gid = 9 + 20 * 7 + 400 * 3          # block indices (9, 7, 3)
    cells = [(np.array([gid], np.uint32), np.zeros((1, 64), np.float32)) for _ in range(n)]
    cells[2][1][:] = 1e-12
    dv = O.lib().vo_cell_size(m, 0)
    out = O.trans_map_1d(m, cells, _periodic_line(n, dx), 0, dt, 1e-15)
    for e in range(64):
        i = e & 3                          # index along x within the block
        v = -1e6 + ((9 * 4 + i) + 0.5) * dv
        frac = abs(v) * dt / dx
        down = 3 if v > 0 else 1
        np.testing.assert_allclose(out[down][1][0, e], 1e-12 * frac, rtol=1e-5)
        np.testing.assert_allclose(out[2][1][0, e], 1e-12 * (1 - frac), rtol=1e-5)
        for c in range(n):
            if c not in (2, down):
                assert out[c][1][0, e] == 0.0


def test_uniform_field_stays_uniform():
    m = _mesh()
    rng = np.random.default_rng(5)
    cells = _line_cells(8, [17, 1500], rng, fill=3e-12)
    for dim in range(3):
        out = O.trans_map_1d(m, cells, _periodic_line(8), dim, 0.7, 1e-15)
        for _, f in out:
            np.testing.assert_allclose(f, 3e-12, rtol=1e-6)


def test_cells_with_zero_ratio_keep_their_values():
    m = _mesh()
    rng = np.random.default_rng(6)
    cells = _line_cells(5, [100, 101], rng)
    plen, pstart, pcell, pdz, pratio = _periodic_line(5)
    pratio = pratio.copy()
    pratio[pcell == 3] = 0.0               # cell 3 is a boundary cell: never written
    out = O.trans_map_1d(m, cells, (plen, pstart, pcell, pdz, pratio), 0, 0.5, 1e-15)
    assert np.array_equal(out[3][1], cells[3][1])
    assert not np.array_equal(out[2][1], cells[2][1])


@pytest.mark.skipif(not HAVE_LIB, reason="libvacc.so not built")
@pytest.mark.parametrize("periodic", [(False, True, False), (True, True, True), (False, False, False)])
def test_uniform_pencil_builder_structure(periodic):
    from vlasiator_amd import trans as T
    dims = (7, 5, 4)
    rng = np.random.default_rng(7)
    kind = np.zeros(np.prod(dims), np.uint8)
    kind[rng.random(kind.size) < 0.12] = T.CELL_BOUNDARY
    kind[rng.random(kind.size) < 0.08] = T.CELL_INACTIVE
    W = T.STENCIL
    coords = np.stack(np.unravel_index(np.arange(kind.size), dims[::-1])[::-1], -1)   # x, y, z
    for dim in range(3):
        L, S, cell, dz, ratio = T.build_pencils_uniform(dims, periodic, kind, (2e6, 3e6, 5e6), dim)
        seen = np.zeros(kind.size, int)
        for p in range(len(L)):
            ids = cell[S[p]:S[p] + L[p]]
            r = ratio[S[p]:S[p] + L[p]]
            run = ids[W:L[p] - W]
            seen[run] += 1
            assert np.all(kind[run] != T.CELL_INACTIVE)
            # contiguous along dim, other coordinates fixed
            c = coords[run]
            other = [d for d in range(3) if d != dim]
            assert np.all(c[:, other] == c[0, other])
            step = (np.diff(c[:, dim]) % dims[dim]) if periodic[dim] else np.diff(c[:, dim])
            assert np.all(step == 1)
            # stencil positions: true neighbour if it exists and is propagated, else the nearest good cell
            for side, base, sgn in ((0, run[0], -1), (1, run[-1], +1)):
                prev = base
                for s in range(1, W + 1):
                    x = coords[base].copy()
                    x[dim] += sgn * s
                    if periodic[dim]:
                        x[dim] %= dims[dim]
                    ok = 0 <= x[dim] < dims[dim]
                    nb = x[0] + dims[0] * (x[1] + dims[1] * x[2]) if ok else None
                    want = nb if (nb is not None and kind[nb] != T.CELL_INACTIVE) else prev
                    pos = W - s if side == 0 else L[p] - W - 1 + s
                    assert ids[pos] == want
                    prev = want
            # target ratios: positions W-1 .. L-W on normal cells
            for q in range(L[p]):
                want = 1.0 if (W - 1 <= q <= L[p] - W and kind[ids[q]] == T.CELL_NORMAL) else 0.0
                assert r[q] == want
            assert np.all(dz[S[p]:S[p] + L[p]] == np.float32((2e6, 3e6, 5e6)[dim]))
        prop = kind != T.CELL_INACTIVE
        if dims[dim] > 1:
            assert np.all(seen[prop] == 1) and np.all(seen[~prop] == 0)

"""Shared test inputs: small seeded cells built with the oracle's initial condition."""
from __future__ import annotations

import numpy as np

import oracle as O
from vlasiator_amd import workloads as W


def oracle_mesh(w: W.Workload):
    return O.make_mesh(w.grid_length, w.vmin, w.vmax)


def oracle_peaks(cell: W.CellSpec):
    return [O.make_peak(p.kind, p.n, p.V, p.T, p.axis, p.vperp, p.vth) for p in cell.peaks]


def build_cells(w: W.Workload):
    """List of (gids, f[n,64]) per cell: sampled peaks over the candidate box, then adjust."""
    m = oracle_mesh(w)
    return [O.init_cell(m, oracle_peaks(c), w.mass, w.min_value, w.add_width) for c in w.cells]


def oracle_intersections(w: W.Workload, map_order: int, pivots=None) -> np.ndarray:
    """12 FP64 intersections per cell from the oracle transform (pivot = cell V or given)."""
    m = oracle_mesh(w)
    out = np.zeros((len(w.cells), 12))
    for i, c in enumerate(w.cells):
        V = c.V if pivots is None else pivots[i]
        fl = O.make_fields(c.B, V)
        T = O.transform(fl, w.mass, w.charge, c.dt, 1.60217653e-19)
        out[i] = O.intersections(m, T, map_order)
    return out


def sort_cell(g, f):
    o = np.argsort(g, kind="stable")
    return g[o], f[o]


def assert_cells_equal(got, want, what=""):
    """Bit-level block-list parity: equal GID sets and equal f per GID (+0 == -0)."""
    assert len(got) == len(want), what
    for i, ((g1, f1), (g2, f2)) in enumerate(zip(got, want)):
        g1, f1 = sort_cell(np.asarray(g1), np.asarray(f1))
        g2, f2 = sort_cell(np.asarray(g2), np.asarray(f2))
        assert np.array_equal(g1, g2), f"{what} cell {i}: GID sets differ ({g1.size} vs {g2.size} blocks)"
        assert not np.isnan(f1).any(), f"{what} cell {i}: NaN"
        if not np.array_equal(f1, f2):
            bad = np.argwhere(f1 != f2)
            b, e = bad[0]
            raise AssertionError(f"{what} cell {i}: {len(bad)} values differ, first block gid {g1[b]} "
                                 f"cell {e}: {f1[b, e]!r} vs {f2[b, e]!r}")

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


def oracle_run(w: W.Workload, cells, isx, steps: int, fp64: bool):
    """The oracle's (accelerate, adjust) sequence on every cell for `steps` steps, map order
    step mod 3, intersections isx[map_order]; FP64 runs the Realf = double build."""
    m = oracle_mesh(w)
    out = []
    for i, (g, f) in enumerate(cells):
        f = f.astype(np.float64 if fp64 else np.float32)
        for s in range(steps):
            g, f, _ = O.accelerate_cell(m, g, f, isx[s % 3][i], s % 3, w.method, w.min_value, fp64=fp64)
            g, f, _ = O.adjust(m, g, f, w.min_value, w.add_width, fp64=fp64)
        out.append((g, f))
    return out


def block_centres(w: W.Workload, gids):
    """Cell-centre velocities (n, 64, 3) of blocks gids (in-block index e = i + 4j + 16k)."""
    L = np.array(w.grid_length, np.int64)
    vmin = np.array(w.vmin, np.float64)
    dv = (np.array(w.vmax, np.float64) - vmin) / L / 4
    g = np.asarray(gids, np.int64)
    b = np.stack([g % L[0], (g // L[0]) % L[1], g // (L[0] * L[1])], -1)
    e = np.arange(64)
    c = np.stack([e & 3, (e >> 2) & 3, e >> 4], -1)
    return vmin + (b[:, None, :] * 4 + c[None, :, :] + 0.5) * dv


def cell_moments(w: W.Workload, gids, f):
    """FP64 density n, bulk velocity V and scalar pressure P = (m/3) sum f |v - V|^2 dV^3
    (arch_moments.h:123-150) of one cell."""
    v = block_centres(w, gids)
    dv3 = float(np.prod((np.array(w.vmax, np.float64) - np.array(w.vmin, np.float64)) / np.array(w.grid_length) / 4))
    f = np.asarray(f, np.float64)
    n = f.sum() * dv3
    V = (f[..., None] * v).sum((0, 1)) * dv3 / n
    P = w.mass / 3.0 * (f * ((v - V) ** 2).sum(-1)).sum() * dv3
    return n, V, P


def compare_runs(w: W.Workload, got, ref):
    """Worst cell of: L1 and L-inf of f relative to ref (blocks missing on one side count as
    zero), relative errors of n, |V| and P, and |V - V_ref| in m/s, in units of the velocity
    cell width dv and of the 1-D thermal speed sqrt(P / (n m)) of ref."""
    dv = float(np.min((np.array(w.vmax, np.float64) - np.array(w.vmin, np.float64)) / np.array(w.grid_length) / 4))
    out = []
    for (g1, f1), (g2, f2) in zip(got, ref):
        allg = np.union1d(g1, g2)
        A = np.zeros((allg.size, 64))
        B = np.zeros((allg.size, 64))
        A[np.searchsorted(allg, g1)] = f1
        B[np.searchsorted(allg, g2)] = f2
        n1, V1, P1 = cell_moments(w, g1, f1)
        n2, V2, P2 = cell_moments(w, g2, f2)
        dV = float(np.linalg.norm(V1 - V2))
        vth = float(np.sqrt(P2 / (n2 * w.mass)))
        out.append({"l1": np.abs(A - B).sum() / np.abs(B).sum(), "linf": np.abs(A - B).max() / np.abs(B).max(),
                    "n": abs(n1 - n2) / n2, "V": dV / float(np.linalg.norm(V2)), "dV_mps": dV,
                    "dV_dv": dV / dv, "dV_vth": dV / vth, "P": abs(P1 - P2) / P2})
    return {k: float(max(o[k] for o in out)) for k in out[0]}

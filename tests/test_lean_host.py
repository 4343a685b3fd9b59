"""CPU checks of the tolerance-form remap arithmetic (options.fast_arith = 1).

vlasiator_amd/csrc/vacc_lean.h is compiled for the host and run exactly as the device kernel
k_remap_lean runs it: the lean reconstructions per source cell and the flux-form deposit per
column of a pencil.  Whole cells are mapped through every pass dimension and compared with
the oracle's map_1d built twice: with FP32 values (the reference's FP32 CPU path) and with
FP64 values (Realf = double, the reference's DPF build).

Stated tolerances, per cell and pass (relative to the cell's maximum f for L-inf, to its L1
norm for L1):
  against the FP64 run:      L-inf <= 2e-6, L1 <= 2e-6, mass <= 1e-7
  against the FP32 oracle:   L-inf <= 3e-5, L1 <= 1.5e-5
The second pair is the FP32 reference's own distance from the FP64 run (this test asserts it
is above 1e-6 L1 on these cells): its Lagrangian coordinates are FP32 numbers of up to
4e6 m/s (ulp 0.25 m/s, 1.25e-5 of a cell), which the flux form does not carry.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from tests import cases
from vlasiator_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "lean_host_check.hip")
OUT = os.path.join(HERE, "_build", "liblean_host_check.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    deps = [SRC, os.path.join(HERE, "..", "vlasiator_amd", "csrc", "vacc_lean.h"),
            os.path.join(HERE, "..", "vlasiator_amd", "csrc", "vacc_common.h")]
    if not os.path.exists(OUT) or os.path.getmtime(OUT) < max(os.path.getmtime(d) for d in deps):
        subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", SRC, "-o", OUT],
                       check=True, capture_output=True)
    L = C.CDLL(OUT)
    L.lean_coeff_batch.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p]
    L.lean_column.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_double, C.c_float, C.c_int, C.c_int, C.c_void_p,
                              C.c_void_p]
    return L


def _dense(g, f, nx, ny, pad):
    """Cell values on a dense cell grid over the blocks' bounding box, padded by `pad` blocks."""
    b = np.stack([g % nx, (g // nx) % ny, g // (nx * ny)], axis=1).astype(np.int64)
    lo = b.min(axis=0) - pad
    shape = tuple(int(x) for x in 4 * (b.max(axis=0) + pad + 1 - lo))
    d = np.zeros(shape[::-1], np.float64)   # [z, y, x]
    occ = np.zeros(tuple(int(x) // 4 for x in shape[::-1]), bool)
    e = np.arange(64)
    ei, ej, ek = e & 3, (e >> 2) & 3, e >> 4
    for q in range(len(g)):
        x0, y0, z0 = 4 * (b[q] - lo)
        d[z0 + ek, y0 + ej, x0 + ei] = f[q]
        occ[(b[q] - lo)[2], (b[q] - lo)[1], (b[q] - lo)[0]] = True
    return d, occ, lo


def lean_map_1d(lib, mesh_len, vmin, vmax, g, f, isx4, dim, method):
    """The lean pass over a whole cell: each column (run of blocks along dim) of each pencil
    through lean_column, from the FP64 intersections.  Like k_remap_lean, each column stores
    its targets; the columns of a pencil must write disjoint target ranges (k_remap_lean keeps
    the blocks they share open across them), which is asserted."""
    nx, ny, nz = mesh_len
    dv = (vmax[dim] - vmin[dim]) / (4 * mesh_len[dim])
    pad = 3
    d, occ, lo = _dense(g, f, nx, ny, pad)
    ax = 2 - dim                                  # numpy axis of dim in [z, y, x]
    dm = np.moveaxis(d, ax, -1)
    om = np.moveaxis(occ, ax, -1)
    out = np.zeros_like(dm)
    tq = [q for q in range(3) if q != dim]        # transverse axes, ascending
    dk = isx4[1 + dim]
    n_lines = dm.shape[:-1]
    for i0 in range(n_lines[0]):
        for i1 in range(n_lines[1]):
            line_occ = om[i0 // 4, i1 // 4]
            if not line_occ.any():
                continue
            # transverse global cell indices: moveaxis keeps the other axes in [z, y, x] order
            t_hi, t_lo = i0, i1                   # t_hi belongs to the larger transverse axis
            gt = {tq[1]: 4 * lo[tq[1]] + t_hi, tq[0]: 4 * lo[tq[0]] + t_lo}
            imin = isx4[0] + gt[tq[0]] * isx4[1 + tq[0]] + gt[tq[1]] * isx4[1 + tq[1]]
            blocks = np.nonzero(line_occ)[0]
            runs = np.split(blocks, np.nonzero(np.diff(blocks) != 1)[0] + 1)
            prev_hi = None
            for r in runs:
                k0, k1 = 4 * int(r[0]), 4 * int(r[-1]) + 4
                col = np.zeros(k1 - k0 + 8, np.float32)
                col[4:-4] = dm[i0, i1, k0:k1]
                x0 = (vmin[dim] + (4 * lo[dim] + k0) * dv - imin) / dk
                o = np.zeros(dm.shape[-1], np.float32)
                base = 4 * int(lo[dim])
                span = np.zeros(2, np.int32)
                lib.lean_column(method, col.ctypes.data, k1 - k0, x0, np.float32(dk / dv), base,
                                base + dm.shape[-1] - 1, o.ctypes.data, span.ctypes.data)
                if span[0] <= span[1]:
                    assert prev_hi is None or span[0] > prev_hi, "columns of a pencil write a shared target"
                    prev_hi = span[1]
                out[i0, i1] += o
    return np.moveaxis(out, -1, ax), lo


def _oracle_dense(g2, f2, nx, ny, lo, shape):
    d = np.zeros(shape, np.float64)
    e = np.arange(64)
    ei, ej, ek = e & 3, (e >> 2) & 3, e >> 4
    b = np.stack([g2 % nx, (g2 // nx) % ny, g2 // (nx * ny)], axis=1).astype(np.int64) - lo
    for q in range(len(g2)):
        x0, y0, z0 = 4 * b[q]
        assert x0 >= 0 and y0 >= 0 and z0 >= 0, "output block outside the padded box"
        d[z0 + ek, y0 + ej, x0 + ei] = f2[q]
    return d


@pytest.mark.parametrize("method", [O.PLM, O.PPM, O.PQM])
def test_lean_reconstruction_close_to_oracle(lib, method):
    """Lean coefficients integrate to the cell average and agree with the reference's to
    float rounding, except where a limiter or filter decision flips on rounding (at most 2 %
    of the cells of these columns)."""
    rng = np.random.default_rng(77 + method)
    st = []
    for _ in range(40):
        n = int(rng.integers(4, 120))
        x = np.linspace(-3, 3, n)
        kind = rng.integers(0, 3)
        if kind == 0:
            v = 1e-10 * np.exp(-((x - rng.uniform(-1, 1)) / rng.uniform(0.3, 2)) ** 2)
        elif kind == 1:
            v = np.where(rng.random(n) < 0.4, 0.0, 10 ** rng.uniform(-16, -10, n))
        else:
            v = 1e-15 * rng.uniform(0.3, 3, n)
        col = np.zeros(n + 8, np.float32)
        col[4:n + 4] = v
        st.extend(col[k:k + 9] for k in range(n))
    st = np.ascontiguousarray(np.stack(st), np.float32)
    nc = O.ncoef(method)
    got = np.zeros((len(st), nc), np.float32)
    lib.lean_coeff_batch(method, st.ctypes.data, len(st), got.ctypes.data)
    want = O.coeff_batch(method, st, 1e-15).astype(np.float64)
    got = got.astype(np.float64)
    scale = np.maximum(np.abs(st).max(axis=1), 1e-30).astype(np.float64)
    assert np.all(np.abs(got.sum(axis=1) - st[:, 4]) <= 2e-5 * scale)
    close = np.all(np.abs(got - want) <= 1e-4 * scale[:, None], axis=1)
    assert (~close).mean() <= 0.02


@pytest.mark.parametrize("cfg,method", [(0, O.PPM), (0, O.PQM), (2, O.PQM), (2, O.PLM)])
def test_lean_pass_within_stated_tolerance(lib, cfg, method):
    w = W.config(cfg, ncells=2)
    cells = cases.build_cells(w)
    m = cases.oracle_mesh(w)
    isx = cases.oracle_intersections(w, 1)
    nx, ny, _ = w.grid_length
    worst = {"lean64": [0.0, 0.0, 0.0], "lean32": [0.0, 0.0], "ref": [0.0, 0.0]}
    for i, (g, f) in enumerate(cells):
        for dim in range(3):
            isx4 = isx[i, 4 * dim:4 * dim + 4]
            lean, lo = lean_map_1d(lib, w.grid_length, w.vmin, w.vmax, g, f, isx4, dim, method)
            g32, f32, _ = O.map_1d(m, g, f, isx4, dim, method, w.min_value)
            g64, f64, _ = O.map_1d(m, g, f.astype(np.float64), isx4, dim, method, w.min_value, fp64=True)
            assert np.array_equal(np.sort(g32), np.sort(g64))
            r32 = _oracle_dense(g32, f32, nx, ny, lo, lean.shape)
            r64 = _oracle_dense(g64, f64, nx, ny, lo, lean.shape)
            # nothing outside the output blocks
            mask = _oracle_dense(g64, np.ones_like(f64), nx, ny, lo, lean.shape) > 0
            assert np.all(lean[~mask] == 0.0)
            mx, l1 = np.abs(r64).max(), np.abs(r64).sum()
            d64, d32, dref = np.abs(lean - r64), np.abs(lean - r32), np.abs(r32 - r64)
            worst["lean64"] = [max(worst["lean64"][0], d64.max() / mx), max(worst["lean64"][1], d64.sum() / l1),
                               max(worst["lean64"][2], abs(lean.sum() / r64.sum() - 1))]
            worst["lean32"] = [max(worst["lean32"][0], d32.max() / mx), max(worst["lean32"][1], d32.sum() / l1)]
            worst["ref"] = [max(worst["ref"][0], dref.max() / mx), max(worst["ref"][1], dref.sum() / l1)]
    print(f"cfg {cfg} method {method}: lean vs FP64 (Linf, L1, mass) {worst['lean64']}, "
          f"lean vs FP32 {worst['lean32']}, FP32 oracle vs FP64 {worst['ref']}", flush=True)
    assert worst["lean64"][0] <= 2e-6 and worst["lean64"][1] <= 2e-6 and worst["lean64"][2] <= 1e-7
    assert worst["lean32"][0] <= 3e-5 and worst["lean32"][1] <= 1.5e-5
    assert worst["ref"][1] > 1e-6   # the FP32 reference's own rounding is what separates the two

"""GPU parity: libvacc (HIP, gfx950) against the CPU oracle on identical inputs.

Block lists and f are compared bit for bit (equal GID sets, equal float values per GID):
the device kernels reproduce the reference's per-operation float rounding, so the
tolerance on f is zero.  FP64 quantities whose summation order differs (moments,
removed mass) and device transcendental functions (transform, initial condition) are
compared with the tolerances written in each test.
"""
import numpy as np
import pytest

import oracle as O
from tests import cases
from vlasiator_amd import vacc as V
from vlasiator_amd import workloads as W

pytestmark = pytest.mark.gpu

HALL_MIN = 1.60217653e-19


def _ctx(w: W.Workload, method=None, wall_margin=0, fast_arith=False):
    mesh = V.make_mesh(w.grid_length, w.vmin, w.vmax)
    sp = V.make_species(w.mass, w.charge, w.min_value, w.add_width)
    op = V.make_options(w.method if method is None else method, wall_margin=wall_margin, fast_arith=fast_arith)
    return V.Context(mesh, sp, op)


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def cfg0():
    w = W.config(0, ncells=8)
    return w, cases.build_cells(w)


@pytest.mark.parametrize("method", [O.PPM, O.PQM, O.PLM])
@pytest.mark.parametrize("dim", [0, 1, 2])
def test_map_1d_bit_exact(cfg0, method, dim):
    w, cells = cfg0
    isx = cases.oracle_intersections(w, map_order=dim)
    _say(f"map_1d dim={dim} method={method} cells={len(cells)}")
    ctx = _ctx(w, method)
    ctx.upload_cells(cells)
    ctx.set_intersections(isx)
    ctx.map_1d(dim)
    got = ctx.download_cells()
    m = cases.oracle_mesh(w)
    want = []
    for i, (g, f) in enumerate(cells):
        g2, f2, _ = O.map_1d(m, g, f, isx[i, 4 * dim:4 * dim + 4], dim, method, w.min_value)
        want.append((g2, f2))
    cases.assert_cells_equal(got, want, f"map_1d dim {dim} method {method}")
    ctx.close()


@pytest.mark.parametrize("method", [O.PPM, O.PQM])
@pytest.mark.parametrize("map_order", [0, 1, 2])
def test_accelerate_bit_exact(cfg0, method, map_order):
    w, cells = cfg0
    isx = cases.oracle_intersections(w, map_order)
    _say(f"accelerate order={map_order} method={method}")
    ctx = _ctx(w, method)
    ctx.upload_cells(cells)
    ctx.set_intersections(isx)
    st = ctx.accelerate(map_order)
    assert st == 0
    got = ctx.download_cells()
    m = cases.oracle_mesh(w)
    want = [O.accelerate_cell(m, g, f, isx[i], map_order, method, w.min_value)[:2] for i, (g, f) in enumerate(cells)]
    cases.assert_cells_equal(got, want, f"accelerate order {map_order} method {method}")
    ctx.close()


def test_steps_with_adjust_bit_exact(cfg0):
    """Five steps of (accelerate, adjust) with the map order cycling, as in the benchmark."""
    w, cells = cfg0
    m = cases.oracle_mesh(w)
    ctx = _ctx(w, O.PPM)
    ctx.upload_cells(cells)
    ref = [(g.copy(), f.copy()) for g, f in cells]
    for step in range(5):
        mo = step % 3
        isx = cases.oracle_intersections(w, mo)
        _say(f"step {step} order {mo}")
        ctx.set_intersections(isx)
        ctx.accelerate(mo)
        loss = ctx.adjust(len(cells))
        new, want_loss = [], []
        for i, (g, f) in enumerate(ref):
            g2, f2, _ = O.accelerate_cell(m, g, f, isx[i], mo, O.PPM, w.min_value)
            g3, f3, l3 = O.adjust(m, g2, f2, w.min_value, w.add_width)
            new.append((g3, f3))
            want_loss.append(l3)
        ref = new
        cases.assert_cells_equal(ctx.download_cells(), ref, f"step {step}")
        # removed mass: FP64 sums in a different order
        assert np.allclose(loss, want_loss, rtol=1e-12, atol=0)
    ctx.close()


def test_multipeak_pqm_step_bit_exact():
    w = W.config(2, ncells=6)
    cells = cases.build_cells(w)
    m = cases.oracle_mesh(w)
    ctx = _ctx(w)
    ctx.upload_cells(cells)
    ref = cells
    for step in range(3):
        mo = (step + 1) % 3
        isx = cases.oracle_intersections(w, mo)
        _say(f"multipeak step {step}")
        ctx.set_intersections(isx)
        ctx.accelerate(mo)
        ctx.adjust(len(cells))
        ref = [O.adjust(m, *O.accelerate_cell(m, g, f, isx[i], mo, O.PQM, w.min_value)[:2], w.min_value,
                        w.add_width)[:2] for i, (g, f) in enumerate(ref)]
        cases.assert_cells_equal(ctx.download_cells(), ref, f"multipeak step {step}")
    ctx.close()


def test_large_rotation_and_many_cells():
    """20 degrees per step (wider deposit spans, larger cross shears) over 64 cells."""
    w = W.config(1, ncells=64)
    for c in w.cells:
        c.dt *= 4.0
    cells = cases.build_cells(w)
    m = cases.oracle_mesh(w)
    isx = cases.oracle_intersections(w, 2)
    _say("large rotation, 64 cells")
    ctx = _ctx(w, O.PQM)
    ctx.upload_cells(cells)
    ctx.set_intersections(isx)
    ctx.accelerate(2)
    want = [O.accelerate_cell(m, g, f, isx[i], 2, O.PQM, w.min_value)[:2] for i, (g, f) in enumerate(cells)]
    cases.assert_cells_equal(ctx.download_cells(), want, "large rotation")
    ctx.close()


def test_edge_cells_empty_single_and_wall():
    w = W.config(0, ncells=4)
    m = cases.oracle_mesh(w)
    cells = cases.build_cells(w)
    # empty cell, single block cell, and a cell pressed against the low x wall
    cells[0] = (np.zeros(0, np.uint32), np.zeros((0, 64), np.float32))
    g, f = cells[1]
    cells[1] = (g[:1].copy(), f[:1].copy() + np.float32(1e-12))
    wall = O.init_cell(m, [O.make_peak(0, 1e6, (-1.93e6, 0, 0), 2e5)], w.mass, w.min_value, 1)
    cells[2] = wall
    isx = cases.oracle_intersections(w, 0)
    _say("edge cells")
    ctx = _ctx(w, O.PPM, wall_margin=1)
    ctx.upload_cells(cells)
    ctx.set_intersections(isx)
    st = ctx.accelerate(0)
    status = ctx.status(len(cells))
    want, wflags = [], []
    for i, (g, f) in enumerate(cells):
        g2, f2, wf = O.accelerate_cell(m, g, f, isx[i], 0, O.PPM, w.min_value, wall_margin=1)
        want.append((g2, f2))
        wflags.append(wf)
    cases.assert_cells_equal(ctx.download_cells(), want, "edge cells")
    assert [(int(s) & V.STATUS_WALL) != 0 for s in status] == [bool(x) for x in wflags]
    assert wflags[2] == 1 and st & V.STATUS_WALL
    ctx.close()


def test_wide_deposit_spans_are_remapped():
    """Cells whose per-cell deposit span (dv + 3|di| + 3|dj|)/dk exceeds the 32-cell ring's
    bound (22 cells) go to the exact kernel's 128-cell-ring instance and match the oracle bit for
    bit, in both arithmetic forms; a span above 118 cells is still passed through and flagged.
    The cross shears di, dj of the x pass are scaled up, with the Lagrangian base moved so that
    the pencil through each cell's block centroid keeps its mapping."""
    w = W.config(0, ncells=4)
    cells = cases.build_cells(w)
    m = cases.oracle_mesh(w)
    nx, ny = w.grid_length[0], w.grid_length[1]
    dv = (w.vmax[0] - w.vmin[0]) / (4 * nx)
    isx = cases.oracle_intersections(w, 0)
    spans = []
    for i, factor in enumerate([6.0, 14.0, 1.0, 30.0]):
        g = cells[i][0]
        # x pass frame: di multiplies the z index of a pencil, dj its y index
        ic, jc = 4 * (g // (nx * ny)).mean() + 2, 4 * (g // nx % ny).mean() + 2   # z, y centroid cells
        I, dk, dj, di = isx[i, 0:4]
        di2, dj2 = di * factor + 2.0 * dv * factor, dj * factor
        isx[i, 0] = I + ic * di + jc * dj - (ic * di2 + jc * dj2)
        isx[i, 2], isx[i, 3] = dj2, di2
        spans.append((dv + 3 * abs(di2) + 3 * abs(dj2)) / dk)
    assert 22 < spans[0] < 118 and 22 < spans[1] < 118 and spans[2] < 22 and spans[3] > 118, spans
    want = []
    for i, (g, f) in enumerate(cells):
        g2, f2, _ = O.map_1d(m, g, f, isx[i, 0:4], 0, O.PPM, w.min_value)
        want.append((g2, f2))
    assert sum(g.size for g, _ in want[:2]) > 100
    for fast in (False, True):
        _say(f"wide spans {[round(x, 1) for x in spans]}, fast_arith={fast}")
        ctx = _ctx(w, O.PPM, fast_arith=fast)
        ctx.upload_cells(cells)
        ctx.set_intersections(isx)
        ctx.map_1d(0)
        st = ctx.status(4)
        got = ctx.download_cells()
        assert [bool(x & V.STATUS_DK) for x in st] == [False, False, False, True]
        cases.assert_cells_equal(got[:2], want[:2], f"wide spans fast={fast}")
        cases.assert_cells_equal([got[3]], [cells[3]], "pass-through beyond the wide ring")
        ctx.close()


def test_nonpositive_dk_passes_cell_through():
    w = W.config(0, ncells=2)
    cells = cases.build_cells(w)
    isx = cases.oracle_intersections(w, 0)
    isx[1, 4 * 0 + 1] = -isx[1, 1]      # x pass of cell 1: negative Lagrangian spacing
    _say("negative dk")
    ctx = _ctx(w, O.PPM)
    ctx.upload_cells(cells)
    ctx.set_intersections(isx)
    ctx.map_1d(0)
    st = ctx.status(2)
    got = ctx.download_cells()
    assert st[1] & V.STATUS_DK and not st[0] & V.STATUS_DK
    cases.assert_cells_equal([got[1]], [cells[1]], "pass-through")
    ctx.close()


def test_device_transform_matches_oracle():
    w = W.config(1, ncells=32)
    ctx = _ctx(w, O.PPM)
    cells = cases.build_cells(w)
    ctx.upload_cells(cells)
    fields = [V.make_fields(c.B, c.V, c.dt) for c in w.cells]
    for mo in range(3):
        _say(f"device transform order {mo}")
        ctx.compute_intersections(fields, mo, use_moments=False)
        got = ctx.get_intersections(len(w.cells))
        want = cases.oracle_intersections(w, mo)
        scale = np.array([4e6, 2e4, 2e4, 2e4] * 3)
        assert np.all(np.abs(got - want) <= 1e-12 * scale)
    ctx.close()


def test_moments_match_oracle(cfg0):
    w, cells = cfg0
    ctx = _ctx(w, O.PPM)
    ctx.upload_cells(cells)
    _say("moments")
    got = ctx.moments(len(cells))
    m = cases.oracle_mesh(w)
    want = np.array([O.moments(m, g, f) for g, f in cells])
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    ctx.close()


def _moment_tol(w, want):
    """FP64 sums in another order: n to 1e-12 relative, n*V to 1e-12 of n * max|v| of the mesh."""
    vmax = float(np.max(np.abs(np.concatenate([np.asarray(w.vmin), np.asarray(w.vmax)]))))
    return 1e-12 * np.abs(want[:, :1]) * np.array([[1.0, vmax, vmax, vmax]]) + 1e-300


def test_moments_follow_the_state(cfg0):
    """Resident moments are recomputed whenever f changes (remaps, adjust, a lone 1-D remap):
    after each, vacc_moments against the oracle's moments of the downloaded state."""
    w, cells = cfg0
    m = cases.oracle_mesh(w)
    ctx = _ctx(w, O.PQM)
    ctx.upload_cells(cells)
    for step in range(3):
        mo = (step + 2) % 3
        ctx.set_intersections(cases.oracle_intersections(w, mo))
        ctx.accelerate(mo)
        for what in ("accelerate", "adjust"):
            if what == "adjust":
                ctx.adjust(len(cells))
            _say(f"moments step {step} after {what}")
            got = ctx.moments(len(cells))
            want = np.array([O.moments(m, g, f) for g, f in ctx.download_cells()])
            assert np.all(np.abs(got - want) <= _moment_tol(w, want)), f"step {step} {what}"
    ctx.map_1d(1)
    got = ctx.moments(len(cells))
    want = np.array([O.moments(m, g, f) for g, f in ctx.download_cells()])
    assert np.all(np.abs(got - want) <= _moment_tol(w, want))
    ctx.close()


def test_step_bit_exact_against_oracle(cfg0):
    """vacc_step (transform from the resident fields, three remaps, adjust, one sync) against
    the oracle's remaps and adjust driven by the intersections the step computed; those
    intersections against the oracle transform."""
    w, cells = cfg0
    m = cases.oracle_mesh(w)
    ctx = _ctx(w, O.PPM)
    ctx.upload_cells(cells)
    ctx.set_fields([V.make_fields(c.B, c.V, c.dt) for c in w.cells])
    ref = cells
    for step in range(4):
        mo = step % 3
        _say(f"vacc_step {step} order {mo}")
        loss = np.zeros(len(cells))
        st = ctx.step(mo, use_moments=False, adjust=True, rho_loss=loss)
        assert st == 0
        isx = ctx.get_intersections(len(cells))
        want_isx = cases.oracle_intersections(w, mo)
        scale = np.abs(want_isx).max(axis=0) + 1e-300
        assert np.all(np.abs(isx - want_isx) <= 1e-12 * scale)
        new, want_loss = [], []
        for i, (g, f) in enumerate(ref):
            g2, f2, _ = O.accelerate_cell(m, g, f, isx[i], mo, O.PPM, w.min_value)
            g3, f3, l3 = O.adjust(m, g2, f2, w.min_value, w.add_width)
            new.append((g3, f3))
            want_loss.append(l3)
        ref = new
        cases.assert_cells_equal(ctx.download_cells(), ref, f"vacc_step {step}")
        assert np.allclose(loss, want_loss, rtol=1e-12, atol=0)
    ctx.close()


def test_step_with_moment_pivot_matches_separate_calls():
    """vacc_step(use_moments=1) equals vacc_compute_cell_intersections(use_moments=1) +
    vacc_accelerate_cells + vacc_adjust_blocks bit for bit, and its pivot is the bulk
    velocity of the oracle's moments of the state the step starts from."""
    w = W.config(1, ncells=16)
    cells = cases.build_cells(w)
    m = cases.oracle_mesh(w)
    fields = [V.make_fields(c.B, c.V, c.dt) for c in w.cells]
    a, b = _ctx(w), _ctx(w)
    a.upload_cells(cells)
    b.upload_cells(cells)
    a.set_fields(fields)
    state = cells
    for step in range(3):
        mo = (step + 1) % 3
        _say(f"moment pivot step {step}")
        a.step(mo, use_moments=True, adjust=True)
        b.compute_intersections(fields, mo, use_moments=True)
        b.accelerate(mo)
        b.adjust(len(cells))
        ia, ib = a.get_intersections(len(cells)), b.get_intersections(len(cells))
        assert np.array_equal(ia, ib)
        mom = np.array([O.moments(m, g, f) for g, f in state])
        want_isx = cases.oracle_intersections(w, mo, pivots=mom[:, 1:] / mom[:, :1])
        scale = np.abs(want_isx).max(axis=0) + 1e-300
        assert np.all(np.abs(ia - want_isx) <= 1e-10 * scale)
        state = a.download_cells()
        cases.assert_cells_equal(state, b.download_cells(), f"step {step}")
    a.close()
    b.close()


def test_init_peaks_matches_oracle_initial_condition():
    w = W.config(2, ncells=8)
    cells = cases.build_cells(w)
    ctx = _ctx(w)
    _say("init_peaks")
    ctx.init_peaks([W.to_vacc_peaks(c) for c in w.cells], W.boxes(w))
    got = ctx.download_cells()
    for i, ((g1, f1), (g2, f2)) in enumerate(zip(got, cells)):
        assert np.array_equal(g1, np.sort(g2)), f"cell {i} block sets differ"
        g2s, f2s = cases.sort_cell(g2, f2)
        # device exp/pow vs host libm: a few ulp on the sampled values
        assert np.allclose(f1, f2s, rtol=1e-6, atol=0), f"cell {i}"
    ctx.close()


@pytest.mark.parametrize("cfg,method", [(0, O.PPM), (2, O.PQM), (1, O.PLM)])
def test_fast_arith_within_tolerance(cfg, method):
    """fast_arith=1 (k_remap_lean: flux-form deposit from the FP64 intersections, lean
    reconstructions): block lists equal to the oracle's, and per cell and step, relative to
    the cell's max f (L-inf) and L1 norm (L1):
      against the FP64 run (oracle with Realf = double):   L-inf <= 5e-6, L1 <= 5e-6,
                                                          density <= 1e-7
      against the FP32 oracle (the reference's FP32 path):  L-inf <= 2e-4, L1 <= 3e-5.
    Three remaps per step.  The FP64 bound is the gate; the FP32 path is itself up to 9.2e-5
    (L-inf) from the FP64 run on the PQM case (its FP32 Lagrangian coordinates carry about
    1e-5 of a cell per pass), so the distance to it is bounded loosely and printed."""
    w = W.config(cfg, ncells=8)
    cells = cases.build_cells(w)
    m = cases.oracle_mesh(w)
    mesh = V.make_mesh(w.grid_length, w.vmin, w.vmax)
    sp = V.make_species(w.mass, w.charge, w.min_value, w.add_width)
    ctx = V.Context(mesh, sp, V.make_options(method, fast_arith=True))
    worst = {"fp64": [0.0, 0.0, 0.0], "fp32": [0.0, 0.0], "ref": [0.0, 0.0]}
    ref = cells
    for step in range(3):
        isx = cases.oracle_intersections(w, step % 3)
        _say(f"fast arith cfg {cfg} method {method} step {step}")
        ctx.upload_cells(ref)   # every step from identical inputs
        ctx.set_intersections(isx)
        ctx.accelerate(step % 3)
        got = ctx.download_cells()
        r32 = [O.accelerate_cell(m, g, f, isx[i], step % 3, method, w.min_value)[:2] for i, (g, f) in enumerate(ref)]
        r64 = [O.accelerate_cell(m, g, f.astype(np.float64), isx[i], step % 3, method, w.min_value, fp64=True)[:2]
               for i, (g, f) in enumerate(ref)]
        for (g1, f1), (ga, fa), (gb, fb) in zip(got, r32, r64):
            ga, fa = cases.sort_cell(ga, fa)
            gb, fb = cases.sort_cell(gb, fb)
            assert np.array_equal(g1, ga) and np.array_equal(g1, gb)
            x, a, b = f1.astype(np.float64), fa.astype(np.float64), fb
            mx, l1 = np.abs(b).max(), np.abs(b).sum()
            worst["fp64"] = [max(worst["fp64"][0], np.abs(x - b).max() / mx),
                             max(worst["fp64"][1], np.abs(x - b).sum() / l1),
                             max(worst["fp64"][2], abs(x.sum() / b.sum() - 1))]
            worst["fp32"] = [max(worst["fp32"][0], np.abs(x - a).max() / mx),
                             max(worst["fp32"][1], np.abs(x - a).sum() / l1)]
            worst["ref"] = [max(worst["ref"][0], np.abs(a - b).max() / mx),
                            max(worst["ref"][1], np.abs(a - b).sum() / l1)]
        ref = [cases.sort_cell(g, f) for g, f in r32]
    print("fast arith worst: vs FP64 (Linf, L1, density)", worst["fp64"], "vs FP32", worst["fp32"],
          "FP32 oracle vs FP64", worst["ref"], flush=True)
    assert worst["fp64"][0] <= 5e-6 and worst["fp64"][1] <= 5e-6 and worst["fp64"][2] <= 1e-7
    assert worst["fp32"][0] <= 2e-4 and worst["fp32"][1] <= 3e-5
    ctx.close()


def test_fast_arith_wide_targets_in_column_sets():
    """fast_arith=1 with target cells 3.5 source cells wide on the x pass (dk scaled by 3.5):
    two columns of one pencil, one block apart, can then deposit into one target cell, which
    k_remap_lean (one store per target) must not take; the column build hands such column sets
    to the exact kernel and single columns stay in k_remap_lean.  Against the oracle with the
    same intersections: equal block lists, density within 1e-6 of the FP64 run (a lost deposit
    would show here), and L-inf within 2e-4 (the exact kernel's FP32 Lagrangian coordinates)."""
    w = W.config(2, ncells=8)
    cells = cases.build_cells(w)
    m = cases.oracle_mesh(w)
    nx, ny = w.grid_length[0], w.grid_length[1]
    gaps = 0
    for g, _ in cells:   # pencils along x with a one-block gap: column sets of the x pass
        key = (g // nx).astype(np.int64)
        for k in np.unique(key):
            gaps += int((np.diff(np.sort(g[key == k] % nx)) == 2).sum())
    assert gaps > 0
    ctx = V.Context(V.make_mesh(w.grid_length, w.vmin, w.vmax),
                    V.make_species(w.mass, w.charge, w.min_value, w.add_width),
                    V.make_options(O.PQM, fast_arith=True))
    ctx.upload_cells(cells)
    isx = cases.oracle_intersections(w, 0)
    isx[:, 1] *= 3.5   # the x pass's dk (map order 0 runs x first)
    ctx.set_intersections(isx)
    _say("fast arith, dk = 3.5 dv on the x pass")
    ctx.accelerate(0)
    got = ctx.download_cells()
    worst = [0.0, 0.0]
    for i, ((g1, f1), (g, f)) in enumerate(zip(got, cells)):
        ga, fa, _ = O.accelerate_cell(m, g, f, isx[i], 0, O.PQM, w.min_value)
        gb, fb, _ = O.accelerate_cell(m, g, f.astype(np.float64), isx[i], 0, O.PQM, w.min_value, fp64=True)
        ga, fa = cases.sort_cell(ga, fa)
        gb, fb = cases.sort_cell(gb, fb)
        assert np.array_equal(g1, ga) and np.array_equal(g1, gb), f"cell {i}"
        x = f1.astype(np.float64)
        worst = [max(worst[0], np.abs(x - fb).max() / np.abs(fb).max()), max(worst[1], abs(x.sum() / fb.sum() - 1))]
    _say(f"wide targets: vs FP64 L-inf {worst[0]:.3g}, density {worst[1]:.3g}")
    assert worst[0] <= 2e-4 and worst[1] <= 1e-6
    ctx.close()


@pytest.mark.parametrize("store", ["bf16", "f16"])
def test_fast_arith_16bit_storage(store):
    """fast_arith=1 with 16-bit storage (k_remap_lean decoding the source blocks and encoding
    each target block), three PQM steps on the multi-peak workload, each step from the device's
    stored state.  Against the oracle's FP32 path with the same storage (blocks re-quantized
    after every pass): equal block lists and no status bit (no ring overflow), and, relative to
    the cell's max f, L-inf <= two code steps (2^-7 BF16, 2^-10 FP16), since the tolerance form's
    ~1e-6 differences can move a value across a rounding boundary of the code.  Against the
    FP64 run: density <= 1e-6 (the per-block scale keeps each block's FP32 mass)."""
    st, ost, linf = {"bf16": (V.STORE_BF16, O.ST_BF16, 2.0 ** -7), "f16": (V.STORE_F16, O.ST_F16, 2.0 ** -10)}[store]
    w = W.config(2, ncells=8)
    m = cases.oracle_mesh(w)
    ctx = V.Context(V.make_mesh(w.grid_length, w.vmin, w.vmax),
                    V.make_species(w.mass, w.charge, w.min_value, w.add_width),
                    V.make_options(O.PQM, fast_arith=True, storage=st))
    ctx.upload_cells(cases.build_cells(w))
    ref = ctx.download_cells()   # the stored (quantized) initial state
    worst = [0.0, 0.0, 0.0]
    for step in range(3):
        mo = step % 3
        isx = cases.oracle_intersections(w, mo)
        _say(f"fast arith {store} step {step}")
        ctx.set_intersections(isx)
        assert not (ctx.accelerate(mo) & V.STATUS_DK), "ring overflow or pass-through reported"
        got = ctx.download_cells()
        for i, ((g1, f1), (g, f)) in enumerate(zip(got, ref)):
            ga, fa, _ = O.accelerate_cell(m, g, f, isx[i], mo, O.PQM, w.min_value, storage=ost)
            gb, fb, _ = O.accelerate_cell(m, g, f.astype(np.float64), isx[i], mo, O.PQM, w.min_value, fp64=True)
            ga, fa = cases.sort_cell(ga, fa)
            gb, fb = cases.sort_cell(gb, fb)
            assert np.array_equal(g1, ga) and np.array_equal(g1, gb), f"step {step} cell {i}"
            x = f1.astype(np.float64)
            mx = np.abs(fb).max()
            worst = [max(worst[0], np.abs(x - fa).max() / mx), max(worst[1], np.abs(x - fa).sum() / np.abs(fa).sum()),
                     max(worst[2], abs(x.sum() / fb.sum() - 1))]
        ref = got
    _say(f"fast arith {store}: vs same-storage oracle (Linf, L1) {worst[:2]}, density vs FP64 {worst[2]}")
    assert worst[0] <= linf and worst[2] <= 1e-6
    ctx.close()


@pytest.mark.parametrize("per_cell,adjust_mode", [(0, 1), (1, 1), (0, 2), (0, 3)])
def test_cpp_host_adapter_driver(tmp_path, per_cell, adjust_mode):
    """The C++ host adapter (vlasiator_amd/host/vacc_host.hpp) through its example driver:
    three steps of intersections + remaps (batched, or one cell per launch) + adjust
    (single-cell, or with ring spatial neighbours, or ring neighbours then translation along
    the ring as a periodic pencil), then a VLSV write and read-back through the adapter;
    checked bit for bit against the oracle run with the intersections the driver computed."""
    import os
    import struct
    import subprocess
    drv = os.path.join(os.path.dirname(V.LIB_PATH), "vacc_acc_driver")
    if not os.path.exists(drv):
        pytest.fail("vacc_acc_driver not built")
    w = W.config(0, ncells=4)
    cells = cases.build_cells(w)
    steps = 3
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as fp:
        fp.write(b"VACC" + struct.pack("<IIiiii", 1, len(cells), O.PPM, steps, per_cell, adjust_mode))
        fp.write(struct.pack("<3I", *w.grid_length) + struct.pack("<3d", *w.vmin) + struct.pack("<3d", *w.vmax))
        fp.write(struct.pack("<3di", w.mass, w.charge, w.min_value, w.add_width))
        for c, (g, f) in zip(w.cells, cells):
            fp.write(struct.pack("<7d", *c.B, *c.V, c.dt) + struct.pack("<I", len(g)))
            fp.write(np.asarray(g, np.uint32).tobytes() + np.asarray(f, np.float32).tobytes())
    out = tmp_path / "out.bin"
    _say(f"driver per_cell={per_cell}")
    subprocess.run([drv, str(inp), str(out)], check=True, timeout=120)
    buf = open(out, "rb").read()
    pos = 0
    isx = []
    for s in range(steps):
        isx.append(np.frombuffer(buf, np.float64, 12 * len(cells), pos).reshape(len(cells), 12))
        pos += 8 * 12 * len(cells)
    got = []
    for _ in cells:
        (n,) = struct.unpack_from("<I", buf, pos)
        pos += 4
        g = np.frombuffer(buf, np.uint32, n, pos)
        pos += 4 * n
        f = np.frombuffer(buf, np.float32, 64 * n, pos).reshape(n, 64)
        pos += 256 * n + 8
        got.append((g, f))
    m = cases.oracle_mesh(w)
    ref = cells
    n = len(cells)
    for s in range(steps):
        acc = [O.accelerate_cell(m, g, f, isx[s][i], s % 3, O.PPM, w.min_value)[:2] for i, (g, f) in enumerate(ref)]
        if adjust_mode == 1:
            ref = [O.adjust(m, g, f, w.min_value, w.add_width)[:2] for g, f in acc]
        else:
            content = [g[O.content_flags(f, w.min_value).astype(bool)] for g, f in acc]
            ref = [O.adjust_spatial(m, g, f, w.min_value, w.add_width,
                                    np.concatenate([content[(i - 1) % n], content[(i + 1) % n]]))[:2]
                   for i, (g, f) in enumerate(acc)]
        if adjust_mode == 3:
            L = n + 4
            pen = (np.array([L], np.uint32), np.array([0], np.uint64),
                   np.array([(q + 2 * n - 2) % n for q in range(L)], np.uint32), np.full(L, 1.0e7, np.float32),
                   np.array([1.0 if 1 <= q <= L - 2 else 0.0 for q in range(L)], np.float32))
            ref = O.trans_map_1d(m, ref, pen, s % 3, 1.5, w.min_value)
    cases.assert_cells_equal(got, ref, f"C++ driver adjust {adjust_mode}")
    # the host intersections agree with the oracle transform
    assert np.allclose(isx[0], cases.oracle_intersections(w, 0), rtol=0,
                       atol=1e-12 * 4e6)


def test_two_population_step_matches_sequential_calls():
    """vacc_accelerate_populations (protons + He2+) against the same step written with
    single-population calls: for each population in order, the plasma pivot
    V = sum(m nV) / sum(m n), RHOQ = sum(q n) from both populations' current moments
    (arch_moments.cpp:406-440), then its transform, remaps and adjust.  Bit for bit."""
    ws = W.populations(4, ncells=4)
    cells = [cases.build_cells(w) for w in ws]
    fields = [[V.make_fields(c.B, c.V, c.dt) for c in w.cells] for w in ws]
    group = [_ctx(w) for w in ws]
    seq = [_ctx(w) for w in ws]
    for p in range(2):
        group[p].upload_cells(cells[p])
        group[p].set_fields(fields[p])
        seq[p].upload_cells(cells[p])
    n = len(ws[0].cells)
    for step in range(2):
        mo = step % 3
        _say(f"two populations step {step}")
        V.accelerate_populations(group, mo, use_moments=True, adjust=True)
        for p in range(2):
            mom = [c.moments(n) for c in seq]
            rhom = np.zeros(n)
            nv = np.zeros((n, 3))
            rhoq = np.zeros(n)
            for q in range(2):
                rhom = rhom + mom[q][:, 0] * ws[q].mass
                nv = nv + mom[q][:, 1:] * ws[q].mass
                rhoq = rhoq + mom[q][:, 0] * ws[q].charge
            piv = nv / rhom[:, None]
            fl = [V.make_fields(c.B, tuple(piv[i]), c.dt, rhoq=rhoq[i]) for i, c in enumerate(ws[p].cells)]
            seq[p].compute_intersections(fl, mo, use_moments=False)
            assert np.array_equal(seq[p].get_intersections(n), group[p].get_intersections(n)), f"pop {p}"
            seq[p].accelerate(mo)
            seq[p].adjust(n)
        for p in range(2):
            cases.assert_cells_equal(group[p].download_cells(), seq[p].download_cells(), f"pop {p} step {step}")
    for c in group + seq:
        c.close()


@pytest.mark.parametrize("corners", [False, True])
def test_spatial_neighbour_adjust_bit_exact(cfg0, corners):
    """vacc_adjust_blocks_spatial after a remap: ring neighbours (i-1, i+1) of the context plus
    caller GIDs, against the oracle's adjust with the union of the neighbours' content GIDs
    (spatial_cell_cpu.cpp:212-222).  One cell starts empty and only receives neighbour blocks.
    corners=True adds GIDs at opposite mesh corners (and one GID outside the mesh, ignored),
    so the cell's box spans the mesh and the LDS bitmap is processed in chunks."""
    w, cells = cfg0
    cells = list(cells)
    cells[3] = (np.zeros(0, np.uint32), np.zeros((0, 64), np.float32))
    n = len(cells)
    m = cases.oracle_mesh(w)
    ntot = int(np.prod(w.grid_length))
    isx = cases.oracle_intersections(w, 1)
    ctx = _ctx(w, O.PPM)
    ctx.upload_cells(cells)
    ctx.set_intersections(isx)
    ctx.accelerate(1)
    state = [O.accelerate_cell(m, g, f, isx[i], 1, O.PPM, w.min_value)[:2] if g.size else (g, f)
             for i, (g, f) in enumerate(cells)]
    nbr_off = np.arange(0, 2 * n + 1, 2, dtype=np.uint32)
    nbr = np.array([[(i - 1) % n, (i + 1) % n] for i in range(n)], np.uint32).reshape(-1)
    rng = np.random.default_rng(5)
    extra = [rng.choice(ntot, 5).astype(np.uint32) for _ in range(n)]
    if corners:
        extra = [np.concatenate([e, np.array([0, ntot - 1, ntot], np.uint32)]) for e in extra]
    extra_off = np.zeros(n + 1, np.uint64)
    extra_off[1:] = np.cumsum([e.size for e in extra])
    _say(f"spatial adjust corners={corners}")
    loss = ctx.adjust_spatial(n, nbr_off, nbr, extra_off, np.concatenate(extra))
    want, want_loss = [], []
    for i, (g, f) in enumerate(state):
        nb = [state[(i - 1) % n], state[(i + 1) % n]]
        ex = [gg[O.content_flags(ff, w.min_value).astype(bool)] for gg, ff in nb] + [extra[i]]
        g2, f2, l2 = O.adjust_spatial(m, g, f, w.min_value, w.add_width, np.concatenate(ex))
        want.append((g2, f2))
        want_loss.append(l2)
    cases.assert_cells_equal(ctx.download_cells(), want, f"spatial adjust corners={corners}")
    assert np.allclose(loss, want_loss, rtol=1e-12, atol=0)
    ctx.close()


def test_step_graph_replays_match_eager_steps():
    """vacc_options.use_graph: 14 steps, so every captured variant (map order x pool parity)
    is replayed, with eager work (moments, downloads) between replays; every compared state
    equals the eager context's bit for bit."""
    w = W.config(2, ncells=6)
    cells = cases.build_cells(w)
    fields = [V.make_fields(c.B, c.V, c.dt) for c in w.cells]
    mesh = V.make_mesh(w.grid_length, w.vmin, w.vmax)
    sp = V.make_species(w.mass, w.charge, w.min_value, w.add_width)
    eager = V.Context(mesh, sp, V.make_options(w.method))
    graph = V.Context(mesh, sp, V.make_options(w.method, use_graph=True))
    for c in (eager, graph):
        c.upload_cells(cells)
        c.set_fields(fields)
    n = len(cells)
    for step in range(14):
        _say(f"graph step {step}")
        le, lg = np.zeros(n), np.zeros(n)
        eager.step(step % 3, use_moments=True, adjust=True, rho_loss=le)
        graph.step(step % 3, use_moments=True, adjust=True, rho_loss=lg)
        assert np.array_equal(le, lg), f"step {step} loss"
        assert eager.total_blocks() == graph.total_blocks()
        if step % 4 == 3:   # eager work between replays
            assert np.array_equal(eager.moments(n), graph.moments(n))
        if step in (6, 13):
            cases.assert_cells_equal(graph.download_cells(), eager.download_cells(), f"graph step {step}")
    eager.close()
    graph.close()


@pytest.mark.parametrize("fast", [False, True])
def test_autotune_keeps_state_and_results(fast):
    """vacc_autotune times trial remaps into scratch buffers: the state is unchanged, and
    steps after tuning equal steps of an untuned context bit for bit, in the exact form and in
    the tolerance form (whose workgroup size, 4, 8 or 16 work units, is tuned as well)."""
    w = W.config(2, ncells=6)
    cells = cases.build_cells(w)
    fields = [V.make_fields(c.B, c.V, c.dt) for c in w.cells]

    def ctx():
        return V.Context(V.make_mesh(w.grid_length, w.vmin, w.vmax),
                         V.make_species(w.mass, w.charge, w.min_value, w.add_width),
                         V.make_options(w.method, fast_arith=fast))

    a, b = ctx(), ctx()
    for c in (a, b):
        c.upload_cells(cells)
        c.set_fields(fields)
    _say(f"autotune fast={fast}")
    mult = a.autotune()
    assert all(m in (8, 32, 128, 512, 1024, 2048, 4096) for m in mult)
    m2, lanes = a.remap_shape()
    assert m2 == mult and all(x in ((64, 128, 256) if fast else (256,)) for x in lanes)
    assert b.remap_shape() == ([128] * 3, [256] * 3)
    cases.assert_cells_equal(a.download_cells(), cells, "state after autotune")
    for step in range(3):
        la, lb = np.zeros(len(cells)), np.zeros(len(cells))
        a.step(step % 3, use_moments=True, adjust=True, rho_loss=la)
        b.step(step % 3, use_moments=True, adjust=True, rho_loss=lb)
        assert np.array_equal(la, lb)
    cases.assert_cells_equal(a.download_cells(), b.download_cells(), "tuned vs untuned")
    a.close()
    b.close()


@pytest.mark.parametrize("method", [O.PPM, O.PQM])
def test_overlapping_column_targets_bit_exact(method):
    """Column sets whose columns' target ranges share blocks (the remap keeps those blocks
    open across the column boundary): every other block along z in a 6x6 block patch, and a
    z-pass with a transverse shear of one cell per cell (di = dj = dv) so that the corner
    spread of a set exceeds the one-block gap between its columns.  The intersection is
    offset by -180 cells so that the sheared targets of the patch (block indices 20..25 in
    x and y, i.e. a shear of 160..206 cells) stay inside the 200-cell z extent."""
    w = W.config(0, ncells=1)
    m = cases.oracle_mesh(w)
    nx, ny, nz = w.grid_length
    rng = np.random.default_rng(9)
    gids = np.array([x + y * nx + z * nx * ny for x in range(20, 26) for y in range(20, 26)
                     for z in range(10, 40, 2)], np.uint32)
    f = (1e-12 * rng.random((gids.size, 64))).astype(np.float32)
    dv = (w.vmax[2] - w.vmin[2]) / nz / 4
    isx = np.zeros((1, 12))
    isx[0, 0:4] = (w.vmin[0], 0.0, 0.0, 1.0)
    isx[0, 4:8] = (w.vmin[1], 0.0, 0.0, 1.0)
    isx[0, 8:12] = (w.vmin[2] + (0.3 - 180.0) * dv, dv, dv, 1.01 * dv)
    _say(f"overlapping column targets method={method}")
    ctx = _ctx(w, method)
    ctx.upload_cells([(gids, f)])
    ctx.set_intersections(isx)
    assert ctx.map_1d(2) == 0
    g2, f2, _ = O.map_1d(m, gids, f, isx[0, 8:12], 2, method, w.min_value)
    got = ctx.download_cells()
    cases.assert_cells_equal(got, [(g2, f2)], "overlapping targets")
    # the case is meaningful: the mass stays inside the mesh, and remapping each column on its
    # own yields more target blocks in total than the joint remap, i.e. columns of a set share
    # target blocks
    assert abs(f2.astype(np.float64).sum() / f.astype(np.float64).sum() - 1.0) < 1e-6
    z = gids // (nx * ny)
    alone = sum(O.map_1d(m, gids[z == k], f[z == k], isx[0, 8:12], 2, method, w.min_value)[0].size
                for k in range(10, 40, 2))
    assert alone > g2.size
    ctx.close()


def test_pressure_tensor_matches_oracle():
    """vacc_pressure (FP64 second moments, arch_moments.h:123-150) against the oracle, about
    the population's own bulk velocity and about a given pivot; sums in another order, so
    within 1e-11 of the diagonal."""
    w = W.config(2, ncells=4)
    cells = cases.build_cells(w)
    m = cases.oracle_mesh(w)
    ctx = _ctx(w)
    ctx.upload_cells(cells)
    piv = np.array([[1e4 * i, -2e4, 3e4] for i in range(len(cells))])
    for pivot in (None, piv):
        _say(f"pressure pivot={'own' if pivot is None else 'given'}")
        got = ctx.pressure(len(cells), pivot)
        for i, (g, f) in enumerate(cells):
            mo = O.moments(m, g, f)
            V = mo[1:] / mo[0] if pivot is None else pivot[i]
            want = O.pressure(m, g, f, V, w.mass)
            scale = np.abs(want[:3]).max()
            assert np.all(np.abs(got[i] - want) <= 1e-11 * scale), (i, got[i], want)
    ctx.close()

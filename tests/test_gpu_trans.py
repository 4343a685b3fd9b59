"""GPU parity of the spatial translation: libvacc's kernels (k_trans_walk for the cells written
from one pencil, k_trans for the others, through vacc_trans_map_1d) against the oracle's
restatement of trans_map_1d_amr (cpu_trans_map_amr.cpp:214) on the same cells and pencils.
f is compared bit for bit; block lists do not change."""
import numpy as np
import pytest

import oracle as O
from tests import cases
from vlasiator_amd import trans as T
from vlasiator_amd import vacc as V
from vlasiator_amd import workloads as W

pytestmark = pytest.mark.gpu

DIMS = (6, 5, 4)
DX = (1.0e7, 8.0e6, 1.2e7)
DT = 1.7                    # |v| <= 2e6 m/s: CFL <= 0.43


def _say(msg):
    print(msg, flush=True)


def _kinds(seed=3):
    rng = np.random.default_rng(seed)
    kind = np.zeros(int(np.prod(DIMS)), np.uint8)
    x = np.arange(kind.size) % DIMS[0]
    kind[x == 0] = T.CELL_INACTIVE
    kind[x == 1] = T.CELL_BOUNDARY
    kind[(x > 1) & (rng.random(kind.size) < 0.06)] = T.CELL_BOUNDARY
    return kind


@pytest.fixture(scope="module")
def grid():
    w = W.config(0, ncells=int(np.prod(DIMS)))
    return w, cases.build_cells(w), cases.oracle_mesh(w), _kinds()


def _ctx(w, storage=V.STORE_F32):
    return V.Context(V.make_mesh(w.grid_length, w.vmin, w.vmax),
                     V.make_species(w.mass, w.charge, w.min_value, w.add_width), V.make_options(w.method, storage=storage))


@pytest.mark.parametrize("periodic", [(False, True, True), (True, False, True)])
@pytest.mark.parametrize("dim", [0, 1, 2])
def test_translation_bit_exact(grid, dim, periodic):
    w, cells, m, kind = grid
    pen = T.build_pencils_uniform(DIMS, periodic, kind, DX, dim)
    _say(f"translation dim={dim} periodic={periodic} pencils={len(pen[0])}")
    ctx = _ctx(w)
    ctx.upload_cells(cells)
    T.set_pencils(ctx, dim, pen)
    T.trans_map_1d(ctx, dim, DT)
    want = O.trans_map_1d(m, cells, pen, dim, DT, w.min_value)
    cases.assert_cells_equal(ctx.download_cells(), want, f"dim {dim}")
    ctx.close()


def test_three_dimensions_after_an_acceleration_step(grid):
    """Acceleration (lists left in column order), then x, y, z translation: the library
    relists before translating; against the oracle run of the same sequence."""
    w, cells, m, kind = grid
    periodic = (False, True, True)
    isx = cases.oracle_intersections(w, 1)
    ctx = _ctx(w)
    ctx.upload_cells(cells)
    ctx.set_intersections(isx)
    _say("accelerate + translate xyz")
    ctx.accelerate(1)
    want = [O.accelerate_cell(m, g, f, isx[i], 1, w.method, w.min_value)[:2] for i, (g, f) in enumerate(cells)]
    for dim in range(3):
        pen = T.build_pencils_uniform(DIMS, periodic, kind, DX, dim)
        T.set_pencils(ctx, dim, pen)
        T.trans_map_1d(ctx, dim, DT)
        want = O.trans_map_1d(m, want, pen, dim, DT, w.min_value)
    cases.assert_cells_equal(ctx.download_cells(), want, "xyz")
    ctx.close()


def test_translation_errors(grid):
    w, cells, m, kind = grid
    ctx = _ctx(w)
    ctx.upload_cells(cells)
    with pytest.raises(V.VaccError, match="no pencils"):
        T.trans_map_1d(ctx, 0, DT)
    pen = T.build_pencils_uniform(DIMS, (True, True, True), kind, DX, 0)
    T.set_pencils(ctx, 0, pen)
    ctx.upload_cells(cells[:10])              # a new cell set retires the pencils
    with pytest.raises(V.VaccError, match="no pencils"):
        T.trans_map_1d(ctx, 0, DT)
    with pytest.raises(V.VaccError, match="outside the context"):
        T.set_pencils(ctx, 0, pen)
    ctx.close()


@pytest.mark.parametrize("storage", [V.STORE_BF16, V.STORE_F16])
def test_translation_16bit_storage(grid, storage):
    """16-bit storage: the kernel computes in FP32 from the decoded blocks and encodes each
    written block; against the oracle run on the decoded state with the written cells'
    blocks requantized (cells that are not written keep their codes)."""
    w, cells, m, kind = grid
    st = {V.STORE_BF16: O.ST_BF16, V.STORE_F16: O.ST_F16}[storage]
    periodic = (False, True, True)
    ctx = _ctx(w, storage)
    ctx.upload_cells(cells)
    held = ctx.download_cells()
    for dim in range(3):
        pen = T.build_pencils_uniform(DIMS, periodic, kind, DX, dim)
        written = set(int(c) for c, r in zip(pen[2], pen[4]) if r > 0)
        _say(f"translation storage={storage} dim={dim}")
        T.set_pencils(ctx, dim, pen)
        T.trans_map_1d(ctx, dim, DT)
        out = O.trans_map_1d(m, held, pen, dim, DT, w.min_value)
        held = [(g, O.requantize(f, st) if i in written else f) for i, (g, f) in enumerate(out)]
        cases.assert_cells_equal(ctx.download_cells(), held, f"storage {storage} dim {dim}")
    ctx.close()


def test_cells_written_by_two_pencils(grid):
    """Refinement-style pencils (the bins of cpu_trans_pencils.hpp:48-50): two pencils share
    their first two cells A, B (width 2 dx, target ratio 1/2 in each) and go on through their
    own cells C, D (width dx, ratio 1).  A and B receive the contributions of both pencils in
    pencil order (k_trans), C1 through the pencil walk, against the oracle."""
    w, cells, m, kind = grid
    six = cells[:6]                       # A, B, C1, D1, C2, D2
    A, B, C1, D1, C2, D2 = range(6)
    dx = 1.0e7
    lengths, starts, ids, dz, ratio = [], [], [], [], []
    for c, d in ((C1, D1), (C2, D2)):
        p = [A, A, A, B, c, d, d, d]
        starts.append(len(ids))
        lengths.append(len(p))
        ids += p
        dz += [2 * dx if x in (A, B) else dx for x in p]
        ratio += [0.0, 0.5, 0.5, 0.5, 1.0, 1.0, 1.0, 0.0]
    pen = (np.array(lengths, np.uint32), np.array(starts, np.uint64), np.array(ids, np.uint32),
           np.array(dz, np.float32), np.array(ratio, np.float32))
    for dim in range(3):
        _say(f"two-pencil cells dim={dim}")
        ctx = _ctx(w)
        ctx.upload_cells(six)
        T.set_pencils(ctx, dim, pen)
        T.trans_map_1d(ctx, dim, DT)
        want = O.trans_map_1d(m, six, pen, dim, DT, w.min_value)
        cases.assert_cells_equal(ctx.download_cells(), want, f"two pencils dim {dim}")
        # the shared cells did receive mass from both pencils
        assert not np.array_equal(want[A][1], six[A][1])
        ctx.close()


def test_graph_steps_interleaved_with_translation(grid):
    """Acceleration steps replayed from captured graphs (vacc_options.use_graph) alternating
    with x, y, z translations: the translation keeps the captured variants (it moves no
    buffer); every state equals that of an eager context doing the same."""
    w, cells, m, kind = grid
    fields = [V.make_fields(c.B, c.V, c.dt) for c in w.cells]
    mesh = V.make_mesh(w.grid_length, w.vmin, w.vmax)
    sp = V.make_species(w.mass, w.charge, w.min_value, w.add_width)
    eager = V.Context(mesh, sp, V.make_options(w.method))
    graph = V.Context(mesh, sp, V.make_options(w.method, use_graph=True))
    pens = [T.build_pencils_uniform(DIMS, (True, True, True), kind, DX, d) for d in range(3)]
    for c in (eager, graph):
        c.upload_cells(cells)
        c.set_fields(fields)
        for d in range(3):
            T.set_pencils(c, d, pens[d])
    for step in range(8):
        _say(f"graph + translation step {step}")
        for c in (eager, graph):
            c.step(step % 3, use_moments=True, adjust=True)
            T.trans_map_1d(c, step % 3, DT)
        if step in (3, 7):
            cases.assert_cells_equal(graph.download_cells(), eager.download_cells(), f"step {step}")
    eager.close()
    graph.close()


@pytest.mark.parametrize("storage", [V.STORE_F32, V.STORE_BF16, V.STORE_F16])
@pytest.mark.parametrize("periodic", [(True, True, True), (False, True, True)])
def test_long_pencils_bit_exact(periodic, storage):
    """Pencils of 130 cells: the pencil walk takes them in segments of at most 58 positions
    whose stencils overlap; boundary and inactive cells split some lines.  In FP32 the wrap
    cells of the periodic lines are walked in two launch levels; in the 16-bit forms they go
    to k_trans, and the oracle's output is requantized for the cells the pencils write."""
    dims = (130, 2, 1)
    n = int(np.prod(dims))
    w = W.config(0, ncells=n)
    cells = cases.build_cells(w)
    m = cases.oracle_mesh(w)
    rng = np.random.default_rng(11)
    kind = np.zeros(n, np.uint8)
    kind[rng.random(n) < 0.03] = T.CELL_BOUNDARY
    kind[67] = T.CELL_INACTIVE
    dx = (1.0e7, 9.0e6, 1.1e7)
    st = {V.STORE_F32: None, V.STORE_BF16: O.ST_BF16, V.STORE_F16: O.ST_F16}[storage]
    for dim in (0, 1):
        pen = T.build_pencils_uniform(dims, periodic, kind, dx, dim)
        _say(f"long pencils dim={dim} periodic={periodic} storage={storage} pencils={len(pen[0])} "
             f"longest={int(pen[0].max())}")
        ctx = _ctx(w, storage)
        ctx.upload_cells(cells)
        held = ctx.download_cells()
        T.set_pencils(ctx, dim, pen)
        T.trans_map_1d(ctx, dim, DT)
        out = O.trans_map_1d(m, held, pen, dim, DT, w.min_value)
        written = set(int(c) for c, r in zip(pen[2], pen[4]) if r > 0 and c != 0xFFFFFFFF)
        want = out if st is None else [(g, O.requantize(f, st) if i in written else f) for i, (g, f) in enumerate(out)]
        cases.assert_cells_equal(ctx.download_cells(), want, f"long dim {dim} storage {storage}")
        ctx.close()


def test_pencil_with_empty_positions(grid):
    """A pencil position without a cell (VACC_INVALID_ID) reads as zero blocks and is never
    written; its neighbours are still walked."""
    w, cells, m, kind = grid
    ten = cells[:10]
    ids = [0, 1, 2, 3, 0xFFFFFFFF, 4, 5, 6, 7, 8, 9, 0xFFFFFFFF]
    ratio = [0.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 0.0]
    pen = (np.array([len(ids)], np.uint32), np.array([0], np.uint64), np.array(ids, np.uint32),
           np.full(len(ids), 1.0e7, np.float32), np.array(ratio, np.float32))
    for dim in range(3):
        _say(f"empty positions dim={dim}")
        ctx = _ctx(w)
        ctx.upload_cells(ten)
        T.set_pencils(ctx, dim, pen)
        T.trans_map_1d(ctx, dim, DT)
        want = O.trans_map_1d(m, ten, pen, dim, DT, w.min_value)
        cases.assert_cells_equal(ctx.download_cells(), want, f"empty positions dim {dim}")
        ctx.close()


def test_shortest_pencils(grid):
    """Pencils of 2*VACC_TRANS_STENCIL+1 positions (one propagated cell, five distinct cells):
    the pencil walk's smallest window, one source; with ratios 1/2 on a second pencil that
    shares the stencil cells but writes another cell."""
    w, cells, m, kind = grid
    seven = cells[:7]
    ids = [0, 1, 2, 3, 4, 1, 2, 5, 3, 6]
    ratio = [0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.5, 0.0, 0.0]
    dz = [1.0e7, 1.1e7, 0.9e7, 1.0e7, 1.2e7, 1.1e7, 0.9e7, 1.3e7, 1.0e7, 0.8e7]
    pen = (np.array([5, 5], np.uint32), np.array([0, 5], np.uint64), np.array(ids, np.uint32),
           np.array(dz, np.float32), np.array(ratio, np.float32))
    for dim in range(3):
        _say(f"shortest pencils dim={dim}")
        ctx = _ctx(w)
        ctx.upload_cells(seven)
        T.set_pencils(ctx, dim, pen)
        T.trans_map_1d(ctx, dim, DT)
        want = O.trans_map_1d(m, seven, pen, dim, DT, w.min_value)
        cases.assert_cells_equal(ctx.download_cells(), want, f"shortest dim {dim}")
        assert not np.array_equal(want[2][1], seven[2][1])
        ctx.close()


def _random_pencils(rng, ncells, npencils):
    """Pencils of 5..24 positions over ncells cells: cells repeat within a pencil (next to each
    other or apart) and across pencils, some positions are empty, ratios are 0, 1/2 or 1 and
    widths vary by +-20 %."""
    lengths, starts, ids, dz, ratio = [], [], [], [], []
    for _ in range(npencils):
        L = int(rng.integers(5, 25))
        starts.append(len(ids))
        lengths.append(L)
        for q in range(L):
            r = rng.random()
            ids.append(0xFFFFFFFF if r < 0.05 else int(rng.integers(0, ncells)))
            dz.append(float(1.0e7 * rng.uniform(0.8, 1.2)))
            edge = q < 1 or q >= L - 1
            ratio.append(0.0 if edge else float(rng.choice([0.0, 0.5, 1.0, 1.0])))
    return (np.array(lengths, np.uint32), np.array(starts, np.uint64), np.array(ids, np.uint32),
            np.array(dz, np.float32), np.array(ratio, np.float32))


def _own_cell_pencils(rng, ncells, npencils):
    """Pencils of 8..16 positions, each over its own ncells // npencils cells: a cell appears
    once, twice (next to itself or apart) or more often in one pencil only, so that the pencil
    walk takes the cells it can (one position, or two at least three apart) and k_trans the
    others."""
    per = ncells // npencils
    lengths, starts, ids, dz, ratio = [], [], [], [], []
    for p in range(npencils):
        L = int(rng.integers(8, 17))
        starts.append(len(ids))
        lengths.append(L)
        for q in range(L):
            ids.append(p * per + int(rng.integers(0, per)))
            dz.append(float(1.0e7 * rng.uniform(0.8, 1.2)))
            ratio.append(0.0 if q < 1 or q >= L - 1 else float(rng.choice([0.0, 1.0, 1.0, 1.0])))
    return (np.array(lengths, np.uint32), np.array(starts, np.uint64), np.array(ids, np.uint32),
            np.array(dz, np.float32), np.array(ratio, np.float32))


@pytest.mark.parametrize("kind_, seed", [("shared", 21), ("shared", 22), ("shared", 23), ("own", 31), ("own", 32),
                                         ("own", 33)])
def test_random_pencils_bit_exact(grid, kind_, seed):
    """Random pencils: "shared" ones write cells at several positions and from several pencils,
    with empty positions and ratios 0, 1/2, 1; "own" ones give each pencil its own cells, so
    that many cells are written at one or two positions of one pencil.  Widths vary by
    +-20 %.  The split between the pencil walk and k_trans, the walk's carries and its
    segments, against the oracle in every dimension."""
    w, cells, m, kind = grid
    rng = np.random.default_rng(seed)
    if kind_ == "shared":
        sub = cells[:12]
        pen = _random_pencils(rng, len(sub), 6)
    else:
        sub = cells[:24]
        pen = _own_cell_pencils(rng, len(sub), 3)
    for dim in range(3):
        _say(f"random pencils {kind_} seed={seed} dim={dim} positions={len(pen[2])}")
        ctx = _ctx(w)
        ctx.upload_cells(sub)
        T.set_pencils(ctx, dim, pen)
        T.trans_map_1d(ctx, dim, DT)
        want = O.trans_map_1d(m, sub, pen, dim, DT, w.min_value)
        cases.assert_cells_equal(ctx.download_cells(), want, f"random {kind_} seed {seed} dim {dim}")
        ctx.close()


def test_coarse_cells_shared_by_finer_pencils(grid):
    """Refinement: coarse cells (width 2 dx) that two pencils of finer cells pass, each once at
    a run position with target ratio 1/2, and a third pencil that writes one of them and a
    fine cell again.
    The pencil walk takes them in launch levels (a later pencil's sum starts from the block
    the earlier one stored); FP32 and BF16 (where k_trans takes them) against the oracle."""
    w, cells, m, kind = grid
    sub = cells[:14]
    C1, C2 = 0, 1
    dx = 1.0e7
    # the third pencil writes C1 and cell 4 again (after the first two); every written cell's
    # stencils span at most 12 cells, so that k_trans can take them in the 16-bit form
    rows = [[2, 3, 4, C1, 5, C2, 6, 7], [8, 9, 10, C1, 11, C2, 12, 13], [3, 9, C1, 4, 10, 5]]
    lengths, starts, ids, dz, ratio = [], [], [], [], []
    for r in rows:
        starts.append(len(ids))
        lengths.append(len(r))
        for q, c in enumerate(r):
            ids.append(c)
            dz.append(2 * dx if c in (C1, C2) else dx)
            edge = q < 2 or q >= len(r) - 2
            ratio.append(0.0 if edge else (0.5 if c in (C1, C2) else 1.0))
    pen = (np.array(lengths, np.uint32), np.array(starts, np.uint64), np.array(ids, np.uint32),
           np.array(dz, np.float32), np.array(ratio, np.float32))
    for storage in (V.STORE_F32, V.STORE_BF16):
        st = {V.STORE_F32: None, V.STORE_BF16: O.ST_BF16}[storage]
        for dim in range(3):
            _say(f"coarse cells storage={storage} dim={dim}")
            ctx = _ctx(w, storage)
            ctx.upload_cells(sub)
            held = ctx.download_cells()
            T.set_pencils(ctx, dim, pen)
            T.trans_map_1d(ctx, dim, DT)
            out = O.trans_map_1d(m, held, pen, dim, DT, w.min_value)
            written = set(int(c) for c, r in zip(pen[2], pen[4]) if r > 0)
            want = out if st is None else [(g, O.requantize(f, st) if i in written else f) for i, (g, f) in enumerate(out)]
            cases.assert_cells_equal(ctx.download_cells(), want, f"coarse storage {storage} dim {dim}")
            assert not np.array_equal(out[C1][1], held[C1][1])
            ctx.close()


def test_wide_stencils_of_coarse_cells():
    """A coarse cell passed by three pencils of its own finer cells has stencils over 14 cells:
    the pencil walk takes it (FP32), the per-cell schedule of the 16-bit forms cannot."""
    w = W.config(0, ncells=18)
    sub = cases.build_cells(w)
    m = cases.oracle_mesh(w)
    C = 0
    rows = [[1, 2, 3, C, 4, 5, 6], [7, 8, 9, C, 10, 11, 12], [13, 14, C, 15, 16, 17]]
    lengths, starts, ids, dz, ratio = [], [], [], [], []
    for r in rows:
        starts.append(len(ids))
        lengths.append(len(r))
        for q, c in enumerate(r):
            ids.append(c)
            dz.append(2.0e7 if c == C else 1.0e7)
            ratio.append(0.0 if (q < 2 or q >= len(r) - 2) else (1.0 / 3.0 if c == C else 1.0))
    pen = (np.array(lengths, np.uint32), np.array(starts, np.uint64), np.array(ids, np.uint32),
           np.array(dz, np.float32), np.array(ratio, np.float32))
    for dim in range(3):
        _say(f"wide stencils dim={dim}")
        ctx = _ctx(w)
        ctx.upload_cells(sub)
        T.set_pencils(ctx, dim, pen)
        T.trans_map_1d(ctx, dim, DT)
        want = O.trans_map_1d(m, sub, pen, dim, DT, w.min_value)
        cases.assert_cells_equal(ctx.download_cells(), want, f"wide dim {dim}")
        ctx.close()
    ctx = _ctx(w, V.STORE_BF16)
    ctx.upload_cells(sub)
    with pytest.raises(V.VaccError, match="more than 12 cells"):
        T.set_pencils(ctx, 0, pen)
    ctx.close()

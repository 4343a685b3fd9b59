"""GPU parity of the 16-bit storage forms (BF16 / FP16 codes with a per-block FP32 scale).

The oracle runs the same FP32 remap and stores every block a pass writes through the same
codec (oracle.requantize: encode, then decode); the library must produce the identical FP32
values after decoding, bit for bit, and identical block lists.
"""
import numpy as np
import pytest

import oracle as O
from tests import cases
from vlasiator_amd import vacc as V
from vlasiator_amd import workloads as W

pytestmark = pytest.mark.gpu

STORES = [(V.STORE_BF16, O.ST_BF16), (V.STORE_F16, O.ST_F16)]


def _ctx(w, method, store):
    mesh = V.make_mesh(w.grid_length, w.vmin, w.vmax)
    sp = V.make_species(w.mass, w.charge, w.min_value, w.add_width)
    return V.Context(mesh, sp, V.make_options(method, storage=store))


def _say(msg):
    print(msg, flush=True)


def _quantized(cells, st):
    return [(g, O.requantize(f, st)) for g, f in cells]


@pytest.fixture(scope="module")
def cfg0():
    w = W.config(0, ncells=8)
    return w, cases.build_cells(w)


@pytest.mark.parametrize("store,st", STORES)
def test_upload_download_is_the_codec(cfg0, store, st):
    w, cells = cfg0
    _say(f"storage {store} upload/download")
    ctx = _ctx(w, O.PPM, store)
    ctx.upload_cells(cells)
    cases.assert_cells_equal(ctx.download_cells(), _quantized(cells, st), f"store {store} round trip")
    ctx.close()


@pytest.mark.parametrize("store,st", STORES)
@pytest.mark.parametrize("method", [O.PPM, O.PQM])
@pytest.mark.parametrize("dim", [0, 1, 2])
def test_map_1d_bit_exact_16bit(cfg0, store, st, method, dim):
    w, cells = cfg0
    isx = cases.oracle_intersections(w, map_order=dim)
    _say(f"storage {store} map_1d dim={dim} method={method}")
    ctx = _ctx(w, method, store)
    ctx.upload_cells(cells)
    ctx.set_intersections(isx)
    ctx.map_1d(dim)
    m = cases.oracle_mesh(w)
    want = []
    for i, (g, f) in enumerate(_quantized(cells, st)):
        g2, f2, _ = O.map_1d(m, g, f, isx[i, 4 * dim:4 * dim + 4], dim, method, w.min_value)
        want.append((g2, O.requantize(f2, st)))
    cases.assert_cells_equal(ctx.download_cells(), want, f"store {store} map_1d dim {dim} method {method}")
    ctx.close()


@pytest.mark.parametrize("store,st", STORES)
def test_pqm_steps_with_adjust_bit_exact_16bit(store, st):
    """Multi-peak PQM: three steps of (accelerate, adjust) with each pass stored 16-bit."""
    w = W.config(2, ncells=6)
    cells = cases.build_cells(w)
    m = cases.oracle_mesh(w)
    ctx = _ctx(w, O.PQM, store)
    ctx.upload_cells(cells)
    ref = _quantized(cells, st)
    for step in range(3):
        mo = (step + 1) % 3
        isx = cases.oracle_intersections(w, mo)
        _say(f"storage {store} step {step}")
        ctx.set_intersections(isx)
        ctx.accelerate(mo)
        loss = ctx.adjust(len(cells))
        new, want_loss = [], []
        for i, (g, f) in enumerate(ref):
            g2, f2, _ = O.accelerate_cell(m, g, f, isx[i], mo, O.PQM, w.min_value, storage=st)
            g3, f3, l3 = O.adjust(m, g2, f2, w.min_value, w.add_width)
            new.append((g3, f3))
            want_loss.append(l3)
        ref = new
        cases.assert_cells_equal(ctx.download_cells(), ref, f"store {store} step {step}")
        assert np.allclose(loss, want_loss, rtol=1e-12, atol=0)
    ctx.close()


@pytest.mark.parametrize("store,st", STORES)
def test_moments_and_init_16bit(store, st):
    w = W.config(2, ncells=8)
    cells = cases.build_cells(w)
    m = cases.oracle_mesh(w)
    ctx = _ctx(w, O.PQM, store)
    _say(f"storage {store} init_peaks + moments")
    ctx.init_peaks([W.to_vacc_peaks(c) for c in w.cells], W.boxes(w))
    got = ctx.download_cells()
    step = 2.0 ** -8 if st == O.ST_BF16 else 2.0 ** -11
    for i, ((g1, f1), (g2, f2)) in enumerate(zip(got, cells)):
        assert np.array_equal(g1, np.sort(g2)), f"cell {i} block sets differ"
        g2s, f2s = cases.sort_cell(g2, f2)
        # device exp/pow vs host libm (a few FP32 ulp) moved through the codec: within two code
        # steps of each block's maximum
        bmax = np.abs(f2s).max(axis=1, keepdims=True)
        assert np.all(np.abs(f1 - f2s) <= 2 * step * bmax + 1e-30), f"cell {i}"
    mom = ctx.moments(len(cells))
    want = np.array([O.moments(m, g, f) for g, f in got])
    assert np.allclose(mom[:, 0], want[:, 0], rtol=1e-12, atol=0)
    ctx.close()


@pytest.mark.parametrize("store", [V.STORE_F32, V.STORE_BF16])
def test_state_beyond_2_to_32_thread_indices(store):
    """About 9e7 blocks: kernels with one wave per block then index past 2^32 threads
    (64 x 67 108 864).  The initial state's density must equal the sum of the sampled
    peaks (FP32 store as the reference for BF16), and an adjust right after the initial
    condition (content flags computed from f for every block) must keep every block (FP32),
    or all but the blocks whose content sat within a code step of the threshold (BF16)."""
    w = W.populations(4, ncells=24000)[0]
    _say(f"storage {store}: large initial state")
    ctx = _ctx(w, O.PQM, store)
    ctx.init_peaks([W.to_vacc_peaks(c) for c in w.cells], W.boxes(w))
    n0 = ctx.total_blocks()
    assert n0 > (1 << 32) // 64
    dens = ctx.moments(len(w.cells))[:, 0]
    want = np.array([sum(p.n for p in c.peaks) for c in w.cells])
    # sampled Maxwellians on the mesh carry their density to well within 1e-3
    assert np.all(np.abs(dens - want) <= 1e-3 * want)
    loss = ctx.adjust(len(w.cells))
    after = ctx.moments(len(w.cells))[:, 0]
    if store == V.STORE_F32:
        assert ctx.total_blocks() == n0
        assert np.allclose(after, dens, rtol=1e-12, atol=0)
    else:
        # the initial keep set follows the FP32 samples; the adjust sees the stored codes, so
        # blocks whose only content was within a code step of the threshold may go (with
        # their halo), a few in 1e4; their mass is the reported loss
        assert n0 >= ctx.total_blocks() >= n0 - n0 // 1000
        assert np.allclose(after + loss, dens, rtol=1e-9, atol=0)
    ctx.close()

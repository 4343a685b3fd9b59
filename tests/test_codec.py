"""CPU checks of the oracle's 16-bit block storage (BF16 / FP16 codes with a per-block FP32
scale) and of its FP64 build, the reference run the storage forms are measured against.

The element conversions are pinned against independent definitions: numpy's IEEE binary16
cast for FP16, and for BF16 the nearest of the two bracketing BF16 values chosen in FP64
(ties to the even code).  The block encoding is pinned through the properties that define
it: exact power-of-two normalisation, kept block mass, per-value error within one code
step of the block maximum, zero blocks and the sign-mixed fallback.
"""
import numpy as np
import pytest

import oracle as O
from tests import cases
from vlasiator_amd import workloads as W


def _edge_values():
    rng = np.random.default_rng(7)
    x = (rng.standard_normal(100000) * 10.0 ** rng.integers(-12, 3, 100000)).astype(np.float32)
    sub = (np.float32(2.0 ** -24) * (np.arange(-4000, 4000, dtype=np.float32) / np.float32(97.0))).astype(np.float32)
    ties = np.float32(2.0 ** -14) * (1 + (2 * np.arange(1, 200, dtype=np.float32) + 1) * np.float32(2.0 ** -11))
    special = np.array([0.0, -0.0, 2.0 ** -25, 3 * 2.0 ** -26, 2.0 ** -24, 65504.0, 65519.0, 65520.0, 1.0, -1.0],
                       np.float32)
    return np.concatenate([x, sub, ties.astype(np.float32), special])


def test_fp16_conversion_is_ieee_round_to_nearest_even():
    x = _edge_values()
    got = O.convert16(x, O.ST_F16)
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)


def test_bf16_conversion_is_nearest_with_ties_to_even():
    x = _edge_values()
    got = O.convert16(x, O.ST_BF16).astype(np.uint32)
    u = x.view(np.uint32)
    lo = u & 0xFFFF0000                      # truncation toward zero (same sign)
    hi = lo + 0x10000                        # next BF16 away from zero
    xl = lo.view(np.float32).astype(np.float64)
    xh = hi.view(np.float32).astype(np.float64)
    xd = x.astype(np.float64)
    dl, dh = np.abs(xd - xl), np.abs(xh - xd)
    pick_hi = (dh < dl) | ((dh == dl) & (((lo >> 16) & 1) == 1))
    want = np.where(pick_hi, hi, lo) >> 16
    assert np.array_equal(got, want)


def _blocks(n=400, seed=3):
    rng = np.random.default_rng(seed)
    # Maxwellian-like blocks: values spanning several decades inside a block
    centre = rng.uniform(-3, 3, (n, 3))
    width = rng.uniform(0.5, 4.0, n)
    i = np.arange(4)
    grid = np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(64, 3)[:, ::-1]   # e = i + 4j + 16k
    r2 = ((grid[None, :, :] - centre[:, None, :]) ** 2).sum(-1)
    amp = 10.0 ** rng.uniform(-14, -10, n)
    return (amp[:, None] * np.exp(-r2 / width[:, None] ** 2)).astype(np.float32)


@pytest.mark.parametrize("storage,step", [(O.ST_BF16, 2.0 ** -8), (O.ST_F16, 2.0 ** -11)])
def test_block_encoding_keeps_mass_and_precision(storage, step):
    for x in _blocks():
        code, sc = O.encode_block(x, storage)
        d = O.decode_block(code, sc, storage)
        m = np.abs(x).max()
        # mass: the decoded sum equals the block's FP32 sum up to the rounding of 64 products
        assert abs(float(d.astype(np.float64).sum()) - float(x.astype(np.float64).sum())) <= 1e-6 * x.sum()
        # each value within a code step of the block maximum (plus the FP16 subnormal floor)
        tol = step * m + (2.0 ** -24 * m if storage == O.ST_F16 else 0.0)
        assert np.all(np.abs(d.astype(np.float64) - x) <= tol)
        # the scale lies within a factor 4 of the block maximum
        assert 0.25 * m <= sc <= 4.0 * m


@pytest.mark.parametrize("storage", [O.ST_BF16, O.ST_F16])
def test_block_encoding_special_cases(storage):
    z = np.zeros(64, np.float32)
    code, sc = O.encode_block(z, storage)
    assert sc == 0.0 and not code.any()
    # a sign-mixed block with zero sum keeps the power-of-two scale of its maximum
    x = np.zeros(64, np.float32)
    x[0], x[1] = np.float32(3e-12), np.float32(-3e-12)
    code, sc = O.encode_block(x, storage)
    _, ex = np.frexp(np.float32(3e-12))
    assert sc == np.float32(np.ldexp(1.0, int(ex)))
    d = O.decode_block(code, sc, storage)
    assert d[0] == -d[1] and abs(d[0] - 3e-12) <= 2.0 ** -8 * 3e-12
    # tiny blocks (FP32 subnormal maximum) normalise exactly instead of overflowing 1/m
    t = np.zeros(64, np.float32)
    t[5] = np.float32(3e-41)
    code, sc = O.encode_block(t, storage)
    d = O.decode_block(code, sc, storage)
    assert np.isfinite(d).all() and d[5] == t[5]


def test_block_csum_matches_exact_sum_to_rounding():
    for x in _blocks(50, seed=5):
        s = float(O.block_csum(x))
        exact = float(np.sum(x.astype(np.float64)))
        assert abs(s - exact) <= 64 * np.finfo(np.float32).eps * np.abs(x).sum()


def test_fp64_oracle_agrees_with_fp32_oracle():
    """The Realf = double build: same block lists for a remap (they depend on the GIDs and
    intersections only), f within FP32 rounding of the FP32 build, and mass kept to FP64
    rounding where the FP32 build keeps it to FP32 rounding."""
    w = W.config(0, ncells=4)
    cells = cases.build_cells(w)
    m = cases.oracle_mesh(w)
    isx = cases.oracle_intersections(w, 1)
    for i, (g, f) in enumerate(cells):
        g32, f32, _ = O.accelerate_cell(m, g, f, isx[i], 1, O.PPM, w.min_value)
        g64, f64, _ = O.accelerate_cell(m, g, f.astype(np.float64), isx[i], 1, O.PPM, w.min_value, fp64=True)
        assert np.array_equal(g32, g64)
        assert f64.dtype == np.float64
        assert np.abs(f32 - f64).sum() <= 1e-4 * np.abs(f64).sum()
        n0 = O.moments(m, g, f)[0]
        n64 = O.moments(m, g64, f64, fp64=True)[0]
        assert abs(n64 - n0) <= 1e-12 * n0


def test_requantized_accelerate_keeps_density():
    """PQM steps with each pass stored as BF16 / FP16: density within 1e-6 of the FP64 run."""
    w = W.config(2, ncells=2)
    cells = cases.build_cells(w)
    m = cases.oracle_mesh(w)
    for i, (g, f) in enumerate(cells):
        runs = {}
        for key, fp64, st in (("f64", True, O.ST_F32), ("bf16", False, O.ST_BF16), ("f16", False, O.ST_F16)):
            gg, ff = g, (f.astype(np.float64) if fp64 else O.requantize(f, st))
            for s in range(3):
                isx = cases.oracle_intersections(w, s % 3)
                gg, ff, _ = O.accelerate_cell(m, gg, ff, isx[i], s % 3, O.PQM, w.min_value, fp64=fp64, storage=st)
                gg, ff, _ = O.adjust(m, gg, ff, w.min_value, w.add_width, fp64=fp64)
            runs[key] = O.moments(m, gg, ff, fp64=fp64)
        for key in ("bf16", "f16"):
            assert abs(runs[key][0] - runs["f64"][0]) <= 1e-6 * runs["f64"][0], key

"""CPU check of the device reconstruction source: vlasiator_amd/csrc/vacc_recon.h is compiled
for the host (same flags as the library: no FP contraction) and streamed over long columns;
every coefficient must equal the oracle's bit for bit (the oracle is itself pinned to the
reference's reconstruction headers by the golden vectors)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "recon_host_check.hip")
OUT = os.path.join(HERE, "_build", "librecon_host_check.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    deps = [SRC, os.path.join(HERE, "..", "vlasiator_amd", "csrc", "vacc_recon.h"),
            os.path.join(HERE, "..", "vlasiator_amd", "csrc", "vacc_common.h")]
    if not os.path.exists(OUT) or os.path.getmtime(OUT) < max(os.path.getmtime(d) for d in deps):
        subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared",
                        "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                        SRC, "-o", OUT], check=True, capture_output=True)
    L = C.CDLL(OUT)
    L.recon_stream.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_float, C.c_void_p]
    return L


def columns(seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(40):
        n = int(rng.integers(4, 120))
        kind = rng.integers(0, 4)
        x = np.linspace(-3, 3, n)
        if kind == 0:
            v = 10 ** rng.uniform(-17, -9, n)
        elif kind == 1:
            v = 1e-10 * np.exp(-((x - rng.uniform(-1, 1)) / rng.uniform(0.3, 2)) ** 2)
        elif kind == 2:
            v = np.where(rng.random(n) < 0.4, 0.0, 10 ** rng.uniform(-16, -10, n))
        else:
            v = 1e-15 * rng.uniform(0.3, 3, n)
        col = np.zeros(n + 8, np.float32)
        col[4:n + 4] = v
        out.append(col)
    return out


@pytest.mark.parametrize("method", [O.PLM, O.PPM, O.PQM])
def test_streamed_device_reconstruction_matches_oracle(lib, method):
    nc = O.ncoef(method)
    for col in columns(31 + method):
        n = col.size - 8
        got = np.zeros(n * nc, np.float32)
        lib.recon_stream(method, col.ctypes.data, n, np.float32(1e-15), got.ctypes.data)
        st = np.stack([col[k:k + 9] for k in range(n)])
        want = O.coeff_batch(method, st, 1e-15).reshape(-1)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def _pqm_stress_columns(rng, count):
    """Columns where the PQM monotonicity test is close to its boundary: smooth profiles with
    relative noise from 1e-7 to 1e-1, plateaus with tiny steps, near-linear ramps, values near
    the FP32 subnormal range, and single spikes."""
    out = []
    for _ in range(count):
        n = int(rng.integers(8, 64))
        x = np.linspace(-1, 1, n)
        kind = int(rng.integers(0, 6))
        if kind == 0:
            base = np.exp(-((x - rng.uniform(-0.5, 0.5)) / rng.uniform(0.2, 1.0)) ** 2)
            v = base * (1 + 10 ** rng.uniform(-7, -1) * rng.standard_normal(n))
        elif kind == 1:
            v = 1.0 + 10 ** rng.uniform(-7, -3) * np.round(rng.standard_normal(n))
        elif kind == 2:
            v = 1.0 + rng.uniform(-1, 1) * x + 10 ** rng.uniform(-8, -4) * rng.standard_normal(n)
        elif kind == 3:
            v = np.exp(-((x / rng.uniform(0.1, 0.5)) ** 2)) * 10 ** rng.uniform(-36, -28)
        elif kind == 4:
            v = np.full(n, 1e-3)
            v[rng.integers(0, n)] = 1.0
        else:
            v = np.abs(np.sin(rng.uniform(1, 20) * x)) + 10 ** rng.uniform(-6, -2)
        col = np.zeros(n + 8, np.float32)
        col[4:n + 4] = np.abs(v) * 10 ** rng.uniform(-14, -9) if kind != 3 else np.abs(v)
        out.append(col)
    return out


def test_pqm_near_the_monotonicity_boundary_keeps_bit_parity(lib):
    """On columns built to sit near the PQM monotonicity test's boundary (where a rounding
    difference would flip the repair), every device coefficient equals the oracle's bit for
    bit, and the reference's own headers' where oracle/_ref is built."""
    rng = np.random.default_rng(2024)
    cols = _pqm_stress_columns(rng, 3000)
    cells = 0
    for col in cols:
        n = col.size - 8
        got = np.zeros(n * 5, np.float32)
        lib.recon_stream(O.PQM, col.ctypes.data, n, np.float32(1e-15), got.ctypes.data)
        st = np.stack([col[k:k + 9] for k in range(n)])
        want = O.coeff_batch(O.PQM, st, 1e-15).reshape(-1)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        if O.ref_lib() is not None:
            ref = O.ref_coeff_batch(O.PQM, st, 1e-15).reshape(-1)
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
        cells += n
    assert cells > 100000


def test_div3_equals_ieee_division_for_every_float(lib):
    """vacc::div3 (FMA-corrected product) against x / 3 over all 2^32 bit patterns."""
    import concurrent.futures as cf
    lib.div3_mismatches.restype = C.c_longlong
    lib.div3_mismatches.argtypes = [C.c_ulonglong, C.c_ulonglong]
    parts = 16
    step = (1 << 32) // parts
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        bad = sum(ex.map(lambda k: lib.div3_mismatches(k * step, (k + 1) * step), range(parts)))
    assert bad == 0


def test_lagrangian_index_equals_truncated_quotient(lib):
    """vacc::lagr_index against (int)(x / dk): random velocities and spacings, plus quotients
    placed within a few ulp of integers (where the product and the quotient can straddle)."""
    lib.lagr_mismatches.restype = C.c_longlong
    lib.lagr_mismatches.argtypes = [C.c_void_p, C.c_void_p, C.c_longlong]
    rng = np.random.default_rng(17)
    n = 2_000_000
    dk = (2e4 * rng.uniform(0.5, 2.0, n)).astype(np.float32)
    x = rng.uniform(-8e6, 8e6, n).astype(np.float32)
    # adversarial: x = (k + tiny) * dk for integers k, nudged by a few ulp either way
    k = rng.integers(-400, 400, n).astype(np.float32)
    xa = (k * dk).astype(np.float32)
    xa = (xa.view(np.int32) + rng.integers(-4, 5, n).astype(np.int32)).view(np.float32)
    for xs in (x, xa):
        xs = np.ascontiguousarray(xs)
        assert lib.lagr_mismatches(xs.ctypes.data, dk.ctypes.data, n) == 0

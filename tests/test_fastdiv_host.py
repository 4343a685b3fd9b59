"""CPU checks of the division-free GID splitting (vlasiator_amd/csrc/vacc_common.h: FastDiv,
gid_xyz) that the column, adjust and moment kernels use: floor(g / d) as a multiply-high and a
shift must equal integer division for every 32-bit g.  Checked against numpy floor division on
every divisor 1..4096 with random and edge dividends (0, 1, multiples of d and their
neighbours, 2^31 +- 1, 2^32 - 1), on large divisors up to 2^32 - 1, including powers of two
and their neighbours, and on whole GID ranges of meshes up to 65535 x 65535 blocks per layer."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "fastdiv_host_check.cpp")
OUT = os.path.join(HERE, "_build", "libfastdiv_host_check.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    deps = [SRC, os.path.join(HERE, "..", "vlasiator_amd", "csrc", "vacc_common.h")]
    if not os.path.exists(OUT) or os.path.getmtime(OUT) < max(os.path.getmtime(d) for d in deps):
        subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared", SRC,
                        "-o", OUT], check=True, capture_output=True)
    L = C.CDLL(OUT)
    L.fastdiv_batch.argtypes = [C.c_uint32, C.c_void_p, C.c_int, C.c_void_p]
    L.gid_xyz_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    return L


def _div(lib, d, g):
    g = np.ascontiguousarray(g, np.uint32)
    q = np.zeros(max(g.size, 1), np.uint32)
    lib.fastdiv_batch(int(d), g.ctypes.data, g.size, q.ctypes.data)
    return q[:g.size]


def _dividends(d, rng, n=2000):
    d = int(d)
    edge = [0, 1, 2, d - 1, d, d + 1, 2 * d - 1, 2 * d, 2 ** 31 - 1, 2 ** 31, 2 ** 31 + 1, 2 ** 32 - 1]
    top = (2 ** 32 - 1) // d
    mult = [k * d for k in (top, top - 1, top // 2, top // 3)] + [k * d - 1 for k in (top, top // 2) if k]
    g = np.array([x for x in edge + mult if 0 <= x < 2 ** 32], np.uint64)
    return np.concatenate([g, rng.integers(0, 2 ** 32, n, dtype=np.uint64)]).astype(np.uint32)


def test_every_small_divisor(lib):
    rng = np.random.default_rng(1)
    for d in range(1, 4097):
        g = _dividends(d, rng, 200)
        assert np.array_equal(_div(lib, d, g), g // np.uint32(d)), d


def test_large_divisors(lib):
    rng = np.random.default_rng(2)
    ds = set(rng.integers(4097, 2 ** 32, 3000, dtype=np.uint64).tolist())
    for p in range(12, 32):
        ds.update({2 ** p - 1, 2 ** p, 2 ** p + 1})
    ds.update({65535, 65535 * 65535, 2 ** 32 - 1, 2 ** 32 - 2, 3 * 2 ** 30})
    for d in sorted(ds):
        d = int(d)
        if not 1 <= d < 2 ** 32:
            continue
        g = _dividends(d, rng, 500)
        assert np.array_equal(_div(lib, d, g), g // np.uint32(d)), d


@pytest.mark.parametrize("len3", [(1, 1, 1), (50, 50, 50), (75, 75, 75), (100, 100, 100), (400, 400, 400),
                                  (300, 7, 13), (65535, 65535, 1), (65535, 3, 21845), (1, 65535, 65535)])
def test_gid_xyz(lib, len3):
    rng = np.random.default_rng(3)
    nx, ny, nz = len3
    total = nx * ny * nz
    g = np.concatenate([np.arange(min(total, 5000)), total - 1 - np.arange(min(total, 5000)),
                        rng.integers(0, total, 20000)]).astype(np.uint32)
    lens = np.array(len3, np.uint32)
    xyz = np.zeros(3 * g.size, np.uint32)
    lib.gid_xyz_batch(lens.ctypes.data, g.ctypes.data, g.size, xyz.ctypes.data)
    xyz = xyz.reshape(-1, 3)
    g64 = g.astype(np.int64)
    assert np.array_equal(xyz[:, 0], g64 % nx)
    assert np.array_equal(xyz[:, 1], (g64 // nx) % ny)
    assert np.array_equal(xyz[:, 2], g64 // (nx * ny))

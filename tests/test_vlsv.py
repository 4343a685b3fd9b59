"""CPU checks of VLSV files: the oracle's restatement of the format against the documented
structure, libvacc's container writer and reader (host code, no GPU calls) against the
oracle, and the oracle's velocity-mesh renumbering against block coordinates.

What pins what: the footer's tags, mandatory attributes and offset values follow
doc/vlsv2/documentation.lyx:85-276 and are checked here with an independent XML parser
(xml.etree); the array list and attributes follow iowrite.cpp:141-318; the renumbering
of a smaller file velocity mesh (ioread.cpp:520-631) is pinned by the physical invariant
it must keep -- every block keeps its velocity coordinates.  The two header words (the
endianness word, then the footer offset) follow the public fmihpc/vlsv library, which the
reference tree does not hold: parity unpinned against files that library wrote.
"""
import os
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from oracle import vlsv as OV
from vlasiator_amd import vacc as V

pytestmark = pytest.mark.skipif(not os.path.exists(V.LIB_PATH), reason="libvacc.so not built")


def _arrays(rng):
    return [
        ("PARAMETER", {"name": "time"}, np.array([12.5])),
        ("VARIABLE", {"mesh": "SpatialGrid", "name": "vg_rho"}, rng.random(17)),
        ("VARIABLE", {"mesh": "SpatialGrid", "name": "vg_v"}, rng.random((17, 3)).astype(np.float32)),
        ("BLOCKIDS", {"mesh": "SpatialGrid", "name": "proton"}, rng.integers(0, 2 ** 31, 40).astype(np.uint32)),
        ("CELLSWITHBLOCKS", {"mesh": "SpatialGrid", "name": "proton"}, np.arange(1, 6, dtype=np.uint64)),
        ("COUNTS", {"name": "i16"}, rng.integers(-300, 300, (9, 2)).astype(np.int16)),
        ("SIGNED", {"name": "i64"}, rng.integers(-2 ** 40, 2 ** 40, 4).astype(np.int64)),
        ("BYTES", {"name": "u8"}, rng.integers(0, 255, 3).astype(np.uint8)),
        ("EMPTY", {"name": "none"}, np.zeros((0, 64), np.float32)),
    ]


def test_oracle_file_follows_the_documented_structure(tmp_path):
    p = tmp_path / "a.vlsv"
    arrs = _arrays(np.random.default_rng(1))
    OV.write(p, arrs)
    raw = p.read_bytes()
    endian, foot = struct.unpack_from("<QQ", raw, 0)
    assert endian == 0
    root = ET.fromstring(raw[foot:].decode())
    assert root.tag == "VLSV" and not root.attrib
    assert len(root) == len(arrs)
    # the footer sits in the file's tail, where tools/vlsvxml.sh looks for it
    assert b"<VLSV>" in raw[-10 * 1024:]
    for el, (tag, attrs, a) in zip(root, arrs):
        assert el.tag == tag
        for k in ("arraysize", "vectorsize", "datatype", "datasize"):
            assert k in el.attrib
        for k, v in attrs.items():
            assert el.attrib[k] == v
        off = int(el.text)
        nbytes = int(el.attrib["arraysize"]) * int(el.attrib["vectorsize"]) * int(el.attrib["datasize"])
        assert raw[off:off + nbytes] == np.ascontiguousarray(a).tobytes()
        assert el.attrib["datatype"] == {"f": "float", "u": "uint", "i": "int"}[a.dtype.kind]
        assert int(el.attrib["datasize"]) == a.dtype.itemsize


def test_library_writer_against_oracle_reader(tmp_path):
    from vlasiator_amd import vlsv as L
    arrs = _arrays(np.random.default_rng(2))
    p = tmp_path / "lib.vlsv"
    with L.Writer(str(p)) as w:
        for tag, attrs, a in arrs:
            w.write_array(tag, attrs, a)
    off, got = OV.read(p)
    assert len(got) == len(arrs)
    for (tag, attrs, a), (t2, at2, a2) in zip(arrs, got):
        assert t2 == tag and all(at2[k] == v for k, v in attrs.items())
        assert a2.dtype == a.dtype.newbyteorder("<") and np.array_equal(a2.reshape(a.shape), a)
    # same data region as the oracle's own file of the same arrays
    q = tmp_path / "oracle.vlsv"
    OV.write(q, arrs)
    assert p.read_bytes()[:off] == q.read_bytes()[:off]


@pytest.mark.parametrize("legacy", [False, True])
def test_library_reader_against_oracle_writer(tmp_path, legacy):
    from vlasiator_amd import vlsv as L
    arrs = _arrays(np.random.default_rng(3))
    p = tmp_path / "o.vlsv"
    OV.write(p, arrs, legacy_header=legacy)
    with L.Reader(str(p)) as r:
        for tag, attrs, a in arrs:
            n, vs, dt, ds = r.array_info(tag, attrs)
            assert (n, ds) == (a.shape[0], a.dtype.itemsize)
            assert vs == (1 if a.ndim == 1 else a.shape[1])
            b = r.read_array(tag, attrs)
            assert np.array_equal(b.reshape(a.shape), a)
        # partial reads
        a = arrs[2][2]
        assert np.array_equal(r.read_array("VARIABLE", {"name": "vg_v"}, 5, 7), a[5:12])
        # lookup by a subset of the attributes finds the first matching tag
        assert np.array_equal(r.read_array("VARIABLE", {"mesh": "SpatialGrid"}), arrs[1][2])


def test_library_reader_errors(tmp_path):
    from vlasiator_amd import vlsv as L
    p = tmp_path / "o.vlsv"
    OV.write(p, _arrays(np.random.default_rng(4)))
    with L.Reader(str(p)) as r:
        with pytest.raises(V.VaccError, match="no array"):
            r.array_info("VARIABLE", {"name": "missing"})
        with pytest.raises(V.VaccError, match="beyond"):
            r.read_array("VARIABLE", {"name": "vg_rho"}, 10, 8)
    bad = tmp_path / "bad.vlsv"
    bad.write_bytes(b"\x00" * 64)
    with pytest.raises(V.VaccError, match="not a VLSV file"):
        L.Reader(str(bad))
    with pytest.raises(V.VaccError):
        L.Reader(str(tmp_path / "absent.vlsv"))
    with L.Writer(str(tmp_path / "w.vlsv")) as w:
        with pytest.raises(V.VaccError, match="set by the writer"):
            w.write_array("X", {"arraysize": "3"}, np.zeros(3))
        with pytest.raises(V.VaccError, match="bad attribute"):
            w.write_array("X", {"name": 'a"b'}, np.zeros(3))


def _block_centres(gids, L, vmin, vmax):
    g = np.asarray(gids, np.int64)
    b = np.stack([g % L[0], (g // L[0]) % L[1], g // (L[0] * L[1])], -1)
    dvb = np.array([OV.cell_size(vmin[d], vmax[d], L[d]) * 4 for d in range(3)])
    return np.asarray(vmin, float) + (b + 0.5) * dvb


def test_oracle_renumbering_keeps_block_velocities(tmp_path):
    """A restart written on a 6 x 5 x 4 block mesh read into a 10 x 9 x 12 mesh with the
    same cell size: every block lands where its velocity is."""
    rng = np.random.default_rng(5)
    dv = 1.0e4
    small_L, small_min = (6, 5, 4), (-2.0e5 + 8 * dv, -2.0e5 + 4 * dv, -2.0e5 + 16 * dv)
    big_L, big_min = (10, 9, 12), (-2.0e5, -2.0e5, -2.0e5)
    small_max = tuple(small_min[d] + small_L[d] * 4 * dv for d in range(3))
    big_max = tuple(big_min[d] + big_L[d] * 4 * dv for d in range(3))
    cells = []
    for _ in range(3):
        g = np.sort(rng.choice(np.prod(small_L), 20, replace=False)).astype(np.uint32)
        cells.append((g, rng.random((20, 64)).astype(np.float32)))
    p = tmp_path / "r.vlsv"
    OV.write(p, OV.velocity_distribution_arrays(small_L, small_min, small_max, "proton", [7, 8, 9], cells))
    ids, got = OV.read_block_data(p, big_L, big_min, big_max, "proton")
    assert list(ids) == [7, 8, 9]
    for (g0, f0), (g1, f1) in zip(cells, got):
        np.testing.assert_allclose(_block_centres(g1, big_L, big_min, big_max),
                                   _block_centres(g0, small_L, small_min, small_max), rtol=0, atol=1e-6 * dv)
        assert np.array_equal(f0, f1)
    with pytest.raises(ValueError, match="shrink"):
        OV.read_block_data(p, (5, 9, 12), big_min, big_max, "proton")
    off_min = (big_min[0] + 2 * dv, big_min[1], big_min[2])   # 2 cells: not a block boundary
    with pytest.raises(ValueError, match="block boundary"):
        OV.read_block_data(p, big_L, off_min, tuple(off_min[d] + big_L[d] * 4 * dv for d in range(3)), "proton")


def test_large_arrays_split_over_io_threads(tmp_path):
    """Arrays above the 8 MiB split size go through concurrent positioned reads and writes."""
    from vlasiator_amd import vlsv as L
    rng = np.random.default_rng(6)
    a = rng.random((170001, 64)).astype(np.float32)          # 43.5 MB, not a multiple of 4 KiB
    b = rng.integers(0, 2 ** 32, 3000001, dtype=np.uint64)     # 24 MB
    p = tmp_path / "big.vlsv"
    with L.Writer(str(p)) as w:
        w.write_array("SMALL", {"name": "s"}, np.arange(3, dtype=np.int32))
        w.write_array("BLOCKVARIABLE", {"name": "a"}, a)
        w.write_array("IDS", {"name": "b"}, b)
    _, got = OV.read(p)
    assert np.array_equal(got[1][2], a) and np.array_equal(got[2][2], b)
    with L.Reader(str(p)) as r:
        assert np.array_equal(r.read_array("BLOCKVARIABLE", {"name": "a"}), a)
        assert np.array_equal(r.read_array("IDS", {"name": "b"}, 17, 2500000), b[17:2500017])


def test_library_reader_survives_corrupted_files(tmp_path):
    """Random byte changes and truncations of a valid file: opening and reading either work or
    fail with an error; the library never crashes and never reads outside the file."""
    from vlasiator_amd import vlsv as L
    arrs = _arrays(np.random.default_rng(8))
    good = tmp_path / "good.vlsv"
    OV.write(good, arrs)
    raw = bytearray(good.read_bytes())
    _, foot = struct.unpack_from("<QQ", raw, 0)
    rng = np.random.default_rng(9)
    bad = tmp_path / "bad.vlsv"
    opened = failed = 0
    for trial in range(400):
        b = bytearray(raw)
        kind = trial % 4
        if kind == 0:      # bytes of the footer
            for _ in range(rng.integers(1, 6)):
                b[rng.integers(foot, len(b))] = rng.integers(0, 256)
        elif kind == 1:    # footer digits (offsets and sizes)
            digits = [i for i in range(foot, len(b)) if chr(b[i]).isdigit()]
            for _ in range(rng.integers(1, 4)):
                b[digits[rng.integers(len(digits))]] = ord(str(rng.integers(0, 10)))
        elif kind == 2:    # the header words
            struct.pack_into("<Q", b, 8 * int(rng.integers(0, 2)), int(rng.integers(0, 2 ** 63)))
        else:              # truncation
            b = b[:rng.integers(0, len(b))]
        bad.write_bytes(bytes(b))
        try:
            r = L.Reader(str(bad))
        except V.VaccError:
            failed += 1
            continue
        opened += 1
        with r:
            for tag, attrs, a in arrs:
                try:
                    r.read_array(tag, attrs)
                except V.VaccError:
                    pass
    assert opened > 0 and failed > 0

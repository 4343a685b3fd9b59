"""ctypes binding of libvacc's VLSV file I/O (include/vacc_vlsv.h).

Writer / Reader mirror the vlsv::Writer / vlsv::ParallelReader calls that fmihpc/vlasiator
makes for velocity block data (iowrite.cpp:141-318, ioread.cpp:371-700); the block
functions move a Context's blocks between HBM and the file.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import vacc as V

INT, UINT, FLOAT = 0, 1, 2
SPATIAL_MESH = "SpatialGrid"

# symbols declared in include/vacc_vlsv.h
EXPORTS = [
    "vacc_vlsv_create", "vacc_vlsv_write_array", "vacc_vlsv_write_velocity_distribution", "vacc_vlsv_close",
    "vacc_vlsv_open", "vacc_vlsv_array_info", "vacc_vlsv_read_array", "vacc_vlsv_read_block_data",
    "vacc_vlsv_reader_close",
]

_bound = False


def lib():
    global _bound
    L = V.lib()
    if not _bound:
        vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32
        L.vacc_vlsv_create.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.vacc_vlsv_write_array.argtypes = [vp, C.c_char_p, u32, vp, u64, u64, i32, u64, vp]
        L.vacc_vlsv_write_velocity_distribution.argtypes = [vp, vp, C.c_char_p, vp]
        L.vacc_vlsv_close.argtypes = [vp]
        L.vacc_vlsv_open.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.vacc_vlsv_array_info.argtypes = [vp, C.c_char_p, u32, vp, C.POINTER(u64), C.POINTER(u64), C.POINTER(i32),
                                           C.POINTER(u64)]
        L.vacc_vlsv_read_array.argtypes = [vp, C.c_char_p, u32, vp, u64, u64, vp]
        L.vacc_vlsv_read_block_data.argtypes = [vp, vp, C.c_char_p, C.c_char_p, u64, u32, vp]
        L.vacc_vlsv_reader_close.argtypes = [vp]
        _bound = True
    return L


def _attrs(attrs: dict | None):
    items = [] if not attrs else [s.encode() for kv in attrs.items() for s in (str(kv[0]), str(kv[1]))]
    arr = (C.c_char_p * max(1, len(items)))(*items)
    return len(items) // 2, C.cast(arr, C.c_void_p), arr


def _dtype_code(dt: np.dtype):
    dt = np.dtype(dt)
    if dt.kind == "f":
        return FLOAT
    if dt.kind == "u":
        return UINT
    if dt.kind == "i":
        return INT
    raise V.VaccError(f"no VLSV datatype for {dt}")


def numpy_dtype(datatype: int, datasize: int) -> np.dtype:
    kind = {INT: "i", UINT: "u", FLOAT: "f"}[datatype]
    return np.dtype(f"<{kind}{datasize}")


class Writer:
    """A VLSV file being written (vlsv::Writer)."""

    def __init__(self, path: str):
        self._h = C.c_void_p()
        V._check(lib().vacc_vlsv_create(str(path).encode(), C.byref(self._h)))

    def write_array(self, tag: str, attrs: dict, data: np.ndarray):
        """data: (arraysize,) or (arraysize, vectorsize) of a numeric dtype."""
        a = np.ascontiguousarray(data)
        if a.dtype.byteorder == ">":
            a = a.astype(a.dtype.newbyteorder("<"))
        n, vs = (a.shape[0], 1) if a.ndim == 1 else (a.shape[0], int(np.prod(a.shape[1:])))
        na, ap, _keep = _attrs(attrs)
        V._check(lib().vacc_vlsv_write_array(self._h, tag.encode(), na, ap, n, vs, _dtype_code(a.dtype),
                                             a.dtype.itemsize, V._ptr(a) if a.size else None))

    def write_velocity_distribution(self, ctx: "V.Context", pop_name: str, cell_ids=None):
        ids = None if cell_ids is None else np.ascontiguousarray(cell_ids, np.uint64)
        V._check(lib().vacc_vlsv_write_velocity_distribution(self._h, ctx.h, pop_name.encode(),
                                                             None if ids is None else V._ptr(ids)))

    def close(self):
        if self._h:
            h, self._h = self._h, C.c_void_p()
            V._check(lib().vacc_vlsv_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Reader:
    """An open VLSV file (vlsv::ParallelReader)."""

    def __init__(self, path: str):
        self._h = C.c_void_p()
        V._check(lib().vacc_vlsv_open(str(path).encode(), C.byref(self._h)))

    def array_info(self, tag: str, attrs: dict | None = None):
        n, vs, ds = C.c_uint64(), C.c_uint64(), C.c_uint64()
        dt = C.c_int32()
        na, ap, _keep = _attrs(attrs)
        V._check(lib().vacc_vlsv_array_info(self._h, tag.encode(), na, ap, C.byref(n), C.byref(vs), C.byref(dt),
                                            C.byref(ds)))
        return n.value, vs.value, dt.value, ds.value

    def read_array(self, tag: str, attrs: dict | None = None, begin: int = 0, amount: int | None = None):
        n, vs, dt, ds = self.array_info(tag, attrs)
        amount = n - begin if amount is None else amount
        out = np.empty((amount, vs), numpy_dtype(dt, ds))
        na, ap, _keep = _attrs(attrs)
        V._check(lib().vacc_vlsv_read_array(self._h, tag.encode(), na, ap, begin, amount,
                                            V._ptr(out) if out.size else None))
        return out[:, 0] if vs == 1 else out

    def read_block_data(self, ctx: "V.Context", pop_name: str, mesh_name: str = SPATIAL_MESH, cell_offset: int = 0,
                        ncells: int | None = None):
        """Loads cells [cell_offset, cell_offset + ncells) into ctx; returns their cell IDs."""
        if ncells is None:
            ncells = self.array_info("CELLSWITHBLOCKS", {"mesh": mesh_name, "name": pop_name})[0] - cell_offset
        ids = np.zeros(max(1, ncells), np.uint64)
        V._check(lib().vacc_vlsv_read_block_data(self._h, ctx.h, pop_name.encode(), mesh_name.encode(), cell_offset,
                                                 ncells, V._ptr(ids)))
        return ids[:ncells]

    def close(self):
        if self._h:
            h, self._h = self._h, C.c_void_p()
            lib().vacc_vlsv_reader_close(h)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

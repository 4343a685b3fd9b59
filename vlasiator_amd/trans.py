"""ctypes binding of libvacc's spatial translation (include/vacc_trans.h).

trans_map_1d mirrors trans_map_1d_amr (vlasovsolver/cpu_trans_map_amr.cpp:214) over pencils
installed per dimension; build_pencils_uniform mirrors the reference's pencil construction
for a grid without refinement (vlasovsolver/cpu_trans_pencils.cpp).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import vacc as V

STENCIL = 2
CELL_NORMAL, CELL_BOUNDARY, CELL_INACTIVE = 0, 1, 2
INVALID = 0xFFFFFFFF

# symbols declared in include/vacc_trans.h
EXPORTS = ["vacc_build_pencils_uniform", "vacc_trans_set_pencils", "vacc_trans_map_1d"]


class Pencils(C.Structure):
    _fields_ = [("npencils", C.c_uint32), ("length", C.c_void_p), ("start", C.c_void_p), ("cell", C.c_void_p),
                ("dz", C.c_void_p), ("ratio", C.c_void_p)]


_bound = False


def lib():
    global _bound
    L = V.lib()
    if not _bound:
        vp = C.c_void_p
        L.vacc_build_pencils_uniform.argtypes = [vp, vp, vp, vp, C.c_uint32, C.POINTER(C.c_uint32),
                                                 C.POINTER(C.c_uint64), vp, vp, vp, vp, vp]
        L.vacc_trans_set_pencils.argtypes = [vp, C.c_uint32, C.POINTER(Pencils)]
        L.vacc_trans_map_1d.argtypes = [vp, C.c_uint32, C.c_double]
        _bound = True
    return L


def build_pencils_uniform(dims, periodic, cell_kind, dx, dim: int):
    """Pencils along dim of a grid without refinement: (lengths, starts, cells, dz, ratios)."""
    d = np.ascontiguousarray(dims, np.uint32)
    per = np.ascontiguousarray([1 if p else 0 for p in periodic], np.int32)
    kind = np.ascontiguousarray(cell_kind, np.uint8)
    assert kind.size == int(np.prod(d))
    h = np.ascontiguousarray(dx, np.float64)
    n, npos = C.c_uint32(), C.c_uint64()
    V._check(lib().vacc_build_pencils_uniform(V._ptr(d), V._ptr(per), V._ptr(kind), V._ptr(h), dim, C.byref(n),
                                              C.byref(npos), None, None, None, None, None))
    length = np.zeros(max(1, n.value), np.uint32)
    start = np.zeros(max(1, n.value), np.uint64)
    cell = np.zeros(max(1, npos.value), np.uint32)
    dz = np.zeros(max(1, npos.value), np.float32)
    ratio = np.zeros(max(1, npos.value), np.float32)
    V._check(lib().vacc_build_pencils_uniform(V._ptr(d), V._ptr(per), V._ptr(kind), V._ptr(h), dim, C.byref(n),
                                              C.byref(npos), V._ptr(length), V._ptr(start), V._ptr(cell),
                                              V._ptr(dz), V._ptr(ratio)))
    k, m = n.value, npos.value
    return length[:k], start[:k], cell[:m], dz[:m], ratio[:m]


def set_pencils(ctx: "V.Context", dim: int, pencils):
    length, start, cell, dz, ratio = (np.ascontiguousarray(pencils[0], np.uint32),
                                      np.ascontiguousarray(pencils[1], np.uint64),
                                      np.ascontiguousarray(pencils[2], np.uint32),
                                      np.ascontiguousarray(pencils[3], np.float32),
                                      np.ascontiguousarray(pencils[4], np.float32))
    p = Pencils(len(length), V._ptr(length) if length.size else None, V._ptr(start) if start.size else None,
                V._ptr(cell) if cell.size else None, V._ptr(dz) if dz.size else None,
                V._ptr(ratio) if ratio.size else None)
    V._check(lib().vacc_trans_set_pencils(ctx.h, dim, C.byref(p)))


def trans_map_1d(ctx: "V.Context", dim: int, dt: float):
    V._check(lib().vacc_trans_map_1d(ctx.h, dim, float(dt)))

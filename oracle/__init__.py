"""ctypes binding of the CPU oracle (test infrastructure only).

The oracle is a plain-C restatement of the reference CPU acceleration path
(fmihpc/vlasiator vlasovsolver/cpu_acc_*.cpp, cpu_1d_*.hpp,
cpu_face_estimates.hpp, cpu_slope_limiters.hpp, spatial_cells/
spatial_cell_cpu.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker or the CPU
baseline; the product path never touches it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvacc_oracle.so")
LIB64_PATH = os.path.join(HERE, "libvacc_oracle64.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_coeffs.so")

PLM, PPM, PQM = 0, 1, 2
WID3 = 64
ST_F32, ST_BF16, ST_F16 = 0, 1, 2


class Mesh(C.Structure):
    _fields_ = [("len", C.c_uint32 * 3), ("vmin", C.c_double * 3), ("vmax", C.c_double * 3)]


class Fields(C.Structure):
    _fields_ = [("B", C.c_double * 3), ("dB", C.c_double * 6), ("V", C.c_double * 3),
                ("rhoq", C.c_double), ("EgradPe", C.c_double * 3)]


class Peak(C.Structure):
    _fields_ = [("kind", C.c_int), ("n", C.c_double), ("V", C.c_double * 3), ("T", C.c_double),
                ("axis", C.c_double * 3), ("vperp", C.c_double), ("vth", C.c_double)]


def build(ref: bool | None = None) -> None:
    """Compile the oracle (and, when the reference checkout exists, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref is None:
        ref = os.path.isdir("/root/reference/vlasovsolver")
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None
_lib64 = None
_ref = None

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(ref=False)
        L = C.CDLL(LIB_PATH)
        L.vo_cell_size.restype = C.c_double
        L.vo_cell_size.argtypes = [C.POINTER(Mesh), C.c_int]
        L.vo_coeff_batch.argtypes = [C.c_int, _f32p, C.c_int, C.c_float, _f32p]
        L.vo_sort_columns.argtypes = [C.POINTER(Mesh), _u32p, C.c_uint32, C.c_int, _u32p, _u32p, _u32p,
                                      C.POINTER(C.c_uint32), _u32p, _u32p, C.POINTER(C.c_uint32)]
        L.vo_map_1d.argtypes = [C.POINTER(Mesh), C.c_uint32, _u32p, _f32p, C.c_double, C.c_double,
                                C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int,
                                C.c_uint32, C.POINTER(C.c_uint32), _u32p, _f32p, C.POINTER(C.c_int)]
        L.vo_adjust.argtypes = [C.POINTER(Mesh), C.c_uint32, _u32p, _f32p, C.c_double, C.c_int,
                                C.c_uint32, C.POINTER(C.c_uint32), _u32p, _f32p, C.POINTER(C.c_double)]
        L.vo_adjust_spatial.argtypes = [C.POINTER(Mesh), C.c_uint32, _u32p, _f32p, C.c_double, C.c_int,
                                        C.c_uint32, _u32p, C.c_uint32, C.POINTER(C.c_uint32), _u32p, _f32p,
                                        C.POINTER(C.c_double)]
        L.vo_content_flags.argtypes = [C.c_uint32, _f32p, C.c_double, _u8p]
        L.vo_moments.argtypes = [C.POINTER(Mesh), C.c_uint32, _u32p, _f32p, _f64p]
        L.vo_transform.argtypes = [C.POINTER(Fields), C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_int, _f64p]
        L.vo_affine_inverse.argtypes = [_f64p, _f64p]
        L.vo_intersections.argtypes = [C.POINTER(Mesh), _f64p, C.c_int, _f64p]
        L.vo_max_v_dt.restype = C.c_double
        L.vo_max_v_dt.argtypes = [_f64p, C.c_double, C.c_double, C.c_double]
        L.vo_peak_value.restype = C.c_double
        L.vo_peak_value.argtypes = [C.POINTER(Peak), C.c_double, C.c_double, C.c_double, C.c_double]
        L.vo_init_cell.argtypes = [C.POINTER(Mesh), C.POINTER(Peak), C.c_int, C.c_double, C.c_double,
                                   C.c_int, C.c_uint32, C.POINTER(C.c_uint32), _u32p, _f32p]
        L.vo_accelerate_cells.restype = C.c_int64
        L.vo_accelerate_cells.argtypes = [C.POINTER(Mesh), C.c_int, _u32p, _u32p, _f32p, _f64p, C.c_int,
                                          C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_uint64, _u32p, _u32p, _f32p, _i32p]
        L.vo_convert16.argtypes = [C.c_int, _f32p, C.c_uint32, _u16p]
        L.vo_encode_block.argtypes = [_f32p, C.c_int, _u16p, C.POINTER(C.c_float)]
        L.vo_decode_block.argtypes = [_u16p, C.c_float, C.c_int, _f32p]
        L.vo_block_csum.restype = C.c_float
        L.vo_block_csum.argtypes = [_f32p]
        L.vo_requantize.argtypes = [C.c_uint32, _f32p, C.c_int]
        L.vo_ppm_nu_batch.argtypes = [_f32p, _f32p, C.c_int, C.c_float, _f32p]
        L.vo_trans_map_1d.argtypes = [C.POINTER(Mesh), C.c_uint32, _u64p, _u32p, _f32p, C.c_uint32, _u32p, _u64p,
                                      _u32p, _f32p, _f32p, C.c_int, C.c_float, C.c_float]
        _lib = L
    return _lib


def lib64():
    """The oracle built with Realf = double (FP64 f): the FP64 reference run."""
    global _lib64
    if _lib64 is None:
        if not os.path.exists(LIB64_PATH):
            build(ref=False)
        L = C.CDLL(LIB64_PATH)
        L.vo_map_1d.argtypes = [C.POINTER(Mesh), C.c_uint32, _u32p, _f64p, C.c_double, C.c_double,
                                C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int,
                                C.c_uint32, C.POINTER(C.c_uint32), _u32p, _f64p, C.POINTER(C.c_int)]
        L.vo_adjust.argtypes = [C.POINTER(Mesh), C.c_uint32, _u32p, _f64p, C.c_double, C.c_int,
                                C.c_uint32, C.POINTER(C.c_uint32), _u32p, _f64p, C.POINTER(C.c_double)]
        L.vo_moments.argtypes = [C.POINTER(Mesh), C.c_uint32, _u32p, _f64p, _f64p]
        _lib64 = L
    return _lib64


def _flib(fp64: bool):
    return (lib64(), np.float64) if fp64 else (lib(), np.float32)


def ref_lib():
    """The reference's own reconstruction headers, compiled (None when not built)."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            return None
        R = C.CDLL(REF_PATH)
        R.ref_coeff_batch.argtypes = [C.c_int, _f32p, C.c_int, C.c_float, _f32p]
        R.ref_coeff_batch.restype = C.c_int
        R.ref_ppm_nu_batch.argtypes = [_f32p, _f32p, C.c_int, C.c_float, _f32p]
        R.ref_ppm_nu_batch.restype = C.c_int
        _ref = R
    return _ref


def make_mesh(len3, vmin3, vmax3) -> Mesh:
    m = Mesh()
    for d in range(3):
        m.len[d] = int(len3[d])
        m.vmin[d] = float(vmin3[d])
        m.vmax[d] = float(vmax3[d])
    return m


def ncoef(method: int) -> int:
    return {PLM: 2, PPM: 3, PQM: 5}[method]


def coeff_batch(method: int, stencils: np.ndarray, threshold: float) -> np.ndarray:
    s = np.ascontiguousarray(stencils, dtype=np.float32)
    n = s.shape[0]
    out = np.zeros((n, ncoef(method)), dtype=np.float32)
    lib().vo_coeff_batch(method, s.reshape(-1), n, np.float32(threshold), out.reshape(-1))
    return out


def ref_coeff_batch(method: int, stencils: np.ndarray, threshold: float) -> np.ndarray:
    R = ref_lib()
    if R is None:
        raise RuntimeError("oracle/_ref not built")
    s = np.ascontiguousarray(stencils, dtype=np.float32)
    n = s.shape[0]
    out = np.zeros((n, ncoef(method)), dtype=np.float32)
    R.ref_coeff_batch(method, s.reshape(-1), n, np.float32(threshold), out.reshape(-1))
    return out


def ppm_nu_batch(h5, stencils: np.ndarray, threshold: float) -> np.ndarray:
    """PPM on non-uniform widths h5 (5 values) for n stencils of 5 values (source at 2)."""
    h = np.ascontiguousarray(h5, dtype=np.float32)
    s = np.ascontiguousarray(stencils, dtype=np.float32)
    out = np.zeros((s.shape[0], 3), dtype=np.float32)
    lib().vo_ppm_nu_batch(h, s.reshape(-1), s.shape[0], np.float32(threshold), out.reshape(-1))
    return out


def ref_ppm_nu_batch(h5, stencils: np.ndarray, threshold: float) -> np.ndarray:
    R = ref_lib()
    if R is None:
        raise RuntimeError("oracle/_ref not built")
    h = np.ascontiguousarray(h5, dtype=np.float32)
    s = np.ascontiguousarray(stencils, dtype=np.float32)
    out = np.zeros((s.shape[0], 3), dtype=np.float32)
    R.ref_ppm_nu_batch(h, s.reshape(-1), s.shape[0], np.float32(threshold), out.reshape(-1))
    return out


TRANS_STENCIL = 2
INVALID = 0xFFFFFFFF


def trans_map_1d(mesh: Mesh, cells, pencils, dim: int, dt: float, threshold: float):
    """Spatial translation of all cells along dim (trans_map_1d_amr).  cells: [(gids, f)];
    pencils: (lengths, starts, cell indices, dz, ratios) in processing order.  Returns the
    cells with the same block lists and new f."""
    plen, pstart, pcell, pdz, pratio = pencils
    counts = [len(g) for g, _ in cells]
    off = np.zeros(len(cells) + 1, np.uint64)
    off[1:] = np.cumsum(counts)
    g = (np.concatenate([np.asarray(g, np.uint32) for g, _ in cells]) if cells else np.zeros(0, np.uint32))
    f = (np.concatenate([np.asarray(f, np.float32).reshape(-1, 64) for _, f in cells]) if cells
         else np.zeros((0, 64), np.float32))
    f = np.ascontiguousarray(f)
    rc = lib().vo_trans_map_1d(C.byref(mesh), len(cells), off, np.ascontiguousarray(g), f.reshape(-1), len(plen),
                               np.ascontiguousarray(plen, np.uint32), np.ascontiguousarray(pstart, np.uint64),
                               np.ascontiguousarray(pcell, np.uint32), np.ascontiguousarray(pdz, np.float32),
                               np.ascontiguousarray(pratio, np.float32), dim, np.float32(dt), np.float32(threshold))
    if rc:
        raise RuntimeError("vo_trans_map_1d failed")
    return [(g[off[i]:off[i + 1]].copy(), f[off[i]:off[i + 1]].copy()) for i in range(len(cells))]


def sort_columns(mesh: Mesh, gids: np.ndarray, dim: int):
    g = np.ascontiguousarray(gids, dtype=np.uint32)
    n = g.size
    sorted_ = np.zeros(max(n, 1), np.uint32)
    col_off = np.zeros(n + 1, np.uint32)
    col_len = np.zeros(n + 1, np.uint32)
    set_off = np.zeros(n + 1, np.uint32)
    set_nc = np.zeros(n + 1, np.uint32)
    nc = C.c_uint32()
    ns = C.c_uint32()
    lib().vo_sort_columns(C.byref(mesh), g, n, dim, sorted_, col_off, col_len, C.byref(nc),
                          set_off, set_nc, C.byref(ns))
    return (sorted_[:n], col_off[:nc.value], col_len[:nc.value], set_off[:ns.value], set_nc[:ns.value])


def map_1d(mesh: Mesh, gids, f, isect4, dim: int, method: int, min_value: float, wall_margin: int = 0,
           fp64: bool = False):
    L, dt = _flib(fp64)
    g = np.ascontiguousarray(gids, dtype=np.uint32)
    ff = np.ascontiguousarray(f, dtype=dt).reshape(-1)
    cap = g.size + g.size // 2 + 64
    while True:
        go = np.zeros(cap, np.uint32)
        fo = np.zeros(cap * WID3, dt)
        nout = C.c_uint32()
        wf = C.c_int(0)
        rc = L.vo_map_1d(C.byref(mesh), g.size, g, ff, *[float(x) for x in isect4], dim, method,
                             float(min_value), int(wall_margin), cap, C.byref(nout), go, fo, C.byref(wf))
        if rc == 0:
            n = nout.value
            return go[:n].copy(), fo[:n * WID3].reshape(n, WID3).copy(), wf.value
        cap = nout.value + 16


def accelerate_cell(mesh: Mesh, gids, f, intersections12, map_order: int, method: int,
                    min_value: float, wall_margin: int = 0, fp64: bool = False, storage: int = ST_F32):
    """Three map_1d passes in map_order.  storage: the blocks written by each pass are stored
    in that form (BF16/FP16 codes with a per-block scale), i.e. re-quantized after each pass."""
    g = np.asarray(gids, np.uint32)
    ff = np.asarray(f, np.float64 if fp64 else np.float32).reshape(-1, WID3)
    order = [(0, 1, 2), (1, 2, 0), (2, 0, 1)][map_order]
    wall = 0
    for dim in order:
        g, ff, wf = map_1d(mesh, g, ff, intersections12[4 * dim:4 * dim + 4], dim, method, min_value, wall_margin,
                           fp64=fp64)
        if storage != ST_F32:
            ff = requantize(ff, storage)
        wall |= wf
    return g, ff, wall


def adjust(mesh: Mesh, gids, f, min_value: float, add_width: int = 1, fp64: bool = False):
    L, dt = _flib(fp64)
    g = np.ascontiguousarray(gids, dtype=np.uint32)
    ff = np.ascontiguousarray(f, dtype=dt).reshape(-1)
    cap = 2 * g.size + 4096
    while True:
        go = np.zeros(cap, np.uint32)
        fo = np.zeros(cap * WID3, dt)
        nout = C.c_uint32()
        loss = C.c_double()
        rc = L.vo_adjust(C.byref(mesh), g.size, g, ff, float(min_value), int(add_width), cap,
                             C.byref(nout), go, fo, C.byref(loss))
        if rc == 0:
            n = nout.value
            return go[:n].copy(), fo[:n * WID3].reshape(n, WID3).copy(), loss.value
        cap = nout.value + 16


def pressure(mesh: Mesh, gids, f, V, mass: float) -> np.ndarray:
    """Pressure tensor (xx, yy, zz, yz, xz, xy) of one cell about V (arch_moments.h:123-150)."""
    L = lib()
    L.vo_pressure.argtypes = [C.POINTER(Mesh), C.c_uint32, _u32p, _f32p, _f64p, C.c_double, _f64p]
    g = np.ascontiguousarray(gids, np.uint32)
    ff = np.ascontiguousarray(f, np.float32).reshape(-1)
    out = np.zeros(6)
    L.vo_pressure(C.byref(mesh), g.size, g, ff, np.ascontiguousarray(V, np.float64), float(mass), out)
    return out


def sparse_min_value(algorithm: int, min_value: float, bulk1: float, bulk2: float, min1: float, min2: float,
                     rho: float, nblocks: int) -> float:
    """updateSparseMinValue (spatial_cell_cpu.cpp:699-727)."""
    L = lib()
    L.vo_sparse_min_value.restype = C.c_double
    L.vo_sparse_min_value.argtypes = [C.c_int] + [C.c_double] * 6 + [C.c_uint32]
    return L.vo_sparse_min_value(int(algorithm), min_value, bulk1, bulk2, min1, min2, float(rho), int(nblocks))


def conserve_mass(f_pre, f_post, fp64: bool = False) -> np.ndarray:
    """sparse_conserve_mass rescale (block_adjust_cpu.cpp:96-113): returns f_post * pre/post."""
    L, dt = _flib(fp64)
    a = np.ascontiguousarray(f_pre, dtype=dt).reshape(-1)
    b = np.array(f_post, dtype=dt).reshape(-1)
    fn = L.vo_conserve_mass
    fn.argtypes = [C.c_uint32, np.ctypeslib.ndpointer(dt, flags="C_CONTIGUOUS"), C.c_uint32,
                   np.ctypeslib.ndpointer(dt, flags="C_CONTIGUOUS")]
    fn(a.size // WID3, a, b.size // WID3, b)
    return b.reshape(-1, WID3)


def adjust_spatial(mesh: Mesh, gids, f, min_value: float, add_width: int, extra_gids):
    """Adjust with the spatial neighbours' content GIDs kept as they are."""
    g = np.ascontiguousarray(gids, dtype=np.uint32)
    ff = np.ascontiguousarray(f, dtype=np.float32).reshape(-1)
    ex = np.ascontiguousarray(extra_gids, dtype=np.uint32)
    exn = ex if ex.size else np.zeros(1, np.uint32)
    cap = 2 * g.size + ex.size + 4096
    while True:
        go = np.zeros(cap, np.uint32)
        fo = np.zeros(cap * WID3, np.float32)
        nout = C.c_uint32()
        loss = C.c_double()
        rc = lib().vo_adjust_spatial(C.byref(mesh), g.size, g, ff, float(min_value), int(add_width), ex.size, exn,
                                     cap, C.byref(nout), go, fo, C.byref(loss))
        if rc == 0:
            n = nout.value
            return go[:n].copy(), fo[:n * WID3].reshape(n, WID3).copy(), loss.value
        cap = nout.value + 16


def content_flags(f, min_value: float) -> np.ndarray:
    ff = np.ascontiguousarray(f, dtype=np.float32).reshape(-1)
    n = ff.size // WID3
    out = np.zeros(max(n, 1), np.uint8)
    lib().vo_content_flags(n, ff, float(min_value), out)
    return out[:n]


def moments(mesh: Mesh, gids, f, fp64: bool = False) -> np.ndarray:
    L, dt = _flib(fp64)
    g = np.ascontiguousarray(gids, dtype=np.uint32)
    ff = np.ascontiguousarray(f, dtype=dt).reshape(-1)
    out = np.zeros(4, np.float64)
    L.vo_moments(C.byref(mesh), g.size, g, ff, out)
    return out


# ---- 16-bit block storage -------------------------------------------------------------

def convert16(x, storage: int) -> np.ndarray:
    """Element-wise round-to-nearest-even FP32 -> BF16 / FP16 bit patterns."""
    a = np.ascontiguousarray(x, dtype=np.float32).reshape(-1)
    out = np.zeros(max(a.size, 1), np.uint16)
    lib().vo_convert16(storage, a, a.size, out)
    return out[:a.size]


def encode_block(x, storage: int):
    a = np.ascontiguousarray(x, dtype=np.float32).reshape(WID3)
    code = np.zeros(WID3, np.uint16)
    sc = C.c_float()
    lib().vo_encode_block(a, storage, code, C.byref(sc))
    return code, np.float32(sc.value)


def decode_block(code, scale, storage: int) -> np.ndarray:
    out = np.zeros(WID3, np.float32)
    lib().vo_decode_block(np.ascontiguousarray(code, np.uint16), float(scale), storage, out)
    return out


def block_csum(x) -> np.float32:
    return np.float32(lib().vo_block_csum(np.ascontiguousarray(x, dtype=np.float32).reshape(WID3)))


def requantize(f, storage: int) -> np.ndarray:
    """The values a 16-bit store holds for these blocks (encode, then decode), as FP32."""
    a = np.array(f, dtype=np.float32, copy=True).reshape(-1, WID3)
    if a.shape[0]:
        lib().vo_requantize(a.shape[0], a.reshape(-1), storage)
    return a


def make_fields(B, V, rhoq=0.0, dB=(0, 0, 0, 0, 0, 0), EgradPe=(0, 0, 0)) -> Fields:
    fl = Fields()
    for i in range(3):
        fl.B[i] = float(B[i])
        fl.V[i] = float(V[i])
        fl.EgradPe[i] = float(EgradPe[i])
    for i in range(6):
        fl.dB[i] = float(dB[i])
    fl.rhoq = float(rhoq)
    return fl


def transform(fields: Fields, mass: float, charge: float, dt: float, hall_min_rhoq: float,
              ohm_gradpe: int = 0) -> np.ndarray:
    out = np.zeros(12, np.float64)
    lib().vo_transform(C.byref(fields), mass, charge, dt, hall_min_rhoq, ohm_gradpe, out)
    return out


def affine_inverse(T) -> np.ndarray:
    t = np.ascontiguousarray(T, dtype=np.float64)
    out = np.zeros(12, np.float64)
    lib().vo_affine_inverse(t, out)
    return out


def intersections(mesh: Mesh, fwd, map_order: int) -> np.ndarray:
    t = np.ascontiguousarray(fwd, dtype=np.float64)
    out = np.zeros(12, np.float64)
    lib().vo_intersections(C.byref(mesh), t, map_order, out)
    return out


def max_v_dt(B, mass, charge, max_rotation_deg) -> float:
    return lib().vo_max_v_dt(np.ascontiguousarray(B, dtype=np.float64), mass, charge, max_rotation_deg)


def make_peak(kind=0, n=1e6, V=(0, 0, 0), T=1e5, axis=(0, 0, 1), vperp=0.0, vth=0.0) -> Peak:
    p = Peak()
    p.kind = int(kind)
    p.n = float(n)
    p.T = float(T)
    p.vperp = float(vperp)
    p.vth = float(vth)
    for i in range(3):
        p.V[i] = float(V[i])
        p.axis[i] = float(axis[i])
    return p


def init_cell(mesh: Mesh, peaks, mass: float, min_value: float, add_width: int = 1, cap: int = 400000):
    arr = (Peak * len(peaks))(*peaks)
    go = np.zeros(cap, np.uint32)
    fo = np.zeros(cap * WID3, np.float32)
    nout = C.c_uint32()
    rc = lib().vo_init_cell(C.byref(mesh), arr, len(peaks), mass, min_value, add_width, cap, C.byref(nout), go, fo)
    if rc != 0:
        raise RuntimeError(f"vo_init_cell capacity {cap} < {nout.value}")
    n = nout.value
    return go[:n].copy(), fo[:n * WID3].reshape(n, WID3).copy()


def accelerate_cells(mesh: Mesh, offsets, gids, f, intersections_all, map_order: int, method: int,
                     min_value: float, wall_margin: int = 0, do_adjust: bool = False, add_width: int = 1,
                     nthreads: int = 0, want_output: bool = True):
    off = np.ascontiguousarray(offsets, np.uint32)
    g = np.ascontiguousarray(gids, np.uint32)
    ff = np.ascontiguousarray(f, np.float32).reshape(-1)
    isx = np.ascontiguousarray(intersections_all, np.float64).reshape(-1)
    ncells = off.size - 1
    wall = np.zeros(ncells, np.int32)
    out_off = np.zeros(ncells + 1, np.uint32)
    if want_output:
        cap = int(g.size * 2 + 1024 * ncells)
        go = np.zeros(cap, np.uint32)
        fo = np.zeros(cap * WID3, np.float32)
    else:
        cap = 0
        go = np.zeros(1, np.uint32)
        fo = np.zeros(1, np.float32)
    tot = lib().vo_accelerate_cells(C.byref(mesh), ncells, off, g, ff, isx, map_order, method, float(min_value),
                                    int(wall_margin), int(do_adjust), int(add_width), int(nthreads), cap,
                                    out_off, go, fo, wall)
    if tot < 0:
        raise RuntimeError("vo_accelerate_cells capacity")
    if not want_output:
        return out_off, None, None, wall
    return out_off, go[:tot].copy(), fo[:tot * WID3].reshape(tot, WID3).copy(), wall

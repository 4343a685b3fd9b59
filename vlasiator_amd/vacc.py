"""ctypes binding of libvacc (include/vacc.h).

This is plumbing for tests and the benchmark: every call goes straight to the
HIP library.  There is no fallback; if the library is missing or no GPU is
visible the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libvacc.so")

PLM, PPM, PQM = 0, 1, 2
STORE_F32, STORE_BF16, STORE_F16 = 0, 1, 2
STATUS_WALL, STATUS_DK, STATUS_CAPACITY = 1, 2, 4
WID3 = 64
INVALID_ID = 0xFFFFFFFF
FIND_NOALLOC_MAX = 65536   # VACC_FIND_NOALLOC_MAX

# symbols declared in include/vacc.h
EXPORTS = [
    "vacc_create", "vacc_destroy", "vacc_last_error", "vacc_upload_cells", "vacc_total_blocks", "vacc_get_counts",
    "vacc_download_cells", "vacc_set_intersections", "vacc_get_intersections",
    "vacc_compute_cell_intersections", "vacc_cell_intersections_host", "vacc_max_v_dt", "vacc_map_1d",
    "vacc_accelerate_cells", "vacc_adjust_blocks", "vacc_moments", "vacc_sync", "vacc_get_status",
    "vacc_init_peaks", "vacc_profile_begin", "vacc_profile_read", "vacc_set_fields", "vacc_step",
    "vacc_accelerate_populations", "vacc_adjust_blocks_spatial", "vacc_autotune",
    "vacc_subcycle_count", "vacc_subcycle_step", "vacc_accelerate_subcycled", "vacc_pressure",
    "vacc_find_blocks", "vacc_find_blocks_noalloc", "vacc_get_remap_shape",
]


class VaccError(RuntimeError):
    pass


class Mesh(C.Structure):
    _fields_ = [("grid_length", C.c_uint32 * 3), ("min_limits", C.c_double * 3), ("max_limits", C.c_double * 3)]


class Species(C.Structure):
    _fields_ = [("mass", C.c_double), ("charge", C.c_double), ("sparse_min_value", C.c_double),
                ("sparse_add_width_v", C.c_int32), ("sparse_dynamic_algorithm", C.c_int32),
                ("sparse_dynamic_bulk1", C.c_double), ("sparse_dynamic_bulk2", C.c_double),
                ("sparse_dynamic_min1", C.c_double), ("sparse_dynamic_min2", C.c_double),
                ("sparse_conserve_mass", C.c_int32), ("reserved", C.c_int32)]


class Options(C.Structure):
    _fields_ = [("method", C.c_int32), ("storage", C.c_int32), ("wall_margin", C.c_int32),
                ("ohm_gradpe_term", C.c_int32), ("hall_minimum_rhoq", C.c_double),
                ("max_rotation_deg", C.c_double), ("fast_arith", C.c_int32), ("use_graph", C.c_int32),
                ("sums_from_f", C.c_int32)]


class CellFields(C.Structure):
    _fields_ = [("B", C.c_double * 3), ("dB", C.c_double * 6), ("V", C.c_double * 3), ("rhoq", C.c_double),
                ("EgradPe", C.c_double * 3), ("dt", C.c_double)]


class Profile(C.Structure):
    _fields_ = [("remap_launches", C.c_uint64), ("remap_ms", C.c_double), ("remap_src_blocks", C.c_uint64),
                ("remap_tgt_blocks", C.c_uint64), ("build_launches", C.c_uint64), ("build_ms", C.c_double),
                ("adjust_launches", C.c_uint64), ("adjust_ms", C.c_double), ("transform_launches", C.c_uint64),
                ("transform_ms", C.c_double), ("trans_launches", C.c_uint64), ("trans_ms", C.c_double),
                ("remap_sum_blocks", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Peak(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("n", C.c_double), ("V", C.c_double * 3),
                ("T", C.c_double), ("axis", C.c_double * 3), ("vperp", C.c_double), ("vth", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VaccError(f"{LIB_PATH} is not built; run vlasiator_amd/build.py")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.vacc_last_error.restype = C.c_char_p
        L.vacc_create.argtypes = [C.POINTER(Mesh), C.POINTER(Species), C.POINTER(Options), C.POINTER(vp)]
        L.vacc_destroy.argtypes = [vp]
        L.vacc_upload_cells.argtypes = [vp, C.c_uint32, vp, vp, vp]
        L.vacc_get_counts.argtypes = [vp, C.POINTER(C.c_uint32), vp]
        L.vacc_total_blocks.argtypes = [vp, C.POINTER(C.c_uint64)]
        L.vacc_download_cells.argtypes = [vp, vp, vp, vp]
        L.vacc_set_intersections.argtypes = [vp, vp]
        L.vacc_get_intersections.argtypes = [vp, vp]
        L.vacc_compute_cell_intersections.argtypes = [vp, vp, C.c_uint32, C.c_int32]
        L.vacc_cell_intersections_host.argtypes = [C.POINTER(Mesh), C.POINTER(Species), C.POINTER(Options),
                                                   C.POINTER(CellFields), C.c_uint32, vp]
        L.vacc_max_v_dt.restype = C.c_double
        L.vacc_max_v_dt.argtypes = [C.POINTER(Species), C.POINTER(Options), vp]
        L.vacc_map_1d.argtypes = [vp, C.c_uint32]
        L.vacc_accelerate_cells.argtypes = [vp, C.c_uint32]
        L.vacc_adjust_blocks.argtypes = [vp, vp]
        L.vacc_moments.argtypes = [vp, vp]
        L.vacc_sync.argtypes = [vp]
        L.vacc_get_status.argtypes = [vp, vp]
        L.vacc_init_peaks.argtypes = [vp, C.c_uint32, C.c_uint32, vp, vp, vp]
        L.vacc_profile_begin.argtypes = [vp]
        L.vacc_profile_read.argtypes = [vp, C.POINTER(Profile)]
        L.vacc_set_fields.argtypes = [vp, vp]
        L.vacc_step.argtypes = [vp, C.c_uint32, C.c_int32, C.c_int32, vp]
        L.vacc_accelerate_populations.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_int32, C.c_int32, vp]
        L.vacc_adjust_blocks_spatial.argtypes = [vp, vp, vp, vp, vp, vp]
        L.vacc_autotune.argtypes = [vp, vp]
        L.vacc_get_remap_shape.argtypes = [vp, vp, vp]
        L.vacc_pressure.argtypes = [vp, vp, vp]
        L.vacc_find_blocks.argtypes = [vp, C.c_uint64, vp, vp, vp]
        L.vacc_find_blocks_noalloc.argtypes = [vp, C.c_uint64, vp, vp, vp]
        L.vacc_subcycle_count.argtypes = [vp, C.c_double, C.POINTER(C.c_uint32)]
        L.vacc_subcycle_step.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_int32, C.c_int32, vp]
        L.vacc_accelerate_subcycled.argtypes = [vp, C.c_double, C.c_uint32, C.c_int32, C.POINTER(C.c_uint32), vp]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _check(rc: int) -> int:
    if rc < 0:
        raise VaccError(f"libvacc error {rc}: {lib().vacc_last_error().decode()}")
    return rc


def make_mesh(grid_length, vmin, vmax) -> Mesh:
    m = Mesh()
    for d in range(3):
        m.grid_length[d] = int(grid_length[d])
        m.min_limits[d] = float(vmin[d])
        m.max_limits[d] = float(vmax[d])
    return m


def make_species(mass, charge, min_value=1e-15, add_width=1, dynamic=0, bulk=(0.0, 1.0), dyn_min=(0.0, 0.0),
                 conserve_mass=False) -> Species:
    s = Species()
    s.mass, s.charge, s.sparse_min_value = float(mass), float(charge), float(min_value)
    s.sparse_add_width_v = int(add_width)
    s.sparse_dynamic_algorithm = int(dynamic)
    s.sparse_dynamic_bulk1, s.sparse_dynamic_bulk2 = float(bulk[0]), float(bulk[1])
    s.sparse_dynamic_min1, s.sparse_dynamic_min2 = float(dyn_min[0]), float(dyn_min[1])
    s.sparse_conserve_mass = 1 if conserve_mass else 0
    return s


def make_options(method=PQM, wall_margin=0, ohm_gradpe=0, hall_min_rhoq=1.60217653e-19, max_rotation_deg=25.0,
                 storage=STORE_F32, fast_arith=False, use_graph=False, sums_from_f=False) -> Options:
    o = Options()
    o.fast_arith = 1 if fast_arith else 0
    o.use_graph = 1 if use_graph else 0
    o.sums_from_f = 1 if sums_from_f else 0
    o.method, o.storage, o.wall_margin, o.ohm_gradpe_term = int(method), int(storage), int(wall_margin), int(ohm_gradpe)
    o.hall_minimum_rhoq, o.max_rotation_deg = float(hall_min_rhoq), float(max_rotation_deg)
    return o


def make_fields(B, V=(0, 0, 0), dt=0.0, rhoq=0.0, dB=(0,) * 6, EgradPe=(0, 0, 0)) -> CellFields:
    f = CellFields()
    for i in range(3):
        f.B[i], f.V[i], f.EgradPe[i] = float(B[i]), float(V[i]), float(EgradPe[i])
    for i in range(6):
        f.dB[i] = float(dB[i])
    f.rhoq, f.dt = float(rhoq), float(dt)
    return f


def cell_intersections_host(mesh: Mesh, species: Species, opts: Options, fields: CellFields, map_order: int):
    out = np.zeros(12, np.float64)
    _check(lib().vacc_cell_intersections_host(C.byref(mesh), C.byref(species), C.byref(opts), C.byref(fields),
                                              int(map_order), _ptr(out)))
    return out


class Context:
    """One population's velocity-space state for a batch of spatial cells on the GPU."""

    def __init__(self, mesh: Mesh, species: Species, opts: Options):
        self.mesh, self.species, self.opts = mesh, species, opts
        h = C.c_void_p()
        _check(lib().vacc_create(C.byref(mesh), C.byref(species), C.byref(opts), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().vacc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state transfer ------------------------------------------------------
    def upload(self, offsets, gids, f):
        off = np.ascontiguousarray(offsets, np.uint64)
        g = np.ascontiguousarray(gids, np.uint32)
        ff = np.ascontiguousarray(f, np.float32).reshape(-1)
        assert ff.size == g.size * WID3 and off[-1] == g.size
        _check(lib().vacc_upload_cells(self.h, off.size - 1, _ptr(off), _ptr(g), _ptr(ff)))

    def upload_cells(self, cells):
        """cells: list of (gids, f[n,64])."""
        counts = np.array([len(g) for g, _ in cells], np.uint64)
        off = np.zeros(len(cells) + 1, np.uint64)
        off[1:] = np.cumsum(counts)
        g = np.concatenate([np.asarray(c[0], np.uint32) for c in cells]) if cells else np.zeros(0, np.uint32)
        f = np.concatenate([np.asarray(c[1], np.float32).reshape(-1, WID3) for c in cells]) if cells else \
            np.zeros((0, WID3), np.float32)
        self.upload(off, g, f)

    def total_blocks(self) -> int:
        n = C.c_uint64()
        _check(lib().vacc_total_blocks(self.h, C.byref(n)))
        return n.value

    def counts(self):
        n = C.c_uint32()
        _check(lib().vacc_get_counts(self.h, C.byref(n), None))
        cnt = np.zeros(max(n.value, 1), np.uint32)
        _check(lib().vacc_get_counts(self.h, C.byref(n), _ptr(cnt)))
        return cnt[:n.value]

    def download(self):
        cnt = self.counts()
        tot = int(cnt.sum())
        off = np.zeros(cnt.size + 1, np.uint64)
        g = np.zeros(max(tot, 1), np.uint32)
        f = np.zeros((max(tot, 1), WID3), np.float32)
        _check(lib().vacc_download_cells(self.h, _ptr(off), _ptr(g), _ptr(f)))
        return off, g[:tot], f[:tot]

    def download_cells(self):
        off, g, f = self.download()
        return [(g[off[i]:off[i + 1]], f[off[i]:off[i + 1]]) for i in range(off.size - 1)]

    # -- intersections ---------------------------------------------------------
    def set_intersections(self, isect):
        a = np.ascontiguousarray(isect, np.float64).reshape(-1)
        _check(lib().vacc_set_intersections(self.h, _ptr(a)))

    def get_intersections(self, ncells):
        a = np.zeros(ncells * 12, np.float64)
        _check(lib().vacc_get_intersections(self.h, _ptr(a)))
        return a.reshape(ncells, 12)

    def compute_intersections(self, fields, map_order: int, use_moments: bool = False):
        """fields=None reuses the resident fields."""
        arr = None if fields is None else C.cast((CellFields * len(fields))(*fields), C.c_void_p)
        _check(lib().vacc_compute_cell_intersections(self.h, arr, int(map_order), 1 if use_moments else 0))

    def set_fields(self, fields):
        arr = (CellFields * len(fields))(*fields)
        _check(lib().vacc_set_fields(self.h, C.cast(arr, C.c_void_p)))

    def step(self, map_order: int, use_moments: bool = True, adjust: bool = True, rho_loss=None) -> int:
        """One subcycle (moments, transform, three remaps, adjust) with one synchronisation.
        rho_loss: optional float64 array of ncells entries that receives the removed mass."""
        ptr = None if rho_loss is None else _ptr(rho_loss)
        return _check(lib().vacc_step(self.h, int(map_order), 1 if use_moments else 0, 1 if adjust else 0, ptr))

    def subcycle_count(self, dt: float) -> int:
        """Per-cell subcycle counts of a step of length dt (kept resident); returns the maximum."""
        n = C.c_uint32()
        _check(lib().vacc_subcycle_count(self.h, float(dt), C.byref(n)))
        return n.value

    def subcycle_step(self, step: int, map_order: int, use_moments: bool = True, adjust: bool = True,
                      rho_loss=None) -> int:
        ptr = None if rho_loss is None else _ptr(rho_loss)
        return _check(lib().vacc_subcycle_step(self.h, int(step), int(map_order), 1 if use_moments else 0,
                                               1 if adjust else 0, ptr))

    def accelerate_subcycled(self, dt: float, map_order: int, use_moments: bool = True, rho_loss=None):
        """Every subcycle of a step of length dt; returns (status, number of subcycles)."""
        n = C.c_uint32()
        ptr = None if rho_loss is None else _ptr(rho_loss)
        st = _check(lib().vacc_accelerate_subcycled(self.h, float(dt), int(map_order), 1 if use_moments else 0,
                                                    C.byref(n), ptr))
        return st, n.value

    # -- work --------------------------------------------------------------------
    def map_1d(self, dim: int) -> int:
        return _check(lib().vacc_map_1d(self.h, int(dim)))

    def accelerate(self, map_order: int) -> int:
        return _check(lib().vacc_accelerate_cells(self.h, int(map_order)))

    def adjust(self, ncells=None):
        if ncells is None:
            ncells = self.counts().size
        loss = np.zeros(max(ncells, 1), np.float64)
        _check(lib().vacc_adjust_blocks(self.h, _ptr(loss)))
        return loss[:ncells]

    def adjust_spatial(self, ncells, nbr_off=None, nbr=None, extra_off=None, extra_gids=None):
        """Adjust with spatial neighbours: nbr_off/nbr (cells of this context, CSR) and/or
        extra_off/extra_gids (content GIDs per cell, CSR).  Returns the removed mass per cell."""
        def arr(a, dt):
            return None if a is None else np.ascontiguousarray(a, dt)
        no, nb = arr(nbr_off, np.uint32), arr(nbr, np.uint32)
        eo, eg = arr(extra_off, np.uint64), arr(extra_gids, np.uint32)
        if nb is not None and nb.size == 0:
            nb = np.zeros(1, np.uint32)
        if eg is not None and eg.size == 0:
            eg = np.zeros(1, np.uint32)
        loss = np.zeros(max(ncells, 1), np.float64)
        p = lambda a: None if a is None else _ptr(a)
        _check(lib().vacc_adjust_blocks_spatial(self.h, p(no), p(nb), p(eo), p(eg), _ptr(loss)))
        return loss[:ncells]

    def autotune(self):
        """Pick the remap launch shape per pass dimension on the current state; returns the
        workgroups per compute unit (remap_shape() gives the lanes per workgroup as well)."""
        out = np.zeros(3, np.uint32)
        _check(lib().vacc_autotune(self.h, _ptr(out)))
        return out.tolist()

    def remap_shape(self):
        """(workgroups per compute unit, lanes per workgroup) of the remap, per pass dimension."""
        m = np.zeros(3, np.uint32)
        w = np.zeros(3, np.uint32)
        _check(lib().vacc_get_remap_shape(self.h, _ptr(m), _ptr(w)))
        return m.tolist(), w.tolist()

    def moments(self, ncells):
        out = np.zeros(max(ncells, 1) * 4, np.float64)
        _check(lib().vacc_moments(self.h, _ptr(out)))
        return out[:ncells * 4].reshape(ncells, 4)

    def find_blocks(self, cells, gids):
        """Position of each (cell, GID) within its cell's ascending-GID block list, or
        INVALID_ID (VelocityMesh::getLocalID)."""
        qc = np.ascontiguousarray(cells, np.uint32)
        qg = np.ascontiguousarray(gids, np.uint32)
        if qc.shape != qg.shape:
            raise ValueError("cells and gids differ in length")
        out = np.zeros(max(qc.size, 1), np.uint32)
        _check(lib().vacc_find_blocks(self.h, qc.size, _ptr(qc), _ptr(qg), _ptr(out)))
        return out[:qc.size]

    def find_blocks_noalloc(self, cells, gids):
        """find_blocks through vacc_find_blocks_noalloc: no device allocation, at most
        FIND_NOALLOC_MAX queries per call."""
        qc = np.ascontiguousarray(cells, np.uint32)
        qg = np.ascontiguousarray(gids, np.uint32)
        if qc.shape != qg.shape:
            raise ValueError("cells and gids differ in length")
        out = np.zeros(max(qc.size, 1), np.uint32)
        _check(lib().vacc_find_blocks_noalloc(self.h, qc.size, _ptr(qc), _ptr(qg), _ptr(out)))
        return out[:qc.size]

    def pressure(self, ncells, pivot=None):
        """Pressure tensor per cell (xx, yy, zz, yz, xz, xy) about pivot[ncells, 3] or, for None,
        the population's own bulk velocity."""
        out = np.zeros(max(ncells, 1) * 6, np.float64)
        pv = None if pivot is None else np.ascontiguousarray(pivot, np.float64).reshape(-1)
        _check(lib().vacc_pressure(self.h, None if pv is None else _ptr(pv), _ptr(out)))
        return out[:ncells * 6].reshape(ncells, 6)

    def profile_begin(self):
        _check(lib().vacc_profile_begin(self.h))

    def profile_read(self) -> dict:
        p = Profile()
        _check(lib().vacc_profile_read(self.h, C.byref(p)))
        return p.as_dict()

    def sync(self) -> int:
        return _check(lib().vacc_sync(self.h))

    def status(self, ncells):
        st = np.zeros(max(ncells, 1), np.uint32)
        _check(lib().vacc_get_status(self.h, _ptr(st)))
        return st[:ncells]

    def init_peaks(self, peaks_per_cell, bboxes, max_peaks: int = 4):
        """Synthetic initial state: sum of peaks sampled over a per-cell candidate box
        bboxes[c] = (x0,y0,z0,x1,y1,z1) (inclusive block indices), then the sparsity adjust."""
        nc = len(peaks_per_cell)
        arr = (Peak * (nc * max_peaks))()
        npk = np.zeros(nc, np.int32)
        for c, pl in enumerate(peaks_per_cell):
            assert len(pl) <= max_peaks
            npk[c] = len(pl)
            for q, p in enumerate(pl):
                arr[c * max_peaks + q] = p
        bb = np.ascontiguousarray(bboxes, np.int32).reshape(nc, 6)
        _check(lib().vacc_init_peaks(self.h, nc, max_peaks, C.cast(arr, C.c_void_p), _ptr(npk), _ptr(bb)))


def accelerate_populations(ctxs, map_order: int, use_moments: bool = True, adjust: bool = True,
                           rho_loss=None) -> int:
    """One step of several populations of the same cells (vacc_accelerate_populations).
    rho_loss: optional list of float64 arrays (ncells each, or None per population)."""
    n = len(ctxs)
    hs = (C.c_void_p * n)(*[c.h for c in ctxs])
    rl = None
    if rho_loss is not None:
        rl = (C.c_void_p * n)(*[None if a is None else a.ctypes.data for a in rho_loss])
    return _check(lib().vacc_accelerate_populations(C.cast(hs, C.c_void_p), n, int(map_order),
                                                    1 if use_moments else 0, 1 if adjust else 0,
                                                    None if rl is None else C.cast(rl, C.c_void_p)))

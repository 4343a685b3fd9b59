"""Synthetic workloads of the acceleration benchmarks (seeded, reproducible).

Each configuration is a set of spatial cells, each with a velocity distribution made
of Maxwellian peaks (and, for the multi-peak case, a field-aligned beam and a
gyrotropic ring), a uniform magnetic field and a bulk velocity; the time step of a
cell is the one that rotates its distribution by a fixed gyro angle.  Random numbers
come from SplitMix64 so the same cells are produced on every machine.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MASS_PROTON = 1.67262158e-27   # common.h physicalconstants
CHARGE = 1.60217653e-19
K_B = 1.3806503e-23


class SplitMix64:
    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self, a=0.0, b=1.0) -> float:
        return a + (b - a) * ((self.next() >> 11) * (1.0 / (1 << 53)))

    def log_uniform(self, a, b) -> float:
        return math.exp(self.uniform(math.log(a), math.log(b)))

    def unit_vector(self):
        z = self.uniform(-1.0, 1.0)
        phi = self.uniform(0.0, 2.0 * math.pi)
        r = math.sqrt(max(0.0, 1.0 - z * z))
        return (r * math.cos(phi), r * math.sin(phi), z)

    def in_ball(self, radius):
        u = self.unit_vector()
        r = radius * self.uniform() ** (1.0 / 3.0)
        return (u[0] * r, u[1] * r, u[2] * r)


@dataclass
class PeakSpec:
    kind: int          # 0 Maxwellian, 1 ring
    n: float
    V: tuple
    T: float = 0.0
    axis: tuple = (0.0, 0.0, 1.0)
    vperp: float = 0.0
    vth: float = 0.0


@dataclass
class CellSpec:
    peaks: list
    B: tuple
    V: tuple           # bulk velocity used as the initial pivot
    dt: float


@dataclass
class Workload:
    name: str
    grid_length: tuple
    vmin: tuple
    vmax: tuple
    method: int        # 0 PLM, 1 PPM, 2 PQM
    min_value: float
    add_width: int
    steps: int
    adjust_every_step: bool
    cells: list = field(default_factory=list)
    mass: float = MASS_PROTON
    charge: float = CHARGE
    rotation_deg: float = 5.0


def gyro_period(B, mass=MASS_PROTON, charge=CHARGE):
    bmag = math.sqrt(B[0] ** 2 + B[1] ** 2 + B[2] ** 2) + 1e-30
    return 2 * math.pi * mass / (charge * bmag)


def maxwellian_radius(n, T, mass, thr):
    fmax = n * (mass / (2 * math.pi * K_B * T)) ** 1.5
    if fmax <= thr:
        return 0.0
    return math.sqrt(2 * K_B * T / mass * math.log(fmax / thr))


def ring_radius(n, vperp, vth, thr):
    fmax = n / ((2 * math.pi) ** 2 * vperp * vth * vth)
    if fmax <= thr:
        return 0.0
    return vperp + vth * math.sqrt(2 * math.log(fmax / thr))


def candidate_box(w: Workload, cell: CellSpec):
    """Inclusive block box that holds every block with content plus add_width+1 blocks."""
    lo = [10 ** 9] * 3
    hi = [-1] * 3
    for p in cell.peaks:
        thr = w.min_value / 16.0
        r = maxwellian_radius(p.n, p.T, w.mass, thr) if p.kind == 0 else ring_radius(p.n, p.vperp, p.vth, thr)
        if r <= 0:
            continue
        for d in range(3):
            bs = (w.vmax[d] - w.vmin[d]) / w.grid_length[d]
            a = math.floor((p.V[d] - r - w.vmin[d]) / bs) - (w.add_width + 1)
            b = math.floor((p.V[d] + r - w.vmin[d]) / bs) + (w.add_width + 1)
            lo[d] = min(lo[d], a)
            hi[d] = max(hi[d], b)
    if hi[0] < 0:
        return None
    return tuple([max(0, lo[d]) for d in range(3)] + [min(w.grid_length[d] - 1, hi[d]) for d in range(3)])


def _field(rng: SplitMix64, bmin_nT, bmax_nT):
    u = rng.unit_vector()
    b = rng.uniform(bmin_nT, bmax_nT) * 1e-9
    return (u[0] * b, u[1] * b, u[2] * b)


def config(index: int, ncells: int | None = None, seed: int = 1000) -> Workload:
    """The five benchmark configurations (0..4); ncells overrides the cell count."""
    if index == 0:
        w = Workload("maxwellian-16c-50^3-ppm", (50, 50, 50), (-2e6,) * 3, (2e6,) * 3, 1, 1e-15, 1, 10, True)
        n = ncells or 16
        for c in range(n):
            rng = SplitMix64(seed + c)
            T = rng.uniform(1e5, 1e6)
            V = tuple(rng.uniform(-6e5, 6e5) for _ in range(3))
            B = _field(rng, 5.0, 20.0)
            w.cells.append(CellSpec([PeakSpec(0, 1e6, V, T)], B, V, gyro_period(B) * 5.0 / 360.0))
        return w
    if index == 1:
        w = Workload("maxwellian-4096c-50^3-ppm", (50, 50, 50), (-2e6,) * 3, (2e6,) * 3, 1, 1e-15, 1, 20, True)
        n = ncells or 4096
        for c in range(n):
            rng = SplitMix64(seed + c)
            T = rng.log_uniform(5e4, 5e6)
            V = rng.in_ball(8e5)
            B = _field(rng, 1.0, 50.0)
            w.cells.append(CellSpec([PeakSpec(0, 1e6, V, T)], B, V, gyro_period(B) * 5.0 / 360.0))
        return w
    if index in (2, 3):
        w = Workload("multipeak-8192c-75^3-pqm", (75, 75, 75), (-3e6,) * 3, (3e6,) * 3, 2, 1e-15, 1, 20, True)
        n = ncells or 8192
        for c in range(n):
            rng = SplitMix64(seed + c)
            B = _field(rng, 1.0, 50.0)
            bm = math.sqrt(sum(x * x for x in B))
            bh = tuple(x / bm for x in B)
            Vc = rng.in_ball(2e5)
            frac = rng.uniform(0.02, 0.10)
            dv = rng.uniform(4e5, 8e5) * (1.0 if rng.uniform() < 0.5 else -1.0)
            Vb = tuple(Vc[d] + dv * bh[d] for d in range(3))
            peaks = [PeakSpec(0, 1e6, Vc, 5e5), PeakSpec(0, 1e6 * frac, Vb, 5e5),
                     PeakSpec(1, 0.05e6, Vc, 0.0, bh, 5e5, 5e4)]
            w.cells.append(CellSpec(peaks, B, Vc, gyro_period(B) * 5.0 / 360.0))
        return w
    if index == 4:
        w = Workload("two-species-32768c-100^3-pqm", (100, 100, 100), (-4e6,) * 3, (4e6,) * 3, 2, 1e-15, 1, 50, True)
        n = ncells or 32768
        for c in range(n):
            rng = SplitMix64(seed + c)
            T = rng.log_uniform(1e5, 1e7)
            V = rng.in_ball(8e5)
            B = _field(rng, 1.0, 50.0)
            w.cells.append(CellSpec([PeakSpec(0, 1e6, V, T)], B, V, gyro_period(B) * 5.0 / 360.0))
        return w
    raise ValueError(index)


HE_MASS = 4.0 * MASS_PROTON   # He2+ as a population with 4x the proton mass-to-charge ratio
HE_CHARGE = CHARGE             # (4 m_p, 1 e); the density fraction is HE_FRACTION
HE_FRACTION = 0.04


def populations(index: int, ncells: int | None = None, seed: int = 1000) -> list:
    """The populations of a configuration, one Workload each over the same cells.  Config 4
    has protons plus a He2+ population at 4 % density with 4x the proton mass-to-charge
    ratio; the He2+ peak has the protons' temperature and bulk velocity, and both
    populations share each cell's field and time step (5 degrees of proton gyration)."""
    w = config(index, ncells, seed)
    if index != 4:
        return [w]
    he = Workload(w.name.replace("two-species", "he2"), w.grid_length, w.vmin, w.vmax, w.method, w.min_value,
                  w.add_width, w.steps, w.adjust_every_step, mass=HE_MASS, charge=HE_CHARGE,
                  rotation_deg=w.rotation_deg * MASS_PROTON / HE_MASS)
    for c in w.cells:
        p = c.peaks[0]
        he.cells.append(CellSpec([PeakSpec(0, p.n * HE_FRACTION, p.V, p.T)], c.B, c.V, c.dt))
    return [w, he]


def to_vacc_peaks(cell: CellSpec):
    from . import vacc as V
    out = []
    for p in cell.peaks:
        q = V.Peak()
        q.kind, q.n, q.T, q.vperp, q.vth = p.kind, p.n, p.T, p.vperp, p.vth
        for i in range(3):
            q.V[i] = p.V[i]
            q.axis[i] = p.axis[i]
        out.append(q)
    return out


def boxes(w: Workload) -> np.ndarray:
    bb = np.zeros((len(w.cells), 6), np.int32)
    for i, c in enumerate(w.cells):
        b = candidate_box(w, c)
        bb[i] = b if b is not None else (0, 0, 0, -1, -1, -1)
    return bb

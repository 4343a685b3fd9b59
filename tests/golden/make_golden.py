"""Writes the golden reconstruction vectors under tests/golden/.

Inputs are seeded stencils (9 cell averages around the source cell) of several
shapes; the expected outputs are produced by the reference's own reconstruction
headers (vlasovsolver/cpu_1d_plm.hpp, cpu_1d_ppm.hpp, cpu_1d_pqm.hpp with
cpu_face_estimates.hpp and cpu_slope_limiters.hpp, VEC_FALLBACK_GENERIC, VECL=8)
compiled by oracle/Makefile into oracle/_ref/libref_coeffs.so.  Run in a
container that holds the reference checkout:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

THRESHOLD = 1e-15


def stencils(n: int, seed: int = 7) -> np.ndarray:
    rng = np.random.default_rng(seed)
    parts = []
    parts.append(10 ** rng.uniform(-18, -9, (n, 9)))                     # rough positive
    x = np.linspace(-3, 3, 9)
    c = rng.uniform(-2, 2, (n, 1))
    w = rng.uniform(0.3, 3, (n, 1))
    a = 10 ** rng.uniform(-14, -9, (n, 1))
    parts.append(a * np.exp(-((x - c) / w) ** 2))                         # smooth peaks
    z = np.zeros((n, 9))
    m = rng.random((n, 9)) < 0.3
    z[m] = 10 ** rng.uniform(-16, -10, m.sum())
    parts.append(z)                                                       # sparse with zeros
    parts.append(np.sort(10 ** rng.uniform(-16, -10, (n, 9)), axis=1))    # monotone
    parts.append(THRESHOLD * rng.uniform(0.5, 2, (n, 9)))                 # near the threshold
    parts.append(np.where(rng.random((n, 9)) < 0.5, 0.0, 1e-12))          # steps
    s = np.concatenate(parts).astype(np.float32)
    s[:4] = 0.0                                                           # all-zero stencils
    return s


def main():
    oracle.build(ref=True)
    s = stencils(700)
    out = {"stencils": s, "threshold": np.float64(THRESHOLD)}
    for m, name in ((oracle.PLM, "plm"), (oracle.PPM, "ppm"), (oracle.PQM, "pqm")):
        out[name] = oracle.ref_coeff_batch(m, s, THRESHOLD)
    np.savez_compressed(os.path.join(HERE, "reconstruction.npz"), **out)
    print("wrote", os.path.join(HERE, "reconstruction.npz"), s.shape)
    # PPM on non-uniform widths (cpu_1d_ppm_nonuniform.hpp), as the spatial translation uses it:
    # 5-cell stencils (source at 2), sets of widths (uniform, refined neighbours, random), and
    # two thresholds (scaled and unscaled limiter branches)
    rng = np.random.default_rng(11)
    s5 = stencils(500, seed=13)[:, 2:7].copy()
    widths = np.array([[1e6] * 5, [2e6, 2e6, 1e6, 1e6, 1e6], [1e6, 1e6, 1e6, 5e5, 5e5], [3.3e5] * 5]
                      + [10 ** rng.uniform(4, 7, 5) for _ in range(4)], np.float32)
    thresholds = np.array([THRESHOLD, 0.0])
    coeffs = np.stack([np.stack([oracle.ref_ppm_nu_batch(h, s5, t) for t in thresholds]) for h in widths])
    np.savez_compressed(os.path.join(HERE, "translation_ppm_nu.npz"), stencils=s5, widths=widths,
                        thresholds=thresholds, coeffs=coeffs)
    print("wrote", os.path.join(HERE, "translation_ppm_nu.npz"), coeffs.shape)


if __name__ == "__main__":
    main()

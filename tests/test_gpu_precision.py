"""Storage forms against an FP64 run: the multi-peak PQM workload, 20 steps of (accelerate,
adjust), with f stored as FP32, BF16 or FP16 on the GPU (FP32 arithmetic, exact or with
fused multiply-adds) and the oracle's Realf = double build as the FP64 run.  All runs use the same FP64 intersections.

Per cell: L1 and L-inf of f relative to the FP64 run (blocks missing on one side count as
zero), and the relative error of density n, bulk velocity |V - V64| / |V64| and scalar
pressure P = (m/3) sum f |v - V|^2 dV^3 (arch_moments.h:123-150).

Bounds asserted (the precision contract of the storage forms): density within 1e-6 relative for every
form.  The other figures are reported (set VACC_PRECISION_OUT to write them as JSON) and
bounded loosely so that a broken codec fails: L1 of f <= 2e-4 (FP32), 3e-3 (FP16), 3e-2 (BF16).
"""
import json
import os

import pytest

import oracle as O
from tests import cases
from vlasiator_amd import vacc as V
from vlasiator_amd import workloads as W

pytestmark = pytest.mark.gpu

NCELLS, STEPS = 8, 20
L1_BOUND = {"f32": 2e-4, "f16": 3e-3, "bf16": 3e-2}


def test_storage_forms_against_fp64_run():
    w = W.config(2, ncells=NCELLS)
    cells = cases.build_cells(w)
    isx = [cases.oracle_intersections(w, mo) for mo in range(3)]
    print("FP64 oracle run", flush=True)
    ref = cases.oracle_run(w, cells, isx, STEPS, fp64=True)
    report = {"workload": w.name, "cells": NCELLS, "steps": STEPS}
    for name, store, fast in (("f32", V.STORE_F32, False), ("bf16", V.STORE_BF16, False), ("f16", V.STORE_F16, False),
                              ("f32-fast", V.STORE_F32, True), ("bf16-fast", V.STORE_BF16, True)):
        print(f"GPU run, {name}", flush=True)
        ctx = V.Context(V.make_mesh(w.grid_length, w.vmin, w.vmax),
                        V.make_species(w.mass, w.charge, w.min_value, w.add_width),
                        V.make_options(w.method, storage=store, fast_arith=fast))
        ctx.upload_cells(cells)
        for s in range(STEPS):
            ctx.set_intersections(isx[s % 3])
            ctx.accelerate(s % 3)
            ctx.adjust(NCELLS)
        got = ctx.download_cells()
        ctx.close()
        report[name] = cases.compare_runs(w, got, ref)
        print(name, report[name], flush=True)
    path = os.environ.get("VACC_PRECISION_OUT")
    if path:
        with open(path, "w") as fh:
            json.dump(report, fh, indent=1)
    for name in ("f32", "bf16", "f16", "f32-fast", "bf16-fast"):
        assert report[name]["n"] <= 1e-6, (name, report[name])
        assert report[name]["l1"] <= L1_BOUND[name.split("-")[0]], (name, report[name])

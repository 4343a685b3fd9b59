#!/usr/bin/env python3
"""Benchmark of the velocity-space acceleration step on one MI355X.

A step is one full acceleration of every spatial cell of the workload: FP64 moments
(bulk-velocity pivot) + transform + SLICE-3D intersections on the device, the three 1-D
remaps in the step's map order, and the sparsity block adjust.  Inputs are resident in
HBM before the timed region.  The metric is velocity-cell updates per second:
sum over cells of 64 * (blocks at step start), per step, over the timed steps.

The JSON line also carries the roofline of the dominant kernel (the remap, k_remap),
measured with HIP events on the library's stream over the timed region, and a CPU
baseline: the C restatement of the reference CPU path (oracle/, OpenMP
schedule(dynamic,1) over cells) timed on a bounded sample of the same cells.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2] [--cells N] [--storage f32|bf16|f16]

Besides the main storage form, the line carries short runs of the other storage forms on the
same workload (storage_forms) unless --no-forms is given.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from vlasiator_amd import vacc as V  # noqa: E402
from vlasiator_amd import trans as TR  # noqa: E402
from vlasiator_amd import vlsv as VL  # noqa: E402
from vlasiator_amd import workloads as W  # noqa: E402
from vlasiator_amd import hbm_probe  # noqa: E402

METRIC = "velocity-cell updates/s per 3-D accel step; % HBM peak; FP32 vs BF16-store"
HBM_PEAK_GBPS = 8000.0
STORES = {"f32": V.STORE_F32, "bf16": V.STORE_BF16, "f16": V.STORE_F16}
# algorithmic bytes moved per block read or written by a remap pass: FP32 values, or 16-bit
# codes plus the block's FP32 scale
BYTES_PER_BLOCK = {"f32": 64 * 4, "bf16": 64 * 2 + 4, "f16": 64 * 2 + 4}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


FAST = True   # tolerance-form remap (fast_arith=1); --exact selects the exact form
AUTOTUNE = True
SUMS_FROM_F = False   # --sums-from-f: moments and removed mass summed from f, not from block sums


def make_context(w: W.Workload, storage: str = "f32", graph: bool = False, fast: bool | None = None):
    mesh = V.make_mesh(w.grid_length, w.vmin, w.vmax)
    sp = V.make_species(w.mass, w.charge, w.min_value, w.add_width)
    op = V.make_options(w.method, fast_arith=FAST if fast is None else fast, storage=STORES[storage],
                        use_graph=graph, sums_from_f=SUMS_FROM_F)
    return V.Context(mesh, sp, op)


def run_step(ctx, step, adjust):
    """One subcycle of every cell: moments pivot, transform, three remaps, adjust."""
    return ctx.step(step % 3, use_moments=True, adjust=adjust)


def host_cpus():
    """The host's CPU model and the threads the CPU baseline may use: the process's CPU share
    as the environment states it (OMP_NUM_THREADS, which the GPU host sets to its per-GPU
    share), else every CPU in the process's affinity mask."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    allowed = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        share, rule = min(int(env), allowed), "OMP_NUM_THREADS (the host's CPU share of this process)"
    else:
        share, rule = allowed, "every CPU of the process's affinity mask"
    return {"model": model, "visible": os.cpu_count(), "allowed": allowed, "share": share, "rule": rule}


def cpu_baseline(w, sample_cells, isect, target_s, nthreads):
    """Oracle (C restatement of the reference CPU path) on a bounded sample of cells."""
    import oracle as O
    O.build(ref=False) if not os.path.exists(O.LIB_PATH) else None
    m = O.make_mesh(w.grid_length, w.vmin, w.vmax)

    def run(ncell):
        cells = sample_cells[:ncell]
        off = np.zeros(ncell + 1, np.uint32)
        off[1:] = np.cumsum([len(g) for g, _ in cells])
        g = np.concatenate([c[0] for c in cells])
        f = np.concatenate([c[1] for c in cells])
        t0 = time.perf_counter()
        O.accelerate_cells(m, off, g, f, isect[:ncell], 0, w.method, w.min_value, do_adjust=w.adjust_every_step,
                           add_width=w.add_width, nthreads=nthreads, want_output=False)
        dt = time.perf_counter() - t0
        return 64.0 * float(off[-1]), dt

    n = min(8, len(sample_cells))
    upd, dt = run(n)
    per_cell = dt / n
    ncell = int(min(len(sample_cells), max(n, target_s / max(per_cell, 1e-9))))
    upd, dt, reps = 0.0, 0.0, 0
    while reps == 0 or (dt < target_s and reps < 16):
        u, d = run(ncell)   # the same step of the same cells, repeated until about target_s of work
        upd, dt, reps = upd + u, dt + d, reps + 1
    return {"value": upd / dt, "unit": "cell-updates/s", "cores": nthreads, "kind": "port",
            "sample": f"first {ncell} cells of the workload, one step (3 remaps"
                      f"{' + adjust' if w.adjust_every_step else ''}) x {reps}, {dt:.1f} s of CPU work"}


def plain_form(w, cells, isect, steps):
    """One cell per launch: every cell in its own single-cell context (cpu_accelerate_cell form)."""
    ctxs = []
    for (g, f), x in zip(cells, isect):
        c = make_context(w)
        c.upload_cells([(g, f)])
        c.set_intersections(x[None, :])
        ctxs.append(c)
    for c in ctxs:   # warm-up
        c.accelerate(0)
    n0 = sum(c.total_blocks() for c in ctxs)
    t0 = time.perf_counter()
    upd = 0
    for s in range(steps):
        for c in ctxs:
            upd += 64 * c.total_blocks()
            c.accelerate(s % 3)
            if w.adjust_every_step:
                c.adjust(1)
    dt = time.perf_counter() - t0
    for c in ctxs:
        c.close()
    return upd / dt, n0


def profiled_traffic(workload: str, storage: str, key: str = "traffic-ratio"):
    """HBM bytes per launch relative to the algorithmic bytes, as measured with rocprofv3
    FETCH_SIZE / WRITE_SIZE (copy-calibrated) for this workload and storage form and recorded
    in a profiles/ summary ("<key> <workload> <storage> <ratio>"; key traffic-ratio for
    k_remap, traffic-ratio-k_trans_walk for the translation)."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(path):
            p = line.split()
            if len(p) == 4 and p[0] == key and p[1] == workload and p[2] == storage:
                return float(p[3]), os.path.relpath(path, ROOT)
    return None, None


def io_form(w, ctx):
    """VLSV block I/O of the state held by ctx (the sample cells' initial state): the restart
    arrays written to a file under the temporary directory (page-cache writes, no fsync),
    the same arrays written to /dev/null (HBM -> host streaming and the host-side packing
    alone), and the file read back into a fresh context."""
    nblocks = ctx.total_blocks()
    nbytes = nblocks * 64 * 4
    d = tempfile.mkdtemp(prefix="vacc_io_")
    path = os.path.join(d, "restart.vlsv")
    try:
        t0 = time.perf_counter()
        with VL.Writer(path) as wr:
            wr.write_velocity_distribution(ctx, "proton")
        t_write = time.perf_counter() - t0
        t0 = time.perf_counter()
        with VL.Writer(os.devnull) as wr:
            wr.write_velocity_distribution(ctx, "proton")
        t_null = time.perf_counter() - t0
        c2 = make_context(w)
        t0 = time.perf_counter()
        with VL.Reader(path) as rd:
            rd.read_block_data(c2, "proton")
        t_read = time.perf_counter() - t0
        ok = c2.total_blocks() == nblocks
        c2.close()
        size = os.path.getsize(path)
    finally:
        if os.path.exists(path):
            os.remove(path)
        os.rmdir(d)
    if not ok:
        raise RuntimeError("VLSV read-back holds a different block count")
    log(f"vlsv io: {nblocks} blocks, write {nbytes / t_write / 1e9:.2f} GB/s, to /dev/null "
        f"{nbytes / t_null / 1e9:.2f} GB/s, read {nbytes / t_read / 1e9:.2f} GB/s")
    return {"cells": len(w.cells), "blocks": nblocks, "f_bytes": nbytes, "file_bytes": size,
            "write_GBps": round(nbytes / t_write / 1e9, 2), "write_devnull_GBps": round(nbytes / t_null / 1e9, 2),
            "read_GBps": round(nbytes / t_read / 1e9, 2), "unit": "GB/s of f (FP32)",
            "file_dir": "tempfile.gettempdir(), page cache"}


def grid_dims(n: int):
    """A near-cubic spatial grid holding n cells (x fastest)."""
    best = (n, 1, 1)
    for a in range(1, n + 1):
        if n % a:
            continue
        for b in range(1, n // a + 1):
            if (n // a) % b:
                continue
            c = n // a // b
            d = sorted((a, b, c), reverse=True)
            if max(d) - min(d) < max(best) - min(best):
                best = tuple(d)
    return best


TRANS_DX = 1.0e7    # m; with dt below, |v| dt / dx <= 0.45 on the +-3000 km/s mesh
TRANS_DT = 1.5      # s


def trans_storage(w, storage: str, steps: int, pencils):
    """x, y, z translation steps of the initial state in another storage form."""
    ctx = make_context(w, storage)
    ctx.init_peaks([W.to_vacc_peaks(c) for c in w.cells], W.boxes(w))
    nb = ctx.total_blocks()
    for d in range(3):
        TR.set_pencils(ctx, d, pencils[d])
        TR.trans_map_1d(ctx, d, TRANS_DT)
    ctx.sync()
    t0 = time.perf_counter()
    for s in range(steps):
        for d in range(3):
            TR.trans_map_1d(ctx, d, TRANS_DT)
    ctx.sync()
    el = time.perf_counter() - t0
    ctx.close()
    return {"ms_per_step": 1e3 * el / steps, "value": 64 * nb * steps / el, "steps": steps}


def trans_form(w, steps: int, warmup: int, cpu_seconds: float):
    """Spatial translation of the workload's initial state on a periodic grid of its cells:
    one step is the x, y and z translation of every cell (vacc_trans_map_1d x 3).  Every cell
    of this grid is written by the pencil walk, so a launch is k_walk_bounds + k_trans_walk
    (HIP events around both).  Roofline: 256 B read + 256 B written per block (the stencil
    neighbours' re-reads are not algorithmic).  CPU baseline: the oracle's restatement (one
    thread) of one dimension on the first 8 x 8 x 4 cells of the grid with periodic pencils."""
    dims = grid_dims(len(w.cells))
    kind = np.zeros(len(w.cells), np.uint8)
    ctx = make_context(w)
    ctx.init_peaks([W.to_vacc_peaks(c) for c in w.cells], W.boxes(w))
    nb = ctx.total_blocks()
    pencils = [TR.build_pencils_uniform(dims, (True, True, True), kind, (TRANS_DX,) * 3, d) for d in range(3)]
    for d in range(3):
        TR.set_pencils(ctx, d, pencils[d])
    for s in range(warmup):
        for d in range(3):
            TR.trans_map_1d(ctx, d, TRANS_DT)
    ctx.sync()
    ctx.profile_begin()
    t0 = time.perf_counter()
    for s in range(steps):
        for d in range(3):
            TR.trans_map_1d(ctx, d, TRANS_DT)
    ctx.sync()
    el = time.perf_counter() - t0
    prof = ctx.profile_read()
    nb_end = ctx.total_blocks()
    sample = ctx.download_cells() if cpu_seconds > 0 else None
    ctx.close()
    launches = max(prof["trans_launches"], 1)
    per_launch_ms = prof["trans_ms"] / launches
    alg = nb * 512
    achieved = alg / (per_launch_ms / 1e3) / 1e9
    out = {"grid": list(dims), "cells": len(w.cells), "blocks": nb, "steps": steps, "dt_s": TRANS_DT,
           "dx_m": TRANS_DX, "ms_per_step": 1e3 * el / steps, "value": 64 * nb * steps / el,
           "unit": "cell-updates/s (x, y, z translation per step)",
           "roofline": {"kernel": "k_walk_bounds + k_trans_walk", "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "launches": prof["trans_launches"],
                        "avg_launch_ms": round(per_launch_ms, 4), "algorithmic_bytes_per_launch": alg}}
    assert nb_end == nb
    ratio, src = profiled_traffic(w.name, "f32", key="traffic-ratio-k_trans_walk")
    if ratio is not None:
        out["roofline"]["traffic"] = int(ratio * alg)
        out["roofline"]["traffic_source"] = f"{src}: rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, x{ratio:.3f} algorithmic"
    out["storage_forms"] = {st: trans_storage(w, st, min(steps, 3), pencils) for st in ("bf16",)}
    log(f"translation: grid {dims}, {nb} blocks, {out['ms_per_step']:.2f} ms per x,y,z step, "
        f"translation launch {per_launch_ms:.3f} ms = {achieved:.0f} GB/s algorithmic")
    if sample is not None:
        import oracle as O
        sd = tuple(min(a, b) for a, b in zip((8, 8, 4), dims))
        ids = [x + dims[0] * (y + dims[1] * z) for z in range(sd[2]) for y in range(sd[1]) for x in range(sd[0])]
        cells = [sample[i] for i in ids]
        pen = TR.build_pencils_uniform(sd, (True, True, True), np.zeros(len(ids), np.uint8), (TRANS_DX,) * 3, 0)
        m = O.make_mesh(w.grid_length, w.vmin, w.vmax)
        blocks = sum(len(g) for g, _ in cells)
        reps, t_cpu = 0, 0.0
        while t_cpu < cpu_seconds:
            t0 = time.perf_counter()
            O.trans_map_1d(m, cells, pen, 0, TRANS_DT, w.min_value)
            t_cpu += time.perf_counter() - t0
            reps += 1
        out["cpu_baseline"] = {"value": 64 * blocks * reps / t_cpu / 3, "unit": out["unit"], "cores": 1,
                               "kind": "port", "sample": f"{len(ids)} cells ({blocks} blocks), x translation x {reps}, "
                                                         f"{t_cpu:.1f} s; per x,y,z step = 3 x that time"}
    return out


def measure(pops, storage: str, steps: int, warmup: int, graph: bool = False, fast: bool | None = None) -> dict:
    """Initial state of every population, W untimed steps, then K timed steps, each step
    ending in the library's one host synchronisation.  Several populations (config 4) step
    through vacc_accelerate_populations with the plasma-moment pivot.  graph=True replays
    captured steps (one population only) and runs without the per-kernel event profile."""
    ctxs = []
    tuned = []
    n_init = 0
    for w in pops:
        ctx = make_context(w, storage, graph, fast)
        t0 = time.perf_counter()
        ctx.init_peaks([W.to_vacc_peaks(c) for c in w.cells], W.boxes(w))
        n = ctx.total_blocks()
        n_init += n
        log(f"[{storage}] {w.name}: initial state {n} blocks ({n * BYTES_PER_BLOCK[storage] / 1e9:.2f} GB) "
            f"in {time.perf_counter() - t0:.1f} s")
        ctx.set_fields([V.make_fields(c.B, c.V, c.dt) for c in w.cells])
        if AUTOTUNE:
            ctx.autotune()   # remap launch shape per pass dimension (untimed)
        tuned.append(ctx.remap_shape())
        ctxs.append(ctx)
    adjust = pops[0].adjust_every_step

    def one_step(s):
        if len(ctxs) == 1:
            return run_step(ctxs[0], s, adjust)
        return V.accelerate_populations(ctxs, s % 3, use_moments=True, adjust=adjust)

    for s in range(warmup):
        one_step(s)
    for c in ctxs:
        c.sync()
        if not graph:
            c.profile_begin()
    updates = 0
    step_ms = []
    t0 = time.perf_counter()
    for s in range(warmup, warmup + steps):
        updates += 64 * sum(c.total_blocks() for c in ctxs)
        ts = time.perf_counter()
        one_step(s)
        step_ms.append(1e3 * (time.perf_counter() - ts))
    status = 0
    for c in ctxs:
        status |= c.sync()
    elapsed = time.perf_counter() - t0
    if graph:
        for c in ctxs:
            c.close()
        return {"value": updates / elapsed, "ms_per_step": 1e3 * elapsed / steps, "remap_wg_per_cu": tuned}
    profs = [c.profile_read() for c in ctxs]
    for c in ctxs:
        c.close()
    prof = {k: sum(p[k] for p in profs) for k in profs[0]}
    value = updates / elapsed
    ms_per_step = 1e3 * elapsed / steps
    log(f"[{storage}] timed {steps} steps: {ms_per_step:.3f} ms/step, {value:.4g} cell-updates/s, status {status}")

    # every pass reads its source blocks and writes its target blocks; the step's last pass also
    # writes 16 B of block sums per target block (tolerance form)
    remap_bytes = ((prof["remap_src_blocks"] + prof["remap_tgt_blocks"]) * BYTES_PER_BLOCK[storage]
                   + 16 * prof.get("remap_sum_blocks", 0))
    remap_s = prof["remap_ms"] / 1e3
    achieved = remap_bytes / remap_s / 1e9 if remap_s > 0 else 0.0
    launches = max(prof["remap_launches"], 1)
    fast_form = FAST if fast is None else fast
    # the tolerance form's pass is k_remap_lean plus k_remap over the few units left to the
    # exact arithmetic (both inside the timed remap interval)
    kernel = "k_remap_lean" if fast_form else "k_remap"
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                "kernel": kernel, "launches": prof["remap_launches"],
                "avg_launch_ms": round(prof["remap_ms"] / launches, 4),
                "algorithmic_bytes_per_launch": int(remap_bytes / launches)}
    breakdown = {k: round(prof[k] / steps, 4) for k in ("remap_ms", "build_ms", "adjust_ms", "transform_ms")}
    breakdown["host_other_ms"] = round(ms_per_step - sum(breakdown.values()), 4)
    wname = " + ".join(p.name for p in pops)
    ratio, src = profiled_traffic(wname, storage, key=f"traffic-ratio-{kernel}")
    if ratio is not None:
        roofline["traffic"] = int(ratio * remap_bytes / launches)
        roofline["traffic_source"] = f"{src}: rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, x{ratio:.3f} algorithmic"
    busy, src = profiled_traffic(wname, storage, key=f"valu-busy-{kernel}")
    if busy is not None:
        # share of SIMD quad-cycles in which at least one VALU instruction issued (gfx950 can
        # issue two in one quad-cycle), at most 1 by construction
        roofline["valu_busy"] = busy
        roofline["valu_busy_source"] = (f"{src}: rocprofv3 (SQ_ACTIVE_INST_VALU - SQ_ACTIVE_INST_VALU2) / "
                                        f"(CUs x GRBM_GUI_ACTIVE), quad-cycles with a VALU issue")
    pct = {f"p{q}": round(float(np.percentile(step_ms, q)), 3) for q in (50, 90, 99)}
    # the same updates over the remap passes' time alone (column build, adjust, moments and
    # transform excluded)
    remap_only = {"value": updates / remap_s if remap_s > 0 else 0.0, "unit": "cell-updates/s",
                  "ms_per_step": round(prof["remap_ms"] / steps, 4)}
    return {"value": value, "ms_per_step": ms_per_step, "blocks_at_start": n_init, "roofline": roofline,
            "breakdown": breakdown, "step_ms_percentiles": pct, "remap_wg_per_cu": tuned, "remap_only": remap_only}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--cells", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-cells", type=int, default=128)
    ap.add_argument("--exact", dest="fast", action="store_false",
                    help="the exact remap arithmetic (the reference's per-operation FP32 rounding, fast_arith=0) "
                         "instead of the tolerance form (fast_arith=1, the default)")
    ap.add_argument("--fast", dest="fast", action="store_true", help="the tolerance form (default)")
    ap.add_argument("--storage", choices=list(STORES), default="f32", help="storage form of f in HBM")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default remap grid (128 workgroups per CU) instead of timing candidates")
    ap.add_argument("--no-forms", dest="forms", action="store_false",
                    help="skip the short runs of the other storage forms")
    ap.add_argument("--sums-from-f", action="store_true",
                    help="tolerance form: take the moments and the adjust's removed mass from f (vacc_options."
                         "sums_from_f=1) instead of from the block sums of each step's last remap pass")
    ap.add_argument("--translation-only", action="store_true",
                    help="print only the spatial-translation measurement of the workload (one JSON line)")
    args = ap.parse_args()
    global FAST, AUTOTUNE, SUMS_FROM_F
    FAST = args.fast
    AUTOTUNE = args.autotune
    SUMS_FROM_F = args.sums_from_f

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    pops = W.populations(args.config, ncells=args.cells or None)
    w = pops[0]
    if args.translation_only:
        print(json.dumps(trans_form(w, args.steps, args.warmup, 0.0 if args.no_cpu else args.cpu_seconds)), flush=True)
        return
    ncells = len(w.cells)
    log(f"workload {w.name}: {ncells} cells, mesh {w.grid_length}, method {w.method}, storage {args.storage}")
    fields = [V.make_fields(c.B, c.V, c.dt) for c in w.cells]

    # bounded samples for the CPU baseline and the plain form, taken from the initial state
    sample_n = min(ncells, 2048)
    sample_cells = None
    isect0 = None
    if not args.no_cpu or args.plain_cells:
        # the same initial state for the first sample_n cells, built in a context of its own
        sw = W.config(args.config, ncells=sample_n)
        sctx = make_context(sw)
        sctx.init_peaks([W.to_vacc_peaks(c) for c in sw.cells], W.boxes(sw))
        sample_cells = sctx.download_cells()
        sctx.compute_intersections(fields[:sample_n], 0, use_moments=True)
        isect0 = sctx.get_intersections(sample_n)
        io = io_form(sw, sctx) if args.forms else None
        sctx.close()
    else:
        io = None
    r = measure(pops, args.storage, args.steps, args.warmup)
    out = {"metric": METRIC, "value": r["value"], "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": " + ".join(p.name for p in pops), "cells": ncells,
                      "populations": len(pops), "grid_length": list(w.grid_length),
                      "method": ["PLM", "PPM", "PQM"][w.method], "storage": args.storage,
                      "arith": "tolerance" if FAST else "exact",
                      "moment_sums": "f" if (SUMS_FROM_F or not FAST) else "block sums of the last remap pass",
                      "adjust_every_step": w.adjust_every_step, "blocks_at_start": r["blocks_at_start"],
                      "rotation_deg_per_step": w.rotation_deg,
                      "remap_wg_per_cu": [t[0] for t in r["remap_wg_per_cu"]],
                      "remap_lanes_per_wg": [t[1] for t in r["remap_wg_per_cu"]]},
           "roofline": r["roofline"], "breakdown_ms_per_step": r["breakdown"],
           "step_ms_percentiles": r["step_ms_percentiles"], "remap_only": r["remap_only"]}
    # the achievable copy bandwidth of this GPU, the second roofline denominator
    copy_gbps = hbm_probe.measure(2.0, 5)["copy_GBps"]
    out["roofline"]["copy_GBps"] = round(copy_gbps, 1)
    out["roofline"]["frac_of_copy"] = round(out["roofline"]["achieved"] / copy_gbps, 4)
    if args.forms:
        # the other storage forms on the same workload (same initial state, fewer steps)
        forms = {}
        for st in STORES:
            if st == args.storage:
                continue
            fr = measure(pops, st, min(args.steps, 10), min(args.warmup, 2))
            forms[st] = {"value": fr["value"], "ms_per_step": fr["ms_per_step"], "steps": min(args.steps, 10),
                         "remap_GBps": fr["roofline"]["achieved"], "remap_frac": fr["roofline"]["frac"]}
        # the other remap arithmetic (exact bit parity, or the tolerance form), FP32 and BF16
        for st in ("f32", "bf16"):
            fr = measure(pops, st, min(args.steps, 10), min(args.warmup, 2), fast=not FAST)
            forms[f"{st}-{'exact' if FAST else 'tolerance'}"] = {
                "value": fr["value"], "ms_per_step": fr["ms_per_step"], "steps": min(args.steps, 10),
                "remap_GBps": fr["roofline"]["achieved"], "remap_frac": fr["roofline"]["frac"]}
        out["storage_forms"] = forms
    if args.forms and len(pops) == 1:
        # the same main form with each step replayed from a captured hipGraph
        gr = measure(pops, args.storage, min(args.steps, 10), min(args.warmup, 3), graph=True)
        out["graph_form"] = {"value": gr["value"], "ms_per_step": gr["ms_per_step"], "steps": min(args.steps, 10)}

    if io is not None:
        out["vlsv_io"] = io
    if args.forms and len(pops) == 1:
        out["translation_form"] = tf = trans_form(w, min(args.steps, 10), 2, 0.0 if args.no_cpu else 5.0)
        tf["roofline"]["frac_of_copy"] = round(tf["roofline"]["achieved"] / copy_gbps, 4)
    if args.plain_cells and sample_cells is not None:
        npl = min(args.plain_cells, len(sample_cells))
        log(f"plain form: {npl} single-cell contexts")
        rate_plain, _ = plain_form(w, sample_cells[:npl], isect0[:npl], 3)
        out["plain_form"] = {"value": rate_plain, "unit": "cell-updates/s", "cells": npl,
                             "batched_over_plain": r["value"] / rate_plain}
    if not args.no_cpu:
        host = host_cpus()
        nthreads = args.cpu_threads or host["share"]
        log(f"cpu baseline: {nthreads} threads on {host['model']}, ~{args.cpu_seconds} s")
        out["cpu_baseline"] = cpu_baseline(w, sample_cells, isect0, args.cpu_seconds, nthreads)
        out["cpu_baseline"].update({"cpu_model": host["model"], "host_cpus": host["visible"],
                                    "cpus_allowed": host["allowed"], "threads_rule": host["rule"]})
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the per-step SPH particle loop on one MI355X.

Default workload: 3D saturated soil-water column, 64x32x64 soil lattice under a 64x32x16 water
layer at dx = 0.005 m plus 3-layer walls and floor (223,228 particles), Drucker-Prager soil,
Darcy drag, plain FP64. Synthetic seeded lattice (no external data).

One step = one whole reference time step (NNPS + all pair passes + constitutive + update) over
all particles, state resident in HBM before the timed region starts. The neighbour batch widths
are autotuned for the deck first (sphg_autotune; the state is restored afterwards). `value` is
particle-steps/s over exactly --steps steps of sphg_step, which replays the step as a hipGraph.
A second timed pass of the same length launches the step kernels eagerly with a HIP event pair
around every kernel on the library's stream: the per-kernel durations (kernels, roofline) are
measured live there, and its rate is reported as ms_per_step_eager_with_events.

Extra fields on the JSON line:
  roofline      dominant kernel: algorithmic bytes per launch / measured average duration vs
                8 TB/s; `traffic` = HBM bytes per launch from the committed rocprofv3 PMC summary
                (profiles/pmc_<workload>.json) when present, else null.
  mixed         the same workload in the mixed arithmetic form (SPHG_MIXED: FP32 kernel values and
                hardware reciprocals in the pair terms, FP64 sums and state), same warm-up and
                steps from the same initial state, same two passes: its graph-replay rate,
                speed-up over FP64 and field-max relative error against the FP64 context after the
                same number of steps.
  cpu_baseline  the reference solver itself (compiled from its sources into oracle/_ref) timed on
                this host's cores on a bounded sample of the same deck; where that build is absent,
                the FP64 oracle (a restatement of the reference loop, validated bit-exact against it).
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "multiphase-sph-solver_amd"))

import numpy as np  # noqa: E402

from sphgpy import column_collapse_2d, soil_water_column_3d, slope_2d  # noqa: E402

HBM_PEAK_GBS = 8000.0

# Algorithmic (compulsory) bytes per particle per kernel launch: every per-particle array the
# kernel reads or writes counted once per particle (perfect reuse of gathered neighbour data),
# plus 4 B per list entry for each pass over the neighbour list; W and grad W are recomputed.
# (F64 reads, F64 writes, int32 reads/writes, neighbour-list passes), 3D / 2D.
KERNEL_TRAFFIC = {
    3: {
        "bin": (3, 0, 2, 0), "scatter": (0, 0, 3, 0), "search": (4, 0, 3, 1),
        "density": (15, 3, 3, 1), "norm_eos": (8, 2, 3, 1), "strain": (20, 13, 3, 1),
        "stress": (31, 35, 3, 0), "accel": (26, 9, 4, 1), "xsph": (12, 6, 3, 1), "update": (21, 13, 2, 1),
    },
    2: {
        "bin": (2, 0, 2, 0), "scatter": (0, 0, 3, 0), "search": (3, 0, 3, 1),
        "density": (11, 3, 3, 1), "norm_eos": (6, 2, 3, 1), "strain": (15, 10, 3, 1),
        "stress": (31, 35, 3, 0), "accel": (22, 7, 4, 1), "xsph": (9, 5, 3, 1), "update": (15, 10, 2, 1),
    },
}


def kernel_bytes(ndim: int, name: str, n: int, nbar: float) -> float:
    fr, fw, ii, passes = KERNEL_TRAFFIC[ndim].get(name, (0, 0, 0, 0))
    return n * (8.0 * (fr + fw) + 4.0 * ii + passes * 4.0 * (nbar + 1))


def step_bytes(ndim: int, n: int, nbar: float) -> float:
    return sum(kernel_bytes(ndim, k, n, nbar) for k in KERNEL_TRAFFIC[ndim])


def make_deck(workload: str):
    if workload == "soilwater3d":
        return soil_water_column_3d(steps=2000)
    if workload == "column2d":
        return column_collapse_2d(steps=5000)
    raise SystemExit(f"unknown workload {workload}")


def make_batch(nreal: int, seed: int = 2024):
    """Monte-Carlo slope batch: nreal 2D slopes, height uniform in [5, 15] m per seed, dx = 0.1 m,
    lognormal cohesion (mean 10 kPa, COV 0.3, correlation lengths 10 m x 2 m) from a 200-mode K-L
    expansion per realisation."""
    rng = np.random.default_rng(seed)
    heights = rng.uniform(5.0, 15.0, nreal)
    decks = [slope_2d(float(h), dx=0.1, steps=3000, seed=seed + r) for r, h in enumerate(heights)]
    xis = rng.standard_normal((nreal, 200))
    return decks, xis


WORKLOAD_DESC = {
    "serve": "serving stream: Poisson arrivals (1/s) of 24 requests drawn by seed from 2D column collapse "
             "(5k-50k particles), 2D random-field slope (20k-80k) and 3D soil-water column (100k-500k + walls), "
             "500-5000 steps each, admitted against a 64 GB budget, advanced in 50-step chunks that overlap",
    "slope_batch": "Monte-Carlo slope batch: 2D slopes of height U[5,15] m at dx=0.1 m (crest 1.5H, 1:2 face, "
                   "toe 1.5H, 4 m base; 17k-89k particles each), DP phi=25, lognormal cohesion (10 kPa, COV 0.3, "
                   "lx=10 m, ly=2 m, 200 K-L modes) per realisation, all realisations in one context",
    "soilwater3d": "3D saturated soil-water column: soil 64x32x64 + water 64x32x16 lattice, dx=0.005 m, "
                   "3-layer walls/floor, DP c=1 kPa phi=25, Darcy drag, water vpu=30 m/s",
    "column2d": "2D dry granular column collapse: soil 60x120 at dx=0.002 m + 3-layer floor/left wall, "
                "DP c=0 phi=30 E=1.8 MPa, AV alpha=0.1",
    "rf_cholesky": "correlation-matrix random field (flspv 3) of a 2D slope H=10 m at dx=0.1 m (46,970 "
                   "particles): cells of 3 dr (11,200 cells), exponential ACF lx=10 m ly=2 m, lognormal "
                   "cohesion (10 kPa, COV 0.3); FP64 Cholesky of the 11,200^2 correlation matrix, then "
                   "E = L zeta and the particle values for 16 realisations",
}


REF_HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_harness")


def _ref_wall(deck, steps: int, threads: int) -> float:
    import subprocess
    import tempfile
    d = deck.copy()
    d.loop, d.fop, d.ntd = steps, 10 ** 8, threads
    with tempfile.TemporaryDirectory(prefix="sphref_") as wd:
        d.write_input_dat(os.path.join(wd, "input.dat"))
        t0 = time.perf_counter()
        r = subprocess.run([REF_HARNESS, wd], capture_output=True)
        el = time.perf_counter() - t0
    if r.returncode != 0:
        raise RuntimeError(f"reference harness failed: rc={r.returncode}")
    return el


def cpu_baseline_reference(deck, budget_s: float, what: str = "the same"):
    """The reference solver itself (its own sources compiled into oracle/_ref, OpenMP with the
    deck's ntd threads). Set-up, the initial and step-1 dumps are the same in every run, so the
    per-step time is the wall-time difference of a short and a long run."""
    threads = min(os.cpu_count() or 1, 16)
    t2 = _ref_wall(deck, 2, threads)
    t4 = _ref_wall(deck, 4, threads)
    est = max((t4 - t2) / 2.0, 1e-4)
    k = int(min(max(budget_s / est, 4), 400))
    tk = _ref_wall(deck, 2 + k, threads)
    per_step = (tk - t2) / k
    return {"value": deck.n / per_step, "unit": "particle-steps/s", "cores": threads, "kind": "reference",
            "sample": f"reference solver (its own sources, g++ -O2 -fopenmp, ntd={threads}) on {what} "
                      f"{deck.n}-particle deck: ({k + 2}-step run - 2-step run) / {k} steps = "
                      f"{per_step * 1e3:.1f} ms/step"}


def cpu_baseline(deck, state, budget_s: float):
    """The oracle restatement of the reference loop (Jacobi density form, OpenMP static loops)."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from helpers import OracleSim, JACOBI  # the checker, used here only as the CPU baseline

    threads = min(os.cpu_count() or 1, 16)
    sim = OracleSim(deck, state, mode=JACOBI, nthreads=threads)
    sim.step(1)  # warm caches / first-touch
    steps = 0
    t0 = time.perf_counter()
    while True:
        rc = sim.step(1)
        if rc:
            raise RuntimeError(f"oracle step failed: {rc}")
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 200:
            break
    return {"value": deck.n * steps / el, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of the same {deck.n}-particle deck after 1 warm-up step, "
                      f"{el:.1f} s wall, OpenMP static schedule"}


def rf_bench(args) -> None:
    """One step = one random-field generation (correlation matrix, factor, E = L zeta, values);
    the figure is the factor's rate, n^3 / 3 flops over its device time (HIP events around the
    factor's launches)."""
    from sphgpy import slope_2d
    from sphgpy.capi import rf_field, rf_generate
    deck = slope_2d(10.0, dx=0.1, seed=1)
    f = rf_field(deck, 2, 10000.0, 3000.0, (10.0, 2.0), field=1, acf=1, cell_ndx=3.0)
    rng = np.random.default_rng(7)
    nreal = 16
    steps, warm = max(1, min(args.steps, 5)), max(0, min(args.warmup, 2))
    zeta = None
    ms = []
    for k in range(warm + steps):
        if zeta is None:
            from sphgpy.capi import rf_cells
            rc, _ = rf_cells(f, deck.xp, deck.type)
            if rc:
                raise RuntimeError(f"sphg_rf_cells: {rc}")
            zeta = rng.standard_normal((nreal, f.cell_total))
        t0 = time.perf_counter()
        rc, value, _, fms = rf_generate(f, deck.xp, deck.type, zeta, fill=0.0)
        wall = time.perf_counter() - t0
        if rc:
            raise RuntimeError(f"sphg_rf_generate: {rc}")
        if k >= warm:
            ms.append((fms, wall))
    n = f.cell_total
    flops = n ** 3 / 3.0
    fms = float(np.mean([m[0] for m in ms]))
    tf = flops / (fms * 1e-3) / 1e12
    line = {"metric": "Cholesky factor rate of the correlation-matrix random field (FP64)", "value": round(tf, 3),
            "unit": "TFLOP/s", "n_gpus": 1, "steps": steps, "warmup": warm, "ms_per_step": round(fms, 3),
            "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic seeded slope deck and standard normals",
            "config": {"workload": "rf_cholesky", "description": WORKLOAD_DESC["rf_cholesky"], "cells": n,
                       "realisations": nreal, "generate_wall_s": round(float(np.mean([m[1] for m in ms])), 3)},
            "roofline": {"bound": "mfma", "achieved": round(tf, 3), "peak": 78.6, "unit": "TFLOP/s",
                         "frac": round(tf / 78.6, 4), "traffic": None, "kernel": "factor (potrf + trsm + mfma update)"},
            "cpu_baseline": None}
    if not args.no_cpu:
        line["cpu_baseline"] = rf_cpu_baseline(deck, f, budget_s=args.cpu_budget)
    print(json.dumps(line), flush=True)


def rf_cpu_baseline(deck, f, budget_s: float):
    """The reference's own Cholesky (ref_harness --chol, its OpenMP loops on up to 16 threads) on a
    leading block of the same correlation matrix, sized to about budget_s; GFLOP/s = m^3 / 3 / t."""
    if not os.path.exists(REF_HARNESS):
        return None
    import struct
    import tempfile
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from helpers import oracle_rf
    _, _, cells, _, _ = oracle_rf(deck.ndim, deck.xp, deck.type, f, np.zeros(1), cells_only=True)
    threads = min(os.cpu_count() or 1, 16)
    m = 1200
    rate = None
    while True:
        sub = rf_field_matrix(deck, f, cells, m)
        wd = tempfile.mkdtemp(prefix="rfcpu_")
        with open(os.path.join(wd, "chol_in.bin"), "wb") as fh:
            fh.write(struct.pack("<i", m) + np.ascontiguousarray(sub, "<f8").tobytes())
        t0 = time.perf_counter()
        r = subprocess.run([REF_HARNESS, "--chol", wd], capture_output=True,
                           env=dict(os.environ, OMP_NUM_THREADS=str(threads)))
        dt = time.perf_counter() - t0
        shutil.rmtree(wd, ignore_errors=True)
        if r.returncode != 0:
            return None
        rate = m ** 3 / 3.0 / dt / 1e9
        if dt > budget_s / 4 or m >= 4000:
            break
        m = min(4000, int(m * 1.6))
    return {"value": round(rate, 3), "unit": "GFLOP/s", "cores": threads, "kind": "reference",
            "sample": f"the reference's Cholesky on the leading {m} x {m} block of the same correlation matrix"}


def rf_field_matrix(deck, f, cells, m):
    """Leading m x m block of the field's correlation matrix (the oracle's restatement of the
    reference's ACF loop)."""
    import ctypes as C
    from helpers import oracle_lib
    L = oracle_lib().lib
    nc = np.ascontiguousarray(cells["nc"], np.int32)
    base = np.ascontiguousarray(cells["base"], np.float64)
    rl = np.array(list(f.rl))
    A = np.zeros((m, m))  # cells 0 .. m-1 of the same grid: the leading block
    dp = lambda a_: a_.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    L.or_rf_corr(deck.ndim, f.acf, dp(rl), C.c_double(f.beta), C.c_double(f.beta_1),
                 nc.ctypes.data_as(C.POINTER(C.c_int)), dp(base), C.c_double(cells["cell_dx"]), m, dp(A))
    return A


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="soilwater3d", choices=sorted(WORKLOAD_DESC))
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nreal", type=int, default=256, help="realisations of the slope_batch workload")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-form leg")
    ap.add_argument("--no-tune", action="store_true", help="keep the default neighbour batch widths")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (every GPU box here has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from sphgpy.capi import GpuSim, GpuBatch, batch_state, kl_field, setting_par
    SPHG_FP64, SPHG_MIXED = 0, 1

    if args.workload == "serve":
        from sphgpy.serve import Server, request_stream
        reqs = request_stream(24, rate=1.0, seed=11, steps=(500, 5000))
        stats = Server(hbm_budget=64e9, chunk=50, max_active=12).run(reqs)
        line = {"metric": "particle-steps/s at 1 GPU, mixed vs FP64 at 1e-4 rel err; p99 time-to-result",
                "value": round(stats["particle_steps_per_s"], 1), "unit": "particle-steps/s", "n_gpus": 1,
                "steps": int(sum(r.steps for r in reqs)), "warmup": 0,
                "ms_per_step": round(stats["makespan_s"] * 1e3 / max(1, sum(r.steps for r in reqs)), 4),
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                "data": "synthetic seeded request stream",
                "config": {"workload": "serve", "description": WORKLOAD_DESC["serve"],
                           "requests": stats["requests"], "failed": stats["failed"],
                           "makespan_s": round(stats["makespan_s"], 3)},
                "p50_time_to_result_s": round(stats["p50_ttr_s"], 3),
                "p99_time_to_result_s": round(stats["p99_ttr_s"], 3),
                "roofline": None, "cpu_baseline": None}
        print(json.dumps(line), flush=True)
        return

    if args.workload == "rf_cholesky":
        rf_bench(args)
        return

    def build(precision):
        """Context of the workload in the given arithmetic form, state resident, step counter 1."""
        extra = {}
        if args.workload == "slope_batch":
            decks, xis = make_batch(args.nreal)
            g = GpuBatch(decks, precision=precision)
            g.load_state(batch_state(decks))
            g.upload_done()
            t_kl = time.perf_counter()
            for r, d in enumerate(decks):
                g.kl_synthesize_part(r, kl_field(d, dist=2, mean=10000.0, sd=3000.0, rl=(10.0, 2.0)), xis[r])
            g.sync()
            extra["realisations"] = len(decks)
            extra["kl_synthesis_s"] = round(time.perf_counter() - t_kl, 3)
            extra["particles_per_realisation"] = [int(min(d.n for d in decks)), int(max(d.n for d in decks))]
            deck, state = decks[0], None
        else:
            deck = make_deck(args.workload)
            state = setting_par(deck)
            g = GpuSim(deck, precision=precision)
            g.load_state(state)
            g.upload_done()
        g.set_step(1)
        if not args.no_tune:
            # neighbour batch widths of this deck (the state is restored after the trials)
            rc, widths = g.autotune(5)
            if rc:
                raise SystemExit(f"autotune failed with error {rc}")
            extra["batch_widths"] = {"density": widths[0], "norm_eos": widths[1], "strain": widths[2],
                                     "xsph": widths[3]}
        return g, deck, state, extra

    def timed(g, per_kernel):
        rc = g.step(args.warmup)
        if rc:
            raise SystemExit(f"warm-up failed with error {rc}")
        g.enable_timing(per_kernel)
        g.sync()
        t0 = time.perf_counter()
        rc = g.step(args.steps)  # returns after the stream is synchronised
        t1 = time.perf_counter()
        if rc:
            raise SystemExit(f"timed steps failed with error {rc}")
        return t1 - t0

    g, deck, state, extra = build(SPHG_FP64)
    # value: the step as sphg_step runs it (hipGraph replay); the per-kernel durations of the
    # roofline come from a second timed pass of the same length with eager launches bracketed by
    # HIP events on the context's stream
    el = timed(g, False)
    snap = {k: g.get(k) for k in ("xp", "vx", "sig", "rho")}  # the FP64 state the mixed form is compared with
    el_events = timed(g, True)
    cnt, _ = g.neighbors()
    nbar = float(cnt.mean())
    n = g.n
    value = n * args.steps / el
    ms_per_step = el * 1e3 / args.steps

    rows = g.kernel_times()
    roof = None
    kernels = {}
    if rows:
        for name, ms, launches in rows:
            kernels[name] = {"avg_us": ms * 1e3 / launches, "launches": launches}
        dom = max((r for r in rows if r[0] in KERNEL_TRAFFIC[deck.ndim]), key=lambda r: r[1])
        dname, dms, dl = dom
        avg_s = dms / dl / 1e3
        alg = kernel_bytes(deck.ndim, dname, n, nbar)
        traffic = None
        pmc_path = os.path.join(REPO, "profiles", f"pmc_{args.workload}.json")
        if os.path.exists(pmc_path):
            with open(pmc_path) as f:
                pmc = json.load(f)
            traffic = pmc.get("kernels", {}).get(dname, {}).get("hbm_bytes_per_launch")
        achieved = alg / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dname,
                "alg_bytes_per_launch": int(alg), "avg_launch_us": round(avg_s * 1e6, 2)}
    step_alg = step_bytes(deck.ndim, n, nbar)
    finite = bool(all(np.isfinite(g.get(k)).all() for k in ("xp", "vx", "sig", "rho")))
    if not finite:
        raise SystemExit("state became non-finite during the timed steps")

    mixed = None
    if not args.no_mixed:
        # the same workload in the mixed form, same warm-up and step count from the same initial
        # state; its error is taken against the FP64 context after the same number of steps
        gm, _, _, mextra = build(SPHG_MIXED)
        elm = timed(gm, False)
        err = {}
        for k in ("xp", "vx", "sig", "rho"):
            a, b = gm.get(k), snap[k]
            err[k] = float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
        timed(gm, True)
        mrows = gm.kernel_times()
        mixed = {"value": round(n * args.steps / elm, 1), "ms_per_step": round(elm * 1e3 / args.steps, 4),
                 "speedup_vs_fp64": round(el / elm, 3),
                 "max_rel_err_vs_fp64": {k: float(f"{v:.3e}") for k, v in err.items()},
                 "err_horizon_steps": args.warmup + args.steps,
                 "batch_widths": mextra.get("batch_widths"),
                 "kernels": {name: {"avg_us": ms * 1e3 / la, "launches": la} for name, ms, la in mrows}}
        gm.close()

    line = {
        "metric": "particle-steps/s at 1 GPU, mixed vs FP64 at 1e-4 rel err; p99 time-to-result",
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic seeded lattice deck",
        "config": dict({"workload": args.workload, "description": WORKLOAD_DESC[args.workload], "particles": n,
                        "precision": "FP64", "mean_neighbours": round(nbar, 2),
                        "timing": "value: hipGraph replay of the step; kernels and roofline: a second timed pass "
                                  "of the same length, eager launches with per-kernel HIP events",
                        "ms_per_step_eager_with_events": round(el_events * 1e3 / args.steps, 4)},
                       **extra),
        "roofline": roof,
        "pair_interactions_per_s": round(float(cnt.sum()) * args.steps / el, 1),
        "step_alg_bytes": int(step_alg),
        "step_roofline_frac": round(step_alg * args.steps / el / 1e9 / HBM_PEAK_GBS, 4),
        "kernels": kernels,
        "mixed": mixed,
    }
    if not args.no_cpu:
        # for the batch: the first realisation alone (the reference runs one deck per process)
        if os.path.exists(REF_HARNESS):
            what = "the first realisation's" if args.workload == "slope_batch" else "the same"
            line["cpu_baseline"] = cpu_baseline_reference(deck, args.cpu_budget, what)
        else:
            line["cpu_baseline"] = cpu_baseline(deck, state if state is not None else setting_par(deck),
                                                args.cpu_budget)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

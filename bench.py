#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the per-step SPH particle loop on one MI355X.

Default workload: the Monte-Carlo slope batch, 256 realisations of a 2D slope of height U[5, 15] m
at dx = 0.1 m (17k-89k particles each, 11.9 M in all) with a lognormal K-L cohesion field each, in
one context, plain FP64 for `value`. Synthetic seeded lattices (no external data). Other workloads:
--workload soilwater3d (3D soil-water column, 223,228 particles), slope_pile3d (3D slope with an
embedded pile, 1.1 M particles), column2d, serve, rf_cholesky.

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
  mixed         the same workload in the mixed arithmetic form (SPHG_MIXED), same warm-up and
                steps, same two passes: its graph-replay rate, speed-up over FP64, roofline and
                pair-interaction rate.
  accuracy      per output field, max |f - f_fp64| / max |f_fp64| after 1, 10, 50 and 200 steps from
                the same initial state, for the mixed form and for the FP64 noise floor (the FP64
                form with every neighbour sum in reversed list order, sphg_set_list_order).
  kl_synthesis  the common-domain K-L synthesis H = Phi[N x 200] Xi[200 x 256] of the tallest
                slope on the matrix cores (FP32): GEMM time, TFLOP/s and the fraction of the
                157.3 TFLOP/s FP32 matrix peak.
  serve         a short serving stream (sphgpy.serve, all four request families at their real
                sizes, 100-400 steps each): p50 / p99 time-to-result and aggregate rate; the same
                p50 / p99 are repeated as p50_time_to_result_s / p99_time_to_result_s.
  cpu_baseline  the oracle (a restatement of the reference loop in its layout: particle records,
                99-slot neighbour tables, merge-sort binning, OpenMP static loops, serial density
                normalisation) on the first realisation's deck, timed on this host's cores (up to
                16 threads), with one thread, and on one socket's cores where the share spans
                sockets; host description from lscpu; the oracle's wall time relative to the
                reference solver, measured in the build container (oracle/ref_ratio.json).
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

from sphgpy import column_collapse_2d, soil_water_column_3d, slope_2d, slope_pile_3d  # noqa: E402
from sphgpy.capi import PARTICLE_FIELDS  # noqa: E402

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
    if workload == "slope_pile3d":
        return slope_pile_3d(steps=1000, v0=(-0.3, 0.1, -0.05))
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
    "serve": "serving stream: Poisson arrivals (1/s) of 24 requests, the four families in turn: 2D column "
             "collapse (5k-50k particles), 2D random-field slope (20k-80k), 3D soil-water column (100k-500k + "
             "walls) and 3D slope with pile (0.5M-2M), 500-5000 steps each, admitted against a 64 GB budget, "
             "advanced in 50-step chunks that overlap",
    "slope_batch": "Monte-Carlo slope batch: 2D slopes of height U[5,15] m at dx=0.1 m (crest 1.5H, 1:2 face, "
                   "toe 1.5H, 4 m base; 17k-89k particles each), DP phi=25, lognormal cohesion (10 kPa, COV 0.3, "
                   "lx=10 m, ly=2 m, 200 K-L modes) per realisation, all realisations in one context",
    "soilwater3d": "3D saturated soil-water column: soil 64x32x64 + water 64x32x16 lattice, dx=0.005 m, "
                   "3-layer walls/floor, DP c=1 kPa phi=25, Darcy drag, water vpu=30 m/s",
    "column2d": "2D dry granular column collapse: soil 60x120 at dx=0.002 m + 3-layer floor/left wall, "
                "DP c=0 phi=30 E=1.8 MPa, AV alpha=0.1",
    "slope_pile3d": "3D slope with an embedded pile: 200x50x100 soil lattice at dx=0.1 m with a 1:2 cut (0.9 M "
                    "AMCC soil particles), 10x12 elastic pile (11.4 k particles, E=30 GPa), deformable soil-pile "
                    "contact (SSI type 2), rigid base and side walls",
    "rf_cholesky": "correlation-matrix random field (flspv 3) of a 2D slope H=10 m at dx=0.1 m (46,970 "
                   "particles): cells of 3 dr (11,200 cells), exponential ACF lx=10 m ly=2 m, lognormal "
                   "cohesion (10 kPa, COV 0.3); FP64 Cholesky of the 11,200^2 correlation matrix, then "
                   "E = L zeta and the particle values for 16 realisations",
}


REF_HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_harness")


def host_description() -> dict:
    """nproc and the lscpu model, socket and core counts of this host."""
    info = {"nproc": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "NUMA node(s)"):
                info[k] = v
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def _oracle_rate(deck, state, threads: int, budget_s: float, max_steps: int = 200):
    """particle-steps/s of the oracle (serial density sweep, the reference's form) on `deck`."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from helpers import OracleSim, GS  # the checker, used here only as the CPU baseline

    sim = OracleSim(deck, state, mode=GS, nthreads=threads)
    sim.step(1)  # first touch
    steps, t0 = 0, time.perf_counter()
    while True:
        rc = sim.step(1)
        if rc:
            raise RuntimeError(f"oracle step failed: {rc}")
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= max_steps:
            break
    return deck.n * steps / el, steps, el


def cpu_baseline(deck, state, budget_s: float, what: str = "the same"):
    """The oracle (the reference loop restated in its own layout) timed on this host: all the cores
    of this process's share (up to 16 threads), one thread, and one socket's cores where the share
    spans more than one socket. OMP_PROC_BIND=close / OMP_PLACES=cores are set before the oracle
    library loads."""
    host = host_description()
    threads = min(os.cpu_count() or 1, 16)
    v_all, s_all, e_all = _oracle_rate(deck, state, threads, budget_s)
    v_one, s_one, e_one = _oracle_rate(deck, state, 1, budget_s / 4, max_steps=20)
    out = {"value": round(v_all, 1), "unit": "particle-steps/s", "cores": threads, "kind": "port",
           "sample": f"oracle (reference layout, serial-form density, OpenMP static, {threads} threads) on {what} "
                     f"{deck.n}-particle deck: {s_all} steps after one warm-up step, {e_all:.1f} s",
           "ntd1": {"value": round(v_one, 1), "steps": s_one, "wall_s": round(e_one, 2)},
           "host": host}
    try:
        cps = int(host.get("Core(s) per socket", "0"))
        sockets = int(host.get("Socket(s)", "1"))
    except ValueError:
        cps, sockets = 0, 1
    if sockets > 1 and 0 < cps < threads:
        v_s, s_s, e_s = _oracle_rate(deck, state, cps, budget_s / 2)
        out["one_socket"] = {"value": round(v_s, 1), "threads": cps, "steps": s_s, "wall_s": round(e_s, 2)}
    ratio = os.path.join(REPO, "oracle", "ref_ratio.json")
    if os.path.exists(ratio):
        with open(ratio) as f:
            out["oracle_vs_reference_wall_time"] = json.load(f)
    return out


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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="slope_batch", choices=sorted(WORKLOAD_DESC))
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nreal", type=int, default=256, help="realisations of the slope_batch workload")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-form leg")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the error-vs-step curves")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-tune", action="store_true", help="keep the default neighbour batch widths")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device three times (scales the HBM byte counters of a profile)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch  # device memory and the runtime's copy kernel only
        a = torch.empty(1 << 30, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (every GPU box here has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    os.environ.setdefault("OMP_PROC_BIND", "close")  # read when the oracle's OpenMP runtime loads
    os.environ.setdefault("OMP_PLACES", "cores")

    from sphgpy.capi import GpuSim, GpuBatch, batch_state, kl_field, setting_par
    SPHG_FP64, SPHG_MIXED = 0, 1
    metric = "particle-steps/s at 1 GPU, mixed vs FP64 at 1e-4 rel err; p99 time-to-result"

    if args.workload == "serve":
        stats = serve_leg(24, (500, 5000), rate=1.0)
        line = {"metric": metric, "value": stats["particle_steps_per_s"], "unit": "particle-steps/s", "n_gpus": 1,
                "steps": stats["steps"], "warmup": 0, "ms_per_step": stats["ms_per_step"],
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                "data": "synthetic seeded request stream",
                "config": {"workload": "serve", "description": WORKLOAD_DESC["serve"], **stats},
                "p50_time_to_result_s": stats["p50_ttr_s"], "p99_time_to_result_s": stats["p99_ttr_s"],
                "roofline": None, "cpu_baseline": None}
        print(json.dumps(line), flush=True)
        return

    if args.workload == "rf_cholesky":
        rf_bench(args)
        return

    cache = {}

    def build(precision, reversed_lists=False, tune=True):
        """Context of the workload in the given arithmetic form, state resident, step counter 1."""
        extra = {}
        if args.workload == "slope_batch":
            if "batch" not in cache:
                decks, xis = make_batch(args.nreal)
                t_kl = time.perf_counter()
                fields = [kl_field(d, dist=2, mean=10000.0, sd=3000.0, rl=(10.0, 2.0)) for d in decks]
                cache["batch"] = (decks, xis, fields, batch_state(decks), time.perf_counter() - t_kl)
            decks, xis, fields, state0, t_modes = cache["batch"]
            g = GpuBatch(decks, precision=precision)
            g.load_state(state0)
            g.upload_done()
            t_kl = time.perf_counter()
            for r in range(len(decks)):
                g.kl_synthesize_part(r, fields[r], xis[r])
            g.sync()
            extra["realisations"] = len(decks)
            extra["kl_modes_host_s"] = round(t_modes, 3)
            extra["kl_synthesis_device_s"] = round(time.perf_counter() - t_kl, 3)
            extra["particles_per_realisation"] = [int(min(d.n for d in decks)), int(max(d.n for d in decks))]
            deck, state = decks[0], None
        else:
            if "deck" not in cache:
                deck = make_deck(args.workload)
                cache["deck"] = (deck, setting_par(deck))
            deck, state = cache["deck"]
            g = GpuSim(deck, precision=precision)
            g.load_state(state)
            g.upload_done()
        if reversed_lists and g.set_list_order(True):
            raise SystemExit("sphg_set_list_order failed")
        g.set_step(1)
        if tune and not args.no_tune:
            # neighbour batch widths of this deck (the state is restored after the trials)
            rc, widths = g.autotune(5)
            if rc:
                raise SystemExit(f"autotune failed with error {rc}")
            extra["batch_widths"] = {"density": widths[0], "norm_eos": widths[1], "strain": widths[2],
                                     "xsph": widths[3], "search_stage": g.search_stage()}
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

    def leg(g, deck):
        """graph-replay rate, then a second pass with per-kernel events: rate, kernels, roofline"""
        el = timed(g, False)
        el_events = timed(g, True)
        cnt, _ = g.neighbors()
        nbar = float(cnt.mean())
        n = g.n
        rows = g.kernel_times()
        kernels = {name: {"avg_us": ms * 1e3 / la, "launches": la} for name, ms, la in rows}
        roof = None
        if rows:
            dname, dms, dl = max((r for r in rows if r[0] in KERNEL_TRAFFIC[deck.ndim]), key=lambda r: r[1])
            avg_s = dms / dl / 1e3
            alg = kernel_bytes(deck.ndim, dname, n, nbar)
            roof = {"bound": "hbm", "achieved": round(alg / avg_s / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(alg / avg_s / 1e9 / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dname,
                    "alg_bytes_per_launch": int(alg), "avg_launch_us": round(avg_s * 1e6, 2)}
        step_alg = step_bytes(deck.ndim, n, nbar)
        for k in ("xp", "vx", "sig", "rho"):
            if not np.isfinite(g.get(k)).all():
                raise SystemExit(f"state became non-finite during the timed steps ({k})")
        return {"value": n * args.steps / el, "ms_per_step": el * 1e3 / args.steps,
                "ms_per_step_eager_with_events": el_events * 1e3 / args.steps, "nbar": nbar,
                "pair_interactions_per_s": float(cnt.sum()) * args.steps / el, "kernels": kernels, "roofline": roof,
                "step_alg_bytes": step_alg, "step_roofline_frac": step_alg * args.steps / el / 1e9 / HBM_PEAK_GBS}

    # ---- accuracy: mixed form and FP64 noise floor against FP64, from the same initial state
    accuracy = None
    if not args.no_accuracy:
        fields = [k for k in PARTICLE_FIELDS if k not in ("divx", "dsig", "rous", "Rij", "x")]
        horizons = (1, 10, 50, 200)
        ctx = {"fp64": build(SPHG_FP64, tune=False)[0], "noise_floor": build(SPHG_FP64, True, tune=False)[0]}
        if not args.no_mixed:
            ctx["mixed"] = build(SPHG_MIXED, tune=False)[0]
        curves = {k: {f: [] for f in fields} for k in ctx if k != "fp64"}
        done = 0
        for h in horizons:
            for g in ctx.values():
                if g.step(h - done):
                    raise SystemExit("accuracy run failed")
            done = h
            ref = {f: ctx["fp64"].get(f) for f in fields}
            for k in curves:
                for f in fields:
                    a, b = ctx[k].get(f), ref[f]
                    den = float(np.max(np.abs(b)))
                    num = float(np.max(np.abs(a - b))) if a.size else 0.0
                    curves[k][f].append(float(f"{(num / den if den > 0 else num):.3e}"))
        for g in ctx.values():
            g.close()
        accuracy = {"metric": "max_i |f - f_fp64| / max_i |f_fp64| per field", "horizons_steps": list(horizons),
                    "mixed_vs_fp64": curves.get("mixed"), "fp64_noise_floor": curves["noise_floor"],
                    "noise_floor_definition": "FP64 with every neighbour sum in reversed list order"}
        # the principal directions (beta, beta_ts) are defined up to sign and order where principal
        # stresses coincide, so they stay out of the maxima (their curves are listed all the same)
        core = [f for f in fields if f not in ("beta", "beta_ts")]
        accuracy["maxima_exclude"] = ["beta", "beta_ts"]
        accuracy["noise_floor_max_at_horizon"] = {str(h): max(curves["noise_floor"][f][i] for f in core)
                                                  for i, h in enumerate(horizons)}
        if "mixed" in curves:
            worst = {f: max(curves["mixed"][f]) for f in core}
            accuracy["mixed_worst_field"] = max(worst, key=worst.get)
            accuracy["mixed_max_err_at_horizon"] = {str(h): max(curves["mixed"][f][i] for f in core)
                                                    for i, h in enumerate(horizons)}

    g, deck, state, extra = build(SPHG_FP64)
    fp = leg(g, deck)
    n = g.n
    g.close()
    pmc_path = os.path.join(REPO, "profiles", f"pmc_{args.workload}.json")
    if fp["roofline"] and os.path.exists(pmc_path):
        with open(pmc_path) as f:
            pmc = json.load(f)
        fp["roofline"]["traffic"] = pmc.get("kernels", {}).get(fp["roofline"]["kernel"], {}).get("hbm_bytes_per_launch")

    mixed = None
    if not args.no_mixed:
        gm, _, _, mextra = build(SPHG_MIXED)
        m = leg(gm, deck)
        gm.close()
        mixed = {"value": round(m["value"], 1), "ms_per_step": round(m["ms_per_step"], 4),
                 "speedup_vs_fp64": round(fp["ms_per_step"] / m["ms_per_step"], 3), "roofline": m["roofline"],
                 "pair_interactions_per_s": round(m["pair_interactions_per_s"], 1),
                 "step_roofline_frac": round(m["step_roofline_frac"], 4),
                 "batch_widths": mextra.get("batch_widths"), "kernels": m["kernels"]}

    kl = None
    if args.workload == "slope_batch":
        kl = kl_mfma_leg(cache["batch"][0], cache["batch"][1])

    serve = None if args.no_serve or args.workload != "slope_batch" else serve_leg(12, (100, 400), rate=4.0)

    line = {
        "metric": metric,
        "value": round(fp["value"], 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fp["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic seeded lattice decks",
        "config": dict({"workload": args.workload, "description": WORKLOAD_DESC[args.workload], "particles": n,
                        "precision": "FP64", "mean_neighbours": round(fp["nbar"], 2),
                        "timing": "value: hipGraph replay of the step; kernels and roofline: a second timed pass "
                                  "of the same length, eager launches with per-kernel HIP events",
                        "ms_per_step_eager_with_events": round(fp["ms_per_step_eager_with_events"], 4)},
                       **extra),
        "roofline": fp["roofline"],
        "pair_interactions_per_s": round(fp["pair_interactions_per_s"], 1),
        "step_alg_bytes": int(fp["step_alg_bytes"]),
        "step_roofline_frac": round(fp["step_roofline_frac"], 4),
        "kernels": fp["kernels"],
        "mixed": mixed,
        "accuracy": accuracy,
        "kl_synthesis": kl,
    }
    if serve:
        line["serve"] = serve
        line["p50_time_to_result_s"] = serve["p50_ttr_s"]
        line["p99_time_to_result_s"] = serve["p99_ttr_s"]
    if not args.no_cpu:
        # for the batch: the first realisation alone (the reference runs one deck per process)
        what = "the first realisation's" if args.workload == "slope_batch" else "the same"
        line["cpu_baseline"] = cpu_baseline(deck, state if state is not None else setting_par(deck),
                                            args.cpu_budget, what)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


def kl_mfma_leg(decks, xis):
    """Common-domain K-L synthesis of the tallest slope for all the batch's standard normals at once:
    H = Phi[N x 200] Xi[200 x nreal] as one FP32 matrix-core GEMM (sphg_kl_synthesize_batch)."""
    from sphgpy.capi import kl_field, kl_synthesize_batch
    d = max(decks, key=lambda x: x.n)
    f = kl_field(d, dist=2, mean=10000.0, sd=3000.0, rl=(10.0, 2.0))
    nreal = len(xis)
    kl_synthesize_batch(f, d.xp, d.type, xis, fill=0.0)  # warm-up
    ms = []
    for _ in range(5):
        _, gms = kl_synthesize_batch(f, d.xp, d.type, xis, fill=0.0)
        ms.append(gms)
    gms = float(np.median(ms))
    flops = 2.0 * d.n * 200 * nreal
    tf = flops / (gms * 1e-3) / 1e12
    return {"mode": "common domain (tallest slope), FP32 MFMA v_mfma_f32_32x32x2_f32", "particles": d.n,
            "realisations": nreal, "gemm_ms": round(gms, 4), "tflops": round(tf, 2), "peak_tflops": 157.3,
            "mfma_utilisation": round(tf / 157.3, 4)}


def serve_leg(nreq: int, steps, rate: float):
    """A serving stream of nreq requests taking the four families in turn, at their real sizes."""
    from sphgpy.serve import Server, request_stream
    reqs = request_stream(nreq, rate=rate, seed=11, steps=steps, cycle=True)
    stats = Server(hbm_budget=64e9, chunk=50, max_active=12).run(reqs)
    total = int(sum(r.steps for r in reqs))
    return {"requests": stats["requests"], "failed": stats["failed"], "errors": stats["errors"],
            "families": sorted(set(r.kind for r in reqs)),
            "particles": [int(min(r.n for r in reqs)), int(max(r.n for r in reqs))],
            "steps_per_request": list(steps), "arrival_rate_per_s": rate, "steps": total,
            "makespan_s": round(stats["makespan_s"], 3),
            "ms_per_step": round(stats["makespan_s"] * 1e3 / max(1, total), 4),
            "particle_steps_per_s": round(stats["particle_steps_per_s"], 1),
            "p50_ttr_s": round(stats["p50_ttr_s"], 3), "p99_ttr_s": round(stats["p99_ttr_s"], 3)}


if __name__ == "__main__":
    main()

"""Measures the wall time per step of the oracle (the CPU restatement timed as the CPU baseline on
the GPU box) against the reference solver built from its own sources (oracle/_ref/ref_harness),
on the same deck with the same thread counts, and writes oracle/ref_ratio.json. Runs only where
the reference checkout and its link harness exist (the build container); the GPU box reads the
committed ratio and reports it beside the CPU baseline.

    python oracle/ref_ratio.py
"""
from __future__ import annotations

import json
import os
import platform
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(REPO, "tests"))
from helpers import OracleSim, GS, have_ref, REF_HARNESS  # noqa: E402
from sphgpy import slope_2d, soil_water_column_3d  # noqa: E402
from sphgpy.capi import setting_par  # noqa: E402


def ref_per_step(deck, threads: int, k: int) -> float:
    def wall(steps):
        d = deck.copy()
        d.loop, d.fop, d.ntd = steps, 10 ** 8, threads
        with tempfile.TemporaryDirectory(prefix="sphref_") as wd:
            d.write_input_dat(os.path.join(wd, "input.dat"))
            t0 = time.perf_counter()
            r = subprocess.run([REF_HARNESS, wd], capture_output=True)
            el = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError(f"reference harness failed: {r.returncode}")
        return el
    return (wall(2 + k) - wall(2)) / k


def oracle_per_step(deck, threads: int, k: int) -> float:
    sim = OracleSim(deck, setting_par(deck), mode=GS, nthreads=threads)
    sim.step(2)
    t0 = time.perf_counter()
    assert sim.step(k) == 0
    return (time.perf_counter() - t0) / k


def main():
    if not have_ref():
        raise SystemExit("reference harness not built (make -C oracle ref)")
    cases = {"slope2d_H10": (slope_2d(10.0, dx=0.1, seed=2024), 30),
             "soilwater3d": (soil_water_column_3d(steps=10), 10)}
    ncpu = os.cpu_count() or 1
    out = {"host": platform.processor() or platform.machine(), "nproc": ncpu, "cases": {}}
    for name, (deck, k) in cases.items():
        rows = {}
        for th in (1, ncpu):
            r = ref_per_step(deck, th, k if th > 1 else max(4, k // 4))
            o = oracle_per_step(deck, th, k if th > 1 else max(4, k // 4))
            rows[f"ntd{th}"] = {"reference_ms": round(r * 1e3, 2), "oracle_ms": round(o * 1e3, 2),
                                "oracle_over_reference": round(o / r, 3)}
            print(name, deck.n, th, rows[f"ntd{th}"], flush=True)
        out["cases"][name] = {"particles": deck.n, **rows}
    with open(os.path.join(HERE, "ref_ratio.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()

"""Writes the golden fixtures tests/golden/*.npz from the reference solver itself.

Runs only where the reference checkout and its link harness (oracle/_ref/ref_harness, built by
`make -C oracle ref`) exist. Each fixture holds a seeded deck's initial state after the
reference's setting_par and the full particle state after selected steps of the reference's
serial (ntd=1) run: inputs and expected outputs, nothing else.

    python tests/golden/make_golden.py [name ...]
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from helpers import run_ref, have_ref  # noqa: E402
from sphgpy import (column_collapse_2d, soil_water_column_3d, model_zoo_2d, moving_box_2d,  # noqa: E402
                    art_stress_block)
from sphgpy.capi import STATE_FIELDS, PARTICLE_FIELDS  # noqa: E402

KEEP = sorted(set(STATE_FIELDS) | set(PARTICLE_FIELDS) | {"type", "matype", "permtype", "etype"})


def decks():
    return {
        "column2d_small": (column_collapse_2d(nx=10, ny=20, wall_h=30, floor_len=30, steps=8), 8, (1, 4, 8)),
        "soilwater3d_small": (soil_water_column_3d(nx=5, ny=4, nz_soil=5, nz_water=3, steps=5), 5, (1, 5)),
        "model_zoo_2d": (model_zoo_2d(steps=9), 9, (1, 3, 6, 9)),
        "moving_box_m1": (moving_box_2d(1, steps=6), 6, (1, 3, 6)),
        "moving_box_m2": (moving_box_2d(2, steps=6), 6, (1, 3, 6)),
        "moving_box_m3": (moving_box_2d(3, steps=6), 6, (1, 3, 6)),
        "art_stress_2d": (art_stress_block(2, steps=6), 6, (1, 2, 4, 6)),
        "art_stress_3d": (art_stress_block(3, nx=4, ny=3, nz=4, steps=6), 6, (1, 3, 6)),
    }


def main(names=None):
    if not have_ref():
        raise SystemExit("reference harness not built (make -C oracle ref)")
    for name, (deck, steps, keep_steps) in decks().items():
        if names and name not in names:
            continue
        dumps = run_ref(deck, steps, 1)
        out = {}
        for s in (0,) + tuple(keep_steps):
            for k, v in dumps[s].items():
                if k in KEEP or (s == 0 and (k.startswith("ps_") or k.startswith("pf_"))):
                    out[f"s{s}/{k}"] = v
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print(name, deck.n, "particles,", len(out), "arrays")


if __name__ == "__main__":
    main(sys.argv[1:])

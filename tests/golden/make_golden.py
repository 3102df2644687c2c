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
from helpers import run_ref, run_ref_kl, have_ref  # noqa: E402
from sphgpy import (column_collapse_2d, soil_water_column_3d, model_zoo_2d, moving_box_2d,  # noqa: E402
                    art_stress_block, slope_2d, ssi_block, tran_box, drag_box, air_water_box,
                    slope_pile_3d, vibration_box, rain_box)
from sphgpy.capi import STATE_FIELDS, PARTICLE_FIELDS  # noqa: E402

KEEP = sorted(set(STATE_FIELDS) | set(PARTICLE_FIELDS) | {"type", "matype", "permtype", "etype", "tssig", "tsqc"})


def with_flknl(deck, flknl):
    """The deck with kernel function flknl (3: B-spline, 4: Yang-Liu, both with the kernel
    gradient correction, CPP_1-SPHLoops.cpp:544-546)."""
    d = deck.copy()
    d.flknl = flknl
    return d


def with_ts(deck, alphats):
    """The deck with the TS transform (flts 1, TSSMP_before / _after around the soil models)."""
    d = deck.copy()
    d.flts, d.alphats = 1, alphats
    return d


def with_permtype(deck, permtype):
    """The deck with the permeability law permtype (2: linear in the saturation between sitar and
    sitas, 3: per-particle layers 1 / 2 / 3 -> pers / perm / perl; CPP_7-2-constitutive.cpp:373-378)."""
    d = deck.copy()
    d.permtype = permtype
    return d


def all_umcc(deck):
    """The deck with every soil particle on the unsaturated Cam-Clay model (matype 3), the only
    soil block that reads the saturation bounds sitas / sitar which permtype 2 interpolates between
    (elsewhere both stay 0 and the reference's linear law divides 0 by 0)."""
    d = deck.copy()
    d.matype = d.matype.copy()
    d.matype[d.type == 2] = 3
    # bounds inside the saturation range, so that permtype 2 takes all three of its branches
    d.soil_layers[3][0].update(sitar=0.5, sitas=0.9)
    return d


def with_bingham(deck, fai=5.0, c=2.0, strain_min=1.0):
    """The deck with Bingham water (bhtype 1): viscosity vis + (pre tan(fai) + c) / max(strain rate,
    strain_min) (CPP_7-2-constitutive.cpp:38-47)."""
    d = deck.copy()
    d.water = [dict(w) for w in d.water]
    for w in d.water:
        w.update(bhtype=1, fai=fai, c=c, strain_min=strain_min)
    return d


# soil velocity of the slope-with-pile fixtures: towards the pile's near face, with a sliding part
PILE_V0 = (-0.3, 0.1, -0.05)


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
        "ssi_2d_t1": (ssi_block(2, 1, steps=6), 6, (1, 3, 6)),
        "ssi_2d_t2": (ssi_block(2, 2, steps=6), 6, (1, 3, 6)),
        "ssi_2d_t3": (ssi_block(2, 3, steps=6), 6, (1, 3, 6)),
        "ssi_3d_t1": (ssi_block(3, 1, nx=5, ny=3, nz=3, steps=5), 5, (1, 5)),
        "ssi_3d_t3": (ssi_block(3, 3, nx=5, ny=3, nz=3, steps=5), 5, (1, 5)),
        "ssi_cpl_2d": (ssi_block(2, 1, steps=6, acc_type=2, water=2), 6, (1, 3, 6)),
        "ssi_cpl_3d": (ssi_block(3, 1, nx=5, ny=3, nz=3, steps=5, acc_type=2, water=2), 5, (1, 5)),
        "tran_2d_b3": (tran_box(2, 3, steps=6), 6, (1, 3, 6)),
        "tran_2d_b4": (tran_box(2, 4, steps=6), 6, (1, 3, 6)),
        "tran_2d_b13": (tran_box(2, 13, steps=6), 6, (1, 3, 6)),
        "tran_2d_b14": (tran_box(2, 14, steps=6), 6, (1, 3, 6)),
        "tran_3d_b3": (tran_box(3, 3, nx=5, ny=3, nz_soil=3, nz_water=2, steps=5), 5, (1, 5)),
        "tran_3d_b14": (tran_box(3, 14, nx=5, ny=3, nz_soil=3, nz_water=2, steps=5), 5, (1, 5)),
        "free_solid_2d": (tran_box(2, 5, steps=6, gap=0.4, vdown=0.3), 6, (1, 3, 6)),
        "free_solid_3d": (tran_box(3, 5, nx=5, ny=3, nz_soil=3, nz_water=2, steps=5, gap=0.4, vdown=0.3), 5, (1, 5)),
        "drag_2d_fli2": (drag_box(2, 2), 6, (1, 3, 6)),
        "drag_2d_fli2_wenyu": (drag_box(2, 2, eps0=5.0), 6, (1, 3, 6)),
        "drag_2d_fli3": (drag_box(2, 3, vis=0.02), 6, (1, 3, 6)),
        "drag_3d_fli2": (drag_box(3, 2, nx=5, ny=3, nz_soil=3, nz_water=2, steps=5, eps0=5.0), 5, (1, 5)),
        "drag_3d_fli3": (drag_box(3, 3, nx=5, ny=3, nz_soil=3, nz_water=2, steps=5, vis=0.02), 5, (1, 5)),
        "kgc_column2d_k3": (with_flknl(column_collapse_2d(nx=10, ny=20, wall_h=30, floor_len=30, steps=8), 3), 8,
                            (1, 4, 8)),
        "kgc_zoo2d_k4": (with_flknl(model_zoo_2d(steps=6), 4), 6, (1, 3, 6)),
        "kgc_ssi2d_k3": (with_flknl(ssi_block(2, 1, steps=6), 3), 6, (1, 3, 6)),
        "kgc_art3d_k3": (with_flknl(art_stress_block(3, nx=4, ny=3, nz=4, steps=6), 3), 6, (1, 3, 6)),
        "ts_column2d": (with_ts(column_collapse_2d(nx=10, ny=20, wall_h=30, floor_len=30, steps=8), 0.5), 8,
                        (1, 4, 8)),
        "ts_art3d": (with_ts(art_stress_block(3, nx=4, ny=3, nz=4, steps=6), 0.3), 6, (1, 3, 6)),
        "ts_drag3d": (with_ts(drag_box(3, 2, nx=5, ny=3, nz_soil=3, nz_water=2, steps=5), 0.4), 5, (1, 5)),
        "kgc_tran3d_k4": (with_flknl(tran_box(3, 14, nx=5, ny=3, nz_soil=3, nz_water=2, steps=5), 4), 5, (1, 5)),
        "air_water_2d_a0": (air_water_box(2, 0), 6, (1, 3, 6)),
        "air_water_2d_a1": (air_water_box(2, 1), 6, (1, 3, 6)),
        "air_water_3d_a0": (air_water_box(3, 0, nx=5, ny=3, steps=5), 5, (1, 5)),
        "air_water_3d_a1": (air_water_box(3, 1, nx=5, ny=3, steps=5), 5, (1, 5)),
        "tm_2d_b1": (tran_box(2, 1, steps=6, vdown=0.3), 6, (1, 3, 6)),
        "tm_3d_b1": (tran_box(3, 1, nx=5, ny=3, nz_soil=3, nz_water=2, steps=5, vdown=0.3), 5, (1, 5)),
        "umcc_perm2": (with_permtype(all_umcc(model_zoo_2d(steps=6)), 2), 6, (1, 3, 6)),
        "zoo_perm3": (with_permtype(model_zoo_2d(steps=6), 3), 6, (1, 3, 6)),
        "ssi_3d_t2": (ssi_block(3, 2, nx=5, ny=3, nz=3, steps=5), 5, (1, 5)),
        "pile3d_amcc": (slope_pile_3d(nx=16, ny=8, nz=10, pile=(3, 3), model=4, steps=5, v0=PILE_V0), 5, (1, 3, 5)),
        "pile3d_sb": (slope_pile_3d(nx=16, ny=8, nz=10, pile=(3, 3), model=1, steps=5, v0=PILE_V0), 5, (1, 3, 5)),
        "bingham_2d": (with_bingham(tran_box(2, 3, steps=6, v0=0.5)), 6, (1, 3, 6)),
        "bingham_3d": (with_bingham(tran_box(3, 14, nx=5, ny=3, nz_soil=3, nz_water=2, steps=5, v0=0.5), fai=8.0,
                                    c=0.5, strain_min=50.0), 5, (1, 5)),
        "vib_2d_acc": (vibration_box(2, 1), 6, (1, 3, 6)),
        "vib_2d_vel": (vibration_box(2, 2), 6, (1, 3, 6)),
        "vib_3d_acc": (vibration_box(3, 1, nx=5, ny=3, nz=3), 6, (1, 3, 6)),
        "vib_3d_vel": (vibration_box(3, 2, nx=5, ny=3, nz=3), 6, (1, 3, 6)),
        "rain_2d_r1": (rain_box(2, 1), 6, (1, 2, 4, 6)),
        "rain_2d_r2": (rain_box(2, 2, steps=9), 9, (1, 3, 9)),
        "rain_2d_r3": (rain_box(2, 3, steps=9), 9, (1, 3, 9)),
        "rain_2d_r4": (rain_box(2, 4), 6, (1, 2, 4, 6)),
        "rain_3d_r1": (rain_box(3, 1, nx=5, ny=3, nz=3), 6, (1, 2, 4, 6)),
        "rain_3d_r4": (rain_box(3, 4, nx=5, ny=3, nz=3), 6, (1, 2, 4, 6)),
    }


def kl_cases():
    """K-L inputs: (deck, dist, mean, sd, rl, coe_depth, field, xi seed, initial value)."""
    return {
        "kl_slope2d": (slope_2d(1.0, dx=0.1, seed=3), 2, 10000.0, 3000.0, (1.0, 0.2, 0.0), 200.0, 1, 11, 10000.0),
        "kl_block3d": (art_stress_block(3, nx=5, ny=4, nz=6, dx=0.2), 1, 20.0, 3.0, (1.0, 1.0, 0.5), 5.0, 0, 12,
                       30.0),
    }


def kl_inputs(name):
    """The K-L case `name` as the arrays and scalars run_ref_kl / oracle_kl take."""
    from sphgpy.capi import kl_field
    deck, dist, mean, sd, rl, cdep, field, seed, v0 = kl_cases()[name]
    f = kl_field(deck, dist=dist, mean=mean, sd=sd, rl=rl, coe_depth=cdep, field=field)
    xi = np.random.default_rng(seed).standard_normal(200)
    return deck, f, (deck.ndim, deck.xp, deck.type, list(f.a), list(f.rl), list(f.x_trans), xi, dist, mean, sd,
                     cdep, f.soil_max, f.water_max, field, np.full(deck.n, v0))


def main_kl(names=None):
    for name in kl_cases():
        if names and name not in names:
            continue
        deck, f, args = kl_inputs(name)
        wi, ev, od, val, rc = run_ref_kl(*args)
        assert rc == 0
        scal = np.array([args[8], args[9], args[10], args[11], args[12]])  # mean sd coe_depth soil_max water_max
        ints = np.array([args[0], args[7], args[13]], np.int32)  # ndim dist field
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), xp=deck.xp, type=deck.type, a=np.array(args[3]),
                            rl=np.array(args[4]), x_trans=np.array(args[5]), xi=args[6], value0=args[14], wi=wi,
                            eigval=ev, order=od, value=val, scal=scal, ints=ints)
        print(name, deck.n, "particles")


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
    args = sys.argv[1:]
    step_names = [a for a in args if not a.startswith("kl_")]
    kl_names = [a for a in args if a.startswith("kl_")]
    if not args or step_names:
        main(step_names or None)
    if not args or kl_names:
        main_kl(kl_names or None)

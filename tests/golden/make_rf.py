"""Writes the correlation-matrix random-field fixtures tests/golden/rf_*.npz from the reference.

- rf_chol.npz: for each case, the reference's own Cholesky (CPP_17-3-Nonstationary-RF.cpp:45-74,
  through `oracle/_ref/ref_harness --chol`, one thread) of the case's correlation matrix, with a
  digest of that matrix so a test knows it factors the same input.
- rf_pipeline.npz: the reference's nonstationary_random_2D/3D (`ref_harness --rf`) on seeded
  decks with sd_value 0, where its clock-seeded normals drop out: the per-particle properties
  and return codes it leaves.
- rf_pair.npz: the same for its correlated friction angle and cohesion
  (nonstationary_correlated_random_2D/3D, `ref_harness --rfc`), normal distribution.
- kl_spv.npz: the reference's whole K-L generators on seeded decks with sd_value 0 -
  SpatialVariables_Generate_2D/3D (flspv 1, `--kln`, every property flag) and
  SpatialVariables_Generate_2D_corr/3D_corr (flspv 2, `--klc`) - and its homogeneous correlated
  random variables (random_variable_generate, flspv 5, `--rv`): the properties and return codes
  they leave. Where sd is 0 the normals drop out of the normal branch; in the lognormal branch the
  correlation coefficient is 0 / 0 (NaN), which the properties fed by the mixed normal carry.

    python tests/golden/make_rf.py          # all
    python tests/golden/make_rf.py spv      # kl_spv.npz only
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from helpers import oracle_rf, run_ref_chol, run_ref_rf  # noqa: E402
from sphgpy import slope_2d, tran_box  # noqa: E402
from sphgpy.capi import rf_field  # noqa: E402


def without_walls(d):
    """The deck's water and soil particles only: in 3D the reference maps every particle to a
    cell of the widened non-boundary box, and wall particles outside it index past the grid."""
    d = d.copy()
    keep = d.type != 0
    for k in ("xp", "vxp", "type", "matype", "permtype_p", "etype"):
        setattr(d, k, getattr(d, k)[keep])
    return d


def rf_decks():
    return {
        "slope": slope_2d(1.0, dx=0.1, seed=3),
        "box2d": tran_box(2, 3, nx=10, ny=3, nz_soil=5, nz_water=4, seed=41),
        "box3d": without_walls(tran_box(3, 3, nx=6, ny=4, nz_soil=4, nz_water=3, seed=43)),
    }


# Cholesky cases: (deck, acf, rl, cell_ndx, beta, beta_1)
CHOL_CASES = {
    "exp2d": ("slope", 1, (2.0, 0.6, 0.0), 5.0, 0.0, 0.0),
    "gauss2d": ("slope", 2, (2.0, 0.6, 0.0), 5.0, 0.0, 0.0),
    "gauss2d_smooth": ("slope", 2, (8.0, 3.0, 0.0), 5.0, 0.0, 0.0),  # nearly singular: the clamp acts
    "rot2d": ("box2d", 3, (0.01, 0.004, 0.0), 1.0, 30.0, 0.0),         # not positive definite: clamps
    "rot3d": ("box3d", 3, (0.008, 0.004, 0.006), 1.0, 20.0, 35.0),
}

# sd = 0 pipeline cases: (deck, field, dist, mean, coe_depth, cell_ndx)
PIPE_CASES = {
    "slope_c_normal": ("slope", 1, 1, 10000.0, 2000.0, 4.0),
    "slope_fai_lognormal": ("slope", 0, 2, 25.0, 1.5, 4.0),
    "box2d_cop_normal": ("box2d", 2, 1, 1e-4, 0.01, 2.0),
    "box2d_ds_lognormal": ("box2d", 3, 2, 0.002, 0.05, 2.0),
    "box2d_faiw_normal": ("box2d", 4, 1, 5.0, 100.0, 2.0),
    "box2d_cw_lognormal": ("box2d", 5, 2, 3.0, 50.0, 2.0),
    "box3d_c_normal": ("box3d", 1, 1, 800.0, 1e4, 1.0),
    "box3d_fai_lognormal": ("box3d", 0, 2, 30.0, 200.0, 1.0),
    "slope_too_many_cells": ("slope", 1, 1, 1.0, 0.0, 0.2),        # error 41
    "slope_crowded_cell": ("slope", 1, 1, 1.0, 0.0, 40.0),          # error 42
}


# correlated pairs at sd = 0, normal: (deck, phase field of the friction angle, means, coe_depths, cell_ndx, coe)
PAIR_CASES = {
    "slope_soil_pair": ("slope", 0, (25.0, 9000.0), (1.5, 2500.0), 4.0, 0.4),
    "box2d_fluid_pair": ("box2d", 4, (3.0, 0.5), (80.0, 20.0), 2.0, -0.3),
    "box3d_soil_pair": ("box3d", 0, (30.0, 700.0), (300.0, 9000.0), 1.0, 0.6),
    "box3d_fluid_pair": ("box3d", 4, (2.0, 0.7), (150.0, 40.0), 1.0, 0.2),
}


def pair_inputs(name):
    dk, ffai, means, cdeps, cndx, coe = PAIR_CASES[name]
    d = rf_decks()[dk]
    fa = rf_field(d, 1, means[0], 0.0, (1.0, 0.5, 0.7), field=ffai, acf=1, cell_ndx=cndx, coe_depth=cdeps[0])
    fc = rf_field(d, 1, means[1], 0.0, (1.0, 0.5, 0.7), field=ffai + 1, acf=1, cell_ndx=cndx, coe_depth=cdeps[1])
    return d, fa, fc, coe


# flspv 1 at sd 0: (deck, {field: (dist, mean, coe_depth)})
KLN_CASES = {
    "kln_slope_c": ("slope", {1: (1, 9000.0, 2500.0)}),
    "kln_box2d_six": ("box2d", {0: (1, 25.0, 1.5), 1: (2, 800.0, 30.0), 2: (1, 1e-4, 0.01), 3: (2, 0.002, 0.05),
                                4: (1, 5.0, 100.0), 5: (2, 3.0, 50.0)}),
    "kln_box3d_three": ("box3d", {0: (2, 30.0, 200.0), 1: (1, 700.0, 9000.0), 5: (1, 2.0, 40.0)}),
}

# flspv 2 at sd 0: (deck, {pair's first field: ((dist, mean, coe_depth) of fai, of c)}, coe)
KLC_CASES = {
    "klc_slope_soil": ("slope", {0: ((1, 25.0, 1.5), (1, 9000.0, 2500.0))}, 0.4),
    "klc_box2d_fluid": ("box2d", {4: ((1, 3.0, 80.0), (1, 0.5, 20.0))}, -0.3),
    "klc_box3d_soil_lognormal": ("box3d", {0: ((2, 30.0, 300.0), (2, 700.0, 9000.0))}, 0.6),
    "klc_box2d_both": ("box2d", {0: ((1, 28.0, 2.0), (2, 600.0, 50.0)), 4: ((2, 4.0, 90.0), (1, 0.8, 10.0))}, 0.25),
}

# flspv 5 at sd 0: (deck, (flag_fai, flag_c, flag_faiw, flag_cw), {psv index: (dist, mean)}, coe)
RV_CASES = {
    "rv_soil_normal": ("box2d", (1, 1, 0, 0), {0: (1, 25.0), 1: (1, 9000.0)}, 0.5),
    "rv_soil_lognormal": ("box2d", (1, 1, 0, 0), {0: (2, 25.0), 1: (2, 9000.3)}, 0.5),
    "rv_fluid_lognormal": ("box3d", (0, 0, 1, 1), {4: (2, 3.3), 5: (2, 0.7)}, -0.2),
    "rv_c_only_default_normal": ("box2d", (0, 1, 0, 0), {1: (2, 9000.3)}, 0.5),
    "rv_all_from_soil": ("box3d", (1, 1, 1, 1), {0: (2, 31.0), 1: (2, 650.7), 4: (1, 5.0), 5: (1, 1.1)}, 0.3),
}


def spv_decks():
    """Decks of the K-L and random-variable cases: the K-L roots need domains that are not small
    against the correlation lengths (the reference's bisection never ends otherwise), so the
    boxes use dx = 0.1 m; the 3D box keeps its walls (the K-L generators skip them)."""
    return {
        "slope": slope_2d(1.0, dx=0.1, seed=3),
        "box2d": tran_box(2, 3, nx=10, ny=3, nz_soil=5, nz_water=4, dx=0.1, seed=41),
        "box3d": tran_box(3, 3, nx=6, ny=4, nz_soil=4, nz_water=3, dx=0.1, seed=43),
    }


def spv_inputs(d):
    """Harness arguments shared by the K-L and random-variable cases: the non-boundary box, the
    phase tops as input_par computes them, and six properties to start from."""
    nb = (d.type != 0) & (d.type != 7)
    box = []
    for k in range(3):
        box += [min(10000.0, float(d.xp[nb, k].min())), max(-10000.0, float(d.xp[nb, k].max()))]
    _, _, ph = d.domain()
    init = [np.random.default_rng(9).uniform(1.0, 2.0, d.n) for _ in range(6)]
    return box, ph["soil"][1], ph["water"][1], init


def chol_matrix(name):
    dk, acf, rl, cndx, beta, beta1 = CHOL_CASES[name]
    d = rf_decks()[dk]
    f = rf_field(d, 1, 1.0, 0.0, rl, acf=acf, cell_ndx=cndx, beta=beta, beta_1=beta1)
    rc, _, cells, L, _ = oracle_rf(d.ndim, d.xp, d.type, f, np.zeros(1), want_l=False)
    assert rc == 0, (name, rc)
    return f, d, cells


def corr_of(d, f, cells):
    import ctypes as C
    from helpers import oracle_lib
    L = oracle_lib().lib
    total = cells["total"]
    A = np.zeros((total, total))
    dp = lambda a_: a_.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    nc = np.ascontiguousarray(cells["nc"], np.int32)
    base = np.ascontiguousarray(cells["base"], np.float64)
    rl = np.array(list(f.rl))
    L.or_rf_corr(d.ndim, f.acf, dp(rl), C.c_double(f.beta), C.c_double(f.beta_1),
                 nc.ctypes.data_as(C.POINTER(C.c_int)), dp(base), C.c_double(cells["cell_dx"]), total, dp(A))
    return A


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.float64).tobytes()).hexdigest()


def pipe_inputs(name):
    dk, field, dist, mean, cdep, cndx = PIPE_CASES[name]
    d = rf_decks()[dk]
    f = rf_field(d, dist, mean, 0.0, (1.0, 0.5, 0.7), field=field, acf=1, cell_ndx=cndx, coe_depth=cdep)
    return d, f


def main() -> None:
    out = {}
    for name in CHOL_CASES:
        f, d, cells = chol_matrix(name)
        A = corr_of(d, f, cells)
        Lr = run_ref_chol(A)
        out[f"{name}/L"] = Lr
        out[f"{name}/digest"] = np.frombuffer(digest(A).encode(), np.uint8)
        print(name, A.shape, "clamped diagonals:", int((np.diag(Lr) == np.sqrt(1e-6)).sum()))
    np.savez_compressed(os.path.join(HERE, "rf_chol.npz"), **out)
    out = {}
    for name in PIPE_CASES:
        d, f = pipe_inputs(name)
        v = 1 if d.ndim == 2 else 2
        _, _, ph = d.domain()
        smax3 = ph["soil"][1]
        wmax3 = ph["water"][1]
        flags = [0] * 6
        flags[f.field] = 1
        psv = [(1, 0.0, 0.0, (1.0, 1.0, 1.0), 0.0)] * 6
        psv[f.field] = (f.dist, f.mean_value, 0.0, tuple(f.rl), f.coe_depth)
        rng = np.random.default_rng(5)
        init = [rng.uniform(1.0, 2.0, d.n) for _ in range(6)]  # the properties before the field
        vals, rc = run_ref_rf(d.ndim, d.xp, d.type, d.dr, list(f.box_nb), smax3, wmax3, flags, f.cell_ndx, 1, 0.0,
                              0.0, psv, init)
        assert float(smax3[v]) == f.soil_max and float(wmax3[v]) == f.water_max
        out[f"{name}/init"] = init[f.field]
        out[f"{name}/value"] = vals[f.field]
        others = [k for k in range(6) if k != f.field]
        assert all(np.array_equal(vals[k], init[k]) for k in others)  # only the flagged property changes
        out[f"{name}/rc"] = np.array([rc], np.int32)
        print(name, "rc", rc)
    np.savez_compressed(os.path.join(HERE, "rf_pipeline.npz"), **out)
    out = {}
    for name in PAIR_CASES:
        d, fa, fc, coe = pair_inputs(name)
        _, _, ph = d.domain()
        flags = [0] * 6
        flags[fa.field] = flags[fc.field] = 1
        psv = [(1, 0.0, 0.0, (1.0, 1.0, 1.0), 0.0)] * 6
        psv[fa.field] = (1, fa.mean_value, 0.0, tuple(fa.rl), fa.coe_depth)
        psv[fc.field] = (1, fc.mean_value, 0.0, tuple(fc.rl), fc.coe_depth)
        rng = np.random.default_rng(6)
        init = [rng.uniform(1.0, 2.0, d.n) for _ in range(6)]
        vals, rc = run_ref_rf(d.ndim, d.xp, d.type, d.dr, list(fa.box_nb), ph["soil"][1], ph["water"][1], flags,
                              fa.cell_ndx, 1, 0.0, 0.0, psv, init, correlated=True, coe=coe)
        out[f"{name}/init_fai"] = init[fa.field]
        out[f"{name}/init_c"] = init[fc.field]
        out[f"{name}/fai"] = vals[fa.field]
        out[f"{name}/c"] = vals[fc.field]
        out[f"{name}/rc"] = np.array([rc], np.int32)
        print(name, "rc", rc)
    np.savez_compressed(os.path.join(HERE, "rf_pair.npz"), **out)
    main_spv()


def main_spv() -> None:
    out = {}
    rl = (1.0, 0.5, 0.7)
    for kind, cases in (("kln", KLN_CASES), ("klc", KLC_CASES), ("rv", RV_CASES)):
        for name, case in cases.items():
            d = spv_decks()[case[0]]
            box, smax3, wmax3, init = spv_inputs(d)
            flags = [0] * 6
            psv = [(1, 0.0, 0.0, rl, 0.0)] * 6
            coe = 0.0
            if kind == "kln":
                for fld, (dist, mean, cdep) in case[1].items():
                    flags[fld] = 1
                    psv[fld] = (dist, mean, 0.0, rl, cdep)
            elif kind == "klc":
                coe = case[2]
                for fld, (pa, pb) in case[1].items():
                    flags[fld] = flags[fld + 1] = 1
                    psv[fld] = (pa[0], pa[1], 0.0, rl, pa[2])
                    psv[fld + 1] = (pb[0], pb[1], 0.0, rl, pb[2])
            else:
                coe = case[3]
                for q, fld in enumerate((0, 1, 4, 5)):
                    flags[fld] = case[1][q]
                for fld, (dist, mean) in case[2].items():
                    psv[fld] = (dist, mean, 0.0, rl, 0.0)
            res = run_ref_rf(d.ndim, d.xp, d.type, d.dr, box, smax3, wmax3, flags, 1.0, 1, 0.0, 0.0, psv, init,
                             coe=coe, kind=kind)
            vals, rc = res[0], res[1]
            if kind != "rv":
                out[f"{name}/a"] = res[2]
            out[f"{name}/init"] = np.array(init)
            out[f"{name}/value"] = vals
            out[f"{name}/rc"] = np.array([rc], np.int32)
            changed = [k for k in range(6) if not np.array_equal(vals[k], init[k])]
            print(name, "rc", rc, "changed", changed, "nan", [k for k in range(6) if np.isnan(vals[k]).any()])
    np.savez_compressed(os.path.join(HERE, "kl_spv.npz"), **out)


if __name__ == "__main__":
    if sys.argv[1:] == ["spv"]:
        main_spv()
    else:
        main()

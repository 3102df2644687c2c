"""Host-side checks that need no GPU: library exports, host initial state, deck format, kernel
closed forms and neighbour-set properties of the oracle."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

from fixtures import load_golden, bitequal
from helpers import OracleSim, GS, REPO, max_rel
from sphgpy import column_collapse_2d, soil_water_column_3d
from sphgpy.capi import LIB_PATH, setting_par, problem_from_deck

HEADER = os.path.join(REPO, "include", "sphg.h")


def _declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"\b(sphg_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB_PATH), "build the library first (__graft_entry__.build())"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (sphg_\w+)", out))
    missing = [s for s in _declared() if s not in exported]
    assert missing == []


def test_create_rejects_unknown_arithmetic_form():
    # argument checks run before any device call
    import ctypes as C
    from sphgpy.capi import product_lib, materials_from_deck
    deck = column_collapse_2d(nx=4, ny=4, wall_h=6, floor_len=8, steps=1, seed=0)
    pro, mat = problem_from_deck(deck), materials_from_deck(deck)
    ctx = C.c_void_p()
    L = product_lib().lib
    assert L.sphg_create(C.byref(ctx), C.byref(pro), C.byref(mat), deck.n, 2) == 100  # SPHG_E_ARG
    assert L.sphg_create(C.byref(ctx), C.byref(pro), C.byref(mat), deck.n, -1) == 100
    assert not ctx.value


def test_time_loop_rejects_bad_arguments():
    # argument checks of sphg_run run before any device call
    import ctypes as C
    from sphgpy.capi import product_lib, OUTPUT_FN
    L = product_lib().lib
    t = C.c_double(0.0)
    assert L.sphg_run(None, 10, 1, 1.0, C.byref(t), OUTPUT_FN(), None) == 100


@pytest.mark.parametrize("name", ["column2d_small", "soilwater3d_small"])
def test_setting_par_matches_reference_init(name):
    deck, states = load_golden(name)
    mine = setting_par(deck)
    ref = states[0]
    for k, v in mine.items():
        assert bitequal(v, ref[k]), k


def test_deck_roundtrip_exact_doubles(tmp_path):
    d = column_collapse_2d(nx=3, ny=4, wall_h=5, floor_len=6, steps=1)
    p = tmp_path / "input.dat"
    d.write_input_dat(str(p))
    txt = p.read_text().splitlines()
    assert all(len(line) >= 3 for line in txt)  # one-character lines break the reference scanner
    i = txt.index("PA section") + 2
    first = txt[i].split()
    assert float(first[0]) == d.xp[0, 0] and float(first[1]) == d.xp[0, 1]


def test_cell_grid_matches_reference_formula():
    d = column_collapse_2d(nx=10, ny=20, wall_h=30, floor_len=30, steps=1)
    p = problem_from_deck(d)
    cs = 2.0 * d.dr
    assert int((p.xmax - p.xmin) / cs + 3.0) >= 3
    assert p.xmin == pytest.approx(d.xp[:, 0].min() - 2 * d.dr)


def test_kernel_partition_of_unity_and_lattice_neighbours():
    # interior of a lattice at spacing dx with support 2*dx: 8 strict neighbours plus the 4
    # axis-aligned ones sitting on the cut-off (kept or not by FP64 rounding); sum m/rho W ~ 1.
    deck, states = load_golden("column2d_small")
    sim = OracleSim(deck, states[0], mode=GS)
    sim.step(1)
    cnt, ids = sim.neighbors()
    xp0 = states[0]["xp"]
    interior = (deck.type == 2) & (xp0[:, 0] > 3 * deck.dr) & (xp0[:, 0] < 7 * deck.dr) & \
               (xp0[:, 1] > 3 * deck.dr) & (xp0[:, 1] < 17 * deck.dr)
    assert interior.sum() > 10
    assert np.all((cnt[interior] >= 8) & (cnt[interior] <= 12))
    # each pair is symmetric (same predicate both ways up to rounding of on-cutoff pairs)
    n = deck.n
    sym = 0
    tot = 0
    for i in np.nonzero(interior)[0]:
        for j in ids[i, : cnt[i]]:
            tot += 1
            sym += i in ids[j, : cnt[j]]
    assert sym / tot > 0.97
    # B-spline normalisation constant: integral of W over the plane is 1
    a1 = 0.682092613250980
    r = np.linspace(0, 2, 200001)
    s = r
    w = np.where(s <= 1, 2 / 3 - s * s + 0.5 * s ** 3, (2 - s) ** 3 / 6)
    integral = np.trapezoid(w * 2 * np.pi * r, r) * a1
    assert integral == pytest.approx(1.0, rel=2e-7)


def test_affine_velocity_gives_exact_strain_rate():
    # CSPM-corrected gradient (CPP_7-1-strain.cpp:153-179) reproduces an affine velocity field
    # v = A x exactly on a symmetric stencil; the seeded 1e-6 dx jitter leaves ~1e-5 relative error.
    deck, states = load_golden("column2d_small")
    st = dict(states[0])
    A = np.array([[1.5e-3, -0.7e-3, 0.0], [2.1e-3, -0.4e-3, 0.0], [0.0, 0.0, 0.0]])
    v = st["xp"] @ A.T
    v[deck.type == 0] = 0.0
    st["vxp"] = v
    st["vx"] = v.copy()
    sim = OracleSim(deck, st, mode=GS)
    sim.step(1)
    deps = sim.get("deps")
    weps = sim.get("weps").reshape(-1, 3, 3)
    xp = st["xp"]
    dr = deck.dr
    inner = (deck.type == 2) & (xp[:, 0] > 3 * dr) & (xp[:, 0] < 7 * dr) & (xp[:, 1] > 3 * dr) & (xp[:, 1] < 17 * dr)
    assert inner.sum() > 10
    np.testing.assert_allclose(deps[inner, 0], A[0, 0], rtol=1e-4)
    np.testing.assert_allclose(deps[inner, 1], A[1, 1], rtol=1e-4)
    np.testing.assert_allclose(deps[inner, 3], A[0, 1] + A[1, 0], rtol=1e-4)
    np.testing.assert_allclose(weps[inner, 0, 1], 0.5 * (A[0, 1] - A[1, 0]), rtol=1e-4)

"""Time step on the device (timestep, CPP_12-2-dt.cpp:13-20).

- SPHG_DT_CONSTANT (default) keeps Para_Pro::dt for every step, as the reference does.
- SPHG_DT_CFL computes dt = min(c_cour dr / (c_wave + max|v|), 0.25 sqrt(dr / max|a|), dt_max) as an
  FP64 min-reduction over the particles before every step. Pinned two ways:
  * at rest (v = a = 0) with c_cour 0.25, c_wave 600 it is the reference's constant 0.25 dr / 600,
    bit for bit;
  * step by step it equals the same formula evaluated in numpy FP64 on the velocity and acceleration
    downloaded before the step (maxima are exact, sqrt and division correctly rounded).
- In the CFL mode the clock advances by each step's dt and sphg_run reports that clock.
"""
import numpy as np
import pytest

from sphgpy import column_collapse_2d, soil_water_column_3d
from sphgpy.capi import GpuSim, setting_par

pytestmark = pytest.mark.gpu


def make(deck, state):
    g = GpuSim(deck)
    g.load_state(state)
    g.upload_done()
    g.set_step(1)
    return g


def cfl_dt(g, deck, cour, cwave, dt_max):
    """The CFL mode's formula in numpy FP64, same operation order, from the downloaded state."""
    v, a, ty = g.get("vxp"), g.get("axp"), g.get("type")
    keep = ty != 7
    v2 = np.zeros(g.n)
    a2 = np.zeros(g.n)
    for k in range(deck.ndim):
        v2 = v2 + v[:, k] * v[:, k]
        a2 = a2 + a[:, k] * a[:, k]
    v2m = float(v2[keep].max()) if keep.any() else 0.0
    a2m = float(a2[keep].max()) if keep.any() else 0.0
    dt = cour * deck.dr / (cwave + np.sqrt(v2m))
    if a2m > 0.0:
        df = 0.25 * np.sqrt(deck.dr / np.sqrt(a2m))
        dt = min(dt, df)
    return min(dt, dt_max)


def test_constant_mode_is_the_default():
    deck = column_collapse_2d(nx=10, ny=20, wall_h=25, floor_len=30, steps=3)
    g = make(deck, setting_par(deck))
    assert g.dt() == deck.dt
    assert g.step(3) == 0
    assert g.dt() == deck.dt
    assert g.set_dt_mode(2) != 0 and g.set_dt_mode(1, c_cour=-1.0) != 0 and g.set_dt_mode(1, dt_max=0.0) != 0


def test_cfl_at_rest_is_the_reference_constant():
    deck = column_collapse_2d(nx=10, ny=20, wall_h=25, floor_len=30, steps=3)
    st = setting_par(deck)
    st["vxp"][:] = 0.0
    st["axp"][:] = 0.0
    g = make(deck, st)
    assert g.set_dt_mode(1, 0.25, 600.0, 1.0) == 0
    assert g.step(1) == 0
    ref = 0.25 * deck.dr / 600.0  # clUpdate_Fun::timestep
    assert g.dt().hex() == ref.hex()


@pytest.mark.parametrize("deck_fn", [
    lambda: column_collapse_2d(nx=12, ny=24, wall_h=30, floor_len=40, steps=40),
    lambda: soil_water_column_3d(nx=8, ny=4, nz_soil=8, nz_water=3, steps=20),
], ids=["column2d", "soilwater3d"])
def test_cfl_matches_formula_step_by_step(deck_fn):
    deck = deck_fn()
    g = make(deck, setting_par(deck))
    cour, cwave, dt_max = 0.3, 40.0, 1.0
    assert g.set_dt_mode(1, cour, cwave, dt_max) == 0
    t = 0.0
    g.set_time(0.0)
    dts = []
    for s in range(12):
        want = cfl_dt(g, deck, cour, cwave, dt_max)
        assert g.step(1) == 0
        got = g.dt()
        assert got.hex() == np.float64(want).hex(), (s, got, want)
        t = t + got
        dts.append(got)
        assert g.time() == t
    assert len(set(dts)) > 1  # the bodies accelerate: dt changes
    for k in ("xp", "vx", "sig", "rho"):
        assert np.isfinite(g.get(k)).all()


def test_cfl_dt_cap_and_switch_back():
    deck = column_collapse_2d(nx=10, ny=20, wall_h=25, floor_len=30, steps=5)
    g = make(deck, setting_par(deck))
    assert g.set_dt_mode(1, 0.25, 600.0, 1e-9) == 0
    assert g.step(2) == 0
    assert g.dt() == 1e-9
    assert g.set_dt_mode(0) == 0
    assert g.dt() == deck.dt
    assert g.step(1) == 0
    assert g.dt() == deck.dt


def test_run_reports_the_device_clock():
    deck = column_collapse_2d(nx=10, ny=20, wall_h=25, floor_len=30, steps=10)
    g = make(deck, setting_par(deck))
    assert g.set_dt_mode(1, 0.3, 40.0, 1.0) == 0
    seen = []
    rc, t = g.run(8, 2, 1.0, 0.0, output=lambda l, tm: seen.append((l, tm)))
    assert rc == 0
    assert [l for l, _ in seen] == [1, 2, 4, 6, 8]
    assert t == g.time() and t == seen[-1][1]
    ts = [tm for _, tm in seen]
    assert all(b > a for a, b in zip(ts, ts[1:]))

"""Mixed arithmetic form (SPHG_MIXED) against the FP64 form of the same library and the oracle.

The FP64 form is pinned bit-exact (or within 1e-9 where device transcendentals enter) to the
oracle and through it to the reference, so the distance mixed -> FP64 is the distance mixed ->
reference up to those bounds. Error metric per field: max_i |f_mixed - f_fp64| / max_i |f_fp64|
(field-max normalised). Bars:
  - one step from an identical state: <= 1e-5 on every state field (kernel values carry FP32
    rounding, ~6e-8 relative, and so do the neighbour fields the momentum sums gather from their
    packed FP32 records; the stress-divergence sums cancel by up to two orders of magnitude, so
    the acceleration carries up to ~2e-6 of its field maximum; sums and state stay FP64);
  - 100 steps of the 2D column collapse: <= 1e-4 on positions, velocities, stresses, densities;
  - neighbour lists after one step: identical (the search is the FP64 code in both forms);
  - a mixed realisation batch equals its realisations run alone in mixed form, bit for bit.
"""
import numpy as np
import pytest

from fixtures import load_golden, bitequal, flush_denormal
from helpers import max_rel
from sphgpy.capi import GpuSim, setting_par

pytestmark = pytest.mark.gpu

FIELDS = ("xp", "vx", "sig", "rho", "eps", "pre", "ax", "vsig", "porosity")
MIXED = 1


def make(deck, state, precision):
    g = GpuSim(deck, precision=precision)
    g.load_state(state)
    g.upload_done()
    g.set_step(1)
    return g


def rel(a, b):
    a, b = flush_denormal(np.asarray(a)), flush_denormal(np.asarray(b))
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return float("inf")
    return max_rel(np.where(na, 0.0, a), np.where(nb, 0.0, b))


def errors(gm, gd, fields=FIELDS):
    return {k: rel(gm.get(k), gd.get(k)) for k in fields}


@pytest.mark.parametrize("name", ["column2d_small", "soilwater3d_small", "model_zoo_2d", "art_stress_3d",
                                  "moving_box_m1", "ssi_2d_t2", "ssi_cpl_3d", "tran_2d_b14", "free_solid_3d",
                                  "kgc_zoo2d_k4", "kgc_tran3d_k4", "drag_3d_fli2", "ts_drag3d"])
def test_mixed_one_step_from_identical_state(name):
    deck, states = load_golden(name)
    gd = make(deck, states[0], 0)
    gm = make(deck, states[0], MIXED)
    assert gd.step(1) == 0 and gm.step(1) == 0
    e = errors(gm, gd)
    print(name, {k: f"{v:.2e}" for k, v in e.items()})
    for k in ("ax", "sig"):
        d = np.abs(gm.get(k) - gd.get(k)).reshape(deck.n, -1).max(axis=1)
        worst = np.argsort(d)[-4:]
        print(k, [(int(i), int(deck.type[i]), float(d[i])) for i in worst])
    assert max(e.values()) <= 1e-5, e
    assert 0.0 < max(e.values())  # the mixed form is in use (FP32 kernel values differ from FP64)
    cd, idd = gd.neighbors()
    cm, im = gm.neighbors()
    assert np.array_equal(cd, cm) and np.array_equal(idd, im)


def test_mixed_column_collapse_100_steps():
    deck, states = load_golden("column2d_small")
    gd = make(deck, states[0], 0)
    gm = make(deck, states[0], MIXED)
    worst = {}
    for s in range(0, 100, 20):
        assert gd.step(20) == 0 and gm.step(20) == 0
        e = errors(gm, gd, ("xp", "vx", "sig", "rho"))
        worst = {k: max(worst.get(k, 0.0), v) for k, v in e.items()}
    print({k: f"{v:.2e}" for k, v in worst.items()})
    assert max(worst.values()) <= 1e-4, worst


def test_mixed_batch_equals_mixed_standalone():
    from sphgpy import slope_2d
    from sphgpy.capi import GpuBatch, batch_state, kl_field
    decks = [slope_2d(h, dx=0.25, seed=s) for h, s in ((1.0, 1), (1.4, 2))]
    rng = np.random.default_rng(9)
    fields = [kl_field(d, dist=2, mean=10000.0, sd=3000.0, rl=(1.0, 0.3)) for d in decks]
    xis = [rng.standard_normal(200) for _ in decks]
    b = GpuBatch(decks, precision=MIXED)
    b.load_state(batch_state(decks))
    b.upload_done()
    for r in range(len(decks)):
        b.kl_synthesize_part(r, fields[r], xis[r])
    b.set_step(1)
    assert b.step(5) == 0
    for r, d in enumerate(decks):
        g = GpuSim(d, precision=MIXED)
        g.load_state(setting_par(d))
        g.upload_done()
        g.kl_synthesize(fields[r], xis[r])
        g.set_step(1)
        assert g.step(5) == 0
        for k in ("xp", "vx", "sig", "rho", "eps"):
            assert bitequal(b.part(k, r), g.get(k)), (r, k)


@pytest.mark.parametrize("name", ["column2d_small", "soilwater3d_small", "model_zoo_2d", "ssi_cpl_3d",
                                  "air_water_2d_a1", "pile3d_amcc"])
def test_mixed_one_step_vs_oracle(name):
    # the mixed form against the FP64 oracle (Jacobi density form) from the same state: the same
    # 1e-5 field-max bar as against the library's FP64 form
    from helpers import OracleSim, JACOBI
    deck, states = load_golden(name)
    gm = make(deck, states[0], MIXED)
    o = OracleSim(deck, states[0], mode=JACOBI)
    assert gm.step(1) == 0 and o.step(1) == 0
    e = {k: rel(gm.get(k), o.get(k)) for k in FIELDS}
    print(name, {k: f"{v:.2e}" for k, v in e.items()})
    assert max(e.values()) <= 1e-5, e
    cd, idd = gm.neighbors()
    co, io = o.neighbors()
    assert np.array_equal(cd, co) and np.array_equal(idd, io)


def test_mixed_column_collapse_100_steps_vs_oracle():
    from helpers import OracleSim, JACOBI
    deck, states = load_golden("column2d_small")
    gm = make(deck, states[0], MIXED)
    o = OracleSim(deck, states[0], mode=JACOBI)
    assert gm.step(100) == 0 and o.step(100) == 0
    e = {k: rel(gm.get(k), o.get(k)) for k in ("xp", "vx", "sig", "rho")}
    print({k: f"{v:.2e}" for k, v in e.items()})
    assert max(e.values()) <= 1e-4, e


@pytest.mark.parametrize("name", ["vib_2d_acc", "vib_3d_vel", "rain_2d_r1", "rain_2d_r2", "rain_2d_r4",
                                  "rain_3d_r4", "eob"])
def test_mixed_dynamic_sets_and_vibration(name):
    # over the fixture horizon (rain rows / cement groups turning reserve particles into water, the
    # water-level decrease, excavation stages, moving walls): the mixed form keeps the particle set
    # of the FP64 form exactly, and its fields within the 1e-4 bar; a packed FP32 record left stale
    # by a particle that changes type would break both
    if name == "eob":
        from sphgpy import eob_box
        deck = eob_box()
        state, nsteps = setting_par(deck), max(s for _, _, s in deck.ecp) + 2
    else:
        deck, states = load_golden(name)
        state, nsteps = states[0], max(states)
    gd = make(deck, state, 0)
    gm = make(deck, state, MIXED)
    assert gd.step(nsteps) == 0 and gm.step(nsteps) == 0
    for k in ("type", "matype", "etype"):
        assert np.array_equal(gd.get(k), gm.get(k)), k
    e = errors(gm, gd, ("xp", "vx", "sig", "rho", "ax"))
    print(name, nsteps, {k: f"{v:.2e}" for k, v in e.items()})
    assert max(e.values()) <= 1e-4, e
    assert 0.0 < max(e.values())

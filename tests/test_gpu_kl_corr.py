"""K-L fields of every property, the correlated K-L pair (flspv 2) and the homogeneous correlated
random variables (flspv 5) on the device, against the oracle and the reference's fixtures; the step
with the per-particle strength they leave (Drucker-Prager friction angle and cohesion, Bingham
water) against the oracle with the same per-particle values.

Bars: device K-L synthesis (OCML cos / sin / log / exp) within 1e-10 field-max relative of the
oracle; at sd 0 the reference's values bit for bit in the normal branch, within 1e-15 in the
lognormal one (exp(log(trend)) on the device), NaN patterns equal; random variables bit-exact
(host arithmetic); steps bit-exact (constants derived on the host with the C library).
"""
import os

import numpy as np
import pytest

from fixtures import bitequal, load_golden
from helpers import GOLDEN, JACOBI, OracleSim, max_rel, oracle_kl, oracle_kl_corr_normals, oracle_random_variable
from golden.make_rf import KLC_CASES, KLN_CASES, spv_decks, spv_inputs

pytestmark = pytest.mark.gpu

RL = (1.0, 0.5, 0.7)
PROP = {0: "ps_fai", 1: "ps_c", 2: "ps_cop", 3: "ps_ds", 4: "pf_fai", 5: "pf_c"}
MEANS = {0: 28.0, 1: 9000.0, 2: 1e-4, 3: 0.002, 4: 6.0, 5: 1.5}


def _gpu(deck):
    from sphgpy.capi import GpuSim, setting_par
    g = GpuSim(deck)
    g.load_state(setting_par(deck))
    g.upload_done()
    return g


def _phase(deck, field):
    return deck.type == (2 if field <= 3 else 1)


def _oracle_values(deck, f, xi, value):
    *_, val, rc = oracle_kl(deck.ndim, deck.xp, deck.type, np.array(list(f.a)), tuple(f.rl),
                            np.array(list(f.x_trans)), xi, f.dist, f.mean_value, f.sd_value, f.coe_depth, f.soil_max,
                            f.water_max, f.field, value)
    assert rc == 0
    return val


@pytest.mark.parametrize("dk", ["box2d", "box3d"])
@pytest.mark.parametrize("dist", [1, 2])
def test_kl_every_property_vs_oracle(dk, dist):
    from sphgpy.capi import kl_field
    d = spv_decks()[dk]
    g = _gpu(d)
    rng = np.random.default_rng(20 + dist)
    for field in range(6):
        before = g.get(PROP[field])
        f = kl_field(d, dist, MEANS[field], 0.2 * MEANS[field], RL, field=field, coe_depth=0.01 * MEANS[field])
        xi = rng.standard_normal(200)
        g.kl_synthesize(f, xi)
        got = g.get(PROP[field])
        want = _oracle_values(d, f, xi, before.copy())
        ph = _phase(d, field)
        assert max_rel(got[ph], want[ph]) <= 1e-10, field
        assert np.std(got[ph]) > 0.01 * MEANS[field], field
        if field != 1:  # the cohesion array holds the slot values only once a field wrote it
            assert bitequal(got[~ph], before[~ph]), field


@pytest.mark.parametrize("dk,first", [("slope", 0), ("box2d", 4), ("box3d", 0), ("box3d", 4)])
def test_kl_corr_pair_vs_oracle(dk, first):
    from sphgpy.capi import kl_field, kl_corr_normals
    d = spv_decks()[dk]
    g = _gpu(d)
    fa = kl_field(d, 2, MEANS[first], 0.25 * MEANS[first], RL, field=first)
    fb = kl_field(d, 1, MEANS[first + 1], 0.3 * MEANS[first + 1], RL, field=first + 1)
    u = np.random.default_rng(31).standard_normal((200, 2))
    for coe in (0.55, -0.3):
        ba, bb = g.get(PROP[first]), g.get(PROP[first + 1])
        assert g.kl_synthesize_corr(fa, fb, coe, u) == 0
        xa, xb, _ = oracle_kl_corr_normals(2, fa.mean_value, fa.sd_value, fb.mean_value, fb.sd_value, coe, u)
        assert bitequal(kl_corr_normals(fa, fb, coe, u)[1], xb)
        ph = _phase(d, first)
        for fld, f, x, b in ((first, fa, xa, ba), (first + 1, fb, xb, bb)):
            got = g.get(PROP[fld])
            want = _oracle_values(d, f, x, b.copy())
            assert max_rel(got[ph], want[ph]) <= 1e-10, (coe, fld)
    # the pair is correlated through the shared normals: coefficient 1 with equal parameters gives
    # equal fields
    fb1 = kl_field(d, 2, MEANS[first], 0.25 * MEANS[first], RL, field=first + 1)
    assert g.kl_synthesize_corr(fa, fb1, 1.0, u) == 0
    ph = _phase(d, first)
    assert max_rel(g.get(PROP[first])[ph], g.get(PROP[first + 1])[ph]) <= 1e-15


def test_kl_corr_refuses_other_field_pairs():
    from sphgpy.capi import kl_field
    d = spv_decks()["box2d"]
    g = _gpu(d)
    fa, fb = kl_field(d, 1, 1.0, 0.1, RL, field=1), kl_field(d, 1, 1.0, 0.1, RL, field=2)
    assert g.kl_synthesize_corr(fa, fb, 0.3, np.zeros((200, 2))) == 100  # SPHG_E_ARG


def _fix(name):
    z = np.load(os.path.join(GOLDEN, "kl_spv.npz"), allow_pickle=False)
    return {k.split("/", 1)[1]: z[k] for k in z.files if k.startswith(name + "/")}


def _close_or_nan(got, want, tol):
    na, nb = np.isnan(got), np.isnan(want)
    assert np.array_equal(na, nb)
    return max_rel(np.where(na, 0.0, got), np.where(nb, 0.0, want)) <= tol


@pytest.mark.parametrize("name", sorted(KLN_CASES) + sorted(KLC_CASES))
def test_kl_generators_through_the_library_are_the_references_at_sd_zero(name):
    from sphgpy.capi import kl_field
    z = _fix(name)
    kln = name in KLN_CASES
    case = KLN_CASES[name] if kln else KLC_CASES[name]
    d = spv_decks()[case[0]]
    g = _gpu(d)
    init = z["init"]
    for fld in range(6):
        if fld != 1:
            g.set(PROP[fld], init[fld])
    g.set("ps_c", init[1])
    u = np.random.default_rng(8).standard_normal((200, 2))
    if kln:
        for fld, (dist, mean, cdep) in case[1].items():
            g.kl_synthesize(kl_field(d, dist, mean, 0.0, RL, field=fld, coe_depth=cdep), u[:, 0].copy())
    else:
        for fld, (pa, pb) in case[1].items():
            fa = kl_field(d, pa[0], pa[1], 0.0, RL, field=fld, coe_depth=pa[2])
            fb = kl_field(d, pb[0], pb[1], 0.0, RL, field=fld + 1, coe_depth=pb[2])
            assert g.kl_synthesize_corr(fa, fb, case[2], u) == 0
    dist_of = {fld: v[0] for fld, v in case[1].items()} if kln else \
        {k: v for fld, (pa, pb) in case[1].items() for k, v in ((fld, pa[0]), (fld + 1, pb[0]))}
    for fld in range(6):
        got, want = g.get(PROP[fld]), z["value"][fld]
        # the normal branch is the trend, bit for bit; the lognormal one exp(log(trend)) on the device
        assert _close_or_nan(got, want, 0.0 if dist_of.get(fld, 1) == 1 else 1e-15), fld


@pytest.mark.parametrize("flags", [(1, 1, 0, 0), (0, 0, 1, 1), (1, 1, 1, 1), (0, 1, 0, 0)])
@pytest.mark.parametrize("dist", [1, 2])
def test_random_variable_vs_oracle(flags, dist):
    d = spv_decks()["box3d"]
    g = _gpu(d)
    before = np.array([g.get(PROP[k]) if k != 1 else np.zeros(d.n) for k in range(6)])
    psv = [(dist, 31.0, 4.0), (dist, 850.0, 120.0), (1, 0.0, 0.0), (1, 0.0, 0.0), (dist, 6.0, 1.5), (dist, 1.7, 0.3)]
    z = (0.45, -1.1)
    assert g.random_variable_generate(flags, 0.4, psv, z) == 0
    want = oracle_random_variable(flags, 0.4, psv, z, d.type, before)
    for q, fld in enumerate((0, 1, 4, 5)):
        got = g.get(PROP[fld])
        ph = _phase(d, fld)
        if flags[q]:
            assert bitequal(got[ph], want[fld][ph]), fld
            assert np.unique(got[ph]).size == 1
        elif fld != 1:
            assert bitequal(got, before[fld]), fld


def test_random_variable_one_realisation_of_a_batch():
    from sphgpy.capi import GpuBatch, batch_state
    decks = [spv_decks()["box2d"], spv_decks()["box2d"]]
    g = GpuBatch(decks)
    g.load_state(batch_state(decks))
    g.upload_done()
    before = g.get("ps_fai")
    assert g.random_variable_generate((1, 1, 0, 0), 0.2, [(1, 33.0, 2.0), (1, 700.0, 80.0)] + [(1, 0.0, 0.0)] * 4,
                                      (1.0, 0.5), r=1) == 0
    after = g.get("ps_fai")
    a, b = int(g.offsets[1]), int(g.offsets[2])
    assert bitequal(after[:a], before[:a])
    soil = decks[1].type == 2
    assert np.all(after[a:b][soil] == 33.0 + 2.0 * (0.2 * 1.0 + np.sqrt(1 - 0.04) * 0.5))


def test_steps_with_random_strength_bit_exact():
    # Bingham water and Drucker-Prager soil with per-particle strength from flspv 5 (soil) and a
    # correlated K-L pair (water): the same per-particle values in the oracle give the same steps,
    # bit for bit, and they differ from the steps with the table strength
    from golden.make_golden import with_bingham
    from sphgpy.capi import GpuSim, kl_field, setting_par
    deck = with_bingham(spv_decks()["box2d"])
    st0 = setting_par(deck)
    g = GpuSim(deck)
    g.load_state(st0)
    g.upload_done()
    g.set_step(1)
    fa = kl_field(deck, 1, 5.0, 1.5, RL, field=4)
    fb = kl_field(deck, 2, 2.0, 0.6, RL, field=5)
    assert g.kl_synthesize_corr(fa, fb, 0.5, np.random.default_rng(4).standard_normal((200, 2))) == 0
    # weak enough that the soil yields within the six steps
    psv = [(1, 5.0, 1.0), (1, 0.1, 0.02)] + [(1, 0.0, 0.0)] * 4
    assert g.random_variable_generate((1, 1, 0, 0), -0.3, psv, (0.8, -0.6)) == 0
    st = dict(st0)
    for k in ("ps_fai", "ps_c", "pf_fai", "pf_c"):
        st[k] = g.get(k)
    water, soil = deck.type == 1, deck.type == 2
    assert np.std(st["pf_fai"][water]) > 0.1 and np.std(st["pf_c"][water]) > 0.01
    o = OracleSim(deck, st, mode=JACOBI)
    base = OracleSim(deck, st0, mode=JACOBI)
    assert g.step(6) == 0 and o.step(6) == 0 and base.step(6) == 0
    for k in ("xp", "vx", "sig", "rho", "eps", "pre"):
        assert bitequal(g.get(k), o.get(k)), k
    assert np.abs(g.get("sig")[water] - base.get("sig")[water]).max() > 1e-3
    assert np.abs(g.get("sig")[soil] - base.get("sig")[soil]).max() > 1e-3

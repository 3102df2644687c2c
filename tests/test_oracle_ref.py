"""Pins the oracle (oracle/sph_oracle.c) to the reference solver.

1. Against the committed golden fixtures (reference output, serial run): bit-equal at every stored
   step in Gauss-Seidel density mode.
2. Where the reference link harness is built (oracle/_ref), against fresh reference runs on larger
   decks and more steps.
"""
import os

import numpy as np
import pytest

from fixtures import load_golden, bitequal, flush_denormal
from helpers import OracleSim, GS, JACOBI, have_ref, run_ref, max_rel
from sphgpy import column_collapse_2d, soil_water_column_3d, model_zoo_2d, moving_box_2d, art_stress_block
from sphgpy.capi import PARTICLE_FIELDS

COMPARE = [k for k in PARTICLE_FIELDS]


def _mismatches(sim, ref):
    bad = []
    for k in COMPARE:
        if k not in ref:
            continue
        got = sim.get(k)
        want = ref[k]
        if k == "adps":
            got, want = flush_denormal(got), flush_denormal(want)
        if not bitequal(got, want):
            bad.append((k, max_rel(got, want)))
    return bad


GOLDEN = ["column2d_small", "soilwater3d_small", "model_zoo_2d", "moving_box_m1", "moving_box_m2", "moving_box_m3",
          "art_stress_2d", "art_stress_3d", "ssi_2d_t1", "ssi_2d_t2", "ssi_2d_t3", "ssi_3d_t1", "ssi_3d_t3",
          "ssi_cpl_2d", "ssi_cpl_3d", "tran_2d_b3", "tran_2d_b4", "tran_2d_b13", "tran_2d_b14", "tran_3d_b3",
          "tran_3d_b14", "free_solid_2d", "free_solid_3d", "drag_2d_fli2", "drag_2d_fli2_wenyu", "drag_2d_fli3",
          "drag_3d_fli2", "drag_3d_fli3", "kgc_column2d_k3", "kgc_zoo2d_k4", "kgc_ssi2d_k3", "kgc_art3d_k3",
          "kgc_tran3d_k4", "ts_column2d", "ts_art3d", "ts_drag3d", "air_water_2d_a0", "air_water_2d_a1",
          "air_water_3d_a0", "air_water_3d_a1", "tm_2d_b1", "tm_3d_b1", "umcc_perm2", "zoo_perm3", "ssi_3d_t2",
          "pile3d_amcc", "pile3d_sb", "bingham_2d", "bingham_3d", "vib_2d_acc", "vib_2d_vel", "vib_3d_acc",
          "vib_3d_vel", "rain_2d_r1", "rain_2d_r2", "rain_2d_r3", "rain_2d_r4", "rain_3d_r1", "rain_3d_r4"]


@pytest.mark.parametrize("name", GOLDEN)
def test_oracle_matches_golden_bitwise(name):
    deck, states = load_golden(name)
    sim = OracleSim(deck, states[0], mode=GS)
    done = 0
    for s in sorted(k for k in states if k > 0):
        assert sim.step(s - done) == 0
        done = s
        assert _mismatches(sim, states[s]) == [], f"step {s}"


def test_oracle_fixture_is_not_trivial():
    # the compared fields actually move over the fixture horizon
    deck, states = load_golden("column2d_small")
    last = max(states)
    assert np.max(np.abs(states[last]["vx"])) > 1e-6
    assert np.max(np.abs(states[last]["sig"] - states[0]["sig"])) > 1.0


def test_jacobi_differs_from_gs_only_slightly():
    # the GPU's Jacobi density normalisation vs the reference's serial in-place sweep (quirk A1)
    deck, states = load_golden("column2d_small")
    a = OracleSim(deck, states[0], mode=GS)
    b = OracleSim(deck, states[0], mode=JACOBI)
    a.step(8)
    b.step(8)
    assert max_rel(b.get("rho"), a.get("rho")) < 1e-6
    assert max_rel(b.get("xp"), a.get("xp")) < 1e-9
    assert not bitequal(a.get("rho"), b.get("rho"))


@pytest.mark.skipif(not have_ref(), reason="reference link harness not built")
@pytest.mark.parametrize("gen,steps", [
    (lambda: column_collapse_2d(nx=24, ny=40, wall_h=50, floor_len=80, steps=40, seed=7), 40),
    (lambda: soil_water_column_3d(nx=6, ny=5, nz_soil=6, nz_water=3, steps=12, seed=3), 12),
    (lambda: model_zoo_2d(nx=14, steps=20, seed=11), 20),
    (lambda: moving_box_2d(1, steps=12), 12),
    (lambda: moving_box_2d(2, steps=12), 12),
    (lambda: moving_box_2d(3, steps=12), 12),
    (lambda: art_stress_block(2, steps=10), 10),
    (lambda: art_stress_block(3, steps=10), 10),
])
def test_oracle_matches_reference_every_step(gen, steps):
    deck = gen()
    dumps = run_ref(deck, steps, 1)
    sim = OracleSim(deck, dumps[0], mode=GS)
    for s in range(1, steps + 1):
        assert sim.step(1) == 0
        assert _mismatches(sim, dumps[s]) == [], f"step {s}"


@pytest.mark.parametrize("name", ["ssi_2d_t1", "ssi_2d_t2", "ssi_3d_t1"])
def test_ssi_contact_force_closed_form(name):
    # Step 1 of the soil-structure decks: the soil block moves at v0 = (0.4 tangential, 0.6 towards
    # the plate(, 0.3 tangential)) onto a resting plate. For a particle in contact, the rigid form's
    # force (interaction_ss_rigid) is m (v_s - v_i) / dt split along the plate normal, with the
    # tangential part capped at myu |f_n| (Coulomb); interfss = zeta f. The deformable form scales
    # v_s - v_i by m_s / (m_s + m_i), m_s the mass interpolated between the structure particles
    # (all 2500 dx^d against the soil's 2000 dx^d). The plate normal is the lattice axis up to the
    # 1e-6 dx jitter. Momentum: the structure particles receive -f in total.
    deck, states = load_golden(name)
    sim = OracleSim(deck, states[0], mode=GS)
    assert sim.step(1) == 0
    f = sim.get("interfss")
    soil, struct = deck.type == 2, deck.type == 4
    ndim = deck.ndim
    nrm = 1 if ndim == 2 else 2
    tan = [0] if ndim == 2 else [0, 1]
    zeta, myu, dt = deck.ss_zeta, deck.ss_myu, deck.dt
    m = 2000.0 * deck.dr ** ndim
    scale = 1.0 if deck.ss_type == 1 else 2500.0 / 4500.0
    contact = soil & (np.abs(f).sum(axis=1) > 0)
    assert contact.sum() >= 4
    vt = np.array([0.4, 0.3])[: len(tan)]
    for i in np.nonzero(contact)[0]:
        fn = zeta * m * scale * 0.6 / dt
        ftan = zeta * m * scale * min(np.linalg.norm(vt), myu * 0.6) / dt
        assert f[i, nrm] == pytest.approx(fn, rel=2e-5)
        assert np.linalg.norm(f[i, tan]) == pytest.approx(ftan, rel=2e-5)
        assert np.all(np.sign(f[i, tan]) == -np.sign(vt))  # friction opposes the sliding
    tot_soil = f[soil].sum(axis=0)
    tot_struct = f[struct].sum(axis=0)
    np.testing.assert_allclose(tot_struct, -tot_soil, rtol=1e-12, atol=1e-12 * np.abs(tot_soil).max())
    # the contact force reaches ax: ax_ssi - ax_without = interfss / m
    deck0 = deck.copy()
    deck0.ss_type = 0
    sim0 = OracleSim(deck0, states[0], mode=GS)
    assert sim0.step(1) == 0
    dax = sim.get("ax") - sim0.get("ax")
    np.testing.assert_allclose(dax[contact], f[contact] / m, rtol=1e-9)


@pytest.mark.parametrize("name", GOLDEN)
def test_golden_fixture_states_are_finite(name):
    # a deck that drives the reference to NaN would make every comparison vacuous
    deck, states = load_golden(name)
    for s, st in states.items():
        for k in ("xp", "vx", "sig", "rho"):
            assert np.isfinite(st[k]).all(), (s, k)


@pytest.mark.parametrize("name", ["free_solid_2d", "free_solid_3d"])
def test_free_solid_wall_contact_closed_form(name):
    # flbndy 5 (boundary_free_solid): the soil and water start 0.4 dx above the floor moving at
    # (0.2 along, 0.3 towards the floor); a particle in contact with the resting floor gets
    # zeta m (v_wall - v) / dt split along the floor normal with the tangential part capped at
    # myu |f_n|, added to ax: the difference to the same deck with zeta = 0 at step 1
    deck, states = load_golden(name)
    a = OracleSim(deck, states[0], mode=GS)
    d0 = deck.copy()
    d0.ss_zeta = 0.0
    b = OracleSim(d0, states[0], mode=GS)
    assert a.step(1) == 0 and b.step(1) == 0
    dax = a.get("ax") - b.get("ax")
    hit = np.abs(dax).sum(axis=1) > 0
    assert hit.sum() >= 4
    nrm = deck.ndim - 1
    zeta, myu, dt = deck.ss_zeta, deck.ss_myu, deck.dt
    # particles next to a side wall may take a tilted wall line (the nearest-two / -three choice
    # follows dist_sorting's id-only swaps); the closed form holds over the floor's interior
    xs = states[0]["xp"][:, : deck.ndim - 1] / deck.dr
    span = xs[deck.type == 2].max(axis=0)
    interior = np.all((xs > 1.0) & (xs < span - 1.0), axis=1)
    assert (hit & interior).sum() >= 1
    for i in np.nonzero(hit & interior)[0]:
        assert deck.type[i] in (1, 2)
        assert dax[i, nrm] == pytest.approx(zeta * 0.3 / dt, rel=2e-5)
        assert dax[i, 0] == pytest.approx(-zeta * min(0.2, myu * 0.3) / dt, rel=2e-5)


@pytest.mark.parametrize("name", ["drag_2d_fli2", "drag_2d_fli2_wenyu", "drag_2d_fli3", "drag_3d_fli2", "drag_3d_fli3"])
def test_drag_fixtures_are_not_trivial(name):
    # the reference's interact2 / interact3 act on both phases and move them apart from the same
    # deck without drag
    deck, states = load_golden(name)
    last = max(states)
    f = states[last]["interf"]
    assert np.abs(f[deck.type == 2]).max() > 0.0 and np.abs(f[deck.type == 1]).max() > 0.0
    d0 = deck.copy()
    d0.fli = 0
    a = OracleSim(d0, states[0], mode=GS)
    a.step(last)
    assert not bitequal(a.get("vx"), states[last]["vx"])


def test_drag_laws_differ():
    # the same deck under fli 1, 2 and 3 gives three different states
    deck, states = load_golden("drag_2d_fli2")
    vx = []
    for fli in (1, 2, 3):
        d = deck.copy()
        d.fli = fli
        s = OracleSim(d, states[0], mode=GS)
        assert s.step(3) == 0
        vx.append(s.get("vx"))
    assert not bitequal(vx[0], vx[1]) and not bitequal(vx[1], vx[2]) and not bitequal(vx[0], vx[2])


def _corrected_moments(name):
    """Per soil particle, M[k][m] = sum_j V_j (x_j - x_i)_k g'_m over the soil / boundary /
    structure neighbours, g' the gradients the oracle holds after the search and the kernel
    gradient correction of the fixture's first step."""
    import ctypes as C
    from helpers import oracle_lib
    deck, states = load_golden(name)
    st = states[0]
    sim = OracleSim(deck, st, mode=GS)
    assert sim.phase(1) == 0
    L = oracle_lib().lib
    L.or_get_neighbors.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int]
    L.or_get_wij.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_int]
    n, nd = deck.n, deck.ndim
    cnt = np.zeros(n, np.int32)
    ids = np.zeros((n, 99), np.int32)
    L.or_get_neighbors(sim.ctx, cnt.ctypes.data_as(C.POINTER(C.c_int)), ids.ctypes.data_as(C.POINTER(C.c_int)), n)
    vol = st["mass"] / st["rho"]
    out = {}
    buf = np.zeros((99, 4))
    for i in np.flatnonzero(st["type"] == 2):
        nc = L.or_get_wij(sim.ctx, int(i), buf.ctypes.data_as(C.POINTER(C.c_double)), 99)
        js = ids[i, :nc]
        keep = np.isin(st["type"][js], (0, 2, 4))
        x = st["xp"][js][:, :nd] - st["xp"][i, :nd]
        out[int(i)] = np.einsum("j,jk,jm->km", vol[js][keep], x[keep], buf[:nc][keep][:, :nd])
    return deck, out


@pytest.mark.parametrize("name", ["kgc_column2d_k3", "kgc_art3d_k3"])
def test_kernel_gradient_correction_restores_first_order_consistency(name):
    # the corrected gradients differentiate linear fields exactly: sum_j V_j x_ij (x) g'_ij = I
    # (the correction matrix is built from the same sum with the uncorrected gradients, which is
    # symmetric up to rounding), wherever the matrix was invertible
    deck, mom = _corrected_moments(name)
    eye = np.eye(deck.ndim)
    err = [np.abs(m - eye).max() for m in mom.values()]
    assert len(err) > 40 and max(err) < 1e-9, max(err)


@pytest.mark.parametrize("name", ["ts_column2d", "ts_art3d", "ts_drag3d"])
def test_ts_transform_state_is_the_references(name):
    # TSSMP_before / _after: the transformed stress (tssig) and qc (tsqc) the reference leaves
    deck, states = load_golden(name)
    sim = OracleSim(deck, states[0], mode=GS)
    done = 0
    for s in sorted(k for k in states if k > 0):
        assert sim.step(s - done) == 0
        done = s
        assert bitequal(sim.get("tssig", 6), states[s]["tssig"]) and bitequal(sim.get("tsqc"), states[s]["tsqc"])
    assert (states[done]["tssig"] != states[done]["sig"]).any()



@pytest.mark.parametrize("name", ["air_water_2d_a0", "air_water_3d_a1"])
def test_air_and_tait_wall_pressures_closed_form(name):
    # Air keeps the reference's comma operator (CPP_7-1-eos_CR.cpp:34): the pow call is never made,
    # so p = 1.403 prep (-1) + prep = -0.403 prep whatever the density; matype-1 walls follow the
    # Tait law prep ((rho / rhop)^7 - 1) + prep against their own rhop (:37-41). Both hold in the
    # reference's values at every stored step.
    deck, states = load_golden(name)
    air = deck.type == 3
    tait = (deck.type == 0) & (deck.matype == 1)
    assert air.sum() > 0 and tait.sum() > 0
    for s in sorted(k for k in states if k > 0):
        st = states[s]
        np.testing.assert_array_equal(st["pre"][air], 1.403 * st["prep"][air] * -1.0 + st["prep"][air])
        r = st["rho"][tait] / st["rhop"][tait]
        np.testing.assert_allclose(st["pre"][tait], st["prep"][tait] * (r ** 7 - 1.0) + st["prep"][tait],
                                   rtol=1e-13)
        assert np.abs(st["interf"][air]).max() > 0.0  # the x2000 air drag acts


def test_permeability_laws_closed_form():
    # permtype 3: per-particle layers 1 / 2 / 3 take pers / perm / perl (Class_particle.cpp:162-173);
    # permtype 2: (perl - pers) satu / (sitas - sitar) + pers between the bounds (:147-158)
    deck, states = load_golden("zoo_perm3")
    soil = deck.type == 2
    st = states[1]
    want = np.select([deck.permtype_p == 1, deck.permtype_p == 2, deck.permtype_p == 3],
                     [deck.pers, deck.perm, deck.perl], np.nan)
    np.testing.assert_array_equal(st["cop"][soil], want[soil])
    assert len(np.unique(st["cop"][soil])) == 3
    deck, states = load_golden("umcc_perm2")
    soil = deck.type == 2
    st = states[3]
    sitas, sitar = states[0]["ps_sitas"][soil], states[0]["ps_sitar"][soil]
    sat = st["satu"][soil]
    lin = (deck.perl - deck.pers) * sat / (sitas - sitar) + deck.pers
    np.testing.assert_allclose(st["cop"][soil], lin, rtol=1e-15)


def test_takeda_morris_wall_slows_approach():
    # flbndy 1 (tmboundary, CPP_13-boundary.cpp:125-177): for a particle approaching a wall particle
    # (v . x < 0 with x = x_i - x_j) the velocity gains (1 - beta) x (v . x) / |x|^2 with
    # beta = min(1.5, 1 + 0.3 dr / (|x| - 0.3 dr)) >= 1, i.e. it loses approach speed; running the
    # same deck with flbndy 2 (no such correction) must give a larger approach speed at the floor
    deck, states = load_golden("tm_2d_b1")
    st1 = OracleSim(deck, states[0], mode=GS)
    assert st1.step(1) == 0
    d2 = deck.copy()
    d2.flbndy = 2
    st2 = OracleSim(d2, states[0], mode=GS)
    assert st2.step(1) == 0
    low = ((deck.type == 1) | (deck.type == 2)) & (deck.xp[:, 1] < 1.5 * deck.dr)
    assert low.sum() > 0
    assert np.all(st1.get("vx")[low, 1] >= st2.get("vx")[low, 1])
    assert np.any(st1.get("vx")[low, 1] > st2.get("vx")[low, 1])


def test_vibration_fixtures_exercise_every_branch():
    # the vibration decks move their walls while the window is open and stop them outside it or
    # under the oversized load; the free-field and damping terms change the soil's acceleration
    from sphgpy.capi import setting_par  # noqa: F401
    for name, wall_mt in (("vib_2d_acc", 3), ("vib_2d_vel", 4)):
        deck, states = load_golden(name)
        base = (deck.type == 0) & (deck.matype == wall_mt)
        side = (deck.type == 0) & (deck.matype == 5)
        assert np.abs(states[3]["interfss"][base]).max() > 0.0
        assert np.abs(states[3]["interfss"][side]).max() > 0.0
        assert np.abs(states[6]["interfss"][base]).max() == 0.0  # t past edtime
        d0 = deck.copy()
        d0.flv = 0
        o = OracleSim(d0, states[0], mode=GS)
        assert o.step(6) == 0
        soil = deck.type == 2
        assert np.abs(o.get("ax")[soil] - states[6]["ax"][soil]).max() > 1.0


def test_rain_fixtures_change_the_particle_set():
    # rain and cement turn the reserve into water one row / group per period; the water-level
    # decrease returns water to the reserve
    for name, before, after in (("rain_2d_r1", 0, 16), ("rain_2d_r4", 0, 16), ("rain_3d_r1", 0, 30),
                                ("rain_2d_r2", 24, 15), ("rain_2d_r3", 24, 15)):
        deck, states = load_golden(name)
        last = max(states)
        assert int((states[0]["type"] == 1).sum()) == before, name
        assert int((states[last]["type"] == 1).sum()) == after, name
    deck, states = load_golden("rain_2d_r1")
    assert int((states[2]["type"] == 1).sum()) == 8 and int((states[4]["type"] == 1).sum()) == 16


@pytest.mark.skipif(not have_ref(), reason="reference link harness not built")
def test_excavation_swap_is_the_references():
    # exca_or_back for one stage through the harness against the oracle's swap (or_eob)
    import struct
    import subprocess
    import tempfile
    from helpers import REF_HARNESS
    from sphgpy import eob_box
    d = eob_box()
    for (a, b, _) in d.ecp:
        wd = tempfile.mkdtemp(prefix="spheob_")
        with open(os.path.join(wd, "eob_in.bin"), "wb") as f:
            f.write(struct.pack("<iii", d.n, a, b))
            f.write(np.ascontiguousarray(d.type, "<i4").tobytes() + np.ascontiguousarray(d.etype, "<i4").tobytes())
        assert subprocess.run([REF_HARNESS, "--eob", wd], capture_output=True).returncode == 0
        raw = np.frombuffer(open(os.path.join(wd, "eob_out.bin"), "rb").read(), "<i4")
        ty, et = raw[:d.n], raw[d.n:]
        want_t, want_e = d.type.copy(), d.etype.copy()
        want_t[a - 1:b], want_e[a - 1:b] = d.etype[a - 1:b], d.type[a - 1:b]
        assert np.array_equal(ty, want_t) and np.array_equal(et, want_e)


def test_oracle_excavation_and_backfill_stages():
    # the oracle fires stage s at its step, one stage per step, in order: the excavated part
    # leaves (type 7) at step 2, the backfill turns into soil at step 4
    from sphgpy import eob_box
    from sphgpy.capi import setting_par
    d = eob_box()
    o = OracleSim(d, setting_par(d), mode=GS)
    (a1, b1, _), (a2, b2, _) = d.ecp
    seen = []
    for _ in range(6):
        assert o.step(1) == 0
        t = o.get_i("type") if hasattr(o, "get_i") else o.get("type")
        seen.append((int((np.asarray(t)[a1 - 1:b1] == 7).all()), int((np.asarray(t)[a2 - 1:b2] == 2).all())))
    assert seen == [(0, 0), (1, 0), (1, 0), (1, 1), (1, 1), (1, 1)]


def test_oracle_particle_deletion_is_the_references():
    # partidel never fires in the fixture runs (no particle reaches 7 dr past the box); its rule is
    # pinned on a particle set through the harness: the water beyond +-7 dr of the box in x and above
    # y = 0.12 stops and leaves, other particles stay
    import ctypes as C
    import struct
    import subprocess
    import tempfile
    from helpers import REF_HARNESS, oracle_lib
    if not have_ref():
        pytest.skip("reference link harness not built")
    deck, states = load_golden("rain_2d_r2")
    st = {k: v.copy() for k, v in states[0].items()}
    bmin, bmax, _ = deck.domain()
    dr = deck.dr
    rng = np.random.default_rng(3)
    n = deck.n
    xs = np.concatenate([bmax[0] + 7.0 * dr + rng.uniform(-1.5, 1.5, n // 2) * dr,
                         bmin[0] - 7.0 * dr + rng.uniform(-1.5, 1.5, n - n // 2) * dr])
    st["xp"][:, 0] = xs
    st["xp"][:, 1] = rng.uniform(0.10, 0.14, n)
    st["vxp"] = rng.uniform(-1, 1, (n, 3))
    st["vx"] = rng.uniform(-1, 1, (n, 3))
    st["ax"] = rng.uniform(-1, 1, (n, 3))
    ty = np.where(rng.uniform(0, 1, n) < 0.7, 1, 2).astype(np.int32)
    st["type"] = ty
    p = problem_of(deck)
    wd = tempfile.mkdtemp(prefix="sphdel_")
    with open(os.path.join(wd, "del_in.bin"), "wb") as f:
        f.write(struct.pack("<i", n) + struct.pack("<3d", dr, p.xmin, p.xmax))
        for k in ("xp", "vxp", "vx", "ax"):
            f.write(np.ascontiguousarray(st[k], "<f8").tobytes())
        f.write(ty.astype("<i4").tobytes())
    assert subprocess.run([REF_HARNESS, "--del", wd], capture_output=True).returncode == 0
    raw = open(os.path.join(wd, "del_out.bin"), "rb").read()
    rty = np.frombuffer(raw, "<i4", n)
    rest = np.frombuffer(raw, "<f8", 9 * n, 4 * n).reshape(3, n, 3)
    o = OracleSim(deck, st, mode=GS)
    oracle_lib().lib.or_partidel_now.argtypes = [C.c_void_p]
    oracle_lib().lib.or_partidel_now(o.ctx)
    assert np.array_equal(np.asarray(o.get("type")), rty)
    for idx, k in enumerate(("vxp", "vx", "ax")):
        assert bitequal(o.get(k), rest[idx]), k
    assert 0 < int((rty == 7).sum()) < int((ty == 1).sum())  # some water left, some stayed


def problem_of(deck):
    from sphgpy.capi import problem_from_deck
    return problem_from_deck(deck)

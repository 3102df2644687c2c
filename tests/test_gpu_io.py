"""Result files and checkpoints of a context on the GPU.

- sphg_write_odb from the device state after l steps writes the same bytes as the host writer
  (byte-identical to the reference's output_total, tests/test_io.py) given the oracle's state
  after the same steps, which the device state equals bit for bit on these decks.
- A realisation of a batch writes the table of the same realisation run alone.
- sphg_save_state / sphg_load_state: a context restored from a checkpoint and stepped on gives
  bit for bit the fields and neighbour lists of the uninterrupted run; a checkpoint is refused
  by a context of another deck or another arithmetic form.
"""
import os

import numpy as np
import pytest

from fixtures import load_golden, bitequal
from helpers import OracleSim, JACOBI
from golden.make_odb import odb_time
from sphgpy.capi import GpuSim, GpuBatch, write_odb_fields, ODB_V3, ODB_V6, ODB_S, batch_state, setting_par

pytestmark = pytest.mark.gpu

ODB_FIELDS = ODB_V3 + ODB_V6 + ODB_S


def make_gpu(deck, state, precision=0):
    g = GpuSim(deck, precision=precision)
    g.load_state(state)
    g.upload_done()
    g.set_step(1)
    return g


@pytest.mark.parametrize("name,l", [("ssi_2d_t1", 6), ("ssi_3d_t1", 5)])
def test_odb_from_device_equals_odb_of_oracle_state(name, l, tmp_path):
    deck, states = load_golden(name)
    g = make_gpu(deck, states[0])
    o = OracleSim(deck, states[0], mode=JACOBI)
    assert g.step(l) == 0 and o.step(l) == 0
    t = odb_time(deck.dt, l)
    dg, do = tmp_path / "gpu", tmp_path / "oracle"
    dg.mkdir()
    do.mkdir()
    pg = g.write_odb(str(dg), l, t)
    st = {k: o.get(k) for k in ODB_FIELDS}
    st["type"], st["matype"] = deck.type, deck.matype
    po = write_odb_fields(str(do), l, t, st)
    assert open(pg, "rb").read() == open(po, "rb").read()


def test_odb_of_batch_realisation_equals_standalone(tmp_path):
    from sphgpy import slope_2d
    decks = [slope_2d(h, dx=0.25, seed=s) for h, s in ((1.0, 1), (1.4, 2))]
    b = GpuBatch(decks)
    b.load_state(batch_state(decks))
    b.upload_done()
    b.set_step(1)
    assert b.step(3) == 0
    first = 0
    for r, d in enumerate(decks):
        g = GpuSim(d)
        g.load_state(setting_par(d))
        g.upload_done()
        g.set_step(1)
        assert g.step(3) == 0
        a, s = tmp_path / f"b{r}", tmp_path / f"s{r}"
        a.mkdir()
        s.mkdir()
        pa = b.write_odb(str(a), 3, 0.5, first, d.n)
        ps = g.write_odb(str(s), 3, 0.5)
        assert open(pa, "rb").read() == open(ps, "rb").read()
        first += d.n


CKPT_FIELDS = ("xp", "vx", "ax", "ux", "sig", "eps", "adps", "rho", "pre", "strep", "interfss", "vsig")


@pytest.mark.parametrize("name,precision", [("column2d_small", 0), ("ssi_cpl_2d", 0), ("tran_3d_b14", 0),
                                            ("soilwater3d_small", 1)])
def test_checkpoint_resume_is_bit_exact(name, precision, tmp_path):
    deck, states = load_golden(name)
    a = make_gpu(deck, states[0], precision)
    assert a.step(3) == 0
    path = str(tmp_path / "ckpt.bin")
    assert a.save_state(path) == 0
    assert a.step(4) == 0
    b = make_gpu(deck, states[0], precision)
    assert b.step(1) == 0  # a different state and buffer parity before the restore
    assert b.restore_state(path) == 0
    assert b.L.lib.sphg_current_step(b.ctx) == 4
    assert b.step(4) == 0
    for k in CKPT_FIELDS:
        assert bitequal(a.get(k), b.get(k)), k
    ca, ia = a.neighbors()
    cb, ib = b.neighbors()
    assert np.array_equal(ca, cb) and np.array_equal(ia, ib)


def test_checkpoint_refused_by_other_deck_or_form(tmp_path):
    deck, states = load_golden("column2d_small")
    a = make_gpu(deck, states[0])
    assert a.step(2) == 0
    path = str(tmp_path / "ckpt.bin")
    assert a.save_state(path) == 0
    other, ostates = load_golden("ssi_2d_t1")
    assert make_gpu(other, ostates[0]).restore_state(path) == 100
    assert make_gpu(deck, states[0], precision=1).restore_state(path) == 100
    assert a.restore_state(str(tmp_path / "missing.bin")) == 100
    with open(path, "r+b") as f:  # a damaged header
        f.write(b"XXXX")
    assert a.restore_state(path) == 100


def test_project_runner_writes_the_reference_outputs(tmp_path):
    # sphg_solve on a project directory (input.dat -> reader -> setting_par -> device loop ->
    # .odb at steps 1 and every fop-th) writes, at each output step, the table the host writer
    # gives for the oracle's state after the same steps, and the same files as the API path
    import subprocess
    from sphgpy.capi import SOLVE_PATH, odb_path
    deck, states = load_golden("ssi_2d_t1")
    deck = deck.copy()
    deck.loop, deck.fop = 6, 3
    proj = tmp_path / "proj"
    proj.mkdir()
    deck.write_input_dat(str(proj / "input.dat"))
    r = subprocess.run([SOLVE_PATH, str(proj)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "last output step 6" in r.stdout
    assert sorted(os.listdir(proj)) == ["data-0000000001.odb", "data-0000000003.odb", "data-0000000006.odb",
                                        "input.dat"]
    o = OracleSim(deck, states[0], mode=JACOBI)
    done = 0
    for l in (1, 3, 6):
        assert o.step(l - done) == 0
        done = l
        st = {k: o.get(k) for k in ODB_FIELDS}
        st["type"], st["matype"] = deck.type, deck.matype
        d = tmp_path / f"o{l}"
        d.mkdir()
        want = open(write_odb_fields(str(d), l, odb_time(deck.dt, l), st), "rb").read()
        assert open(odb_path(str(proj), l), "rb").read() == want, l
    # the API path: same deck through GpuSim and sphg_run with the writer as output callback
    api = tmp_path / "api"
    api.mkdir()
    g = GpuSim(deck)
    g.load_state(setting_par(deck))
    g.upload_done()
    g.set_step(1)
    rc, t = g.run(deck.loop, deck.fop, deck.dt * (deck.loop + 1),
                  output=lambda l, tm: g.write_odb(str(api), l, tm) and 0)
    assert rc == 0
    for l in (1, 3, 6):
        assert open(odb_path(str(api), l), "rb").read() == open(odb_path(str(proj), l), "rb").read()


@pytest.mark.parametrize("name", ["vib_2d_vel", "rain_2d_r1"])
def test_project_runner_with_vibration_and_rain(name, tmp_path):
    # flv 2 / flr 1 projects: the runner reads the VL section and vibra.dat, or the RP section, and
    # its tables at the output steps are those of the oracle with the same parameters
    import subprocess
    from sphgpy.capi import SOLVE_PATH, odb_path
    deck, states = load_golden(name)
    deck = deck.copy()
    deck.loop, deck.fop = 6, 3
    proj = tmp_path / "proj"
    proj.mkdir()
    deck.write_input_dat(str(proj / "input.dat"))
    r = subprocess.run([SOLVE_PATH, str(proj)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    o = OracleSim(deck, states[0], mode=JACOBI)
    done = 0
    for l in (1, 3, 6):
        assert o.step(l - done) == 0
        done = l
        st = {k: o.get(k) for k in ODB_FIELDS}
        # the particle set changes with rain: the table lists the types of the step
        st["type"], st["matype"] = np.asarray(o.get("type")), np.asarray(o.get("matype"))
        d = tmp_path / f"o{l}"
        d.mkdir()
        want = open(write_odb_fields(str(d), l, odb_time(deck.dt, l), st), "rb").read()
        assert open(odb_path(str(proj), l), "rb").read() == want, l

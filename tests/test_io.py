"""Result files on the host: the library's .odb writer against the reference's own writer.

The fixtures tests/golden/odb_*.odb.gz are what the reference's output_total
(CPP_14-2-output.cpp:15-135) writes for a golden particle state (tests/golden/make_odb.py runs it
through oracle/_ref/ref_odb). sphg_write_odb_fields, given the same state, must write the same
bytes. Host only: no device is touched.
"""
import gzip
import os

import numpy as np
import pytest

from fixtures import load_golden
from golden.make_odb import CASES, fixture_path, odb_time
from sphgpy.capi import write_odb_fields, odb_path, product_lib


@pytest.mark.parametrize("name,l", CASES)
def test_odb_writer_matches_reference_writer(name, l, tmp_path):
    deck, states = load_golden(name)
    path = write_odb_fields(str(tmp_path), l, odb_time(deck.dt, l), states[l])
    assert os.path.basename(path) == f"data-{l:010d}.odb"
    with gzip.open(fixture_path(name, l), "rb") as f:
        want = f.read()
    got = open(path, "rb").read()
    assert got == want
    # the table has a row per particle and boundary rows differ from the others in layout only
    assert got.count(b"\n") == deck.n + 3  # rule, "l t", rows, rule
    assert (np.asarray(states[l]["type"]) == 0).any()


def test_odb_writer_rejects_bad_arguments(tmp_path):
    import ctypes as C
    from sphgpy.capi import OdbFields
    L = product_lib().lib
    assert L.sphg_write_odb_fields(str(tmp_path).encode(), 1, 0.0, 1, None) == 100
    assert L.sphg_write_odb_fields(str(tmp_path).encode(), 1, 0.0, 1, C.byref(OdbFields())) == 100  # null arrays
    assert L.sphg_write_odb(None, str(tmp_path).encode(), 1, 0.0, 0, 0) == 100


def test_odb_file_name_past_eight_digits(tmp_path):
    # output_total pads l to 10 digits below 1e8 and leaves the number out from 1e8 on
    deck, states = load_golden("ssi_2d_t1")
    assert os.path.basename(write_odb_fields(str(tmp_path), 7, 0.0, states[6])) == "data-0000000007.odb"
    p = write_odb_fields(str(tmp_path), 100000000, 0.0, states[6])
    assert os.path.basename(p) == "data-.odb" and odb_path(str(tmp_path), 100000000) == p
    assert open(p).read().splitlines()[1] == f"{100000000:10d} {0.0:14.6e}"


def test_odb_writer_every_column_and_boundary_rows():
    # synthetic state: every field distinct and nonzero over many decades, all particle types
    from golden.make_odb import SYNTH, HERE as GOLDEN
    name, l = SYNTH
    z = np.load(os.path.join(GOLDEN, f"{name}_state.npz"), allow_pickle=False)
    st = {k: z[k] for k in z.files}
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        got = open(write_odb_fields(tmp, l, 0.125, st), "rb").read()
    with gzip.open(os.path.join(GOLDEN, f"{name}.odb.gz"), "rb") as f:
        want = f.read()
    assert got == want
    # and a planted column swap is seen: the fixture pins the order of every column
    st2 = dict(st)
    st2["divq"], st2["divr"] = st["divr"], st["divq"]
    with tempfile.TemporaryDirectory() as tmp:
        assert open(write_odb_fields(tmp, l, 0.125, st2), "rb").read() != want


# ------------------------------------------------------------------------------ input.dat reader

def _flat(obj, prefix=""):
    """(name, value) leaves of a ctypes struct, arrays and nested structs expanded."""
    import ctypes as C
    if isinstance(obj, C.Structure):
        for name, _ in obj._fields_:
            yield from _flat(getattr(obj, name), f"{prefix}{name}.")
    elif isinstance(obj, C.Array):
        for i, v in enumerate(obj):
            yield from _flat(v, f"{prefix}{i}.")
    else:
        yield prefix.rstrip("."), obj


def _same_struct(a, b):
    import struct as st
    bad = []
    for (na, va), (nb, vb) in zip(_flat(a), _flat(b)):
        assert na == nb
        if isinstance(va, float):
            if st.pack("<d", va) != st.pack("<d", vb):
                bad.append((na, va, vb))
        elif va != vb:
            bad.append((na, va, vb))
    return bad


def _reader_decks():
    from golden.make_golden import decks
    from sphgpy import column_collapse_2d
    out = {name: d for name, (d, _, _) in decks().items()}
    # flags no golden deck sets: air phase, TS-transform material line, moving walls
    d = column_collapse_2d(nx=4, ny=5, wall_h=7, floor_len=9, steps=3, seed=4)
    d.fla = 1
    d.air = dict(d.air, dens=1.2, paa0=101325.0, vis=1.8e-5, vpu=34.0, press_type=1, acc_type=1, pre_coe=0.5)
    d.flts = 1
    d.alphats, d.damp_coe = 0.25, 0.125
    d.if_move, d.move_vx = 1, (0.1, -0.2, 0.3)
    out["air_ts_move"] = d
    return out


@pytest.mark.parametrize("name", sorted(_reader_decks()))
def test_read_input_gives_the_structs_the_reference_runs_on(name, tmp_path):
    # The deck file written for the reference and read back by sphg_read_input gives exactly the
    # problem/material structs that problem_from_deck / materials_from_deck build, on which the
    # library's runs match the reference's runs of the same file (the GPU parity tests), and the
    # same particles.
    from sphgpy.capi import read_input, problem_from_deck, materials_from_deck
    deck = _reader_decks()[name]
    path = str(tmp_path / "input.dat")
    deck.write_input_dat(path)
    pro, mat, parts = read_input(path)
    assert _same_struct(pro, problem_from_deck(deck)) == []
    assert _same_struct(mat, materials_from_deck(deck)) == []
    xp = np.zeros((deck.n, 3))
    xp[:, :deck.ndim] = deck.xp[:, :deck.ndim]
    assert np.array_equal(parts["xp"].view(np.int64), xp.view(np.int64))
    assert np.array_equal(parts["vxp"][:, :deck.ndim], deck.vxp[:, :deck.ndim])
    for k, v in (("type", deck.type), ("matype", deck.matype), ("permtype", deck.permtype_p),
                 ("etype", deck.etype)):
        assert np.array_equal(parts[k], v), k


def test_read_input_section_order_and_errors(tmp_path):
    import ctypes as C
    from sphgpy.capi import read_input, problem_from_deck, Problem, Materials
    from sphgpy import column_collapse_2d
    deck = column_collapse_2d(nx=3, ny=3, wall_h=4, floor_len=5, steps=2, seed=9)
    path = tmp_path / "input.dat"
    deck.write_input_dat(str(path))
    lines = path.read_text().splitlines()
    # sections are found from the top of the file, so their order does not matter
    pa = lines.index("PA section")
    gf = lines.index("GF section")
    moved = lines[:gf] + lines[gf + 4:pa] + lines[gf:gf + 4] + lines[pa:]
    p2 = tmp_path / "moved.dat"
    p2.write_text("\n".join(moved) + "\n")
    pro, _, _ = read_input(str(p2))
    assert _same_struct(pro, problem_from_deck(deck)) == []
    L = product_lib().lib
    pr, mt, n = Problem(), Materials(), C.c_int(0)
    assert L.sphg_read_input(str(tmp_path / "none.dat").encode(), C.byref(pr), C.byref(mt), C.byref(n), 0,
                             None, None, None, None, None, None) == 100
    # ntotal 0: the reference's error 2
    zero = [ln for ln in lines]
    i = zero.index("PR section") + 3
    f = zero[i].split()
    f[3] = "0"
    zero[i] = " ".join(f)
    p3 = tmp_path / "zero.dat"
    p3.write_text("\n".join(zero) + "\n")
    assert L.sphg_read_input(str(p3).encode(), C.byref(pr), C.byref(mt), C.byref(n), 0,
                             None, None, None, None, None, None) == 2
    # capacity below ntotal, then a truncated particle section
    arr = np.zeros((deck.n, 3))
    ints = [np.zeros(deck.n, np.int32) for _ in range(4)]
    ip = [i_.ctypes.data_as(C.POINTER(C.c_int)) for i_ in ints]
    dp = arr.ctypes.data_as(C.POINTER(C.c_double))
    assert L.sphg_read_input(str(path).encode(), C.byref(pr), C.byref(mt), C.byref(n), deck.n - 1, dp, dp,
                             *ip) == 103
    p4 = tmp_path / "short.dat"
    p4.write_text("\n".join(lines[:-3]) + "\n")
    assert L.sphg_read_input(str(p4).encode(), C.byref(pr), C.byref(mt), C.byref(n), deck.n, dp, dp, *ip) == 100


def test_project_runner_argument_errors(tmp_path):
    # checks that end before any device call: usage, a directory without input.dat, and a
    # random-field deck (flspv), whose SF/SV sections the reader does not take
    import ctypes as C
    import subprocess
    from sphgpy.capi import SOLVE_PATH
    from sphgpy import column_collapse_2d
    assert subprocess.run([SOLVE_PATH], capture_output=True, timeout=60).returncode == 2
    L = product_lib().lib
    L.sphg_solve_project.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int)]
    last = C.c_int(-1)
    assert L.sphg_solve_project(str(tmp_path).encode(), 0, C.byref(last)) == 100 and last.value == 0
    d = column_collapse_2d(nx=3, ny=3, wall_h=4, floor_len=5, steps=2, seed=9)
    d.flspv = 1
    d.write_input_dat(str(tmp_path / "input.dat"))
    assert L.sphg_solve_project(str(tmp_path).encode(), 0, None) == 102


def test_read_vibration_section_and_table(tmp_path):
    # sphg_read_vibration: the VL section of input.dat and vibra.dat as the deck writer wrote them
    import ctypes as C
    from sphgpy import vibration_box
    from sphgpy.capi import product_lib
    L = product_lib().lib
    deck = vibration_box(2, 1)
    deck.write_input_dat(str(tmp_path / "input.dat"))
    tt, st, et, ns = C.c_double(), C.c_double(), C.c_double(), C.c_int()
    D = C.POINTER(C.c_double)
    assert L.sphg_read_vibration(str(tmp_path).encode(), C.byref(tt), C.byref(ns), C.byref(st), C.byref(et),
                                 C.cast(None, D), 0) == 0
    assert (tt.value, ns.value, st.value, et.value) == (deck.vib_ttime, deck.vib_nsteps, deck.vib_sttime,
                                                        deck.vib_edtime)
    loads = np.zeros((ns.value, 3))
    assert L.sphg_read_vibration(str(tmp_path).encode(), C.byref(tt), C.byref(ns), C.byref(st), C.byref(et),
                                 loads.ctypes.data_as(D), ns.value) == 0
    assert np.array_equal(loads, deck.vib_loads)
    assert L.sphg_read_vibration(str(tmp_path).encode(), C.byref(tt), C.byref(ns), C.byref(st), C.byref(et),
                                 loads.ctypes.data_as(D), ns.value - 1) == 103  # SPHG_E_CAPACITY
    (tmp_path / "vibra.dat").unlink()
    assert L.sphg_read_vibration(str(tmp_path).encode(), C.byref(tt), C.byref(ns), C.byref(st), C.byref(et),
                                 loads.ctypes.data_as(D), ns.value) == 100


def test_read_rain_and_excavation_sections(tmp_path):
    import ctypes as C
    from sphgpy import rain_box, eob_box
    from sphgpy.capi import Rain
    L = product_lib().lib
    d = rain_box(2, 4)
    d.write_input_dat(str(tmp_path / "input.dat"))
    r = Rain()
    assert L.sphg_read_rain(str(tmp_path).encode(), C.byref(r)) == 0
    assert (r.rtype, r.nrain, r.nhor, r.press_factor, r.drop_velocity) == d.rain
    e = eob_box()
    (tmp_path / "e").mkdir()
    e.write_input_dat(str(tmp_path / "e" / "input.dat"))
    n = C.c_int()
    assert L.sphg_read_excavation(str(tmp_path / "e").encode(), C.byref(n), None, 0) == 0
    ecp = np.full((n.value, 3), -1, np.int32)
    assert L.sphg_read_excavation(str(tmp_path / "e").encode(), C.byref(n),
                                  ecp.ctypes.data_as(C.POINTER(C.c_int)), n.value) == 0
    # first and last particle as written; the step is not in the file (the reference never sets it)
    assert ecp.tolist() == [[a, b, 0] for a, b, _ in e.ecp]
    assert L.sphg_read_excavation(str(tmp_path).encode(), C.byref(n), None, 0) == 8  # no EC section

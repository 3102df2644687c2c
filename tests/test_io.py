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
